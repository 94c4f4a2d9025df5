// Shared device helpers for the GFN1 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/xtb_hip.h"

#define XTB_MAXPRIM 8
#define XTB_PI 3.14159265358979323846
// primitive pairs with mu R^2 above this are skipped: their Gaussian factor
// exp(-mu R^2) < 2e-35 keeps every contribution under 1e-29 for the H, C,
// N, O basis (contraction coefficients below 2, (pi/p)^1.5 below 50 with
// the smallest exponent 0.128, distance polynomials below 2500 within the
// 50 bohr cutoff), far below the FP64 resolution of the O(1) elements
#define XTB_PRIM_SCREEN 80.0

namespace xtb {

// Contracted overlap block <i|j> between an s/p shell i on atom A and an
// s/p shell j on atom B (McMurchie-Davidson closed forms for l <= 1),
// R = R_A - R_B.  Cartesian p order is (y, z, x).  With GRAD, dS[k] holds
// d<i|j>/dR_A,k.  S and dS are [3][3] (only the leading li/lj part is used).
template <bool GRAD>
__device__ inline void shell_overlap(int li, int lj, int npi, const double* ai, const double* ci,
                                     int npj, const double* aj, const double* cj,
                                     const double R[3], double S[3][3], double dS[3][3][3]) {
    const int ni = 2 * li + 1, nj = 2 * lj + 1;
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y) {
            S[x][y] = 0.0;
            if (GRAD)
                for (int k = 0; k < 3; ++k) dS[k][x][y] = 0.0;
        }
    const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
    for (int u = 0; u < npi; ++u) {
        const double a = ai[u];
        for (int v = 0; v < npj; ++v) {
            const double b = aj[v];
            const double p = a + b;
            const double ip = 1.0 / p;
            const double mu = a * b * ip;
            if (mu * r2 > XTB_PRIM_SCREEN) continue;
            const double s00 = ci[u] * cj[v] * (XTB_PI * ip) * sqrt(XTB_PI * ip) * exp(-mu * r2);
            const double pa = -b * ip;  // (P - A) = pa R
            const double pb = a * ip;   // (P - B) = pb R
            for (int mi = 0; mi < ni; ++mi) {
                const int x = li ? (mi + 1) % 3 : -1;  // p order y, z, x
                for (int mj = 0; mj < nj; ++mj) {
                    const int y = lj ? (mj + 1) % 3 : -1;
                    double f;
                    if (x < 0 && y < 0) f = 1.0;
                    else if (y < 0) f = pa * R[x];
                    else if (x < 0) f = pb * R[y];
                    else f = pa * pb * R[x] * R[y] + (x == y ? 0.5 * ip : 0.0);
                    S[mi][mj] += f * s00;
                    if (GRAD) {
                        for (int k = 0; k < 3; ++k) {
                            double df = 0.0;
                            if (x >= 0 && y < 0) df = (k == x) ? pa : 0.0;
                            else if (x < 0 && y >= 0) df = (k == y) ? pb : 0.0;
                            else if (x >= 0 && y >= 0) {
                                if (k == x) df += pa * pb * R[y];
                                if (k == y) df += pa * pb * R[x];
                            }
                            dS[k][mi][mj] += (df - 2.0 * mu * R[k] * f) * s00;
                        }
                    }
                }
            }
        }
    }
}

// Distance polynomial Pi(R) = (1 + k_i sqrt(R/Rab)) (1 + k_j sqrt(R/Rab))
// and (dPi/dR)/R; xtb/base.py:298-312, xtb/gfn1.py:333-345 upstream.
__device__ inline void dist_poly(double dist, double rab, double ki, double kj, double& pi,
                                 double& dpi_over_r) {
    const double rr = sqrt(dist / rab);
    const double ta = 1.0 + ki * rr, tb = 1.0 + kj * rr;
    pi = ta * tb;
    dpi_over_r = (ta * kj + tb * ki) * rr * 0.5 / (dist * dist);
}

}  // namespace xtb
