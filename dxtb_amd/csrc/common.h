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

template <int V>
struct IntC {
    static constexpr int value = V;
};

// Calls fn(IntC<li>, IntC<lj>) for runtime angular momenta li, lj in {0, 1},
// so that the shell-pair code it holds is compiled for each (li, lj) with
// its blocks in registers (runtime-sized blocks were indexed through
// scratch memory).
template <class Fn>
__device__ inline void dispatch_l(int li, int lj, Fn&& fn) {
    if (li == 0) {
        if (lj == 0) fn(IntC<0>{}, IntC<0>{});
        else fn(IntC<0>{}, IntC<1>{});
    } else {
        if (lj == 0) fn(IntC<1>{}, IntC<0>{});
        else fn(IntC<1>{}, IntC<1>{});
    }
}

// Contracted overlap block <i|j> between an s/p shell i (angular momentum
// LI) on atom A and an s/p shell j (LJ) on atom B (McMurchie-Davidson closed
// forms for l <= 1), R = R_A - R_B.  Cartesian p order is (y, z, x).  With
// GRAD, dS[k] holds d<i|j>/dR_A,k.  Primitive pairs with mu R^2 above
// XTB_PRIM_SCREEN are skipped.
template <bool GRAD, int LI, int LJ>
__device__ inline void shell_overlap_t(int npi, const double* ai, const double* ci, int npj,
                                       const double* aj, const double* cj, const double R[3],
                                       double (&S)[2 * LI + 1][2 * LJ + 1],
                                       double (&dS)[3][2 * LI + 1][2 * LJ + 1]) {
    constexpr int NI = 2 * LI + 1, NJ = 2 * LJ + 1;
#pragma unroll
    for (int x = 0; x < NI; ++x)
#pragma unroll
        for (int y = 0; y < NJ; ++y) {
            S[x][y] = 0.0;
            if (GRAD)
#pragma unroll
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
#pragma unroll
            for (int mi = 0; mi < NI; ++mi) {
                const int x = LI ? (mi + 1) % 3 : -1;  // p order y, z, x
#pragma unroll
                for (int mj = 0; mj < NJ; ++mj) {
                    const int y = LJ ? (mj + 1) % 3 : -1;
                    double f;
                    if (x < 0 && y < 0) f = 1.0;
                    else if (y < 0) f = pa * R[x];
                    else if (x < 0) f = pb * R[y];
                    else f = pa * pb * R[x] * R[y] + (x == y ? 0.5 * ip : 0.0);
                    S[mi][mj] += f * s00;
                    if (GRAD) {
#pragma unroll
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

// Runtime-l form of shell_overlap_t: S and dS are [3][3] (only the leading
// (2 li + 1) x (2 lj + 1) part is written; the rest is zero).
template <bool GRAD>
__device__ inline void shell_overlap(int li, int lj, int npi, const double* ai, const double* ci,
                                     int npj, const double* aj, const double* cj,
                                     const double R[3], double S[3][3], double dS[3][3][3]) {
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y) {
            S[x][y] = 0.0;
            if (GRAD)
                for (int k = 0; k < 3; ++k) dS[k][x][y] = 0.0;
        }
    dispatch_l(li, lj, [&](auto LIc, auto LJc) {
        constexpr int LI = decltype(LIc)::value, LJ = decltype(LJc)::value;
        double s[2 * LI + 1][2 * LJ + 1], ds[3][2 * LI + 1][2 * LJ + 1];
        shell_overlap_t<GRAD, LI, LJ>(npi, ai, ci, npj, aj, cj, R, s, ds);
        for (int x = 0; x < 2 * LI + 1; ++x)
            for (int y = 0; y < 2 * LJ + 1; ++y) {
                S[x][y] = s[x][y];
                if (GRAD)
                    for (int k = 0; k < 3; ++k) dS[k][x][y] = ds[k][x][y];
            }
    });
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
