// Per-atom-pair code of the analytic gradient, shared by the batched
// kernels (k_grad_pairs, k_grad_cn: one thread per atom pair of 16x16
// tiles, atomics) and the fused single-point kernel (one workgroup per
// molecule, partial sums added in a fixed order), so that both compute
// every pair's terms with the same operations.
#pragma once

#include "common.h"

namespace xtb {

#define XTB_MAX_US_LDS_F 64

// One shell pair's part of the gradient, contracted inside the primitive
// loop (no dS block is formed): with w_xy = 2 (P_xy ht - W_xy) - P_xy
// (v_x + v_y) over the pair's block (rows of shell i on A, columns of shell
// j on B, leading dimension n), gvec += sum_xy w_xy d<x|y>/dR_A and
// phs += sum_xy P_xy <x|y>.  The primitive-pair terms are those of
// shell_overlap_t (McMurchie-Davidson, p order y, z, x).
template <int LI, int LJ>
__device__ inline void grad_shell_pair(int npi, const double* ai, const double* ci, int npj,
                                       const double* aj, const double* cj, const double R[3],
                                       const double* __restrict__ Pb, const double* __restrict__ Wb,
                                       int n, const double* vi, const double* vj, double ht,
                                       double gvec[3], double& phs) {
    constexpr int NI = 2 * LI + 1, NJ = 2 * LJ + 1;
    double pm[NI][NJ], wv[NI][NJ];
#pragma unroll
    for (int x = 0; x < NI; ++x)
#pragma unroll
        for (int y = 0; y < NJ; ++y) {
            const double pmn = Pb[(size_t)x * n + y];
            pm[x][y] = pmn;
            wv[x][y] = 2.0 * (pmn * ht - Wb[(size_t)x * n + y]) - pmn * (vi[x] + vj[y]);
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
                const int x = LI ? (mi + 1) % 3 : -1;
#pragma unroll
                for (int mj = 0; mj < NJ; ++mj) {
                    const int y = LJ ? (mj + 1) % 3 : -1;
                    double f;
                    if (x < 0 && y < 0) f = 1.0;
                    else if (y < 0) f = pa * R[x];
                    else if (x < 0) f = pb * R[y];
                    else f = pa * pb * R[x] * R[y] + (x == y ? 0.5 * ip : 0.0);
                    phs = fma(pm[mi][mj], f * s00, phs);
                    const double wf = wv[mi][mj] * s00;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        double df = 0.0;
                        if (x >= 0 && y < 0) df = (k == x) ? pa : 0.0;
                        else if (x < 0 && y >= 0) df = (k == y) ? pb : 0.0;
                        else if (x >= 0 && y >= 0) {
                            if (k == x) df += pa * pb * R[y];
                            if (k == y) df += pa * pb * R[x];
                        }
                        gvec[k] = fma(wf, df - 2.0 * mu * R[k] * f, gvec[k]);
                    }
                }
            }
        }
    }
}

// On-site part of dE/dCN_A: sum_{i in A} (-kcn_i) sum_{mu in i} P_mumu.
// P: the molecule's density (leading dimension n); sh0: its first shell.
__device__ inline double onsite_dedcn(const xtb_batch& b, const xtb_params& par, const double* P, int n,
                                      int sh0, int ga) {
    const int sa0 = b.atom_sh0[ga], nsa = b.atom_nsh[ga];
    double d = 0.0;
    for (int i = sa0; i < sa0 + nsa; ++i) {
        const int us = b.sh_us[sh0 + i];
        const int o0 = b.sh_ao[sh0 + i], nl = 2 * par.us_l[us] + 1;
        double pp = 0.0;
        for (int x = 0; x < nl; ++x) pp += P[(size_t)(o0 + x) * n + (o0 + x)];
        d -= par.us_kcn[us] * pp;
    }
    return d;
}

// Atom pair A != B of one molecule (global atom indices ga, gb): g is the
// pair's part of dE/dR_A (dE/dR_B gets -g), dA and dB its parts of dE/dCN_A
// and dE/dCN_B.  P, W: the molecule's density and energy-weighted density
// (leading dimension n); v: its orbital potential; g_m: its gamma (nsh x
// nsh); q_m: its shell charges; s_alpha, s_coeff: staged primitives.
__device__ inline void atom_pair_grad(const xtb_batch& b, const xtb_params& par, const double* __restrict__ pos,
                                      const double* __restrict__ cn, const double* __restrict__ g_m,
                                      const double* __restrict__ q_m, const double* __restrict__ P,
                                      const double* __restrict__ W, const double* __restrict__ v, int n,
                                      int nsh, int sh0, int ga, int gb, const double* s_alpha,
                                      const double* s_coeff, double g[3], double& dA, double& dB) {
    const int sa0 = b.atom_sh0[ga], nsa = b.atom_nsh[ga];
    const int sb0 = b.atom_sh0[gb], nsb = b.atom_nsh[gb];
    const double R[3] = {pos[3 * ga] - pos[3 * gb], pos[3 * ga + 1] - pos[3 * gb + 1],
                         pos[3 * ga + 2] - pos[3 * gb + 2]};
    const double dist = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    const int spa = b.atom_sp[ga], spb = b.atom_sp[gb];
    const double rab = par.sp_rad[spa] + par.sp_rad[spb];
    double gvec[3] = {0.0, 0.0, 0.0};
    double radial = 0.0;  // coefficient of (R_A - R_B)
    dA = 0.0;
    dB = 0.0;
    const bool in_cut = dist < par.int_cutoff && dist > 0.1;
    for (int i = sa0; i < sa0 + nsa; ++i) {
        const int usi = b.sh_us[sh0 + i];
        const int li = par.us_l[usi], oi = b.sh_ao[sh0 + i];
        const double hi = par.us_level[usi] - par.us_kcn[usi] * cn[ga];
        for (int j = sb0; j < sb0 + nsb; ++j) {
            const int usj = b.sh_us[sh0 + j];
            const int lj = par.us_l[usj], oj = b.sh_ao[sh0 + j];
            // ES2 with fixed charges: d gamma / dR_A = -gamma^3 (R_A - R_B)
            const double gij = g_m[(size_t)i * nsh + j];
            radial -= gij * gij * gij * q_m[i] * q_m[j];
            if (!in_cut) continue;
            const double hj = par.us_level[usj] - par.us_kcn[usj] * cn[gb];
            double pi, dpi;
            dist_poly(dist, rab, par.us_shpoly[usi], par.us_shpoly[usj], pi, dpi);
            const double kk = par.us_K[usi * par.nus + usj];
            const double ht = 0.5 * (hi + hj) * kk * pi;
            double phs = 0.0;
            dispatch_l(li, lj, [&](auto LIc, auto LJc) {
                constexpr int LI = decltype(LIc)::value, LJ = decltype(LJc)::value;
                grad_shell_pair<LI, LJ>(par.us_nprim[usi], s_alpha + usi * XTB_MAXPRIM,
                                        s_coeff + usi * XTB_MAXPRIM, par.us_nprim[usj],
                                        s_alpha + usj * XTB_MAXPRIM, s_coeff + usj * XTB_MAXPRIM, R,
                                        P + (size_t)oi * n + oj, W + (size_t)oi * n + oj, n, v + oi, v + oj,
                                        ht, gvec, phs);
            });
            radial += 2.0 * phs * ht * dpi / pi;
            const double fac = kk * pi;
            dA -= phs * fac * par.us_kcn[usi];
            dB -= phs * fac * par.us_kcn[usj];
        }
    }
    // repulsion (repulsion/base.py:337-406)
    if (dist <= par.rep_cutoff) {
        const double alpha = sqrt(par.sp_arep[spa] * par.sp_arep[spb] + 2.2250738585072014e-308);
        const double rk = pow(dist, par.kexp);
        const double e = par.sp_zeff[spa] * par.sp_zeff[spb] * exp(-alpha * rk) / dist;
        const double dedr = -e * (1.0 / dist + alpha * par.kexp * rk / dist);
        radial += dedr / dist;
    }
    for (int k = 0; k < 3; ++k) g[k] = gvec[k] + radial * R[k];
}

// CN chain rule for the atom pair A, B (ncoord/utils.py:30-52): the pair's
// part of dE/dR_A, f (R_A - R_B) with f = (dE/dCN_A + dE/dCN_B) dcount/dR / R
// (dE/dR_B gets the negative); false (g untouched) beyond the CN cutoff.
__device__ inline bool cn_pair_grad(const xtb_batch& b, const xtb_params& par, const double* __restrict__ pos,
                                    const double* __restrict__ dedcn, int ga, int gb, double g[3]) {
    const double R[3] = {pos[3 * ga] - pos[3 * gb], pos[3 * ga + 1] - pos[3 * gb + 1],
                         pos[3 * ga + 2] - pos[3 * gb + 2]};
    const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    if (r > par.cn_cutoff) return false;
    const double rc = par.sp_rcov[b.atom_sp[ga]] + par.sp_rcov[b.atom_sp[gb]];
    const double e = exp(-par.cn_steep * (rc / r - 1.0));
    const double dcf = -par.cn_steep * rc * e / (r * r * (1.0 + e) * (1.0 + e));
    const double f = (dedcn[ga] + dedcn[gb]) * dcf / r;
    for (int k = 0; k < 3; ++k) g[k] = f * R[k];
    return true;
}

}  // namespace xtb
