// Per-element code of the integral pass, shared by the batched kernels (one
// thread per shell or atom pair of 16x16 tiles) and the fused single-point
// kernel (one workgroup per molecule), so that both compute every element
// with the same operations in the same order.
#pragma once

#include "common.h"

namespace xtb {

// unique shells whose primitives fit the LDS staging of the kernels
#define XTB_MAX_US_LDS 64

// ---------------------------------------------------------------------------
// coordination number: D3 counting with exp_count,
// CN_A = sum_{B != A, R <= cutoff} 1 / (1 + exp(-k (Rcov_AB / R - 1)))
// (tad_mctc cn_d3 as called at src/dxtb/_src/xtb/gfn1.py:57-64).
// CN_PARTS consecutive lanes share one atom: lane `part` sums the partners
// B = atom0 + part, atom0 + part + CN_PARTS, ... in order, and a fixed
// butterfly adds the parts.  Every kernel that needs CN uses this same
// order, so all of them get the same bits.
// ---------------------------------------------------------------------------
constexpr int CN_PARTS = 8;

__device__ inline double cn_part(const xtb_batch& b, const xtb_params& par,
                                 const double* __restrict__ pos, int a, int part) {
    const int m = b.atom_mol[a];
    const int a0 = b.mol_atom0[m], na = b.mol_nat[m];
    const double ra = par.sp_rcov[b.atom_sp[a]];
    const double x = pos[3 * a], y = pos[3 * a + 1], z = pos[3 * a + 2];
    double s = 0.0;
    for (int j = a0 + part; j < a0 + na; j += CN_PARTS) {
        if (j == a) continue;
        const double dx = x - pos[3 * j], dy = y - pos[3 * j + 1], dz = z - pos[3 * j + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (r > par.cn_cutoff) continue;
        const double rc = ra + par.sp_rcov[b.atom_sp[j]];
        s += 1.0 / (1.0 + exp(-par.cn_steep * (rc / r - 1.0)));
    }
    return s;
}

// sum over the CN_PARTS lanes of a group (aligned to CN_PARTS); the same
// bits in every lane of the group (addition is commutative)
__device__ inline double cn_combine(double v) {
#pragma unroll
    for (int o = 1; o < CN_PARTS; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// one upper-triangle entry (i <= j < nsh) of molecule m's gamma, mirrored
__device__ inline void gamma_pair(const xtb_batch& b, const xtb_params& par,
                                  const double* __restrict__ pos, double* __restrict__ gam,
                                  int m, int i, int j, int nsh) {
    const int sh0 = b.mol_sh0[m], at0 = b.mol_atom0[m];
    const double eps = 2.220446049250313e-16;
    const int ga = at0 + b.sh_atom[sh0 + i], gb = at0 + b.sh_atom[sh0 + j];
    double d;
    if (ga == gb) d = eps;
    else {
        const double dx = pos[3 * ga] - pos[3 * gb], dy = pos[3 * ga + 1] - pos[3 * gb + 1],
                     dz = pos[3 * ga + 2] - pos[3 * gb + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz) + eps;
        d = r * r;
    }
    const double hi = par.us_hard[b.sh_us[sh0 + i]] + eps;
    const double hj = par.us_hard[b.sh_us[sh0 + j]] + eps;
    const double avg = 2.0 / (1.0 / hi + 1.0 / hj);
    const double g = 1.0 / sqrt(d + 1.0 / (avg * avg));
    const size_t g0 = (size_t)b.mol_gam0[m];
    gam[g0 + (size_t)i * nsh + j] = g;
    gam[g0 + (size_t)j * nsh + i] = g;
}

// ---------------------------------------------------------------------------
// S and H0 (src/dxtb/_src/integral/driver/pytorch/impls/overlap.py:147-243,
// src/dxtb/_src/xtb/base.py:251-360).  Off-site blocks:
//   S_ij from the contracted s/p overlap (cutoff 50 bohr, same-centre mask),
//   H0_ij = S_ij * 1/2 (h_i + h_j) * K_ij * Pi_ij,  h = level - kcn * CN.
// On-site blocks: S = identity on the shell's own orbitals, H0 = h on the
// diagonal, zero between different shells of the same atom.
// ---------------------------------------------------------------------------
__device__ inline void shell_pair_s_h0(const xtb_batch& b, const xtb_params& par,
                                       const double* __restrict__ pos, double* __restrict__ S,
                                       double* __restrict__ H0, int m, int i, int j, double cni,
                                       double cnj, const double* s_alpha, const double* s_coeff) {
    const int n = b.mol_n[m];
    const int sh0 = b.mol_sh0[m], at0 = b.mol_atom0[m];
    const size_t mat0 = (size_t)b.mol_mat0[m];
    const int usi = b.sh_us[sh0 + i], usj = b.sh_us[sh0 + j];
    const int ati = b.sh_atom[sh0 + i], atj = b.sh_atom[sh0 + j];
    const int li = par.us_l[usi], lj = par.us_l[usj];
    const int ni = 2 * li + 1, nj = 2 * lj + 1;
    const int oi = b.sh_ao[sh0 + i], oj = b.sh_ao[sh0 + j];
    const double hi = par.us_level[usi] - par.us_kcn[usi] * cni;
    const double hj = par.us_level[usj] - par.us_kcn[usj] * cnj;
    (void)ni;
    (void)nj;
    dispatch_l(li, lj, [&](auto LIc, auto LJc) {
        constexpr int LI = decltype(LIc)::value, LJ = decltype(LJc)::value;
        constexpr int NI = 2 * LI + 1, NJ = 2 * LJ + 1;
        double blk[NI][NJ], hblk[NI][NJ];
#pragma unroll
        for (int x = 0; x < NI; ++x)
#pragma unroll
            for (int y = 0; y < NJ; ++y) { blk[x][y] = 0.0; hblk[x][y] = 0.0; }
        if (ati == atj) {
            if constexpr (LI == LJ) {
                if (i == j)
#pragma unroll
                    for (int x = 0; x < NI; ++x) { blk[x][x] = 1.0; hblk[x][x] = hi; }
            }
        } else {
            const int ga = at0 + ati, gb = at0 + atj;
            double R[3] = {pos[3 * ga] - pos[3 * gb], pos[3 * ga + 1] - pos[3 * gb + 1],
                           pos[3 * ga + 2] - pos[3 * gb + 2]};
            const double dist = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
            if (dist < par.int_cutoff && dist > 0.1) {
                double dummy[3][NI][NJ];
                shell_overlap_t<false, LI, LJ>(par.us_nprim[usi], s_alpha + usi * XTB_MAXPRIM,
                                               s_coeff + usi * XTB_MAXPRIM, par.us_nprim[usj],
                                               s_alpha + usj * XTB_MAXPRIM, s_coeff + usj * XTB_MAXPRIM, R,
                                               blk, dummy);
                const double rab = par.sp_rad[b.atom_sp[ga]] + par.sp_rad[b.atom_sp[gb]];
                double pi, dpi;
                dist_poly(dist, rab, par.us_shpoly[usi], par.us_shpoly[usj], pi, dpi);
                const double fac = 0.5 * (hi + hj) * par.us_K[usi * par.nus + usj] * pi;
#pragma unroll
                for (int x = 0; x < NI; ++x)
#pragma unroll
                    for (int y = 0; y < NJ; ++y) hblk[x][y] = blk[x][y] * fac;
            }
        }
#pragma unroll
        for (int x = 0; x < NI; ++x)
#pragma unroll
            for (int y = 0; y < NJ; ++y) {
                const size_t e1 = mat0 + (size_t)(oi + x) * n + (oj + y);
                const size_t e2 = mat0 + (size_t)(oj + y) * n + (oi + x);
                S[e1] = blk[x][y];
                H0[e1] = hblk[x][y];
                S[e2] = blk[x][y];
                H0[e2] = hblk[x][y];
            }
    });
}

// ---------------------------------------------------------------------------
// repulsion energy (repulsion/base.py:269-334, rep.py:65-98):
//   E = sum_{A<B, R<=25} Z_A Z_B exp(-sqrt(a_A a_B) R^kexp) / R
// of molecule m by one workgroup, written (not added): thread t < REP_LANES
// takes the atom pairs t, t + REP_LANES, ... of the upper triangle in order
// and the partial sums are added in a fixed order, so the energy has the
// same bits on every run and in every kernel that calls this (workgroups of
// 256 threads in the integral pass, 512 in the fused single point).
// ---------------------------------------------------------------------------
__device__ inline void repulsion_molecule(const xtb_batch& b, const xtb_params& par,
                                   const double* __restrict__ pos, double* out, int m,
                                   double* red) {
    constexpr int REP_LANES = 256;  // <= blockDim.x of every caller
    const int na = b.mol_nat[m], a0 = b.mol_atom0[m];
    const long np = (long)na * (na - 1) / 2;
    double e = 0.0;
    for (long p = threadIdx.x; p < np && (int)threadIdx.x < REP_LANES; p += REP_LANES) {
        // row A of the upper triangle holding pair index p (row A starts at
        // A (2 na - A - 1) / 2)
        int A = (int)((2.0 * na - 1.0 - sqrt((2.0 * na - 1.0) * (2.0 * na - 1.0) - 8.0 * (double)p)) * 0.5);
        if (A < 0) A = 0;
        while (A > 0 && (long)A * (2 * na - A - 1) / 2 > p) --A;
        while ((long)(A + 1) * (2 * na - A - 2) / 2 <= p) ++A;
        const int B = A + 1 + (int)(p - (long)A * (2 * na - A - 1) / 2);
        const int ga = a0 + A, gb = a0 + B;
        const double dx = pos[3 * ga] - pos[3 * gb], dy = pos[3 * ga + 1] - pos[3 * gb + 1],
                     dz = pos[3 * ga + 2] - pos[3 * gb + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (r <= par.rep_cutoff) {
            const int sa = b.atom_sp[ga], sb = b.atom_sp[gb];
            const double alpha = sqrt(par.sp_arep[sa] * par.sp_arep[sb] + 2.2250738585072014e-308);
            e += par.sp_zeff[sa] * par.sp_zeff[sb] * exp(-alpha * pow(r, par.kexp)) / r;
        }
    }
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < REP_LANES / 64; ++w) tot += red[w];
        out[(size_t)m * XTB_OUT_STRIDE + XTB_OUT_EREP] = tot;
    }
}

}  // namespace xtb
