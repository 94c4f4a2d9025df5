// Overlap, core Hamiltonian, coordination number, ES2 Coulomb matrix and
// classical repulsion for a ragged batch of molecules (gfx950, FP64).
//
// Work decomposition: the shell-pair (atom-pair) space of every molecule is
// cut into 16x16 tiles of its upper triangle; one 256-thread workgroup per
// tile, one thread per shell (atom) pair.  The tile list is built on the
// host once per batch.  Primitive exponents and contraction coefficients of
// all unique shells are staged in LDS at workgroup start.
#include "common.h"

namespace xtb {

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

__global__ void __launch_bounds__(256) k_cn(xtb_batch b, xtb_params par, const double* __restrict__ pos,
                                            double* __restrict__ cn) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = t / CN_PARTS, part = t % CN_PARTS;
    const bool ok = a < b.nat_total;  // the butterfly needs every lane of the group
    const double s = cn_combine(ok ? cn_part(b, par, pos, a, part) : 0.0);
    if (ok && part == 0) cn[a] = s;
}

__device__ inline void stage_prims(const xtb_params& par, double* s_alpha, double* s_coeff) {
    const int nn = par.nus * XTB_MAXPRIM;
    for (int i = threadIdx.x; i < nn; i += blockDim.x) {
        s_alpha[i] = par.us_alpha[i];
        s_coeff[i] = par.us_coeff[i];
    }
    __syncthreads();
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

__global__ void __launch_bounds__(256) k_s_h0(xtb_batch b, xtb_params par,
                                              const double* __restrict__ pos,
                                              const double* __restrict__ cn,
                                              double* __restrict__ S, double* __restrict__ H0,
                                              double* __restrict__ gam) {
    __shared__ double s_alpha[XTB_MAX_US_LDS * XTB_MAXPRIM];
    __shared__ double s_coeff[XTB_MAX_US_LDS * XTB_MAXPRIM];
    stage_prims(par, s_alpha, s_coeff);
    const int* tl = b.tile_sh + 3 * blockIdx.x;
    const int m = tl[0];
    const int i = tl[1] + threadIdx.x / 16;
    const int j = tl[2] + threadIdx.x % 16;
    const int nsh = b.mol_nsh[m];
    if (i >= nsh || j >= nsh || j < i) return;
    const int sh0 = b.mol_sh0[m], at0 = b.mol_atom0[m];
    const int ati = b.sh_atom[sh0 + i], atj = b.sh_atom[sh0 + j];
    shell_pair_s_h0(b, par, pos, S, H0, m, i, j, cn[at0 + ati], cn[at0 + atj], s_alpha, s_coeff);
    // the same shell pair's ES2 entry (one launch fewer per batch)
    if (gam) gamma_pair(b, par, pos, gam, m, i, j, nsh);
}

// Padding orbitals of the padded layout: unit diagonal in S (as the
// eigensolver sees it upstream, scf/unrolling/base.py:122-139), zero H0.
__global__ void k_pad_fill(xtb_batch b, double* __restrict__ S, double* __restrict__ H0) {
    const int m = blockIdx.y;
    const int n = b.mol_n[m], nao = b.mol_nao[m];
    if (n == nao) return;
    const size_t mat0 = (size_t)b.mol_mat0[m];
    const size_t tot = (size_t)n * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / n), c = (int)(e % n);
        if (r < nao && c < nao) continue;
        S[mat0 + e] = (r == c) ? 1.0 : 0.0;
        H0[mat0 + e] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// ES2 shell-resolved Coulomb matrix (secondorder.py:799-870, average.py:41-56)
//   gamma_ij = 1 / sqrt(d_AB + eta_ij^-2),  eta_ij = 2 / (1/eta_i + 1/eta_j)
//   d_AB = (R_AB + eps)^2 off-site, eps on-site.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gamma(xtb_batch b, xtb_params par,
                                               const double* __restrict__ pos,
                                               double* __restrict__ gam) {
    const int* tl = b.tile_sh + 3 * blockIdx.x;
    const int m = tl[0];
    const int i = tl[1] + threadIdx.x / 16;
    const int j = tl[2] + threadIdx.x % 16;
    const int nsh = b.mol_nsh[m];
    if (i >= nsh || j >= nsh || j < i) return;
    gamma_pair(b, par, pos, gam, m, i, j, nsh);
}

// ---------------------------------------------------------------------------
// repulsion energy (repulsion/base.py:269-334, rep.py:65-98):
//   E = sum_{A<B, R<=25} Z_A Z_B exp(-sqrt(a_A a_B) R^kexp) / R
// of molecule m by one workgroup, written (not added): thread t takes the
// atom pairs t, t + blockDim, ... of the upper triangle in order and the
// partial sums are added in a fixed order, so the energy has the same bits
// on every run.
// ---------------------------------------------------------------------------
__device__ void repulsion_molecule(const xtb_batch& b, const xtb_params& par,
                                   const double* __restrict__ pos, double* out, int m,
                                   double* red) {
    const int na = b.mol_nat[m], a0 = b.mol_atom0[m];
    const long np = (long)na * (na - 1) / 2;
    double e = 0.0;
    for (long p = threadIdx.x; p < np; p += blockDim.x) {
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
        for (int w = 0; w < (int)blockDim.x / 64; ++w) tot += red[w];
        out[(size_t)m * XTB_OUT_STRIDE + XTB_OUT_EREP] = tot;
    }
}

__global__ void __launch_bounds__(256) k_repulsion_mol(xtb_batch b, xtb_params par,
                                                       const double* __restrict__ pos, double* out) {
    __shared__ double red[4];
    repulsion_molecule(b, par, pos, out, blockIdx.x, red);
}

// ---------------------------------------------------------------------------
// The whole integral pass in one launch: workgroups [0, ntile_sh) are the
// shell-pair tiles of k_s_h0 (S, H0 and gamma), each first forming the
// coordination numbers of its tile's atoms in LDS (the same summation as
// k_cn, so the same bits; the diagonal tile holding an atom's first shell
// also stores it); workgroups [ntile_sh, ntile_sh + nmol) each write one
// molecule's repulsion energy.  One launch instead of four (CN, S/H0/gamma,
// zeroing, repulsion): under a busy serving loop every launch waits for
// free compute units.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_integrals(xtb_batch b, xtb_params par,
                                                   const double* __restrict__ pos,
                                                   double* __restrict__ cn, double* __restrict__ S,
                                                   double* __restrict__ H0, double* __restrict__ gam,
                                                   double* __restrict__ out) {
    __shared__ double s_alpha[XTB_MAX_US_LDS * XTB_MAXPRIM];
    __shared__ double s_coeff[XTB_MAX_US_LDS * XTB_MAXPRIM];
    __shared__ double s_cn[32];
    if ((int)blockIdx.x >= b.ntile_sh) {
        repulsion_molecule(b, par, pos, out, blockIdx.x - b.ntile_sh, s_cn);
        return;
    }
    const int* tl = b.tile_sh + 3 * blockIdx.x;
    const int m = tl[0], i0 = tl[1], j0 = tl[2];
    const int nsh = b.mol_nsh[m];
    const int sh0 = b.mol_sh0[m], at0 = b.mol_atom0[m];
    // atoms of the tile's row shells [i0, i0+16) and column shells [j0, j0+16)
    const int ra0 = b.sh_atom[sh0 + i0], ra1 = b.sh_atom[sh0 + min(i0 + 15, nsh - 1)];
    const int ca0 = b.sh_atom[sh0 + j0], ca1 = b.sh_atom[sh0 + min(j0 + 15, nsh - 1)];
    {
        const int slot = threadIdx.x / CN_PARTS, part = threadIdx.x % CN_PARTS;
        const int la = slot < 16 ? ra0 + slot : ca0 + (slot - 16);
        const bool ok = slot < 16 ? la <= ra1 : la <= ca1;
        const double v = cn_combine(ok ? cn_part(b, par, pos, at0 + la, part) : 0.0);
        if (ok && part == 0) {
            s_cn[slot] = v;
            // the diagonal tile whose rows hold the atom's first shell stores it
            if (slot < 16 && i0 == j0 && b.atom_sh0[at0 + la] >= i0) cn[at0 + la] = v;
        }
    }
    __syncthreads();
    stage_prims(par, s_alpha, s_coeff);
    const int i = i0 + threadIdx.x / 16;
    const int j = j0 + threadIdx.x % 16;
    if (i >= nsh || j >= nsh || j < i) return;
    const int ati = b.sh_atom[sh0 + i], atj = b.sh_atom[sh0 + j];
    shell_pair_s_h0(b, par, pos, S, H0, m, i, j, s_cn[ati - ra0], s_cn[16 + atj - ca0], s_alpha, s_coeff);
    gamma_pair(b, par, pos, gam, m, i, j, nsh);
}

// repulsion gradient (repulsion/base.py:337-406), one thread per atom A
// summing over the other atoms of its molecule in order (no atomics, so the
// result is the same bit for bit on every run):
//   dE/dR_A = sum_B dE_AB/dr (R_A - R_B) / r,
//   dE_AB/dr = -E_AB (1/r + a kexp r^(kexp-1))
__global__ void __launch_bounds__(256) k_repulsion_grad(xtb_batch b, xtb_params par,
                                                        const double* __restrict__ pos,
                                                        double* __restrict__ grad) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= b.nat_total) return;
    const int m = b.atom_mol[a];
    const int a0 = b.mol_atom0[m], na = b.mol_nat[m];
    const int sa = b.atom_sp[a];
    const double x = pos[3 * a], y = pos[3 * a + 1], z = pos[3 * a + 2];
    double g[3] = {0.0, 0.0, 0.0};
    for (int j = a0; j < a0 + na; ++j) {
        if (j == a) continue;
        const double R[3] = {x - pos[3 * j], y - pos[3 * j + 1], z - pos[3 * j + 2]};
        const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        if (r > par.rep_cutoff) continue;
        const int sb = b.atom_sp[j];
        const double alpha = sqrt(par.sp_arep[sa] * par.sp_arep[sb] + 2.2250738585072014e-308);
        const double rk = pow(r, par.kexp);
        const double e = par.sp_zeff[sa] * par.sp_zeff[sb] * exp(-alpha * rk) / r;
        const double f = -e * (1.0 / r + alpha * par.kexp * rk / r) / r;
        for (int k = 0; k < 3; ++k) g[k] += f * R[k];
    }
    for (int k = 0; k < 3; ++k) grad[3 * a + k] = g[k];
}

}  // namespace xtb

using namespace xtb;

static int check_launch() { return hipGetLastError() == hipSuccess ? XTB_OK : XTB_EHIP; }

extern "C" int32_t xtb_abi_version(void) { return XTB_ABI_VERSION; }

extern "C" int32_t xtb_abi_struct_sizes(int64_t* out4) {
    if (!out4) return XTB_EINVAL;
    out4[0] = sizeof(xtb_batch);
    out4[1] = sizeof(xtb_params);
    out4[2] = sizeof(xtb_scf_opts);
    out4[3] = sizeof(xtb_scf_state);
    return XTB_OK;
}

extern "C" int32_t xtb_workspace_doubles(const xtb_batch* batch, int32_t generations,
                                         int64_t* out4) {
    if (!batch || !out4 || generations < 0) return XTB_EINVAL;
    out4[0] = batch->nmat_total;
    out4[1] = batch->ngam_total;
    out4[2] = batch->nvec_total;
    out4[3] = (int64_t)(generations + 1) * batch->nvec_total;
    return XTB_OK;
}

static int32_t build_s_h0(const xtb_batch* batch, const xtb_params* par, const double* positions,
                          double* S, double* H0, double* cn, double* gamma, void* stream) {
    if (!batch || !par || !positions || !S || !H0 || !cn) return XTB_EINVAL;
    if (batch->nmol <= 0 || batch->nat_total <= 0) return XTB_EINVAL;
    if (par->nus > XTB_MAX_US_LDS) return XTB_ETOOBIG;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_cn, dim3((batch->nat_total * CN_PARTS + 255) / 256), dim3(256), 0, s, *batch,
                       *par, positions, cn);
    if (batch->ntile_sh > 0)
        hipLaunchKernelGGL(k_s_h0, dim3(batch->ntile_sh), dim3(256), 0, s, *batch, *par, positions,
                           cn, S, H0, gamma);
    // padding orbitals (padded layout only: every workgroup of a launch waits
    // for a free CU, which under a busy serving loop costs milliseconds)
    if (batch->has_padding)
        hipLaunchKernelGGL(k_pad_fill, dim3(8, batch->nmol), dim3(256), 0, s, *batch, S, H0);
    return check_launch();
}

extern "C" int32_t xtb_build_s_h0(const xtb_batch* batch, const xtb_params* par,
                                  const double* positions, double* S, double* H0, double* cn,
                                  void* stream) {
    return build_s_h0(batch, par, positions, S, H0, cn, nullptr, stream);
}

extern "C" int32_t xtb_build_s_h0_gamma(const xtb_batch* batch, const xtb_params* par,
                                        const double* positions, double* S, double* H0,
                                        double* cn, double* gamma, void* stream) {
    if (!gamma) return XTB_EINVAL;
    return build_s_h0(batch, par, positions, S, H0, cn, gamma, stream);
}

extern "C" int32_t xtb_build_gamma(const xtb_batch* batch, const xtb_params* par,
                                   const double* positions, double* gamma, void* stream) {
    if (!batch || !par || !positions || !gamma) return XTB_EINVAL;
    if (batch->nmol <= 0) return XTB_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    if (batch->ntile_sh > 0)
        hipLaunchKernelGGL(k_gamma, dim3(batch->ntile_sh), dim3(256), 0, s, *batch, *par, positions,
                           gamma);
    return check_launch();
}

extern "C" int32_t xtb_repulsion(const xtb_batch* batch, const xtb_params* par,
                                 const double* positions, double* out, void* stream) {
    if (!batch || !par || !positions || !out) return XTB_EINVAL;
    if (batch->nmol <= 0) return XTB_EINVAL;
    hipLaunchKernelGGL(k_repulsion_mol, dim3(batch->nmol), dim3(256), 0, (hipStream_t)stream, *batch,
                       *par, positions, out);
    return check_launch();
}

extern "C" int32_t xtb_build_integrals(const xtb_batch* batch, const xtb_params* par,
                                       const double* positions, double* S, double* H0, double* cn,
                                       double* gamma, double* out, void* stream) {
    if (!batch || !par || !positions || !S || !H0 || !cn || !gamma || !out) return XTB_EINVAL;
    if (batch->nmol <= 0 || batch->nat_total <= 0 || batch->ntile_sh <= 0) return XTB_EINVAL;
    if (par->nus > XTB_MAX_US_LDS) return XTB_ETOOBIG;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_integrals, dim3(batch->ntile_sh + batch->nmol), dim3(256), 0, s, *batch, *par,
                       positions, cn, S, H0, gamma, out);
    if (batch->has_padding)
        hipLaunchKernelGGL(k_pad_fill, dim3(8, batch->nmol), dim3(256), 0, s, *batch, S, H0);
    return check_launch();
}

extern "C" int32_t xtb_repulsion_gradient(const xtb_batch* batch, const xtb_params* par,
                                          const double* positions, double* grad, void* stream) {
    if (!batch || !par || !positions || !grad) return XTB_EINVAL;
    if (batch->nmol <= 0 || batch->nat_total <= 0) return XTB_EINVAL;
    hipLaunchKernelGGL(k_repulsion_grad, dim3((batch->nat_total + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, *batch, *par, positions, grad);
    return check_launch();
}
