// Overlap, core Hamiltonian, coordination number, ES2 Coulomb matrix and
// classical repulsion for a ragged batch of molecules (gfx950, FP64).
//
// Work decomposition: the shell-pair (atom-pair) space of every molecule is
// cut into 16x16 tiles of its upper triangle; one 256-thread workgroup per
// tile, one thread per shell (atom) pair.  The tile list is built on the
// host once per batch.  Primitive exponents and contraction coefficients of
// all unique shells are staged in LDS at workgroup start.
#include "integral_terms.h"

namespace xtb {

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
