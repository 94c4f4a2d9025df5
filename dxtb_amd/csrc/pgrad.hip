// Parameter gradients dE/dtheta of the converged GFN1 energy
// E = E_h0 + E_es2 + E_es3 + (-TS) + E_rep with respect to the Hamiltonian,
// Coulomb and repulsion parameters.  The SCF energy is stationary in the
// density (the Fock matrix is its derivative), so dE/dtheta is the partial
// derivative at the converged density and shell charges; no response term
// (the reason upstream's analytic forces need none either,
// src/dxtb/_src/calculators/types/analytical.py:61-222).  Upstream obtains
// these derivatives by autograd through the SCF (examples/optim/param.py,
// src/dxtb/_src/param/module/utils.py:123-240).
//
// xtb_param_grad writes per molecule a row of 4 nus + 3 nsp doubles:
//   [0, nus)          dE/d level_s      (Hartree per Hartree)
//   [nus, 2nus)       dE/d kcn_s
//   [2nus, 3nus)      dE/d shpoly_s
//   [3nus, 4nus)      dE/d eta_s        (ES2 shell hardness, lgam_s * gam_A)
//   [4nus, 4nus+nsp)  dE/d gam3_A
//   then dE/d zeff_A, dE/d arep_A, and dE/d K[a][b] (nus x nus, the H0
//   pair scaling K_ab as each off-site shell pair uses it).
// xtb_param_grad_basis adds dE/dzeta of one Slater exponent to a column.
#include <hip/hip_runtime.h>

#include "common.h"
#include "xtb_hip.h"

namespace xtb {

constexpr int PG_MAX_US = 32;
constexpr int PG_MAX_SP = 16;

__device__ __host__ inline int pg_stride(int nus, int nsp) { return 4 * nus + 3 * nsp + nus * nus; }

__device__ inline double* pg_row(double* out, int m, int nus, int nsp) {
    return out + (size_t)m * pg_stride(nus, nsp);
}

// One thread per shell pair (i <= j) of a 16 x 16 tile of one molecule:
//   H0: off-site E_ij = 2 OP_ij * 1/2 (h_i + h_j) K_ij Pi_ij with the block
//       population OP_ij = sum_{mu in i, nu in j} P S (k_s_h0's formula,
//       src/dxtb/_src/xtb/base.py:251-360); on-site E_ii = h_i sum_mu P_mumu.
//   ES2: E = 1/2 sum_ij q_i q_j gamma_ij, gamma = (d + avg^-2)^-1/2,
//       avg = 2 / (1/eta_i + 1/eta_j) (coulomb/secondorder.py:799-870):
//       d gamma / d eta_i = gamma^3 / (2 avg eta_i^2).
__global__ void __launch_bounds__(256) k_pgrad_shells(xtb_batch b, xtb_params par,
                                                      const double* __restrict__ pos,
                                                      const double* __restrict__ cn,
                                                      const double* __restrict__ S,
                                                      const double* __restrict__ P,
                                                      const double* __restrict__ qsh,
                                                      double* __restrict__ out) {
    __shared__ double acc[4 * PG_MAX_US];
    __shared__ double acck[PG_MAX_US * PG_MAX_US];
    const int nus = par.nus;
    for (int x = threadIdx.x; x < 4 * PG_MAX_US; x += blockDim.x) acc[x] = 0.0;
    for (int x = threadIdx.x; x < nus * nus; x += blockDim.x) acck[x] = 0.0;
    __syncthreads();
    const int* tl = b.tile_sh + 3 * blockIdx.x;
    const int m = tl[0];
    const int i = tl[1] + threadIdx.x / 16;
    const int j = tl[2] + threadIdx.x % 16;
    const int nsh = b.mol_nsh[m];
    if (i < nsh && j < nsh && j >= i) {
        const int n = b.mol_n[m];
        const int sh0 = b.mol_sh0[m], at0 = b.mol_atom0[m];
        const size_t mat0 = (size_t)b.mol_mat0[m];
        const int usi = b.sh_us[sh0 + i], usj = b.sh_us[sh0 + j];
        const int ati = b.sh_atom[sh0 + i], atj = b.sh_atom[sh0 + j];
        const int ni = 2 * par.us_l[usi] + 1, nj = 2 * par.us_l[usj] + 1;
        const int oi = b.sh_ao[sh0 + i], oj = b.sh_ao[sh0 + j];
        const double cni = cn[at0 + ati], cnj = cn[at0 + atj];
        const double eps = 2.220446049250313e-16;
        if (ati == atj) {
            if (i == j) {
                double pd = 0.0;
                for (int x = 0; x < ni; ++x) pd += P[mat0 + (size_t)(oi + x) * n + oi + x];
                atomicAdd(&acc[usi], pd);
                atomicAdd(&acc[PG_MAX_US + usi], -cni * pd);
            }
        } else {
            const int ga = at0 + ati, gb = at0 + atj;
            const double dx = pos[3 * ga] - pos[3 * gb], dy = pos[3 * ga + 1] - pos[3 * gb + 1],
                         dz = pos[3 * ga + 2] - pos[3 * gb + 2];
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            if (dist < par.int_cutoff && dist > 0.1) {
                double op = 0.0;
                for (int x = 0; x < ni; ++x)
                    for (int y = 0; y < nj; ++y) {
                        const size_t e = mat0 + (size_t)(oi + x) * n + oj + y;
                        op = fma(P[e], S[e], op);
                    }
                const double rab = par.sp_rad[b.atom_sp[ga]] + par.sp_rad[b.atom_sp[gb]];
                const double ki = par.us_shpoly[usi], kj = par.us_shpoly[usj];
                const double rr = sqrt(dist / rab);
                const double pi = (1.0 + ki * rr) * (1.0 + kj * rr);
                const double K = par.us_K[usi * nus + usj];
                const double hi = par.us_level[usi] - par.us_kcn[usi] * cni;
                const double hj = par.us_level[usj] - par.us_kcn[usj] * cnj;
                // both triangles of the symmetric blocks: factor 2
                const double g = 2.0 * op * 0.5 * K;  // d E / d h_i = d E / d h_j
                atomicAdd(&acc[usi], g * pi);
                atomicAdd(&acc[usj], g * pi);
                atomicAdd(&acc[PG_MAX_US + usi], -cni * g * pi);
                atomicAdd(&acc[PG_MAX_US + usj], -cnj * g * pi);
                const double gh = g * (hi + hj);
                atomicAdd(&acc[2 * PG_MAX_US + usi], gh * rr * (1.0 + kj * rr));
                atomicAdd(&acc[2 * PG_MAX_US + usj], gh * rr * (1.0 + ki * rr));
                atomicAdd(&acck[usi * nus + usj], op * (hi + hj) * pi);
            }
        }
        // ES2 (k_gamma's definitions)
        const int ga = at0 + ati, gb = at0 + atj;
        double d;
        if (ga == gb) d = eps;
        else {
            const double dx = pos[3 * ga] - pos[3 * gb], dy = pos[3 * ga + 1] - pos[3 * gb + 1],
                         dz = pos[3 * ga + 2] - pos[3 * gb + 2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz) + eps;
            d = r * r;
        }
        const double ei = par.us_hard[usi] + eps, ej = par.us_hard[usj] + eps;
        const double avg = 2.0 / (1.0 / ei + 1.0 / ej);
        const double gm = 1.0 / sqrt(d + 1.0 / (avg * avg));
        const double g3 = gm * gm * gm;
        const double qq = qsh[sh0 + i] * qsh[sh0 + j];
        if (i == j) {
            atomicAdd(&acc[3 * PG_MAX_US + usi], 0.5 * qq * g3 / (ei * ei * ei));
        } else {
            atomicAdd(&acc[3 * PG_MAX_US + usi], qq * g3 / (2.0 * avg * ei * ei));
            atomicAdd(&acc[3 * PG_MAX_US + usj], qq * g3 / (2.0 * avg * ej * ej));
        }
    }
    __syncthreads();
    double* row = pg_row(out, m, nus, par.nsp);
    for (int x = threadIdx.x; x < 4 * nus; x += blockDim.x) {
        const double v = acc[(x / nus) * PG_MAX_US + x % nus];
        if (v != 0.0) atomicAdd(row + x, v);
    }
    double* rowk = row + 4 * nus + 3 * par.nsp;
    for (int x = threadIdx.x; x < nus * nus; x += blockDim.x)
        if (acck[x] != 0.0) atomicAdd(rowk + x, acck[x]);
}

// One thread per atom pair (A <= B) of a 16 x 16 tile of one molecule:
//   ES3: E = sum_A gam3_A q_A^3 / 3 (thirdorder.py:182-357), A == B;
//   repulsion: E = sum_{A<B} z_A z_B exp(-sqrt(a_A a_B) R^kexp) / R for
//   R <= cutoff (repulsion/base.py:269-334).
__global__ void __launch_bounds__(256) k_pgrad_atoms(xtb_batch b, xtb_params par,
                                                     const double* __restrict__ pos,
                                                     const double* __restrict__ qsh,
                                                     double* __restrict__ out) {
    __shared__ double acc[3 * PG_MAX_SP];
    const int nsp = par.nsp;
    for (int x = threadIdx.x; x < 3 * PG_MAX_SP; x += blockDim.x) acc[x] = 0.0;
    __syncthreads();
    const int* tl = b.tile_at + 3 * blockIdx.x;
    const int m = tl[0];
    const int A = tl[1] + threadIdx.x / 16;
    const int B = tl[2] + threadIdx.x % 16;
    const int nat = b.mol_nat[m];
    if (A < nat && B < nat && B >= A) {
        const int at0 = b.mol_atom0[m], sh0 = b.mol_sh0[m];
        const int ga = at0 + A, gb = at0 + B;
        const int sa = b.atom_sp[ga], sb = b.atom_sp[gb];
        if (A == B) {
            const int s0 = b.atom_sh0[ga], ns = b.atom_nsh[ga];
            double qa = 0.0;
            for (int x = 0; x < ns; ++x) qa += qsh[sh0 + s0 + x];
            atomicAdd(&acc[sa], qa * qa * qa / 3.0);
        } else {
            const double dx = pos[3 * ga] - pos[3 * gb], dy = pos[3 * ga + 1] - pos[3 * gb + 1],
                         dz = pos[3 * ga + 2] - pos[3 * gb + 2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r <= par.rep_cutoff) {
                const double aa = par.sp_arep[sa], ab = par.sp_arep[sb];
                const double alpha = sqrt(aa * ab + 2.2250738585072014e-308);
                const double rk = pow(r, par.kexp);
                const double ex = exp(-alpha * rk) / r;
                const double za = par.sp_zeff[sa], zb = par.sp_zeff[sb];
                atomicAdd(&acc[PG_MAX_SP + sa], zb * ex);
                atomicAdd(&acc[PG_MAX_SP + sb], za * ex);
                const double de = -za * zb * ex * rk / (2.0 * alpha);  // d E / d alpha * d alpha / d a
                atomicAdd(&acc[2 * PG_MAX_SP + sa], de * ab);
                atomicAdd(&acc[2 * PG_MAX_SP + sb], de * aa);
            }
        }
    }
    __syncthreads();
    double* row = pg_row(out, m, par.nus, nsp) + 4 * par.nus;
    for (int x = threadIdx.x; x < 3 * nsp; x += blockDim.x) {
        const double v = acc[(x / nsp) * PG_MAX_SP + x % nsp];
        if (v != 0.0) atomicAdd(row + x, v);
    }
}

// Basis exponents: dE/dzeta = sum over off-site shell-pair blocks of
// [2 (P H~ - W) - P (v_mu + v_nu)] dS/dzeta, the expression of the analytic
// forces with dS/dR replaced (H0 = S H~ with H~ free of zeta; on-site S is
// the identity).  dS/dzeta by central differences of the contracted
// overlaps with the shells' primitives rebuilt at zeta +- h on the host
// (the tables (alpha, c)+-; `mask` marks the unique shells they change: the
// shell itself and a shell orthogonalised against it).  One thread per
// atom pair A < B of a 16 x 16 tile; result to column `col` of `out`.
__global__ void __launch_bounds__(256) k_pgrad_basis(xtb_batch b, xtb_params par,
                                                     const double* __restrict__ ap,
                                                     const double* __restrict__ cp,
                                                     const double* __restrict__ am,
                                                     const double* __restrict__ cm, uint32_t mask,
                                                     double inv2h, const double* __restrict__ pos,
                                                     const double* __restrict__ cn, xtb_scf_state st,
                                                     double* __restrict__ out, int col, int ncol) {
    __shared__ double s_ap[PG_MAX_US * XTB_MAXPRIM], s_cp[PG_MAX_US * XTB_MAXPRIM];
    __shared__ double s_am[PG_MAX_US * XTB_MAXPRIM], s_cm[PG_MAX_US * XTB_MAXPRIM];
    __shared__ double red[4];
    for (int x = threadIdx.x; x < par.nus * XTB_MAXPRIM; x += blockDim.x) {
        s_ap[x] = ap[x];
        s_cp[x] = cp[x];
        s_am[x] = am[x];
        s_cm[x] = cm[x];
    }
    __syncthreads();
    const int* tl = b.tile_at + 3 * blockIdx.x;
    const int m = tl[0];
    const int A = tl[1] + threadIdx.x / 16;
    const int B = tl[2] + threadIdx.x % 16;
    const int nat = b.mol_nat[m];
    double acc = 0.0;
    if (A < nat && B < nat && B > A) {
        const int n = b.mol_n[m];
        const int at0 = b.mol_atom0[m], sh0 = b.mol_sh0[m], vec0 = b.mol_vec0[m];
        const size_t mat0 = (size_t)b.mol_mat0[m];
        const double* P = st.P + mat0;
        const double* W = st.T + mat0;
        const double* v = st.pot + vec0;
        const int ga = at0 + A, gb = at0 + B;
        const double R[3] = {pos[3 * ga] - pos[3 * gb], pos[3 * ga + 1] - pos[3 * gb + 1],
                             pos[3 * ga + 2] - pos[3 * gb + 2]};
        const double dist = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        if (dist < par.int_cutoff && dist > 0.1) {
            const double rab = par.sp_rad[b.atom_sp[ga]] + par.sp_rad[b.atom_sp[gb]];
            const int sa0 = b.atom_sh0[ga], nsa = b.atom_nsh[ga];
            const int sb0 = b.atom_sh0[gb], nsb = b.atom_nsh[gb];
            for (int i = sa0; i < sa0 + nsa; ++i) {
                const int usi = b.sh_us[sh0 + i];
                const int li = par.us_l[usi], ni = 2 * li + 1, oi = b.sh_ao[sh0 + i];
                const double hi = par.us_level[usi] - par.us_kcn[usi] * cn[ga];
                for (int j = sb0; j < sb0 + nsb; ++j) {
                    const int usj = b.sh_us[sh0 + j];
                    if (!(((mask >> usi) | (mask >> usj)) & 1u)) continue;
                    const int lj = par.us_l[usj], nj = 2 * lj + 1, oj = b.sh_ao[sh0 + j];
                    const double hj = par.us_level[usj] - par.us_kcn[usj] * cn[gb];
                    double Sp[3][3], Sm[3][3], dummy[3][3][3];
                    shell_overlap<false>(li, lj, par.us_nprim[usi], s_ap + usi * XTB_MAXPRIM,
                                         s_cp + usi * XTB_MAXPRIM, par.us_nprim[usj],
                                         s_ap + usj * XTB_MAXPRIM, s_cp + usj * XTB_MAXPRIM, R, Sp, dummy);
                    shell_overlap<false>(li, lj, par.us_nprim[usi], s_am + usi * XTB_MAXPRIM,
                                         s_cm + usi * XTB_MAXPRIM, par.us_nprim[usj],
                                         s_am + usj * XTB_MAXPRIM, s_cm + usj * XTB_MAXPRIM, R, Sm, dummy);
                    double pi, dpi;
                    dist_poly(dist, rab, par.us_shpoly[usi], par.us_shpoly[usj], pi, dpi);
                    const double ht = 0.5 * (hi + hj) * par.us_K[usi * par.nus + usj] * pi;
                    for (int x = 0; x < ni; ++x)
                        for (int y = 0; y < nj; ++y) {
                            const size_t e = (size_t)(oi + x) * n + (oj + y);
                            const double sval = 2.0 * (P[e] * ht - W[e]) - P[e] * (v[oi + x] + v[oj + y]);
                            acc += sval * (Sp[x][y] - Sm[x][y]) * inv2h;
                        }
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double v = red[0] + red[1] + red[2] + red[3];
        if (v != 0.0) atomicAdd(out + (size_t)m * ncol + col, v);
    }
}

__global__ void k_pgrad_zero(double* __restrict__ a, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        a[e] = 0.0;
}

}  // namespace xtb

using namespace xtb;

extern "C" int32_t xtb_param_grad(const xtb_batch* batch, const xtb_params* par,
                                  const double* positions, const double* cn, const double* S,
                                  const xtb_scf_state* state, double* grad, void* stream) {
    if (!batch || !par || !positions || !cn || !S || !state || !grad) return XTB_EINVAL;
    if (!state->P || !state->qsh) return XTB_EINVAL;
    if (par->nus > PG_MAX_US || par->nsp > PG_MAX_SP || batch->nmol <= 0) return XTB_ETOOBIG;
    hipStream_t s = (hipStream_t)stream;
    const int64_t tot = (int64_t)batch->nmol * pg_stride(par->nus, par->nsp);
    hipLaunchKernelGGL(k_pgrad_zero, dim3(64), dim3(256), 0, s, grad, tot);
    if (batch->ntile_sh > 0)
        hipLaunchKernelGGL(k_pgrad_shells, dim3(batch->ntile_sh), dim3(256), 0, s, *batch, *par,
                           positions, cn, S, state->P, state->qsh, grad);
    if (batch->ntile_at > 0)
        hipLaunchKernelGGL(k_pgrad_atoms, dim3(batch->ntile_at), dim3(256), 0, s, *batch, *par,
                           positions, state->qsh, grad);
    return hipGetLastError() == hipSuccess ? XTB_OK : XTB_EHIP;
}

extern "C" int32_t xtb_param_grad_basis(const xtb_batch* batch, const xtb_params* par,
                                        const double* alpha_plus, const double* coeff_plus,
                                        const double* alpha_minus, const double* coeff_minus,
                                        uint32_t mask, double h, const double* positions,
                                        const double* cn, const xtb_scf_state* state, double* grad,
                                        int32_t col, int32_t ncol, void* stream) {
    if (!batch || !par || !alpha_plus || !coeff_plus || !alpha_minus || !coeff_minus || !positions ||
        !cn || !state || !grad || !(h > 0.0) || col < 0 || col >= ncol)
        return XTB_EINVAL;
    if (!state->P || !state->T || !state->pot) return XTB_EINVAL;
    if (par->nus > PG_MAX_US || batch->nmol <= 0) return XTB_ETOOBIG;
    if (batch->ntile_at > 0)
        hipLaunchKernelGGL(k_pgrad_basis, dim3(batch->ntile_at), dim3(256), 0, (hipStream_t)stream, *batch,
                           *par, alpha_plus, coeff_plus, alpha_minus, coeff_minus, mask, 0.5 / h,
                           positions, cn, *state, grad, col, ncol);
    return hipGetLastError() == hipSuccess ? XTB_OK : XTB_EHIP;
}
