// Analytic nuclear gradient of the GFN1 energy (gfx950, FP64).
//
// Because the SCF energy is variational, dE/dR needs no SCF response: it is
// the contraction of the converged density P, the energy-weighted density W
// and the orbital potential v with derivatives of the integrals, which are
// recomputed here per atom pair instead of being stored
// (upstream stores dS/dR as an (n, n, 3) tensor,
//  src/dxtb/_src/integral/driver/pytorch/impls/overlap.py:246-349).
//
//   dE/dR_A = sum_{mu in A, nu not in A} [2 P H~ - 2 W - P (v_mu + v_nu)] dS/dR_A
//           + sum_B 2 (P H~ S)_AB (dPi/dR)/(R Pi) (R_A - R_B)
//           + sum_C dE/dCN_C dCN_C/dR_A
//           - sum_B q_a q_b gamma_ab^3 (R_A - R_B)            (ES2, fixed q)
//           + repulsion
// (src/dxtb/_src/xtb/gfn1.py:185-408, src/dxtb/_src/calculators/types/
//  analytical.py:61-222, coulomb/secondorder.py:873-926,
//  classicals/repulsion/base.py:337-406, src/dxtb/_src/ncoord/utils.py:30-52).
#include "common.h"

#include <mutex>

namespace xtb {

// zero two arrays in one small launch: every workgroup of a launch waits for
// a free CU, which under a busy serving loop (CUs held by SCF workgroups)
// costs milliseconds per workgroup spread, so this is one workgroup
__global__ void __launch_bounds__(1024) k_zero2(double* __restrict__ a, int64_t na,
                                                double* __restrict__ b, int64_t nb) {
    for (int64_t i = threadIdx.x; i < na; i += blockDim.x) a[i] = 0.0;
    for (int64_t i = threadIdx.x; i < nb; i += blockDim.x) b[i] = 0.0;
}

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

__global__ void __launch_bounds__(256, 3) k_grad_pairs(xtb_batch b, xtb_params par,
                                                    const double* __restrict__ pos,
                                                    const double* __restrict__ cn,
                                                    const double* __restrict__ gam,
                                                    xtb_scf_state st, double* __restrict__ dedcn,
                                                    double* __restrict__ grad) {
    __shared__ double s_alpha[XTB_MAX_US_LDS_F * XTB_MAXPRIM];
    __shared__ double s_coeff[XTB_MAX_US_LDS_F * XTB_MAXPRIM];
    for (int i = threadIdx.x; i < par.nus * XTB_MAXPRIM; i += blockDim.x) {
        s_alpha[i] = par.us_alpha[i];
        s_coeff[i] = par.us_coeff[i];
    }
    __syncthreads();
    const int* tl = b.tile_at + 3 * blockIdx.x;
    const int m = tl[0];
    const int A = tl[1] + threadIdx.x / 16;
    const int B = tl[2] + threadIdx.x % 16;
    const int nat = b.mol_nat[m];
    if (A >= nat || B >= nat || B < A) return;
    const int n = b.mol_n[m];
    const int at0 = b.mol_atom0[m], sh0 = b.mol_sh0[m], vec0 = b.mol_vec0[m];
    const size_t mat0 = (size_t)b.mol_mat0[m];
    const double* P = st.P + mat0;
    const double* W = st.T + mat0;
    const double* v = st.pot + vec0;
    const int ga = at0 + A, gb = at0 + B;
    const int sa0 = b.atom_sh0[ga], nsa = b.atom_nsh[ga];
    if (A == B) {
        // on-site part of dE/dCN: sum_{i in A} (-kcn_i) sum_{mu in i} P_mumu
        double d = 0.0;
        for (int i = sa0; i < sa0 + nsa; ++i) {
            const int us = b.sh_us[sh0 + i];
            const int o0 = b.sh_ao[sh0 + i], nl = 2 * par.us_l[us] + 1;
            double pp = 0.0;
            for (int x = 0; x < nl; ++x) pp += P[(size_t)(o0 + x) * n + (o0 + x)];
            d -= par.us_kcn[us] * pp;
        }
        atomicAdd(dedcn + ga, d);
        return;
    }
    const int sb0 = b.atom_sh0[gb], nsb = b.atom_nsh[gb];
    const double R[3] = {pos[3 * ga] - pos[3 * gb], pos[3 * ga + 1] - pos[3 * gb + 1],
                         pos[3 * ga + 2] - pos[3 * gb + 2]};
    const double dist = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    const int spa = b.atom_sp[ga], spb = b.atom_sp[gb];
    const double rab = par.sp_rad[spa] + par.sp_rad[spb];
    const int nsh = b.mol_nsh[m];
    const double* g = gam + (size_t)b.mol_gam0[m];
    const double* qsh = st.qsh + sh0;
    double gvec[3] = {0.0, 0.0, 0.0};
    double radial = 0.0;  // coefficient of (R_A - R_B)
    double dA = 0.0, dB = 0.0;
    const bool in_cut = dist < par.int_cutoff && dist > 0.1;
    for (int i = sa0; i < sa0 + nsa; ++i) {
        const int usi = b.sh_us[sh0 + i];
        const int li = par.us_l[usi], ni = 2 * li + 1, oi = b.sh_ao[sh0 + i];
        const double hi = par.us_level[usi] - par.us_kcn[usi] * cn[ga];
        for (int j = sb0; j < sb0 + nsb; ++j) {
            const int usj = b.sh_us[sh0 + j];
            const int lj = par.us_l[usj], nj = 2 * lj + 1, oj = b.sh_ao[sh0 + j];
            // ES2 with fixed charges: d gamma / dR_A = -gamma^3 (R_A - R_B)
            const double gij = g[(size_t)i * nsh + j];
            radial -= gij * gij * gij * qsh[i] * qsh[j];
            if (!in_cut) continue;
            const double hj = par.us_level[usj] - par.us_kcn[usj] * cn[gb];
            double pi, dpi;
            dist_poly(dist, rab, par.us_shpoly[usi], par.us_shpoly[usj], pi, dpi);
            const double kk = par.us_K[usi * par.nus + usj];
            const double ht = 0.5 * (hi + hj) * kk * pi;
            double phs = 0.0;
            (void)ni;
            (void)nj;
            dispatch_l(li, lj, [&](auto LIc, auto LJc) {
                constexpr int LI = decltype(LIc)::value, LJ = decltype(LJc)::value;
                grad_shell_pair<LI, LJ>(par.us_nprim[usi], s_alpha + usi * XTB_MAXPRIM,
                                        s_coeff + usi * XTB_MAXPRIM, par.us_nprim[usj],
                                        s_alpha + usj * XTB_MAXPRIM, s_coeff + usj * XTB_MAXPRIM, R,
                                        P + (size_t)oi * n + oj, W + (size_t)oi * n + oj, n, v + oi, v + oj,
                                        ht, gvec, phs);
            });
            const double ps = phs;
            radial += 2.0 * phs * ht * dpi / pi;
            const double fac = kk * pi;
            dA -= ps * fac * par.us_kcn[usi];
            dB -= ps * fac * par.us_kcn[usj];
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
    for (int k = 0; k < 3; ++k) {
        const double gk = gvec[k] + radial * R[k];
        atomicAdd(grad + 3 * ga + k, gk);
        atomicAdd(grad + 3 * gb + k, -gk);
    }
    atomicAdd(dedcn + ga, dA);
    atomicAdd(dedcn + gb, dB);
}

// CN chain rule: dE/dR += sum_A dE/dCN_A dCN_A/dR  (ncoord/utils.py:30-52)
__global__ void __launch_bounds__(256) k_grad_cn(xtb_batch b, xtb_params par,
                                                 const double* __restrict__ pos,
                                                 const double* __restrict__ dedcn,
                                                 double* __restrict__ grad) {
    const int* tl = b.tile_at + 3 * blockIdx.x;
    const int m = tl[0];
    const int A = tl[1] + threadIdx.x / 16;
    const int B = tl[2] + threadIdx.x % 16;
    const int nat = b.mol_nat[m];
    if (A >= nat || B >= nat || B <= A) return;
    const int at0 = b.mol_atom0[m];
    const int ga = at0 + A, gb = at0 + B;
    const double R[3] = {pos[3 * ga] - pos[3 * gb], pos[3 * ga + 1] - pos[3 * gb + 1],
                         pos[3 * ga + 2] - pos[3 * gb + 2]};
    const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    if (r > par.cn_cutoff) return;
    const double rc = par.sp_rcov[b.atom_sp[ga]] + par.sp_rcov[b.atom_sp[gb]];
    const double e = exp(-par.cn_steep * (rc / r - 1.0));
    const double dcf = -par.cn_steep * rc * e / (r * r * (1.0 + e) * (1.0 + e));
    const double f = (dedcn[ga] + dedcn[gb]) * dcf / r;
    for (int k = 0; k < 3; ++k) {
        atomicAdd(grad + 3 * ga + k, f * R[k]);
        atomicAdd(grad + 3 * gb + k, -f * R[k]);
    }
}

}  // namespace xtb

using namespace xtb;

static int32_t launch_forces(const xtb_batch* batch, const xtb_params* par,
                             const double* positions, const double* cn, const double* gamma,
                             const xtb_scf_state* state, double* dedcn, double* grad,
                             void* stream) {
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero2, dim3(1), dim3(1024), 0, s, grad, (int64_t)3 * batch->nat_total, dedcn,
                       (int64_t)batch->nat_total);
    if (batch->ntile_at > 0) {
        hipLaunchKernelGGL(k_grad_pairs, dim3(batch->ntile_at), dim3(256), 0, s, *batch, *par,
                           positions, cn, gamma, *state, dedcn, grad);
        hipLaunchKernelGGL(k_grad_cn, dim3(batch->ntile_at), dim3(256), 0, s, *batch, *par,
                           positions, dedcn, grad);
    }
    return hipGetLastError() == hipSuccess ? XTB_OK : XTB_EHIP;
}

static int32_t check_forces_args(const xtb_batch* batch, const xtb_params* par,
                                 const double* positions, const double* cn, const double* gamma,
                                 const xtb_scf_state* state, const double* grad) {
    if (!batch || !par || !positions || !cn || !gamma || !state || !grad) return XTB_EINVAL;
    if (batch->nmol <= 0 || batch->nat_total <= 0) return XTB_EINVAL;
    if (!state->P || !state->T || !state->pot || !state->qsh) return XTB_EINVAL;
    if (par->nus > XTB_MAX_US_LDS_F) return XTB_ETOOBIG;
    return XTB_OK;
}

extern "C" int32_t xtb_forces(const xtb_batch* batch, const xtb_params* par,
                              const double* positions, const double* cn, const double* gamma,
                              const xtb_scf_state* state, double* dedcn, double* grad,
                              void* stream) {
    const int32_t rc = check_forces_args(batch, par, positions, cn, gamma, state, grad);
    if (rc != XTB_OK) return rc;
    if (!dedcn) return XTB_EINVAL;
    return launch_forces(batch, par, positions, cn, gamma, state, dedcn, grad, stream);
}

// ---------------------------------------------------------------------------
// xtb_forces without the caller's scratch argument, for callers that cannot
// allocate device memory.  dE/dCN (one double per atom, accumulated with
// atomics by k_grad_pairs and read by k_grad_cn) lives in a file-scope
// device array; its device address is looked up once per device with
// hipGetSymbolAddress (the array's name in host code is only a host-side
// stand-in, never a device address).  Every call on one device shares the
// array: calls that overlap on one device race.  k_zero2 clears it on the
// caller's stream before each use.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[XTB_NOALLOC_MAX_ATOMS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static int32_t scratch_for_current_device(double** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return XTB_EHIP;
    if (dev < 0 || dev >= kMaxDevices) return XTB_ETOOBIG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return XTB_EHIP;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return XTB_OK;
}

extern "C" int32_t xtb_forces_noalloc(const xtb_batch* batch, const xtb_params* par,
                                      const double* positions, const double* cn,
                                      const double* gamma, const xtb_scf_state* state,
                                      double* grad, void* stream) {
    int32_t rc = check_forces_args(batch, par, positions, cn, gamma, state, grad);
    if (rc != XTB_OK) return rc;
    if (batch->nat_total > XTB_NOALLOC_MAX_ATOMS) return XTB_ETOOBIG;
    double* dedcn = nullptr;
    rc = scratch_for_current_device(&dedcn);
    if (rc != XTB_OK) return rc;
    return launch_forces(batch, par, positions, cn, gamma, state, dedcn, grad, stream);
}
