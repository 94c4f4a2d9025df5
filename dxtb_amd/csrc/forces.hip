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
#include "gradient_terms.h"

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
    const int ga = at0 + A, gb = at0 + B;
    if (A == B) {
        atomicAdd(dedcn + ga, onsite_dedcn(b, par, P, n, sh0, ga));
        return;
    }
    double g[3], dA, dB;
    atom_pair_grad(b, par, pos, cn, gam + (size_t)b.mol_gam0[m], st.qsh + sh0, P, st.T + mat0,
                   st.pot + vec0, n, b.mol_nsh[m], sh0, ga, gb, s_alpha, s_coeff, g, dA, dB);
    for (int k = 0; k < 3; ++k) {
        atomicAdd(grad + 3 * ga + k, g[k]);
        atomicAdd(grad + 3 * gb + k, -g[k]);
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
    double g[3];
    if (!cn_pair_grad(b, par, pos, dedcn, ga, gb, g)) return;
    for (int k = 0; k < 3; ++k) {
        atomicAdd(grad + 3 * ga + k, g[k]);
        atomicAdd(grad + 3 * gb + k, -g[k]);
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
