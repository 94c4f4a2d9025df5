// The per-molecule SCF kernel k_scf (one workgroup runs one molecule's whole
// self-consistent cycle; the device code is in scf_kernel.h) and its host
// entry points xtb_scf_lds_bytes and xtb_scf.
#include "scf_kernel.h"

namespace xtb {

__global__ void __launch_bounds__(NT) k_scf(xtb_batch b, xtb_params par, const double* S,
                                            const double* H0, const double* gam,
                                            xtb_scf_opts o, xtb_scf_state st, int big_cap) {
    constexpr bool FUSED = false;
    const FusedArgs fa{};
#include "scf_body.inc"
}

}  // namespace xtb

using namespace xtb;

extern "C" int32_t xtb_scf_lds_bytes(const xtb_batch* batch, int32_t* bytes, int32_t* big_cap) {
    if (!batch || !bytes || !big_cap) return XTB_EINVAL;
    const Lay L = make_lay(batch->max_n, batch->max_nsh, batch->max_nat);
    const int budget = 160 * 1024 / 8;
    const int n = batch->max_n;
    const int want_c = n * (n + 1) / 2 + n * n;
    const int want_a = n * (n + 1) / 2;
    const int avail = budget - L.total_fixed;
    int cap = NGEMM_MFMA_LDS;
    if (want_c <= avail) cap = want_c > cap ? want_c : cap;
    else if (want_a <= avail) cap = want_a > cap ? want_a : cap;
    else cap = NBLOCK_LDS > cap ? NBLOCK_LDS : cap;  // block Jacobi
    if (L.total_fixed + cap > budget) return XTB_ETOOBIG;
    *big_cap = cap;
    *bytes = (L.total_fixed + cap) * 8;
    return XTB_OK;
}


extern "C" int32_t xtb_scf(const xtb_batch* batch, const xtb_params* par, const double* S,
                           const double* H0, const double* gamma, const xtb_scf_opts* opts,
                           const xtb_scf_state* state, void* stream) {
    if (!batch || !par || !S || !H0 || !gamma || !opts || !state) return XTB_EINVAL;
    int rc = check_scf_args(batch, opts, state);
    if (rc != XTB_OK) return rc;
    int32_t bytes = 0, cap = 0;
    rc = xtb_scf_lds_bytes(batch, &bytes, &cap);
    if (rc != XTB_OK) return rc;
    if (hipFuncSetAttribute((const void*)k_scf, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) !=
        hipSuccess)
        return XTB_EHIP;
    hipLaunchKernelGGL(k_scf, dim3(batch->nmol + opts->wg_skip), dim3(NT), bytes, (hipStream_t)stream, *batch, *par,
                       S, H0, gamma, *opts, *state, cap);
    return hipGetLastError() == hipSuccess ? XTB_OK : XTB_EHIP;
}


