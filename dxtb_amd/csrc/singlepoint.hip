// Fused single point on gfx950: one workgroup per molecule computes its
// integral pass, runs its SCF (the body of k_scf, scf_kernel.h) and writes
// its analytic gradient, so that a batch is one launch.  A translation unit
// of its own: sharing the SCF's out-of-line device functions with this
// kernel in one code object makes the compiler give those functions more
// conservative register conventions, which grew k_scf's stack frame.
#include "scf_kernel.h"
#include "integral_terms.h"
#include "gradient_terms.h"

namespace xtb {

// One molecule's integral pass by its own workgroup, with the per-element
// code of k_integrals (so the same bits): coordination numbers (CN_PARTS
// lanes per atom), S, H0 and gamma over the shell pairs i <= j (a wave per
// row i), the padding orbitals of a padded layout, and the repulsion energy
// into the molecule's record.
__device__ __noinline__ void molecule_integrals(const Mol& M, const xtb_batch& b, const xtb_params& par,
                                                const FusedArgs& fa, double* outs, const Shared& sh) {
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int a0 = 0; a0 < M.nat; a0 += NT / CN_PARTS) {
        // every lane of a CN_PARTS group takes part in the butterfly
        const int a = a0 + (int)threadIdx.x / CN_PARTS, part = threadIdx.x % CN_PARTS;
        const bool ok = a < M.nat;
        const double v = cn_combine(ok ? cn_part(b, par, fa.pos, M.at0 + a, part) : 0.0);
        if (ok && part == 0) fa.cn[M.at0 + a] = v;
    }
    double* s_alpha = (double*)sh.big;
    double* s_coeff = s_alpha + XTB_MAX_US_LDS * XTB_MAXPRIM;
    for (int i = threadIdx.x; i < par.nus * XTB_MAXPRIM; i += NT) {
        s_alpha[i] = par.us_alpha[i];
        s_coeff[i] = par.us_coeff[i];
    }
    __syncthreads();
    for (int i = w; i < M.nsh; i += NWAVE) {
        const double cni = fa.cn[M.at0 + b.sh_atom[M.sh0 + i]];
        for (int j = i + lane; j < M.nsh; j += WAVE) {
            shell_pair_s_h0(b, par, fa.pos, fa.S, fa.H0, M.m, i, j, cni, fa.cn[M.at0 + b.sh_atom[M.sh0 + j]],
                            s_alpha, s_coeff);
            gamma_pair(b, par, fa.pos, fa.gam, M.m, i, j, M.nsh);
        }
    }
    if (M.n > M.nao) {
        // padded layout: unit overlap diagonal and zero H0 on the padding
        // orbitals (as k_pad_fill)
        const size_t tot = (size_t)M.n * M.n;
        for (size_t e = threadIdx.x; e < tot; e += NT) {
            const int r = (int)(e / M.n), c = (int)(e % M.n);
            if (r < M.nao && c < M.nao) continue;
            fa.S[M.mat0 + e] = (r == c) ? 1.0 : 0.0;
            fa.H0[M.mat0 + e] = 0.0;
        }
    }
    repulsion_molecule(b, par, fa.pos, outs, M.m, (double*)sh.red);
    __syncthreads();
}

// One molecule's analytic gradient by its own workgroup, with the per-pair
// code of k_grad_pairs and k_grad_cn but no atomics: each atom pair A < B
// writes its terms to `scr` (the molecule's F buffer, free once W is formed:
// n^2 doubles, and every H, C, N, O atom carries at least two orbitals, so
// n^2 >= 4 nat^2 > 5 nat (nat - 1) / 2), then each atom adds its pairs in a
// fixed order; so the gradient has the same bits on every run.  Then the
// CN chain rule the same way.  dE/dCN goes to fa.dedcn, dE/dR to fa.grad.
__device__ __noinline__ void molecule_forces(const Mol& M, const xtb_batch& b, const xtb_params& par,
                                             const FusedArgs& fa, const Shared& sh) {
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int nat = M.nat, n = M.n;
    double* scr = (double*)M.F;
    if (5 * ((size_t)nat * (nat - 1) / 2) > (size_t)n * n) {
        for (int e = threadIdx.x; e < 3 * nat; e += NT) fa.grad[3 * (size_t)M.at0 + e] = NAN;
        return;
    }
    double* s_alpha = (double*)sh.big;
    double* s_coeff = s_alpha + XTB_MAX_US_LDS * XTB_MAXPRIM;
    for (int i = threadIdx.x; i < par.nus * XTB_MAXPRIM; i += NT) {
        s_alpha[i] = par.us_alpha[i];
        s_coeff[i] = par.us_coeff[i];
    }
    __syncthreads();
    auto pidx = [nat](int A, int B) { return (size_t)A * (2 * nat - A - 1) / 2 + (size_t)(B - A - 1); };
    const double* P = (const double*)M.P;
    for (int A = w; A < nat; A += NWAVE)
        for (int B = A + 1 + lane; B < nat; B += WAVE) {
            double g[3], dA, dB;
            atom_pair_grad(b, par, fa.pos, fa.cn, (const double*)M.gam, (const double*)M.qsh_out, P,
                           (const double*)M.T, (const double*)M.pot, n, M.nsh, M.sh0, M.at0 + A, M.at0 + B,
                           s_alpha, s_coeff, g, dA, dB);
            double* e = scr + 5 * pidx(A, B);
            e[0] = g[0];
            e[1] = g[1];
            e[2] = g[2];
            e[3] = dA;
            e[4] = dB;
        }
    __syncthreads();
    for (int A = threadIdx.x; A < nat; A += NT) {
        double g[3] = {0.0, 0.0, 0.0};
        double d = onsite_dedcn(b, par, P, n, M.sh0, M.at0 + A);
        for (int B = 0; B < A; ++B) {
            const double* e = scr + 5 * pidx(B, A);
            g[0] -= e[0];
            g[1] -= e[1];
            g[2] -= e[2];
            d += e[4];
        }
        for (int B = A + 1; B < nat; ++B) {
            const double* e = scr + 5 * pidx(A, B);
            g[0] += e[0];
            g[1] += e[1];
            g[2] += e[2];
            d += e[3];
        }
        fa.dedcn[M.at0 + A] = d;
        for (int k = 0; k < 3; ++k) fa.grad[3 * (size_t)(M.at0 + A) + k] = g[k];
    }
    __syncthreads();
    for (int A = w; A < nat; A += NWAVE)
        for (int B = A + 1 + lane; B < nat; B += WAVE) {
            double g[3] = {0.0, 0.0, 0.0};
            cn_pair_grad(b, par, fa.pos, fa.dedcn, M.at0 + A, M.at0 + B, g);
            double* e = scr + 3 * pidx(A, B);
            e[0] = g[0];
            e[1] = g[1];
            e[2] = g[2];
        }
    __syncthreads();
    for (int A = threadIdx.x; A < nat; A += NT) {
        double g[3] = {0.0, 0.0, 0.0};
        for (int B = 0; B < A; ++B)
            for (int k = 0; k < 3; ++k) g[k] -= scr[3 * pidx(B, A) + k];
        for (int B = A + 1; B < nat; ++B)
            for (int k = 0; k < 3; ++k) g[k] += scr[3 * pidx(A, B) + k];
        for (int k = 0; k < 3; ++k) fa.grad[3 * (size_t)(M.at0 + A) + k] += g[k];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(NT) k_singlepoint(xtb_batch b, xtb_params par, xtb_scf_opts o,
                                                    xtb_scf_state st, int big_cap, FusedArgs fa) {
    constexpr bool FUSED = true;
    const double* S = fa.S;
    const double* H0 = fa.H0;
    const double* gam = fa.gam;
#include "scf_body.inc"
}

}  // namespace xtb

using namespace xtb;

extern "C" int32_t xtb_singlepoint(const xtb_batch* batch, const xtb_params* par, const double* positions,
                                   double* S, double* H0, double* cn, double* gamma, const xtb_scf_opts* opts,
                                   const xtb_scf_state* state, double* dedcn, double* grad, void* stream) {
    if (!batch || !par || !positions || !S || !H0 || !cn || !gamma || !opts || !state) return XTB_EINVAL;
    if (opts->want_w && (!dedcn || !grad)) return XTB_EINVAL;
    int rc = check_scf_args(batch, opts, state);
    if (rc != XTB_OK) return rc;
    // the primitives of every unique shell are staged in the eigensolver's
    // LDS region (at least NGEMM_MFMA_LDS doubles)
    if (par->nus > XTB_MAX_US_LDS || 2 * XTB_MAX_US_LDS * XTB_MAXPRIM > NGEMM_MFMA_LDS) return XTB_ETOOBIG;
    int32_t bytes = 0, cap = 0;
    rc = xtb_scf_lds_bytes(batch, &bytes, &cap);
    if (rc != XTB_OK) return rc;
    if (hipFuncSetAttribute((const void*)k_singlepoint, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) !=
        hipSuccess)
        return XTB_EHIP;
    const FusedArgs fa{positions, cn, S, H0, gamma, dedcn, grad};
    hipLaunchKernelGGL(k_singlepoint, dim3(batch->nmol + opts->wg_skip), dim3(NT), bytes, (hipStream_t)stream,
                       *batch, *par, *opts, *state, cap, fa);
    return hipGetLastError() == hipSuccess ? XTB_OK : XTB_EHIP;
}
