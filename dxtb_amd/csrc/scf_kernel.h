// Persistent per-molecule GFN1 SCF on gfx950: one workgroup owns one
// molecule and runs its whole self-consistent cycle without returning to
// the host.  Follows upstream's default ("full") SCF in potential mode
// (src/dxtb/_src/scf/unrolling/default.py:71-140, scf/base.py:651-907):
//
//   v -> F = H0 - 1/2 S o (v_i + v_j)           (Fock build)
//     -> F C = S C e                             (generalized eigensolve)
//     -> f = Fermi(e), P = C diag(f) C^T         (smearing, density)
//     -> q = n0 - diag(P S), e_orb = diag(P H0)  (Mulliken)
//     -> v' = spread(gamma q_sh + hd q_at^2)     (ES2 + ES3 potential)
//   mixed with upstream's Anderson mixer (scf/mixer/anderson.py:163-317).
//
// Eigensolver: the generalized problem is reduced once per geometry with the
// Loewdin orthogonaliser X = S^-1/2 (coupled Newton-Schulz on matrix cores).  Every
// SCF step transforms F into the basis of the previous step's orbitals,
// A = C^T F C, which is close to diagonal, and finishes it with a few sweeps
// of parallel-ordered cyclic Jacobi whose rotations are applied directly to
// the stored orbitals.  A lives in LDS (packed upper triangle) whenever it
// fits, and so do the orbitals when both fit.
#pragma once

#include "scf_core.h"

namespace xtb {

struct Lay {
    // LDS layout in doubles (offsets), identical on host and device
    int red, cs, skey, sidx, ev, ftot, qsh, vsh, gq, qat, big, total_fixed;
};

__host__ __device__ inline int next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

__host__ __device__ inline Lay make_lay(int nmax, int max_nsh, int max_nat) {
    Lay L;
    const int npow = next_pow2(nmax < 2 ? 2 : nmax);
    int o = 0;
    L.red = o; o += 256;
    L.cs = o; o += 3 * (nmax / 2 + 2);
    L.skey = o; o += npow;
    L.sidx = o; o += (npow + 1) / 2;  // ints
    L.ev = o; o += nmax;
    L.ftot = o; o += nmax;
    L.qsh = o; o += max_nsh;
    L.vsh = o; o += max_nsh;
    L.gq = o; o += max_nsh;
    L.qat = o; o += max_nat;
    o = (o + 1) & ~1;
    L.big = o;
    L.total_fixed = o;
    return L;
}

// ---------------------------------------------------------------------------
// eigen step: A (full, in mol.F) -> eigenvalues in sh.ev, rotations applied
// to mol.Ct.  Returns Jacobi sweeps (-1 on failure).
// ---------------------------------------------------------------------------
__device__ __noinline__ int eig_step(const Mol& M, const Shared& sh, int max_sweeps, double rel_tol,
                                     double loose, int inner, int cb_lo, int cb_hi, bool intra) {
    const int n = M.n;
    // scale for the absolute tolerance
    double dmax = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) dmax = fmax(dmax, fabs(M.F[(size_t)i * n + i]));
    dmax = wg_max(dmax, sh.red);
    const double tol = fmax(rel_tol * fmax(dmax, 1.0), loose);
    const int npk = n * (n + 1) / 2;
    int sweeps;
    if (npk <= sh.big_cap) {
        SymPacked A{sh.big, n};
        for (int i = threadIdx.x / WAVE; i < n; i += NWAVE)
            for (int j = i + threadIdx.x % WAVE; j < n; j += WAVE) A.at(i, j) = M.F[(size_t)i * n + j];
        const bool cl = (npk + n * n) <= sh.big_cap;
        ldsf* cbuf = sh.big + npk;
        if (cl) {
            for (int idx = threadIdx.x; idx < n * n; idx += NT) cbuf[idx] = M.Ct[idx];
        }
        __syncthreads();
        if (cl) sweeps = wg_jacobi(A, n, RowsLds{cbuf, n}, max_sweeps, tol, sh.cs, sh.red, cb_lo, cb_hi, intra);
        else sweeps = wg_jacobi(A, n, RowsGlobal{M.Ct, n}, max_sweeps, tol, sh.cs, sh.red, cb_lo, cb_hi, intra);
        for (int i = threadIdx.x; i < n; i += NT) sh.ev[i] = A.at(i, i);
        if (cl) {
            for (int idx = threadIdx.x; idx < n * n; idx += NT) M.Ct[idx] = cbuf[idx];
        }
    } else if (sh.big_cap >= NBLOCK_LDS) {
        // block Jacobi; T is free scratch here and holds the subproblem rotations
        sweeps = wg_block_jacobi(M.F, n, RowsGlobal{M.Ct, n}, M.T, max_sweeps, tol, sh.big, sh.red,
                                 sh.tacc, inner, cb_lo, cb_hi, intra);
        for (int i = threadIdx.x; i < n; i += NT) sh.ev[i] = M.F[(size_t)i * n + i];
    } else {
        SymFull A{M.F, n};
        sweeps = wg_jacobi(A, n, RowsGlobal{M.Ct, n}, max_sweeps, tol, sh.cs, sh.red, cb_lo, cb_hi, intra);
        for (int i = threadIdx.x; i < n; i += NT) sh.ev[i] = A.at(i, i);
    }
    __syncthreads();
    return sweeps;
}

// Loewdin orthogonaliser Ct <- X = S^-1/2 (symmetric, so rows == columns) by
// the coupled Newton-Schulz iteration on matrix cores:
//   Y0 = S/c, Z0 = I;  T = (3I - Z Y)/2;  Y <- Y T;  Z <- T Z
// Y -> (S/c)^1/2 and Z -> (S/c)^-1/2 quadratically once ||I - S/c|| < 1; c is a
// Gershgorin bound on the largest eigenvalue of S.  X = Z / sqrt(c).
__device__ __noinline__ int ortho_setup(const Mol& M, const Shared& sh, int* status) {
    const int n = M.n;
    const size_t nn = (size_t)n * n;
    double rs = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) {
        double a = 0.0;
        for (int j = 0; j < n; ++j) a += fabs(M.S[(size_t)i * n + j]);
        rs = fmax(rs, a);
    }
    const double c = wg_max(rs, sh.red) * 1.0001;
    glbf* Y = M.F;
    glbf* Z = M.Ct;
    glbf* T = M.T;
    glbf* W = M.P;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int r = w; r < n; r += NWAVE)
        for (int q = lane; q < n; q += WAVE) {
            const size_t e = (size_t)r * n + q;
            Y[e] = M.S[e] / c;
            Z[e] = (r == q) ? 1.0 : 0.0;
        }
    __syncthreads();
    int it = 0;
    int extra = -1;  // iterations left once in the quadratic regime
    for (; it < 60; ++it) {
        gemm<false, false>(n, n, n, -0.5, Z, n, Y, n, 0.0, T, n, sh.big);  // -(Z Y)/2
        double err = 0.0;
        for (int r = w; r < n; r += NWAVE)
            for (int q = lane; q < n; q += WAVE) {
                const size_t e = (size_t)r * n + q;
                const double v = T[e] + ((r == q) ? 1.5 : 0.0);
                err = fmax(err, fabs(v - ((r == q) ? 1.0 : 0.0)));
                T[e] = v;
            }
        err = wg_max(err, sh.red);
        if (extra == 0 || err < 1e-15) break;
        if (extra > 0) --extra;
        else if (err < 1e-9) extra = 1;  // quadratic: one more step reaches rounding level
        gemm<false, false>(n, n, n, 1.0, Y, n, T, n, 0.0, W, n, sh.big);  // Y T
        gemm<false, false>(n, n, n, 1.0, T, n, Z, n, 0.0, Y, n, sh.big);  // T Z (into Y's buffer)
        // rotate buffers: Y <- W, Z <- (T Z) held in Y's buffer
        glbf* oldZ = Z;
        Z = Y;
        Y = W;
        W = oldZ;
    }
    if (it >= 60) *status |= XTB_MOL_OVERLAP;  // S numerically singular
    const double sc = 1.0 / sqrt(c);
    for (size_t e = threadIdx.x; e < nn; e += NT) M.Ct[e] = Z[e] * sc;  // Z may alias Ct
    __syncthreads();
    // symmetrise (rounding)
    for (int r = w; r < n; r += NWAVE)
        for (int q = r + 1 + lane; q < n; q += WAVE) {
            const size_t e = (size_t)r * n + q;
            const double a = 0.5 * (M.Ct[e] + M.Ct[(size_t)q * n + r]);
            M.Ct[e] = a;
            M.Ct[(size_t)q * n + r] = a;
        }
    __syncthreads();
    return 0;
}

// ---------------------------------------------------------------------------
// Orthogonaliser from the Cholesky factor S = L L^T: the rows of X = L^-1 are
// S-orthonormal (X S X^T = I), so Ct <- X starts the SCF like the Loewdin
// X = S^-1/2 does (eigenpairs of the generalised problem do not depend on the
// choice), at about n^3 flops instead of Newton-Schulz's ~10 x 6 n^3.
// Blocked right-looking factorisation with 32-wide panels in T (lower
// triangle): each diagonal block is factored and inverted by one wave in
// LDS; the panel below (x D^T = b, i.e. x = b D^-T) and the trailing update
// are matrix-core GEMMs.  X follows by block forward substitution L X = I:
// X_i = D_i^-1 (E_i - L_i,<i X_<i), again with GEMMs.  A block padded beyond
// n is the identity.  Scratch: F (up to 32 x n, or mt x 32 with stride n),
// P (the nb x nb inverse, stride nb).
// Returns 0; a non-positive pivot (S not positive definite, where upstream's
// torch.linalg.cholesky raises) sets XTB_MOL_OVERLAP in *status.
// ---------------------------------------------------------------------------
// In-place blocked right-looking Cholesky L L^T = A of the n x n symmetric
// positive definite A (lower triangle read and overwritten by L; the strict
// upper triangle is left as it was), 32-wide panels: one wave factors and
// inverts each diagonal block in LDS, the panel solve and the trailing
// update are matrix-core GEMMs.  R (>= n rows x 32, leading dimension n) and
// Dg (32 x 32) are global scratch.  Returns false when a pivot is not
// positive.
__device__ __noinline__ bool chol_factor(glbf* L, int n, glbf* R, glbf* Dg, const Shared& sh) {
    constexpr int NB = 32;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    // LDS: two 32 x 32 blocks at the end of the GEMM staging area; every GEMM
    // below runs after the blocks were consumed (the inverse goes through Dg)
    ldsf* D = sh.big + NGEMM_MFMA_LDS - 2 * NB * NB;
    ldsf* Di = D + NB * NB;
    bool bad = false;
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int nb = n - j0 < NB ? n - j0 : NB;
        for (int e = threadIdx.x; e < NB * NB; e += NT) {
            const int r = e / NB, c = e % NB;
            D[e] = (r < nb && c < nb) ? L[(size_t)(j0 + r) * n + j0 + c] : (r == c ? 1.0 : 0.0);
        }
        __syncthreads();
        if (w == 0) {
            if (!chol_block_factor(D, lane)) bad = true;
            chol_block_inverse(D, Di, Dg, nb, lane);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nb * nb; e += NT) {
            const int r = e / nb, c = e % nb;
            if (c <= r) L[(size_t)(j0 + r) * n + j0 + c] = D[r * NB + c];
        }
        const int mt = n - j0 - nb;
        if (mt > 0) {
            glbf* Pn = L + (size_t)(j0 + nb) * n + j0;  // panel below the block
            gemm<false, true>(mt, nb, nb, 1.0, Pn, n, Dg, nb, 0.0, R, n, sh.big);  // b D^-T
            for (int r = w; r < mt; r += NWAVE)
                for (int c = lane; c < nb; c += WAVE) Pn[(size_t)r * n + c] = R[(size_t)r * n + c];
            __syncthreads();
            gemm<false, true>(mt, mt, nb, -1.0, Pn, n, Pn, n, 1.0, Pn + nb, n, sh.big);
        } else {
            __syncthreads();
        }
    }
    return !__syncthreads_or(bad);
}

// Load the lower-triangular diagonal block i0 of L (nb x nb, NB-padded with
// the identity) and its inverse into Dg (leading dimension nb).
__device__ inline void chol_diag_inverse(const glbf* L, int n, int i0, int nb, glbf* Dg, const Shared& sh) {
    constexpr int NB = 32;
    ldsf* D = sh.big + NGEMM_MFMA_LDS - 2 * NB * NB;
    ldsf* Di = D + NB * NB;
    for (int e = threadIdx.x; e < NB * NB; e += NT) {
        const int r = e / NB, c = e % NB;
        D[e] = (r < nb && c < nb) ? (c <= r ? L[(size_t)(i0 + r) * n + i0 + c] : 0.0) : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    if (threadIdx.x < WAVE) chol_block_inverse(D, Di, Dg, nb, threadIdx.x);
    __syncthreads();
}

// Cholesky orthogonaliser: L L^T = S, Ct <- X = L^-1 (rows S-orthonormal).
// X by block rows, X_i = D_i^-1 (E_i - L_i,<i X_<i), where only the first
// i0 + nb columns of block row i are non-zero.  L in T, R in F, Dg in P.
__device__ __noinline__ int chol_setup(const Mol& M, const Shared& sh, int* status) {
    constexpr int NB = 32;
    const int n = M.n;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    glbf* L = M.T;
    glbf* X = M.Ct;
    glbf* R = M.F;   // NB x n (and mt x NB) scratch
    glbf* Dg = M.P;  // NB x NB inverse of the diagonal block
    for (int r = w; r < n; r += NWAVE)
        for (int c = lane; c < n; c += WAVE) L[(size_t)r * n + c] = M.S[(size_t)r * n + c];
    __syncthreads();
    if (!chol_factor(L, n, R, Dg, sh)) *status |= XTB_MOL_OVERLAP;
    for (int r = w; r < n; r += NWAVE)
        for (int c = lane; c < n; c += WAVE) X[(size_t)r * n + c] = 0.0;
    __syncthreads();
    for (int i0 = 0; i0 < n; i0 += NB) {
        const int nb = n - i0 < NB ? n - i0 : NB;
        if (i0 > 0) gemm<false, false>(nb, i0, i0, -1.0, L + (size_t)i0 * n, n, X, n, 0.0, R, n, sh.big);
        for (int e = threadIdx.x; e < nb * nb; e += NT) {
            const int r = e / nb, c = e % nb;
            R[(size_t)r * n + i0 + c] = (r == c) ? 1.0 : 0.0;
        }
        chol_diag_inverse(L, n, i0, nb, Dg, sh);
        gemm<false, false>(nb, i0 + nb, nb, 1.0, Dg, nb, R, n, 0.0, X + (size_t)i0 * n, n, sh.big);
    }
    return 0;
}

// Warm start: the kept orbitals (rows of Ct, S_old-orthonormal, ascending
// energy) are re-orthonormalised against the new overlap by Cholesky
// Gram-Schmidt in their order, C <- L^-1 C with G = C S C^T = L L^T, so the
// lowest orbitals stay within the span of the lowest kept ones.  G and L in
// P, C S and Dg in T, R in F; the forward substitution C_i <- D_i^-1 (C_i -
// L_i,<i C_<i) runs in place by block rows.  Returns false (Ct unusable,
// start cold) when G is not numerically positive definite.
__device__ __noinline__ bool warm_orthonormalise(const Mol& M, const Shared& sh) {
    constexpr int NB = 32;
    const int n = M.n;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    glbf* G = M.P;
    glbf* R = M.F;
    glbf* Dg = M.T;
    gemm<false, false>(n, n, n, 1.0, M.Ct, n, M.S, n, 0.0, M.T, n, sh.big);  // C S
    gemm_sym<false, true>(n, n, 1.0, M.T, n, M.Ct, n, G, n, sh.big);         // C S C^T
    if (!chol_factor(G, n, R, Dg, sh)) return false;
    for (int i0 = 0; i0 < n; i0 += NB) {
        const int nb = n - i0 < NB ? n - i0 : NB;
        glbf* Ci = M.Ct + (size_t)i0 * n;
        for (int r = w; r < nb; r += NWAVE)
            for (int c = lane; c < n; c += WAVE) R[(size_t)r * n + c] = Ci[(size_t)r * n + c];
        __syncthreads();
        if (i0 > 0) gemm<false, false>(nb, n, i0, -1.0, G + (size_t)i0 * n, n, M.Ct, n, 1.0, R, n, sh.big);
        chol_diag_inverse(G, n, i0, nb, Dg, sh);
        gemm<false, false>(nb, n, nb, 1.0, Dg, nb, R, n, 0.0, Ci, n, sh.big);
    }
    return true;
}

// ---------------------------------------------------------------------------
// Orbitals in ascending energy order for the density products.  The
// orbitals with f > 0 form a prefix of length kocc; G[j] = C_{sidx[j]} (row
// j, for j < kocc, or every j when `all_rows`) and H[j] = w_j G[j] (j < kocc)
// with w = f (or f e when `with_e`).  H may alias Ct: it is written only
// after every row of G has been gathered.  Returns kocc.
// ---------------------------------------------------------------------------
__device__ __noinline__ int gather_occupied(const Mol& M, const Shared& sh, glbf* G, glbf* H,
                                            bool with_e, bool all_rows = false) {
    const int n = M.n;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    int k = 0;
    for (int j = threadIdx.x; j < n; j += NT)
        if (sh.ftot[sh.sidx[j]] > 0.0) k = j + 1;
    const int kocc = (int)wg_max((double)k, sh.red);
    for (int j = w; j < (all_rows ? n : kocc); j += NWAVE) {
        const glbf* src = M.Ct + (size_t)sh.sidx[j] * n;
        for (int c = lane; c < n; c += WAVE) G[(size_t)j * n + c] = src[c];
    }
    __syncthreads();
    for (int j = w; j < kocc; j += NWAVE) {
        const int o = sh.sidx[j];
        const double f = with_e ? sh.ftot[o] * sh.ev[o] : sh.ftot[o];
        for (int c = lane; c < n; c += WAVE) H[(size_t)j * n + c] = f * G[(size_t)j * n + c];
    }
    __syncthreads();
    return kocc;
}

// ---------------------------------------------------------------------------
// Fock build F = H0 - 1/2 S o (v_i + v_j) (scf/base.py:817-859)
// ---------------------------------------------------------------------------
// Streaming passes over rows run four rows per wave at a time, so that each
// lane has the loads of four rows in flight (one row per wave left a CU's
// waves waiting on memory latency at a fraction of its share of HBM);
// rows past the end are clamped to the last row for the loads and skipped
// for the stores.
constexpr int ROWS_PER_WAVE = 4;

__device__ inline void fock_build(const Mol& M, const glbf* vin) {
    const int n = M.n;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int r0 = w * ROWS_PER_WAVE; r0 < n; r0 += NWAVE * ROWS_PER_WAVE) {
        int rr[ROWS_PER_WAVE];
        double vr[ROWS_PER_WAVE];
#pragma unroll
        for (int u = 0; u < ROWS_PER_WAVE; ++u) {
            rr[u] = r0 + u < n ? r0 + u : n - 1;
            vr[u] = vin[rr[u]];
        }
        for (int c = lane; c < n; c += WAVE) {
            const double vc = vin[c];
            double h[ROWS_PER_WAVE], sv[ROWS_PER_WAVE];
#pragma unroll
            for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                h[u] = M.H0[(size_t)rr[u] * n + c];
                sv[u] = M.S[(size_t)rr[u] * n + c];
            }
#pragma unroll
            for (int u = 0; u < ROWS_PER_WAVE; ++u)
                if (r0 + u < n) M.F[(size_t)rr[u] * n + c] = h[u] - 0.5 * sv[u] * (vr[u] + vc);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Perturbative occupied-virtual decoupling for intermediate SCF steps of
// closed shells with a clear gap (pseudo-diagonalisation in the sense of
// Stewart, Csaszar and Pulay, J. Comput. Chem. 3, 227 (1982), done with
// matrix-core GEMMs).  The rows of Ct are in ascending energy order, the
// first k doubly occupied.  With A = C^T F C (only its occupied rows and the
// diagonal are formed) and e = diag(A), the first-order amplitudes
// K_ia = A_ia / (e_a - e_i) define the antisymmetric generator
// G = [[0, K], [-K^T, 0]]; the orbitals are rotated by exp(G), which is
// exactly orthogonal:
//   C_occ <- cos(sqrt X) C_occ - sinc(sqrt X) Z
//   C_vir <- C_vir + K^T [sinc(sqrt X) C_occ + phi(X) Z]
// with X = K K^T (k x k), Z = K C_vir (k x n), phi(x) = (cos sqrt x - 1)/x;
// the functions of X are power series evaluated by Horner's rule, truncated
// once the next term is below 1e-17, or 1e-13 in inexact steps (||X||_F
// bounds the spectral norm).  Another pass measures the coupling left and
// rotates again (pass limits and exits: PD_INEXACT_PASSES, PD_EXACT_PASSES).
// Returns 1 when done (rotated, or already decoupled to the tolerance), 0 to
// ask for the Jacobi eigensolve with Ct and F untouched, -1 to ask for it
// after F was used as scratch (the caller rebuilds F).  Jacobi is asked for
// when an amplitude exceeds pd_kmax, a denominator is not positive, k > n/2,
// a series would need more than 9 terms, or (exact calls, loose = 0) the
// last measured coupling is above the tolerance.  On return 1, sh.ev holds the
// diagonal of C F C^T in the rows' order (occupied rows: of the last pass).
// Scratch: T holds C F, then X and Z; P the couplings/amplitudes; F the two
// Horner accumulators (k x n each, k <= n/2).
// ---------------------------------------------------------------------------
// Coefficient of term j of the series in X = K K^T used by the rotation:
// which 0: cos, (-1)^j / (2j)!; 1: sinc, (-1)^j / (2j+1)!; 2: phi,
// -(-1)^j / (2j+2)!  ((2j)! by the running product (2t-1)(2t), t <= j)
__device__ inline double horner_coef(int j, int which) {
    double f = 1.0;
    for (int t = 1; t <= j; ++t) f *= (2.0 * t - 1.0) * (2.0 * t);
    const double sg = (j & 1) ? -1.0 : 1.0;
    if (which == 0) return sg / f;
    if (which == 1) return sg / (f * (2.0 * j + 1.0));
    return -sg / (f * (2.0 * j + 1.0) * (2.0 * j + 2.0));
}

__device__ __noinline__ int pd_step(const Mol& M, const Shared& sh, const xtb_scf_opts& o, int k,
                                    double loose, const glbf* vin) {
    const int n = M.n, m = n - k;
    if (k <= 0 || m <= 0 || 2 * k > n) return 0;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    glbf* Ct = M.Ct;
    glbf* Cv = M.Ct + (size_t)k * n;
    glbf* Kb = M.P;  // k x m, leading dimension m
    // Virtual orbital energies only scale the amplitudes: an error of a few
    // SCF steps in them leaves a second-order coupling, so they are formed
    // (all rows of C F) every XTB_PD_REFRESH-th perturbative step only; the
    // others reuse the last values (sh.skey, in the rows' order) and form
    // the occupied rows of C F.  Every 4th step costs no SCF iterations on
    // 500-orbital molecules; never refreshing them costs two.
#ifndef XTB_PD_REFRESH
#define XTB_PD_REFRESH 4
#endif
    // sh.tacc[7] counts the steps since the last refresh (>= 1)
    const bool refresh = sh.tacc[7] == 0.0 || sh.tacc[7] >= XTB_PD_REFRESH;
    sh.tacc[7] = refresh ? 1.0 : sh.tacc[7] + 1.0;
    const int rows = refresh ? n : k;
    double tm = now_us();
    (void)tm;
    const int npass = loose > 0.0 ? PD_INEXACT_PASSES : PD_EXACT_PASSES;
    for (int pass = 0; pass < npass; ++pass) {
        if (pass > 0) fock_build(M, vin);
        // T = C F (rows: all, or the occupied ones); A_ov = T_occ C_vir^T
        gemm<false, false>(rows, n, n, 1.0, Ct, n, M.F, n, 0.0, M.T, n, sh.big);
        PD_MARK(sh, 0, tm);
        gemm<false, true>(k, m, n, 1.0, M.T, n, Cv, n, 0.0, Kb, m, sh.big);
        PD_MARK(sh, 1, tm);
        for (int i0 = w * ROWS_PER_WAVE; i0 < rows; i0 += NWAVE * ROWS_PER_WAVE) {
            int ii[ROWS_PER_WAVE];
            double d[ROWS_PER_WAVE];
#pragma unroll
            for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                ii[u] = i0 + u < rows ? i0 + u : rows - 1;
                d[u] = 0.0;
            }
            for (int c = lane; c < n; c += WAVE) {
                double tv[ROWS_PER_WAVE], cv[ROWS_PER_WAVE];
#pragma unroll
                for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                    tv[u] = M.T[(size_t)ii[u] * n + c];
                    cv[u] = Ct[(size_t)ii[u] * n + c];
                }
#pragma unroll
                for (int u = 0; u < ROWS_PER_WAVE; ++u) d[u] = fma(tv[u], cv[u], d[u]);
            }
#pragma unroll
            for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                const double v = wave_sum(d[u]);
                if (lane == 0 && i0 + u < rows) sh.ev[i0 + u] = v;
            }
        }
        for (int a = rows + threadIdx.x; a < n; a += NT) sh.ev[a] = sh.skey[a];
        __syncthreads();
        double dmax = 0.0;
        for (int i = threadIdx.x; i < n; i += NT) dmax = fmax(dmax, fabs(sh.ev[i]));
        dmax = wg_max(dmax, sh.red);
        const double tol = fmax(o.jacobi_tol * fmax(dmax, 1.0), loose > 0.0 ? loose : PD_FINAL_REL * o.x_atol_max);
        double aov = 0.0, kmax = 0.0, dmin = INFINITY;
        for (int i0 = w * ROWS_PER_WAVE; i0 < k; i0 += NWAVE * ROWS_PER_WAVE) {
            int ii[ROWS_PER_WAVE];
            double ei[ROWS_PER_WAVE];
#pragma unroll
            for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                ii[u] = i0 + u < k ? i0 + u : k - 1;
                ei[u] = sh.ev[ii[u]];
            }
            for (int a = lane; a < m; a += WAVE) {
                const double ea = sh.ev[k + a];
                double v[ROWS_PER_WAVE];
#pragma unroll
                for (int u = 0; u < ROWS_PER_WAVE; ++u) v[u] = Kb[(size_t)ii[u] * m + a];
#pragma unroll
                for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                    if (i0 + u >= k) continue;
                    const double den = ea - ei[u];
                    const double amp = v[u] / den;
                    aov = fmax(aov, fabs(v[u]));
                    dmin = fmin(dmin, den);
                    kmax = fmax(kmax, fabs(amp));
                    Kb[(size_t)ii[u] * m + a] = amp;
                }
            }
        }
        aov = wg_max(aov, sh.red);
        kmax = wg_max(kmax, sh.red);
        dmin = -wg_max(-dmin, sh.red);
        if (aov <= tol) return 1;  // decoupled: orbitals stay
        const int fail = pass == 0 ? 0 : -1;
        PD_MARK(sh, 2, tm);
        if (!(dmin > 0.0) || !(kmax <= o.pd_kmax)) return fail;
        // an exact call's last pass only measures: a rotation after it
        // would go unchecked
        if (loose == 0.0 && pass == npass - 1) return fail;
        glbf* X = M.T;
        glbf* Z = M.T + (size_t)k * k;
        gemm_sym<false, true>(k, m, 1.0, Kb, m, Kb, m, X, k, sh.big);
        PD_MARK(sh, 3, tm);
        gemm<false, false>(k, n, m, 1.0, Kb, m, Cv, n, 0.0, Z, n, sh.big);
        PD_MARK(sh, 4, tm);
        double xf = 0.0;
        for (int e = threadIdx.x; e < k * k; e += NT) xf = fma(X[e], X[e], xf);
        xf = sqrt(wg_sum(xf, sh.red));
        // highest power J kept: the first omitted cosine term (the largest of
        // the three series) ||X||^(J+1) / (2J+2)! is below 1e-17 in an exact
        // step and below min(1e-13, 1e-3 loose) in an inexact one.  What is
        // left out perturbs orthonormality by that much; over an SCF run it
        // adds up to ~1e-12, far below the results' tolerances
        const double tolj = loose > 0.0 ? fmin(1e-13, fmax(1e-17, 1e-3 * loose)) : 1e-17;
        int J = 0;
        double term = 0.5 * xf;
        while (term > tolj) {
            ++J;
            term *= xf / ((2.0 * J + 1.0) * (2.0 * J + 2.0));
            if (J > 8) return fail;  // F and Ct untouched in this pass
        }
        // series coefficients of term j (computed where used: arrays indexed
        // by the run-time j would live in scratch memory)
        auto cc = [](int j) { return horner_coef(j, 0); };
        auto ss = [](int j) { return horner_coef(j, 1); };
        auto ph = [](int j) { return horner_coef(j, 2); };
        // Horner steps acc <- X acc + a_j C_occ + b_j Z, the update fused into
        // the GEMM epilogue; the two k x n halves of F alternate
        glbf* cur = M.F;
        glbf* nxt = M.F + (size_t)k * n;
        // R = sinc(X) C_occ + phi(X) Z
        wg_axpby((size_t)k * n, ss(J), Ct, ph(J), Z, cur);
        __syncthreads();
        for (int j = J - 1; j >= 0; --j) {
            PD_MARK(sh, 6, tm);
            gemm_epi(k, n, k, X, k, cur, n, nxt, n, Ct, ss(j), Z, ph(j), sh.big);
            PD_MARK(sh, 5, tm);
            glbf* t = cur;
            cur = nxt;
            nxt = t;
        }
        PD_MARK(sh, 6, tm);
        // C_vir += K^T R
        gemm<true, false>(m, n, k, 1.0, Kb, m, cur, n, 1.0, Cv, n, sh.big);
        PD_MARK(sh, 7, tm);
        // C_occ <- cos(X) C_occ - sinc(X) Z.  The last Horner term is written
        // over C_occ itself: each element of C_occ is read (as E1) and
        // written by the same thread, and no product takes C_occ as an
        // operand, so no copy pass follows
        wg_axpby((size_t)k * n, cc(J), Ct, -ss(J), Z, J == 0 ? Ct : cur);
        __syncthreads();
        for (int j = J - 1; j >= 0; --j) {
            PD_MARK(sh, 6, tm);
            gemm_epi(k, n, k, X, k, cur, n, j == 0 ? Ct : nxt, n, Ct, cc(j), Z, -ss(j), sh.big);
            PD_MARK(sh, 5, tm);
            glbf* t = cur;
            cur = nxt;
            nxt = t;
        }
        PD_MARK(sh, 6, tm);
#ifdef XTB_PD_TIMERS
        sh.ptacc[9] += (double)J;
#endif
        // inexact steps: the second-order estimate of what is left
        if (loose > 0.0 && kmax * aov <= tol) return 1;
    }
    // the passes spent: good enough for an inexact step (an exact one has
    // returned from its last, measuring pass)
    return 1;
}

// ---------------------------------------------------------------------------
// one SCF map v_in -> v_out (scf/base.py:651-675 iterate_potential);
// leaves q, sorted e/occupations and shell/atom charges behind, and in the
// last step (`last`) also P and e_orb (the other steps never store P).
// ---------------------------------------------------------------------------
__device__ __noinline__ int scf_step(const Mol& M, const Shared& sh, const xtb_batch& b,
                        const xtb_params& par, const xtb_scf_opts& o, const glbf* vin,
                        glbf* vout, int* status, double* ef, double loose, bool allow_bd,
                        bool sort_rows, bool final_step = false, int* decoupled = nullptr,
                        bool last = false) {
    const int n = M.n;
    // Occupation classes for this step's eigensolve.  With integer alpha/beta
    // counts and frontier gaps of at least bd_gap_kt x kT in the previous
    // step's spectrum (sh.skey, ascending; the rows of Ct are in that order),
    // only the coupling between doubly, singly and un-occupied orbitals is
    // removed: that fixes the density of 0/1 occupations, whose Fermi tails
    // at such gaps are below 1/(1 + e^(bd_gap_kt/2)).  The padding orbitals
    // of a padded layout (unit overlap diagonal, zero H0 and potential) are
    // exact eigenvectors at energy 0 with no coupling to any other orbital;
    // they sit in whichever class the sorted spectrum puts them, and the gap
    // test below covers them like any other level.
    int cb_lo = -1, cb_hi = -1;
    if ((allow_bd || final_step) && o.bd_gap_kt > 0.0) {
        const double na = M.nel_a, nb = M.nel_b;
        const int ia = (int)rint(na), ib = (int)rint(nb);
        bool ok = fabs(na - ia) < 1e-12 && fabs(nb - ib) < 1e-12 && ib <= ia;
        // (at kT -> 0 a gap of bd_gap_kt x kT(300 K) is still asked for:
        // the decoupling needs a clear gap to converge, not to be exact)
        const double need = o.bd_gap_kt * fmax(o.kt, 9.5e-4);
        if (ok && ia > 0 && ia < n) ok = sh.skey[ia] - sh.skey[ia - 1] >= need;
        if (ok && ib > 0 && ib < n) ok = sh.skey[ib] - sh.skey[ib - 1] >= need;
        if (ok) { cb_lo = ib; cb_hi = ia; }
    }
    const size_t nn = (size_t)n * n;
    // Fock build (scf/base.py:817-859)
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    double t0 = now_us();
    fock_build(M, vin);
    {
        const double t = now_us();
        sh.tacc[8] += t - t0;
        sh.tacc[9] += 1.0;
        t0 = t;
    }
    int sw = 0;
    bool pd = false, intra = false;
    if (final_step) {
        // Final (exact) step of a gapped closed shell: remove the last
        // occupied-virtual coupling perturbatively (it is second order after
        // that) to the full tolerance.  That fixes the density of the 0/1
        // occupations, the energies and W (rotations inside the occupied
        // block leave C^T f C and C_occ^T A_oo C_occ unchanged); only the
        // orbital energies need the diagonalisation inside the two classes.
        if (cb_lo >= 0 && cb_lo == cb_hi && o.pd_kmax > 0.0) {
            const int r = pd_step(M, sh, o, cb_lo, 0.0, vin);
            intra = r > 0;
            if (r != 0) fock_build(M, vin);  // F may have served as scratch
        }
        if (!intra) cb_lo = cb_hi = -1;
        pd = intra && !o.orbitals;
        if (pd && decoupled) *decoupled = cb_lo;
    } else if (cb_lo >= 0 && cb_lo == cb_hi && o.pd_kmax > 0.0) {
        const int r = pd_step(M, sh, o, cb_lo, loose, vin);
        pd = r > 0;
        if (r < 0) fock_build(M, vin);  // F served as scratch
    }
    if (pd) {
        const double dt = now_us() - t0;
        sh.tacc[2] += dt;
        sh.tacc[3] += dt;
        sh.tacc[6] += 1.0;
    } else {
        // T = F C ; A = C^T T  -> F
        gemm<false, true>(n, n, n, 1.0, M.F, n, M.Ct, n, 0.0, M.T, n, sh.big);
        gemm_sym<false, false>(n, n, 1.0, M.Ct, n, M.T, n, M.F, n, sh.big);
        double t1 = now_us();
        sw = eig_step(M, sh, o.jacobi_max_sweeps, o.jacobi_tol, loose, o.jacobi_inner, cb_lo, cb_hi, intra);
        sh.tacc[7] = 1.0;  // eigenvalues fresh for the next perturbative step
        double t2 = now_us();
        sh.tacc[1] += t1 - t0;
        sh.tacc[2] += t2 - t1;
    }
    // an intermediate step's eigensolve that ran out of sweeps only slows
    // the SCF down; the final step's results would be unreliable
    if (sw < 0 && final_step) *status |= XTB_MOL_JACOBI;
    // sort eigenvalues
    for (int j = threadIdx.x; j < sh.npow; j += NT) {
        sh.skey[j] = (j < n) ? sh.ev[j] : INFINITY;
        sh.sidx[j] = j;
    }
    for (int j = threadIdx.x; j < n; j += NT) sh.ftot[j] = 0.0;
    __syncthreads();
    wg_bitonic_sort(sh.skey, sh.sidx, sh.npow);
    // Fermi occupations, alpha then beta
    bool ok = fermi_channel(M, sh, M.nel_a, o.kt, o.fermi_thresh, o.fermi_maxiter, o.padded != 0,
                            M.occ, &ef[0]);
    ok = fermi_channel(M, sh, M.nel_b, o.kt, o.fermi_thresh, o.fermi_maxiter, o.padded != 0,
                       M.occ + M.nvec_total, &ef[1]) && ok;
    if (!ok) *status |= XTB_MOL_FERMI;
    {
        // ef[2]: the largest f (1 - f) of either channel.  Above ~1e-6 some
        // levels lie within ~14 kT of the Fermi level, where the density
        // depends on rotations among near-degenerate orbitals of different
        // occupation: the caller then tightens inexact eigensolves by 2 kT,
        // the bound on |f_i - f_j| / |e_i - e_j| over both channels
        double fs = 0.0;
        for (int c = 0; c < 2; ++c)
            for (int j = threadIdx.x; j < n; j += NT) {
                const double f = M.occ[c * M.nvec_total + j];
                fs = fmax(fs, f * (1.0 - f));
            }
        ef[2] = wg_max(fs, sh.red);
    }
    // P = C diag(f) C^T over the occupied orbitals (F and T are free here).
    // When the rows of Ct are already in ascending order (the usual case once
    // the SCF settles) they are used in place: no gather into F and no copy
    // back, 4 n^2 doubles of traffic less per step
    t0 = now_us();
    bool moved = false;
    for (int j = threadIdx.x; j < n; j += NT) moved |= sh.sidx[j] != j;
    const bool perm = __syncthreads_or(moved) != 0;
    int kocc;
    if (perm) {
        kocc = gather_occupied(M, sh, M.F, M.T, false, sort_rows);
    } else {
        int k = 0;
        for (int j = threadIdx.x; j < n; j += NT)
            if (sh.ftot[j] > 0.0) k = j + 1;
        kocc = (int)wg_max((double)k, sh.red);
        for (int j0 = w * ROWS_PER_WAVE; j0 < kocc; j0 += NWAVE * ROWS_PER_WAVE) {
            int jj[ROWS_PER_WAVE];
            double f[ROWS_PER_WAVE];
#pragma unroll
            for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                jj[u] = j0 + u < kocc ? j0 + u : kocc - 1;
                f[u] = sh.ftot[jj[u]];
            }
            for (int c = lane; c < n; c += WAVE) {
                double cv[ROWS_PER_WAVE];
#pragma unroll
                for (int u = 0; u < ROWS_PER_WAVE; ++u) cv[u] = M.Ct[(size_t)jj[u] * n + c];
#pragma unroll
                for (int u = 0; u < ROWS_PER_WAVE; ++u)
                    if (j0 + u < kocc) M.T[(size_t)jj[u] * n + c] = f[u] * cv[u];
            }
        }
        __syncthreads();
    }
    if (last) {
        gemm_sym<true, false>(n, kocc, 1.0, perm ? M.F : M.Ct, n, M.T, n, M.P, n, sh.big);
    } else {
        // an intermediate step needs the charges, not P: the Mulliken sums
        // a_i = sum_j P_ij S_ij come from the density product's tiles (P is
        // never stored; its buffer holds the tiles' partial sums), into q
        gemm_sym_mull<true, false>(n, kocc, perm ? M.F : M.Ct, n, M.T, n, M.S, M.nao, M.P, M.q, sh.big);
    }
#ifdef XTB_PD_TIMERS
    sh.ptacc[8] += now_us() - t0;
#endif
    if (sort_rows && perm) {
        // the next step starts from the orbitals in ascending energy order
        wg_axpby((size_t)n * n, 1.0, M.F, 0.0, nullptr, M.Ct);
    }
    sh.tacc[1] += now_us() - t0;
    const double tm0 = now_us();
    // Mulliken charges and orbital energies (scf/base.py:764-815), one wave
    // per row; the orbital energies only enter the final energy (last step)
    if (!last) {
        for (int i = threadIdx.x; i < n; i += NT) M.q[i] = (i < M.nao) ? b.n0[M.vec0 + i] - M.q[i] : 0.0;
    } else {
        const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
        for (int i0 = w * ROWS_PER_WAVE; i0 < n; i0 += NWAVE * ROWS_PER_WAVE) {
            int ii[ROWS_PER_WAVE];
            double sq[ROWS_PER_WAVE], se[ROWS_PER_WAVE];
#pragma unroll
            for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                ii[u] = i0 + u < n ? i0 + u : n - 1;
                sq[u] = 0.0;
                se[u] = 0.0;
            }
            if (last) {
                for (int j = l; j < M.nao; j += WAVE) {
                    double pv[ROWS_PER_WAVE], sv[ROWS_PER_WAVE], hv[ROWS_PER_WAVE];
#pragma unroll
                    for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                        pv[u] = M.P[(size_t)ii[u] * n + j];
                        sv[u] = M.S[(size_t)ii[u] * n + j];
                        hv[u] = M.H0[(size_t)ii[u] * n + j];
                    }
#pragma unroll
                    for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                        sq[u] = fma(pv[u], sv[u], sq[u]);
                        se[u] = fma(pv[u], hv[u], se[u]);
                    }
                }
            } else {
                for (int j = l; j < M.nao; j += WAVE) {
                    double pv[ROWS_PER_WAVE], sv[ROWS_PER_WAVE];
#pragma unroll
                    for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                        pv[u] = M.P[(size_t)ii[u] * n + j];
                        sv[u] = M.S[(size_t)ii[u] * n + j];
                    }
#pragma unroll
                    for (int u = 0; u < ROWS_PER_WAVE; ++u) sq[u] = fma(pv[u], sv[u], sq[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                const double a = wave_sum(sq[u]), e = wave_sum(se[u]);
                const int i = i0 + u;
                if (l == 0 && i < n) {
                    M.q[i] = (i < M.nao) ? b.n0[M.vec0 + i] - a : 0.0;
                    M.eorb[i] = (i < M.nao) ? e : 0.0;
                }
            }
        }
    }
    __syncthreads();
    // shell and atom charges
    for (int s = threadIdx.x; s < M.nsh; s += NT) {
        const int o0 = b.sh_ao[M.sh0 + s];
        const int nl = 2 * par.us_l[b.sh_us[M.sh0 + s]] + 1;
        double acc = 0.0;
        for (int x = 0; x < nl; ++x) acc += M.q[o0 + x];
        sh.qsh[s] = acc;
    }
    __syncthreads();
    for (int a = threadIdx.x; a < M.nat; a += NT) {
        const int s0 = b.atom_sh0[M.at0 + a], ns = b.atom_nsh[M.at0 + a];
        double acc = 0.0;
        for (int x = 0; x < ns; ++x) acc += sh.qsh[s0 + x];
        sh.qat[a] = acc;
    }
    __syncthreads();
    // ES2 (gamma q_sh) + ES3 (hd q_at^2) shell potential, one wave per row
    {
        const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
        const int m = M.nsh;
        for (int s0 = w * ROWS_PER_WAVE; s0 < m; s0 += NWAVE * ROWS_PER_WAVE) {
            int ss[ROWS_PER_WAVE];
            double acc[ROWS_PER_WAVE];
#pragma unroll
            for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                ss[u] = s0 + u < m ? s0 + u : m - 1;
                acc[u] = 0.0;
            }
            for (int t = l; t < m; t += WAVE) {
                const double qt = sh.qsh[t];
                double g[ROWS_PER_WAVE];
#pragma unroll
                for (int u = 0; u < ROWS_PER_WAVE; ++u) g[u] = M.gam[(size_t)ss[u] * m + t];
#pragma unroll
                for (int u = 0; u < ROWS_PER_WAVE; ++u) acc[u] = fma(g[u], qt, acc[u]);
            }
#pragma unroll
            for (int u = 0; u < ROWS_PER_WAVE; ++u) {
                const double v = wave_sum(acc[u]);
                const int s = s0 + u;
                if (l == 0 && s < m) {
                    const int a = b.sh_atom[M.sh0 + s];
                    const double hd = par.sp_gam3[b.atom_sp[M.at0 + a]];
                    sh.gq[s] = v;
                    sh.vsh[s] = v + hd * sh.qat[a] * sh.qat[a];
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += NT) {
        const int s = (i < M.nao) ? b.ao_sh[M.vec0 + i] : -1;
        vout[i] = (s >= 0) ? sh.vsh[s] : 0.0;
    }
    __syncthreads();
    sh.tacc[10] += now_us() - tm0;
    // (P, S) reads of this stage in units of 16 n^2 bytes: the last step
    // reads P, S and H0 (orbital energies); the others form q inside the
    // density product (timed with it) and only add the gamma products here
    sh.tacc[11] += last ? 1.5 : 0.0;
    return sw < 0 ? 0 : sw;
}

// Inputs and outputs of the fused single point (k_singlepoint) beyond the
// SCF's: positions, and the buffers the integral pass and the gradient fill.
struct FusedArgs {
    const double* pos;
    double* cn;
    double* S;
    double* H0;
    double* gam;
    double* dedcn;
    double* grad;
};

// Defined with the fused kernel (singlepoint.hip); the plain SCF kernel
// never calls them.
__device__ __noinline__ void molecule_integrals(const Mol& M, const xtb_batch& b, const xtb_params& par,
                                                const FusedArgs& fa, double* outs, const Shared& sh);
__device__ __noinline__ void molecule_forces(const Mol& M, const xtb_batch& b, const xtb_params& par,
                                             const FusedArgs& fa, const Shared& sh);

}  // namespace xtb

// Validation shared by xtb_scf and xtb_singlepoint (host).
static inline int32_t check_scf_args(const xtb_batch* batch, const xtb_scf_opts* opts, const xtb_scf_state* state) {
    if (batch->nmol <= 0 || batch->max_n <= 0) return XTB_EINVAL;
    if (batch->max_n > 1024) return XTB_ETOOBIG;
    if (opts->mixer != 0 && opts->mixer != 1) return XTB_EINVAL;
    if (opts->generations < 1 || opts->generations > (opts->mixer == 1 ? xtb::BROYDEN_MAXR : xtb::MIX_MAXG))
        return XTB_EINVAL;
    if (opts->wg_skip < 0 || opts->wg_skip > 7) return XTB_EINVAL;
    if (opts->jacobi_inner < 1 || opts->jacobi_inner > 20 || opts->eig_inexact < 0.0) return XTB_EINVAL;
    if (!state->F || !state->C || !state->T || !state->P || !state->pot || !state->vnew ||
        !state->q || !state->eorb || !state->emo || !state->occ || !state->hist_x ||
        !state->hist_f || !state->qsh || !state->out || !state->order)
        return XTB_EINVAL;
    return XTB_OK;
}
