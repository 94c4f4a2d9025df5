// Whole-GPU SCF for one large molecule (thousands of orbitals) on gfx950.
//
// The persistent per-molecule kernel (scf.hip) gives one CU to each molecule,
// the right shape for streams of molecules and the wrong one for a single
// system of n ~ 5000 orbitals.  Here every stage of the same SCF is a grid
// kernel that spreads over all CUs, and the SCF driver runs on the host,
// stream ordered, reading a few control scalars back once or twice per step:
//
//   setup   Loewdin orthogonaliser X = S^-1/2 by the coupled Newton-Schulz
//           iteration on matrix-core GEMMs; the orbitals start as the rows
//           of X
//   step    Fock build F = H0 - 1/2 S o (v_i + v_j)  (scf/base.py:817-859)
//           -> gapped closed shell: perturbative occupied-virtual rotation
//              of the previous step's orbitals (GEMMs), as in scf.hip
//           -> otherwise A = C F C^T and parallel-ordered block Jacobi: one
//              launch solves the 32 x 32 subproblems of a round (a wave
//              each), one applies their rotations to A and to the orbitals
//              on matrix cores
//           -> sort, Fermi smearing (filling.py:201-366), density
//              P = C^T diag(f) C, Mulliken charges, ES2/ES3 potential,
//              Anderson mixing (mixer/anderson.py:163-317), all on device
//   final   the exact step, energies, -TS and the energy-weighted density W
//           for the analytic gradient (types/analytical.py:186-190)
//
// Upstream runs the same SCF (src/dxtb/_src/scf/unrolling/default.py:71-140)
// with LAPACK's eigh per step (scf/unrolling/base.py:142-175).
#include <chrono>
#include <vector>

#include "scf_core.h"

namespace xtb {
namespace big {

// ---------------------------------------------------------------------------
// Grid GEMM: one 128 x 128 output tile per workgroup (gemm_tile: 8 waves of
// 64 x 32, 16-deep k panels double-buffered in LDS).  Workgroup b runs on XCD
// b mod 8; each XCD takes a contiguous run of tiles in row-major order, so
// the tiles sharing one XCD's L2 share the row panels of op(A).  With
// `upper` only tiles on and above the diagonal are formed (a product known
// to be symmetric; k_mirror fills the rest).
// ---------------------------------------------------------------------------
// control values the host reads back (slots of the `ctl` array)
enum {
    C_EFA = 0, C_EFB, C_FOK, C_KOCC, C_PERM, C_GAPA, C_GAPB, C_DMAX,
    C_AOV, C_KMAX, C_DMIN, C_XF, C_OFF, C_ERR, C_GERSH, C_CONV, C_RESID, C_FSOFT, C_BRANK, C_BDX,
    C_NCTL = 24
};

constexpr int GT = 128;
// k-panel depth of the grid GEMM (one workgroup per CU either way: its
// register use allows three waves per SIMD, fewer than a second workgroup);
// 32 deep panels keep more loads in flight than 16 (the 600-water single
// point 577 -> 569 ms)
#ifndef XTB_GRID_KP
#define XTB_GRID_KP 32
#endif
constexpr int GRID_KP = XTB_GRID_KP;
constexpr int GEMM_LDS_BYTES = 2 * GRID_KP * (GT + 1 + GT + 1) * 8;

__device__ inline void tile_of(int idx, int tn, int upper, int& ti, int& tj) {
    if (!upper) {
        ti = idx / tn;
        tj = idx % tn;
        return;
    }
    int r = 0, rem = idx;
    while (rem >= tn - r) {
        rem -= tn - r;
        ++r;
    }
    ti = r;
    tj = r + rem;
}

// square tiles of TE x TE (128, or 64 for products too small to give a
// quarter of the CUs a 128-tile)
template <bool TA, bool TB, bool EPI, int TE>
__global__ void __launch_bounds__(NT) k_gemm(int M, int N, int K, double alpha, const double* A, int lda,
                                             const double* B, int ldb, double beta, double* C, int ldc,
                                             int upper, int ntiles, const double* E1, const double* E2,
                                             double s1, double s2) {
    extern __shared__ double lds[];
    const int per = gridDim.x / 8;  // the grid is a multiple of 8
    const int idx = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (idx >= ntiles) return;
    const int tn = (N + TE - 1) / TE;
    int ti, tj;
    tile_of(idx, tn, upper, ti, tj);
    gemm_tile<TA, TB, TE / 32, TE / 64, GRID_KP, EPI>(ti * TE, tj * TE, M, N, K, alpha, to_glb(A), lda, to_glb(B),
                                                      ldb, beta, to_glb(C), ldc, to_lds(lds), to_glb(E1),
                                                      to_glb(E2), s1, s2);
}

// C[j][i] = C[i][j] for the 64 x 64 blocks below the diagonal whose tile
// (tile edge `per` x 64) was not formed by an `upper` GEMM; one workgroup
// per block, through LDS.
__global__ void __launch_bounds__(256) k_mirror(double* C, int n, int ldc, int per) {
    __shared__ double t[64 * 65];
    const int nb = (n + 63) / 64;
    // block index -> (bi, bj), bi > bj, lower triangle row-major
    int bi = 1, rem = blockIdx.x;
    while (rem >= bi) {
        rem -= bi;
        ++bi;
    }
    const int bj = rem;
    if (bi >= nb || bi / per <= bj / per) return;  // formed by the GEMM
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int r = e / 64, c = e % 64;  // source row bj*64 + r, column bi*64 + c
        const int gr = bj * 64 + r, gc = bi * 64 + c;
        t[r * 65 + c] = (gr < n && gc < n) ? C[(size_t)gr * ldc + gc] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int c = e / 64, r = e % 64;  // destination row bi*64 + c, column bj*64 + r
        const int gr = bi * 64 + c, gc = bj * 64 + r;
        if (gr < n && gc < n) C[(size_t)gr * ldc + gc] = t[r * 65 + c];
    }
}

// ---------------------------------------------------------------------------
// Elementwise and row kernels.  Row kernels: one wave per row, 8 rows per
// 512-thread workgroup.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fock(int n, const double* H0, const double* S, const double* v,
                                              double* F) {
    const int r = blockIdx.y;
    const double vr = v[r];
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256) {
        const size_t e = (size_t)r * n + c;
        F[e] = H0[e] - 0.5 * S[e] * (vr + v[c]);
    }
}

// out[i] = sum_c X[i][c] Y[i][c], i < rows
__global__ void __launch_bounds__(NT) k_rowdot(int rows, int n, const double* X, const double* Y, double* out) {
    const int i = blockIdx.x * NWAVE + threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    if (i >= rows) return;
    double d = 0.0;
    for (int c = lane; c < n; c += WAVE) d = fma(X[(size_t)i * n + c], Y[(size_t)i * n + c], d);
    d = wave_sum(d);
    if (lane == 0) out[i] = d;
}

// per-workgroup partials of max |X_ij| over j > i (rows i of this
// workgroup's waves): part[blockIdx.x]
__global__ void __launch_bounds__(NT) k_offmax(int n, const double* A, double* part) {
    __shared__ double red[NWAVE];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int i = blockIdx.x * NWAVE + w;
    double m = 0.0;
    if (i < n)
        for (int j = i + 1 + lane; j < n; j += WAVE) m = fmax(m, fabs(A[(size_t)i * n + j]));
    m = wave_max(m);
    if (lane == 0) red[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int k = 1; k < NWAVE; ++k) r = fmax(r, red[k]);
        part[blockIdx.x] = r;
    }
}

// per-workgroup partials of max |d_i| (diagonal of A, or a vector) and of
// the Gershgorin radius max_i sum_j |A_ij| (when A is given)
__global__ void __launch_bounds__(NT) k_rowabs(int n, const double* A, double* part) {
    __shared__ double red[NWAVE];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int i = blockIdx.x * NWAVE + w;
    double s = 0.0;
    if (i < n)
        for (int j = lane; j < n; j += WAVE) s += fabs(A[(size_t)i * n + j]);
    s = wave_sum(s);
    if (lane == 0) red[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int k = 1; k < NWAVE; ++k) r = fmax(r, red[k]);
        part[blockIdx.x] = r;
    }
}

// Reduction of per-workgroup partials: slot s of ctl[s] = op_s over
// part[b * nslot + s], b < nblk, in a fixed order (deterministic).  op:
// bits 2s..2s+1 of `ops`: 0 sum, 1 max, 2 min.
__global__ void __launch_bounds__(NT) k_finish(const double* part, int nblk, int nslot, int ops, double* ctl,
                                               int ctl0) {
    __shared__ double red[NWAVE];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int s = 0; s < nslot; ++s) {
        const int op = (ops >> (2 * s)) & 3;
        auto comb = [op](double x, double y) { return op == 0 ? x + y : (op == 1 ? fmax(x, y) : fmin(x, y)); };
        double a = op == 0 ? 0.0 : (op == 1 ? -INFINITY : INFINITY);
        for (int b = threadIdx.x; b < nblk; b += NT) a = comb(a, part[(size_t)b * nslot + s]);
        // xor butterfly (each pair combined in the same order in both
        // lanes: every lane ends with the same bits), then the waves in
        // index order
        for (int o = 1; o < WAVE; o <<= 1) a = comb(a, __shfl_xor(a, o, WAVE));
        if (lane == 0) red[w] = a;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = red[0];
            for (int k = 1; k < NWAVE; ++k) r = comb(r, red[k]);
            ctl[ctl0 + s] = r;
        }
        __syncthreads();
    }
}

// Perturbative amplitudes K_ia <- A_ia / (e_a - e_i) (k x m, row-major);
// partials per workgroup (one wave per row i): max |A_ia|, max |K_ia|,
// min (e_a - e_i)
__global__ void __launch_bounds__(NT) k_amplitudes(int k, int m, double* Kb, const double* ev, double* part) {
    __shared__ double red[3 * NWAVE];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int i = blockIdx.x * NWAVE + w;
    double aov = 0.0, kmax = 0.0, dmin = INFINITY;
    if (i < k) {
        const double ei = ev[i];
        for (int a = lane; a < m; a += WAVE) {
            const double v = Kb[(size_t)i * m + a];
            const double den = ev[k + a] - ei;
            const double amp = v / den;
            aov = fmax(aov, fabs(v));
            dmin = fmin(dmin, den);
            kmax = fmax(kmax, fabs(amp));
            Kb[(size_t)i * m + a] = amp;
        }
    }
    aov = wave_max(aov);
    kmax = wave_max(kmax);
    dmin = -wave_max(-dmin);
    if (lane == 0) {
        red[w] = aov;
        red[NWAVE + w] = kmax;
        red[2 * NWAVE + w] = dmin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red[0], b = red[NWAVE], c = red[2 * NWAVE];
        for (int x = 1; x < NWAVE; ++x) {
            a = fmax(a, red[x]);
            b = fmax(b, red[NWAVE + x]);
            c = fmin(c, red[2 * NWAVE + x]);
        }
        part[(size_t)blockIdx.x * 3 + 0] = a;
        part[(size_t)blockIdx.x * 3 + 1] = b;
        part[(size_t)blockIdx.x * 3 + 2] = c;
    }
}

// partials of sum x_e^2 over a contiguous array
__global__ void __launch_bounds__(NT) k_sumsq(size_t len, const double* x, double* part) {
    __shared__ double red[NWAVE];
    double s = 0.0;
    for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < len; e += (size_t)gridDim.x * NT) s = fma(x[e], x[e], s);
    s = wave_sum(s);
    if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int k = 0; k < NWAVE; ++k) r += red[k];
        part[blockIdx.x] = r;
    }
}

// Y = a X + b Z over `len` contiguous elements (Y may alias X or Z)
__global__ void __launch_bounds__(256) k_axpby(size_t len, double a, const double* X, double b, const double* Z,
                                               double* Y) {
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < len; e += (size_t)gridDim.x * 256)
        Y[e] = a * X[e] + (Z ? b * Z[e] : 0.0);
}

// Newton-Schulz helpers: T <- T + d I; X <- s (X + X^T) / 2 into Y (n x n)
__global__ void __launch_bounds__(256) k_add_diag(int n, double d, double* T) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) T[(size_t)i * n + i] += d;
}

__global__ void __launch_bounds__(256) k_identity(int n, double* T) {
    const int r = blockIdx.y;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256)
        T[(size_t)r * n + c] = (r == c) ? 1.0 : 0.0;
}

__global__ void __launch_bounds__(256) k_symmetrise(int n, double s, const double* X, double* Y) {
    const int r = blockIdx.y;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256)
        Y[(size_t)r * n + c] = 0.5 * s * (X[(size_t)r * n + c] + X[(size_t)c * n + r]);
}

// partials of max |T_ij - delta_ij|
__global__ void __launch_bounds__(NT) k_dev_identity(int n, const double* T, double* part) {
    __shared__ double red[NWAVE];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int i = blockIdx.x * NWAVE + w;
    double m = 0.0;
    if (i < n)
        for (int j = lane; j < n; j += WAVE) m = fmax(m, fabs(T[(size_t)i * n + j] - (i == j ? 1.0 : 0.0)));
    m = wave_max(m);
    if (lane == 0) red[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int k = 1; k < NWAVE; ++k) r = fmax(r, red[k]);
        part[blockIdx.x] = r;
    }
}

// ---------------------------------------------------------------------------
// Cholesky orthogonaliser X = L^-1, S = L L^T (the per-workgroup version is
// chol_setup in scf.hip), 32-wide panels.  k_chol_diag: one wave factors the
// diagonal block j0 of the lower triangle of L in place and writes its
// inverse to Dg (nb x nb, leading dimension nb); a non-positive pivot sets
// ctl[C_ERR].  With factor = 0 it only inverts the (already factored) block.
// ---------------------------------------------------------------------------
__device__ inline void chol_diag_block(ldsf* D, ldsf* Di, glbf* L, int n, int j0, int factor, glbf* Dg,
                                       double* ctl, int lane) {
    constexpr int NB = 32;
    const int nb = n - j0 < NB ? n - j0 : NB;
    for (int e = lane; e < NB * NB; e += WAVE) {
        const int r = e / NB, c = e % NB;
        D[e] = (r < nb && c < nb) ? ((factor || c <= r) ? L[(size_t)(j0 + r) * n + j0 + c] : 0.0)
                                  : (r == c ? 1.0 : 0.0);
    }
    wave_sync();
    if (factor) {
        const bool ok = chol_block_factor(D, lane);
        if (!ok && lane == 0) ctl[C_ERR] = 1.0;
        for (int e = lane; e < nb * nb; e += WAVE) {
            const int r = e / nb, c = e % nb;
            if (c <= r) L[(size_t)(j0 + r) * n + j0 + c] = D[r * NB + c];
        }
        wave_sync();
    }
    chol_block_inverse(D, Di, Dg, nb, lane);
}

__global__ void __launch_bounds__(WAVE) k_chol_diag(double* L_, int n, int j0, int factor, double* Dg_,
                                                    double* ctl) {
    __shared__ double dl[32 * 32], dil[32 * 32];
    chol_diag_block(to_lds(dl), to_lds(dil), to_glb(L_), n, j0, factor, to_glb(Dg_), ctl, threadIdx.x);
}

// The inverses of all diagonal blocks of a finished factor L at once (they
// are independent): block j (rows 32 j ..) to Dg + 32 j * 32, nb x nb with
// leading dimension nb; one wave per block.  Dg holds 32 n doubles (n >= 32)
// or n^2.
__global__ void __launch_bounds__(WAVE) k_chol_inv_all(double* L_, int n, double* Dg_) {
    __shared__ double dl[32 * 32], dil[32 * 32];
    const int j0 = blockIdx.x * 32;
    chol_diag_block(to_lds(dl), to_lds(dil), to_glb(L_), n, j0, 0, to_glb(Dg_ + (size_t)j0 * 32), nullptr,
                    threadIdx.x);
}

// Y[r][c] = X[r][c] for r < rows, c < cols (leading dimensions ldx, ldy)
__global__ void __launch_bounds__(256) k_copy_block(int rows, int cols, const double* X, int ldx, double* Y,
                                                    int ldy) {
    const int r = blockIdx.y;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < cols; c += gridDim.x * 256)
        Y[(size_t)r * ldy + c] = X[(size_t)r * ldx + c];
}

__global__ void k_set_slot(double* ctl, int slot, double v) {
    if (threadIdx.x == 0) ctl[slot] = v;
}

// ---------------------------------------------------------------------------
// Block Jacobi on the whole GPU (the per-workgroup version is
// wg_block_jacobi).  Round r of the round-robin order pairs the 16-index
// blocks; k_jsub solves each pair's 32 x 32 subproblem with one wave
// (wave_jacobi_sub), stores its rotation Q_k and the rotated diagonal
// block, and flags Q_k == I; k_japply then applies every Q_k of the round
// to the off-diagonal tiles of A (Q_k^T A_kl Q_l) and to the orbital rows.
// ---------------------------------------------------------------------------
__device__ inline void build_schedules_wave(ldsi* tab) {
    const int lane = threadIdx.x % WAVE;
    for (int e = lane; e < (JS - 1) * NPAIR32; e += WAVE) {
        int p, q;
        rr32(e / NPAIR32, e % NPAIR32, p, q);
        tab[e] = p | (q << 8);
    }
    for (int e = lane; e < NBLK32; e += WAVE) {
        int P = 0, rem = e;
        while (rem >= NPAIR32 - P) {
            rem -= NPAIR32 - P;
            ++P;
        }
        tab[(JS - 1) * NPAIR32 + e] = P | ((P + rem) << 8);
    }
}

// one wave (64-thread workgroup) per block pair k of round r
__global__ void __launch_bounds__(WAVE) k_jsub(double* A_, int n, int r, int nb2, double* Qbuf, int* ident,
                                               double tol, int inner) {
    __shared__ double buf[JWAVE_LDS];
    __shared__ int sched[SCHED_INTS];
    glbf* A = to_glb(A_);
    const int k = blockIdx.x, lane = threadIdx.x;
    ldsf* M = to_lds(buf);
    ldsf* Q = M + JPK;
    ldsf* cs = Q + JS * JS;
    ldsi* sc = to_lds(sched);
    build_schedules_wave(sc);
    int I, J;
    rr_pair(r, k, nb2, I, J);
    for (int e = lane; e < JS * JS; e += WAVE) {
        const int i = e / JS, j = e % JS;
        if (j >= i) {
            const int gi = slot_index(I, J, i, n), gj = slot_index(I, J, j, n);
            double v = 0.0;
            if (gi >= 0 && gj >= 0) v = A[(size_t)(gi < gj ? gi : gj) * n + (gi < gj ? gj : gi)];
            pk(M, i, j) = v;
        }
    }
    wave_sync();
    const int rot = wave_jacobi_sub(M, Q, cs, sc, tol, inner, 0, 0u, 0u);
    glbf* qg = to_glb(Qbuf) + (size_t)k * JS * JS;
    if (lane == 0) ident[k] = rot ? 0 : 1;
    if (rot) {
        for (int e = lane; e < JS * JS; e += WAVE) {
            qg[e] = Q[(e % JS) * JS + e / JS];  // Q[i][j] = Qt[j][i]
            const int i = e / JS, j = e % JS;
            const int gi = slot_index(I, J, i, n), gj = slot_index(I, J, j, n);
            if (gi >= 0 && gj >= 0) A[(size_t)gi * n + gj] = pk(M, i, j);
        }
    }
}

// one wave per work item: the ntile off-diagonal tiles (k < l) of A, then
// npair x nch row chunks (64 columns) of the orbitals
__global__ void __launch_bounds__(NT) k_japply(double* A_, double* Ct_, int n, int r, int nb2,
                                               const double* Qbuf_, const int* ident) {
    glbf* A = to_glb(A_);
    glbf* Ct = to_glb(Ct_);
    const glbf* Qbuf = to_glb(Qbuf_);
    const int npair = nb2 / 2;
    const int ntile = npair * (npair - 1) / 2;
    const int nch = (n + WAVE - 1) / WAVE;
    const int it = __builtin_amdgcn_readfirstlane(blockIdx.x * NWAVE + threadIdx.x / WAVE);
    if (it >= ntile + npair * nch) return;
    if (it < ntile) {
        int k = 0, rem = it;
        while (rem >= npair - 1 - k) {
            rem -= npair - 1 - k;
            ++k;
        }
        const int l = k + 1 + rem;
        if (ident[k] && ident[l]) return;
        int Ik, Jk, Il, Jl;
        rr_pair(r, k, nb2, Ik, Jk);
        rr_pair(r, l, nb2, Il, Jl);
        tile_update(A, n, Qbuf + (size_t)k * JS * JS, Qbuf + (size_t)l * JS * JS, Ik, Jk, Il, Jl, ident[k],
                    ident[l]);
    } else {
        const int v = it - ntile;
        const int k = v / nch, ch = v % nch;
        if (ident[k]) return;
        int Ik, Jk;
        rr_pair(r, k, nb2, Ik, Jk);
        rows_update(Ct, n, (size_t)n, n, Qbuf + (size_t)k * JS * JS, Ik, Jk, ch * WAVE);
    }
}

// ---------------------------------------------------------------------------
// Sort and Fermi smearing (one workgroup): eigenvalue estimates ev (in the
// rows' order) -> skey (ascending), sidx, ftot[row] = f_alpha + f_beta,
// occ (alpha, beta, ascending order), and control values:
//   ctl[C_EFA], ctl[C_EFB]   Fermi levels
//   ctl[C_FOK]               1 if both channels converged
//   ctl[C_KOCC]              orbitals with f > 0 (a prefix in sorted order)
//   ctl[C_PERM]              1 if the sort moved any row
//   ctl[C_GAPA], ctl[C_GAPB] frontier gaps skey[ia] - skey[ia-1] (ib), +inf
//                            when the index is outside 1..n-1
//   ctl[C_DMAX]              max |ev|
//   ctl[C_FSOFT]             largest f (1 - f) of either channel
// ---------------------------------------------------------------------------

struct SharedG {
    ldsf *skey, *red;
    glbf* ftot;
    ldsi* sidx;
};

__global__ void __launch_bounds__(NT) k_sort_fermi(int n, int nao, int npow, const double* ev, double* skey_out,
                                                   int* sidx_out, double* ftot, double* occ, double nel_a,
                                                   double nel_b, int ia, int ib, double kt, double thr,
                                                   int maxiter, double* ctl) {
    extern __shared__ double lds[];
    ldsf* ls = to_lds(lds);
    SharedG sh;
    sh.red = ls;
    sh.skey = ls + 256;
    sh.sidx = reinterpret_cast<ldsi*>(ls + 256 + npow);
    sh.ftot = to_glb(ftot);
    double dmax = 0.0;
    for (int j = threadIdx.x; j < npow; j += NT) {
        const double e = (j < n) ? ev[j] : INFINITY;
        if (j < n) dmax = fmax(dmax, fabs(e));
        sh.skey[j] = e;
        sh.sidx[j] = j;
    }
    for (int j = threadIdx.x; j < n; j += NT) ftot[j] = 0.0;
    dmax = wg_max(dmax, sh.red);
    __syncthreads();
    wg_bitonic_sort(sh.skey, sh.sidx, npow);
    Mol M;
    M.n = n;
    M.nao = nao;
    double efa = 0.0, efb = 0.0;
    bool ok = fermi_channel(M, sh, nel_a, kt, thr, maxiter, false, to_glb(occ), &efa);
    ok = fermi_channel(M, sh, nel_b, kt, thr, maxiter, false, to_glb(occ) + n, &efb) && ok;
    int kk = 0, moved = 0;
    double fs = 0.0;
    for (int j = threadIdx.x; j < 2 * n; j += NT) fs = fmax(fs, occ[j] * (1.0 - occ[j]));
    fs = wg_max(fs, sh.red);
    for (int j = threadIdx.x; j < n; j += NT) {
        if (ftot[sh.sidx[j]] > 0.0) kk = j + 1;
        moved |= sh.sidx[j] != j;
        skey_out[j] = sh.skey[j];
        sidx_out[j] = sh.sidx[j];
    }
    const double kocc = wg_max((double)kk, sh.red);
    const int perm = __syncthreads_or(moved);
    if (threadIdx.x == 0) {
        ctl[C_EFA] = efa;
        ctl[C_EFB] = efb;
        ctl[C_FOK] = ok ? 1.0 : 0.0;
        ctl[C_KOCC] = kocc;
        ctl[C_PERM] = perm ? 1.0 : 0.0;
        ctl[C_GAPA] = (ia > 0 && ia < n) ? sh.skey[ia] - sh.skey[ia - 1] : INFINITY;
        ctl[C_GAPB] = (ib > 0 && ib < n) ? sh.skey[ib] - sh.skey[ib - 1] : INFINITY;
        ctl[C_DMAX] = dmax;
        ctl[C_FSOFT] = fs;
    }
}

// G[j] = Ct[sidx[j]] (j < rows); one workgroup per row
__global__ void __launch_bounds__(256) k_gather(int n, const double* Ct, const int* sidx, double* G) {
    const int j = blockIdx.x;
    const double* src = Ct + (size_t)sidx[j] * n;
    for (int c = threadIdx.x; c < n; c += 256) G[(size_t)j * n + c] = src[c];
}

// H[j] = w_j G[j], w_j = ftot[sidx[j]] (x ev[sidx[j]] when ev is given;
// sidx null: the identity); one workgroup per row j < k
__global__ void __launch_bounds__(256) k_scale_rows(int n, const double* G, const int* sidx, const double* ftot,
                                                    const double* ev, double* H) {
    const int j = blockIdx.x;
    const int o = sidx ? sidx[j] : j;
    const double w = ev ? ftot[o] * ev[o] : ftot[o];
    for (int c = threadIdx.x; c < n; c += 256) H[(size_t)j * n + c] = w * G[(size_t)j * n + c];
}

// A'_ij = (f_i + f_j)/2 A_ij for the k x k occupied block, f of row i =
// ftot[sidx[i]] (sidx null: ftot[i])
__global__ void __launch_bounds__(256) k_scale_pairs(int k, double* A, const int* sidx, const double* ftot) {
    const int i = blockIdx.x;
    const double fi = ftot[sidx ? sidx[i] : i];
    for (int j = threadIdx.x; j < k; j += 256) A[(size_t)i * k + j] *= 0.5 * (fi + ftot[sidx ? sidx[j] : j]);
}

// Mulliken (scf/base.py:764-815): q_i = n0_i - (P S)_ii, eorb_i = (P H0)_ii
// (last step only); one wave per row
__global__ void __launch_bounds__(NT) k_mulliken(int n, int nao, const double* P, const double* S, const double* H0,
                                                 const double* n0, int last, double* q, double* eorb) {
    const int i = blockIdx.x * NWAVE + threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    if (i >= n) return;
    double sq = 0.0, se = 0.0;
    if (i < nao) {
        const double* pr = P + (size_t)i * n;
        const double* sr = S + (size_t)i * n;
        const double* hr = H0 + (size_t)i * n;
        for (int j = lane; j < nao; j += WAVE) {
            sq = fma(pr[j], sr[j], sq);
            if (last) se = fma(pr[j], hr[j], se);
        }
    }
    sq = wave_sum(sq);
    se = wave_sum(se);
    if (lane == 0) {
        q[i] = (i < nao) ? n0[i] - sq : 0.0;
        eorb[i] = (i < nao) ? se : 0.0;
    }
}

// shell and atom charges (one workgroup)
__global__ void __launch_bounds__(NT) k_shell_atom(int nsh, int nat, const int* sh_ao, const int* sh_us,
                                                   const int* us_l, const int* atom_sh0, const int* atom_nsh,
                                                   const double* q, double* qsh, double* qat) {
    for (int s = threadIdx.x; s < nsh; s += NT) {
        const int o0 = sh_ao[s];
        const int nl = 2 * us_l[sh_us[s]] + 1;
        double acc = 0.0;
        for (int x = 0; x < nl; ++x) acc += q[o0 + x];
        qsh[s] = acc;
    }
    __syncthreads();
    for (int a = threadIdx.x; a < nat; a += NT) {
        const int s0 = atom_sh0[a], ns = atom_nsh[a];
        double acc = 0.0;
        for (int x = 0; x < ns; ++x) acc += qsh[s0 + x];
        qat[a] = acc;
    }
}

// ES2 (gamma q_sh) + ES3 (hd q_at^2) shell potential, one wave per shell
__global__ void __launch_bounds__(NT) k_shell_pot(int nsh, const double* gam, const double* qsh, const double* qat,
                                                  const int* sh_atom, const int* atom_sp, const double* sp_gam3,
                                                  double* gq, double* vsh) {
    const int s = blockIdx.x * NWAVE + threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    if (s >= nsh) return;
    const double* gr = gam + (size_t)s * nsh;
    double acc = 0.0;
    for (int t = lane; t < nsh; t += WAVE) acc = fma(gr[t], qsh[t], acc);
    acc = wave_sum(acc);
    if (lane == 0) {
        const int a = sh_atom[s];
        const double hd = sp_gam3[atom_sp[a]];
        gq[s] = acc;
        vsh[s] = acc + hd * qat[a] * qat[a];
    }
}

__global__ void __launch_bounds__(256) k_spread(int n, int nao, const int* ao_sh, const double* vsh, double* v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int s = (i < nao) ? ao_sh[i] : -1;
    v[i] = (s >= 0) ? vsh[s] : 0.0;
}

// Mixing (one workgroup; the same mixers as the per-molecule kernel):
// Anderson, or Broyden when o.mixer == 1.  ctl[C_CONV] = converged,
// ctl[C_RESID] = the residual's max norm; the Broyden memory's rank and last
// step norm go to ctl[C_BRANK], ctl[C_BDX] for the host's copy of its state
__global__ void __launch_bounds__(NT) k_mix(Mixer mx, Broyden bm, int n, xtb_scf_opts o, double* hx, double* hf,
                                            const double* xnew, const double* xold, double* xout, double* ctl) {
    __shared__ double red[256];
    struct { ldsf* red; } sh{to_lds(red)};
    Mol M;
    M.n = n;
    M.vec0 = 0;
    M.nvec_total = n;
    double resid = 0.0;
    const bool c = o.mixer == 1
        ? broyden(bm, M, sh, o, to_glb(hx), to_glb(hf), to_glb(xnew), to_glb(xold), to_glb(xout), &resid)
        : anderson(mx, M, sh, o, to_glb(hx), to_glb(hf), to_glb(xnew), to_glb(xold), to_glb(xout), &resid);
    if (threadIdx.x == 0) {
        ctl[C_CONV] = c ? 1.0 : 0.0;
        ctl[C_RESID] = resid;
        ctl[C_BRANK] = (double)bm.rank;
        ctl[C_BDX] = bm.dxnorm;
    }
}

// energies (scf/base.py:514-623) into the result record (one workgroup)
__global__ void __launch_bounds__(NT) k_energies(int n, int nao, int nsh, int nat, const double* eorb,
                                                 const double* qsh, const double* gq, const double* qat,
                                                 const int* atom_sp, const double* sp_gam3, const double* occ,
                                                 double kt, double* out) {
    __shared__ double red[256];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nao; i += NT) acc[0] += eorb[i];
    for (int s = threadIdx.x; s < nsh; s += NT) acc[1] += 0.5 * qsh[s] * gq[s];
    for (int a = threadIdx.x; a < nat; a += NT) {
        const double hd = sp_gam3[atom_sp[a]];
        acc[2] += hd * qat[a] * qat[a] * qat[a] / 3.0;
    }
    const double eps = 2.220446049250313e-16;
    for (int j = threadIdx.x; j < 2 * n; j += NT) {
        const double f = occ[j];
        const double f1 = fmax(f, eps), f2 = fmax(1.0 - f, eps);
        acc[3] += f1 * log(f1) + f2 * log(f2);
    }
    wg_sum_n<4>(acc, to_lds(red));
    if (threadIdx.x == 0) {
        out[XTB_OUT_ESCF] = acc[0] + acc[1] + acc[2];
        out[XTB_OUT_EH0] = acc[0];
        out[XTB_OUT_EES2] = acc[1];
        out[XTB_OUT_EES3] = acc[2];
        out[XTB_OUT_FENERGY] = acc[3] * kt;
    }
}

__global__ void k_copy_emo(int n, const double* skey, int nan_out, double* emo) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) emo[i] = nan_out ? NAN : skey[i];
}

}  // namespace big
}  // namespace xtb

// ===========================================================================
// Host driver
// ===========================================================================
namespace {

using namespace xtb;
using namespace xtb::big;

// result-record slots written from the host (mask bit i: slot i)
struct Rec {
    double v[XTB_OUT_STRIDE];
    uint32_t mask;
};

__global__ void k_put(double* out, Rec r) {
    const int i = threadIdx.x;
    if (i < XTB_OUT_STRIDE && ((r.mask >> i) & 1u)) out[i] = r.v[i];
}

__global__ void k_diag(int n, const double* A, double* ev) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) ev[i] = A[(size_t)i * n + i];
}

// ev[a] = src[a] for rows <= a < n
__global__ void k_copy_tail(int rows, int n, const double* src, double* ev) {
    const int a = rows + blockIdx.x * 256 + threadIdx.x;
    if (a < n) ev[a] = src[a];
}

// per-workgroup max |x_i|
__global__ void __launch_bounds__(NT) k_vec_absmax(int n, const double* x, double* part) {
    __shared__ double red[NWAVE];
    double m = 0.0;
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) m = fmax(m, fabs(x[i]));
    m = wave_max(m);
    if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int k = 1; k < NWAVE; ++k) r = fmax(r, red[k]);
        part[blockIdx.x] = r;
    }
}

inline size_t al32(size_t d) { return (d + 31) & ~(size_t)31; }

struct Layout {
    size_t E, Q, ev, skey, ftot, sidx, ident, part, ctl, qat, gq, vsh, total;
};

Layout make_layout(int n, int nsh, int nat) {
    const int nb = (n + JB - 1) / JB, nb2 = nb + (nb & 1), npair = nb2 / 2;
    Layout L;
    size_t o = 0;
    L.E = o; o += al32((size_t)n * n);
    L.Q = o; o += al32((size_t)npair * JS * JS);
    L.ev = o; o += al32(n);
    L.skey = o; o += al32(n);
    L.ftot = o; o += al32(n);
    L.sidx = o; o += al32((n + 1) / 2);
    L.ident = o; o += al32((npair + 1) / 2);
    L.part = o; o += al32(4 * (size_t)n + 4096);
    L.ctl = o; o += al32(C_NCTL);
    L.qat = o; o += al32(nat);
    L.gq = o; o += al32(nsh);
    L.vsh = o; o += al32(nsh);
    L.total = o;
    return L;
}

// largest orbital count: the sort keeps keys and indices of the next power
// of two in one workgroup's LDS
constexpr int LARGE_MAX_N = 8192;

struct Ctx {
    hipStream_t s;
    const xtb_batch* b;
    const xtb_params* par;
    xtb_scf_opts o;
    const xtb_scf_state* st;
    int n, nao, nsh, nat, npow, nb2, npair;
    double nel_a, nel_b;
    int ia, ib;
    const double *S, *H0, *gam;
    double *F, *Ct, *T, *P, *E;  // n x n; the roles of F, Ct and E rotate
    double *ev, *skey, *ftot, *Qbuf, *part, *ctl, *qat, *gq, *vsh;
    int *sidx, *ident;
    double h[C_NCTL];
    double gap_a = INFINITY, gap_b = INFINITY;  // frontier gaps of the last sort
    int refresh = 0;                            // perturbative steps since fresh eigenvalues
    hipError_t err = hipSuccess;
    int pd_steps = 0;
    double t_eig = 0.0, t_pd = 0.0;
};

inline void chk(Ctx& c, hipError_t e) {
    if (e != hipSuccess && c.err == hipSuccess) c.err = e;
}
inline void chk_launch(Ctx& c) { chk(c, hipGetLastError()); }

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool read_ctl(Ctx& c) {
    chk(c, hipMemcpyAsync(c.h, c.ctl, sizeof(c.h), hipMemcpyDeviceToHost, c.s));
    chk(c, hipStreamSynchronize(c.s));
    return c.err == hipSuccess;
}

void finish(Ctx& c, int nblk, int nslot, int ops, int slot) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(NT), 0, c.s, c.part, nblk, nslot, ops, c.ctl, slot);
    chk_launch(c);
}

template <bool TA, bool TB, bool EPI>
void gemm_t(Ctx& c, int M, int N, int K, double alpha, const double* A, int lda, const double* B, int ldb,
            double beta, double* C, int ldc, bool upper, const double* E1 = nullptr, const double* E2 = nullptr,
            double s1 = 0.0, double s2 = 0.0) {
    if (M <= 0 || N <= 0) return;
    static bool attr = false;
    if (!attr) {
        chk(c, hipFuncSetAttribute((const void*)k_gemm<TA, TB, EPI, GT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   GEMM_LDS_BYTES));
        chk(c, hipFuncSetAttribute((const void*)k_gemm<TA, TB, EPI, 64>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   GEMM_LDS_BYTES));
        attr = true;
    }
    auto count = [&](int te) {
        const int tm = (M + te - 1) / te, tn = (N + te - 1) / te;
        return upper ? tm * (tm + 1) / 2 : tm * tn;
    };
    // fewer 128-tiles than a quarter of the CUs: 64-tiles, four times as many
    const int te = count(GT) < 64 ? 64 : GT;
    const int ntiles = count(te);
    const int grid = (ntiles + 7) / 8 * 8;
    if (te == GT)
        hipLaunchKernelGGL((k_gemm<TA, TB, EPI, GT>), dim3(grid), dim3(NT), GEMM_LDS_BYTES, c.s, M, N, K, alpha, A,
                           lda, B, ldb, beta, C, ldc, upper ? 1 : 0, ntiles, E1, E2, s1, s2);
    else
        hipLaunchKernelGGL((k_gemm<TA, TB, EPI, 64>), dim3(grid), dim3(NT), GEMM_LDS_BYTES, c.s, M, N, K, alpha, A,
                           lda, B, ldb, beta, C, ldc, upper ? 1 : 0, ntiles, E1, E2, s1, s2);
    chk_launch(c);
    if (upper) {
        const int nb = (M + 63) / 64;
        if (nb > 1) {
            hipLaunchKernelGGL(k_mirror, dim3(nb * (nb - 1) / 2), dim3(256), 0, c.s, C, M, ldc, te / 64);
            chk_launch(c);
        }
    }
}

// C (n x n, symmetric) = op(A) op(B): the upper tiles, then the mirror
template <bool TA, bool TB>
void gemm_sym(Ctx& c, int n, int K, const double* A, int lda, const double* B, int ldb, double* C, int ldc) {
    gemm_t<TA, TB, false>(c, n, n, K, 1.0, A, lda, B, ldb, 0.0, C, ldc, true);
}

void fock(Ctx& c, const double* v) {
    hipLaunchKernelGGL(k_fock, dim3((c.n + 255) / 256, c.n), dim3(256), 0, c.s, c.n, c.H0, c.S, v, c.F);
    chk_launch(c);
}

void axpby(Ctx& c, size_t len, double a, const double* X, double b, const double* Z, double* Y) {
    const size_t g = (len + 255) / 256;
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)(g < 8192 ? g : 8192)), dim3(256), 0, c.s, len, a, X, b, Z, Y);
    chk_launch(c);
}

// Loewdin orthogonaliser by coupled Newton-Schulz (as ortho_setup in
// scf.hip): Ct <- X = S^-1/2.  Returns false when the iteration did not
// converge (S numerically singular).
bool ortho(Ctx& c) {
    const int n = c.n;
    const size_t nn = (size_t)n * n;
    const int rb = (n + NWAVE - 1) / NWAVE;
    hipLaunchKernelGGL(k_rowabs, dim3(rb), dim3(NT), 0, c.s, n, c.S, c.part);
    chk_launch(c);
    finish(c, rb, 1, 1, C_GERSH);
    if (!read_ctl(c)) return false;
    const double cg = c.h[C_GERSH] * 1.0001;
    double* Y = c.F;
    double* Z = c.Ct;
    double* W = c.P;
    double* T = c.T;
    axpby(c, nn, 1.0 / cg, c.S, 0.0, nullptr, Y);
    hipLaunchKernelGGL(k_identity, dim3((n + 255) / 256, n), dim3(256), 0, c.s, n, Z);
    chk_launch(c);
    int it = 0, extra = -1;
    for (; it < 60; ++it) {
        gemm_t<false, false, false>(c, n, n, n, -0.5, Z, n, Y, n, 0.0, T, n, false);
        hipLaunchKernelGGL(k_add_diag, dim3((n + 255) / 256), dim3(256), 0, c.s, n, 1.5, T);
        chk_launch(c);
        hipLaunchKernelGGL(k_dev_identity, dim3(rb), dim3(NT), 0, c.s, n, T, c.part);
        chk_launch(c);
        finish(c, rb, 1, 1, C_ERR);
        if (!read_ctl(c)) return false;
        const double err = c.h[C_ERR];
        if (extra == 0 || err < 1e-15) break;
        if (extra > 0) --extra;
        else if (err < 1e-9) extra = 1;
        gemm_t<false, false, false>(c, n, n, n, 1.0, Y, n, T, n, 0.0, W, n, false);  // Y T
        gemm_t<false, false, false>(c, n, n, n, 1.0, T, n, Z, n, 0.0, Y, n, false);  // T Z
        double* oldZ = Z;
        Z = Y;
        Y = W;
        W = oldZ;
    }
    // X = Z / sqrt(c), symmetrised, into the orbital buffer (Z is F, Ct or P)
    double* dst = (Z == c.Ct) ? c.E : c.Ct;
    hipLaunchKernelGGL(k_symmetrise, dim3((n + 255) / 256, n), dim3(256), 0, c.s, n, 1.0 / sqrt(cg), Z, dst);
    chk_launch(c);
    if (dst == c.E) {
        double* t = c.Ct;
        c.Ct = c.E;
        c.E = t;
    }
    return it < 60;
}

// Warm start (as in k_scf): the orbitals a previous call left in the state
// (rows in the orbital buffer) become S-orthonormal for the new geometry,
// C <- (C S C^T)^-1/2 C by coupled Newton-Schulz (C S C^T is close to I
// after a small displacement, so a few steps; the symmetric form keeps each
// orbital as close to its old self as possible), and are sorted by their
// energies in the Fock matrix of the kept potential `pot`, so that the first
// step can decouple occupied from virtual orbitals without an eigensolve.
// Returns false, with the orbitals untouched, when C S C^T is too far from
// I (then the caller starts cold).
bool warm_orthonormalise(Ctx& c, const double* pot) {
    const int n = c.n;
    const int rb = (n + NWAVE - 1) / NWAVE;
    // G = C S C^T, the iteration's Y
    gemm_t<false, false, false>(c, n, n, n, 1.0, c.Ct, n, c.S, n, 0.0, c.T, n, false);
    gemm_sym<false, true>(c, n, n, c.T, n, c.Ct, n, c.F, n);
    double* Y = c.F;
    double* Z = c.P;
    double* W = c.E;
    double* T = c.T;
    hipLaunchKernelGGL(k_identity, dim3((n + 255) / 256, n), dim3(256), 0, c.s, n, Z);
    chk_launch(c);
    int it = 0, extra = -1;
    for (; it < 30; ++it) {
        gemm_t<false, false, false>(c, n, n, n, -0.5, Z, n, Y, n, 0.0, T, n, false);
        hipLaunchKernelGGL(k_add_diag, dim3((n + 255) / 256), dim3(256), 0, c.s, n, 1.5, T);
        chk_launch(c);
        hipLaunchKernelGGL(k_dev_identity, dim3(rb), dim3(NT), 0, c.s, n, T, c.part);
        chk_launch(c);
        finish(c, rb, 1, 1, C_ERR);
        if (!read_ctl(c)) return false;
        const double err = c.h[C_ERR];
        // |I - T| = |I - G|/2 at the start: beyond 1/4 the kept orbitals
        // are too far off to be worth keeping
        if (it == 0 && !(err < 0.25)) return false;
        if (extra == 0 || err < 1e-15) break;
        if (extra > 0) --extra;
        else if (err < 1e-9) extra = 1;
        gemm_t<false, false, false>(c, n, n, n, 1.0, Y, n, T, n, 0.0, W, n, false);  // Y T
        gemm_t<false, false, false>(c, n, n, n, 1.0, T, n, Z, n, 0.0, Y, n, false);  // T Z
        double* oldZ = Z;
        Z = Y;
        Y = W;
        W = oldZ;
    }
    if (it >= 30) return false;
    // C <- Z C (Z = G^-1/2), through T
    gemm_t<false, false, false>(c, n, n, n, 1.0, Z, n, c.Ct, n, 0.0, T, n, false);
    chk(c, hipMemcpyAsync(c.Ct, T, sizeof(double) * n * n, hipMemcpyDeviceToDevice, c.s));
    // energy estimates e_i = (C F C^T)_ii, rows sorted by them; the sort's
    // frontier gaps steer the first step's occupied-virtual decoupling
    fock(c, pot);
    gemm_t<false, false, false>(c, n, n, n, 1.0, c.Ct, n, c.F, n, 0.0, T, n, false);
    hipLaunchKernelGGL(k_rowdot, dim3(rb), dim3(NT), 0, c.s, n, n, T, c.Ct, c.ev);
    chk_launch(c);
    const size_t lds = (256 + (size_t)c.npow) * 8 + (size_t)c.npow * 4;
    const xtb_scf_opts& o = c.o;
    hipLaunchKernelGGL(k_sort_fermi, dim3(1), dim3(NT), lds, c.s, n, c.nao, c.npow, c.ev, c.skey, c.sidx, c.ftot,
                       c.st->occ, c.nel_a, c.nel_b, c.ia, c.ib, o.kt, o.fermi_thresh, o.fermi_maxiter, c.ctl);
    chk_launch(c);
    hipLaunchKernelGGL(k_gather, dim3(n), dim3(256), 0, c.s, n, c.Ct, c.sidx, c.F);
    chk_launch(c);
    chk(c, hipMemcpyAsync(c.Ct, c.F, sizeof(double) * n * n, hipMemcpyDeviceToDevice, c.s));
    if (!read_ctl(c)) return false;
    c.gap_a = c.h[C_GAPA];
    c.gap_b = c.h[C_GAPB];
    c.refresh = 1;  // the virtual orbitals' estimates are fresh
    return true;
}

// Cholesky orthogonaliser (as chol_setup in scf.hip): S = L L^T by
// right-looking 32-wide panels (a wave factors each diagonal block, the panel
// solve and the trailing update are grid GEMMs), then X = L^-1 by the
// right-looking triangular inverse: block row j of X <- D_j^-1 X_j, rows
// below -= L_(>j, j) X_j, on the identity.  L in T, X in the orbital buffer,
// R (rows x 32, leading dimension n) in F, the block inverse in P.  Returns
// false when S is not numerically positive definite.
bool ortho_chol(Ctx& c) {
    constexpr int NB = 32;
    const int n = c.n;
    double* L = c.T;
    double* X = c.Ct;
    double* R = c.F;
    double* Dg = c.P;
    chk(c, hipMemcpyAsync(L, c.S, sizeof(double) * n * n, hipMemcpyDeviceToDevice, c.s));
    hipLaunchKernelGGL(k_set_slot, dim3(1), dim3(1), 0, c.s, c.ctl, (int)C_ERR, 0.0);
    chk_launch(c);
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int nb = n - j0 < NB ? n - j0 : NB;
        hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(WAVE), 0, c.s, L, n, j0, 1, Dg, c.ctl);
        chk_launch(c);
        const int mt = n - j0 - nb;
        if (mt > 0) {
            double* Pn = L + (size_t)(j0 + nb) * n + j0;  // panel below the block
            gemm_t<false, true, false>(c, mt, nb, nb, 1.0, Pn, n, Dg, nb, 0.0, R, n, false);  // b D^-T
            hipLaunchKernelGGL(k_copy_block, dim3(1, mt), dim3(256), 0, c.s, mt, nb, R, n, Pn, n);
            chk_launch(c);
            gemm_t<false, true, false>(c, mt, mt, nb, -1.0, Pn, n, Pn, n, 1.0, Pn + nb, n, false);
        }
    }
    if (!read_ctl(c) || c.h[C_ERR] != 0.0) return false;
    hipLaunchKernelGGL(k_identity, dim3((n + 255) / 256, n), dim3(256), 0, c.s, n, X);
    chk_launch(c);
    hipLaunchKernelGGL(k_chol_inv_all, dim3((n + NB - 1) / NB), dim3(WAVE), 0, c.s, L, n, Dg);
    chk_launch(c);
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int nb = n - j0 < NB ? n - j0 : NB;
        const int w = j0 + nb;  // non-zero columns of block row j
        double* Xj = X + (size_t)j0 * n;
        gemm_t<false, false, false>(c, nb, w, nb, 1.0, Dg + (size_t)j0 * NB, nb, Xj, n, 0.0, R, n,
                                    false);  // D^-1 X_j
        hipLaunchKernelGGL(k_copy_block, dim3((w + 255) / 256, nb), dim3(256), 0, c.s, nb, w, R, n, Xj, n);
        chk_launch(c);
        const int mt = n - w;
        if (mt > 0)
            gemm_t<false, false, false>(c, mt, w, nb, -1.0, L + (size_t)w * n + j0, n, Xj, n, 1.0,
                                        X + (size_t)w * n, n, false);
    }
    return c.err == hipSuccess;
}

// Block Jacobi on A = F (full symmetric); rotations applied to the orbital
// rows.  Tolerance as eig_step in scf.hip.  Eigenvalue estimates -> ev.
// Returns sweeps, or -1 when max_sweeps were spent.
int eig(Ctx& c, double rel_tol, double loose) {
    const double t0 = now_s();
    const int n = c.n;
    const int rb = (n + NWAVE - 1) / NWAVE;
    const int ntile = c.npair * (c.npair - 1) / 2;
    const int nch = (n + WAVE - 1) / WAVE;
    const int nitem = ntile + c.npair * nch;
    // scale of the tolerance: the largest diagonal element of A
    hipLaunchKernelGGL(k_diag, dim3((n + 255) / 256), dim3(256), 0, c.s, n, c.F, c.ev);
    chk_launch(c);
    hipLaunchKernelGGL(k_vec_absmax, dim3(1), dim3(NT), 0, c.s, n, c.ev, c.part);
    chk_launch(c);
    finish(c, 1, 1, 1, C_DMAX);
    double tol = 0.0;
    int sweep = 0;
    for (;;) {
        hipLaunchKernelGGL(k_offmax, dim3(rb), dim3(NT), 0, c.s, n, c.F, c.part);
        chk_launch(c);
        finish(c, rb, 1, 1, C_OFF);
        if (!read_ctl(c)) return -1;
        if (sweep == 0) tol = fmax(rel_tol * fmax(c.h[C_DMAX], 1.0), loose);
        if (c.h[C_OFF] <= tol) break;
        if (sweep >= c.o.jacobi_max_sweeps) {
            sweep = -1;
            break;
        }
        for (int r = 0; r < c.nb2 - 1; ++r) {
            hipLaunchKernelGGL(k_jsub, dim3(c.npair), dim3(WAVE), 0, c.s, c.F, n, r, c.nb2, c.Qbuf, c.ident, tol,
                               c.o.jacobi_inner);
            chk_launch(c);
            hipLaunchKernelGGL(k_japply, dim3((nitem + NWAVE - 1) / NWAVE), dim3(NT), 0, c.s, c.F, c.Ct, n, r,
                               c.nb2, c.Qbuf, c.ident);
            chk_launch(c);
        }
        ++sweep;
    }
    hipLaunchKernelGGL(k_diag, dim3((n + 255) / 256), dim3(256), 0, c.s, n, c.F, c.ev);
    chk_launch(c);
    c.refresh = 1;  // eigenvalues fresh for the next perturbative step
    c.t_eig += now_s() - t0;
    return sweep;
}

// Perturbative occupied-virtual rotation of a closed shell with k doubly
// occupied rows first (pd_step in scf.hip, with grid GEMMs).  Returns 1
// when done, 0 to ask for Jacobi with Ct and F untouched, -1 to ask for it
// after F served as scratch.
int pd(Ctx& c, int k, double loose, const double* vin) {
    const int n = c.n, m = n - k;
    if (k <= 0 || m <= 0 || 2 * k > n) return 0;
    const double t0 = now_s();
    double* Cv = c.Ct + (size_t)k * n;
    double* Kb = c.P;  // k x m
    const bool refresh = c.refresh == 0 || c.refresh >= 4;
    c.refresh = refresh ? 1 : c.refresh + 1;
    const int rows = refresh ? n : k;
    int ret = 1;
    const int npass = loose > 0.0 ? PD_INEXACT_PASSES : PD_EXACT_PASSES;
    for (int pass = 0; pass < npass; ++pass) {
        if (pass > 0) fock(c, vin);
        gemm_t<false, false, false>(c, rows, n, n, 1.0, c.Ct, n, c.F, n, 0.0, c.T, n, false);  // C F
        gemm_t<false, true, false>(c, k, m, n, 1.0, c.T, n, Cv, n, 0.0, Kb, m, false);        // A_ov
        hipLaunchKernelGGL(k_rowdot, dim3((rows + NWAVE - 1) / NWAVE), dim3(NT), 0, c.s, rows, n, c.T, c.Ct, c.ev);
        chk_launch(c);
        if (rows < n) {
            hipLaunchKernelGGL(k_copy_tail, dim3((n - rows + 255) / 256), dim3(256), 0, c.s, rows, n, c.skey, c.ev);
            chk_launch(c);
        }
        hipLaunchKernelGGL(k_vec_absmax, dim3(1), dim3(NT), 0, c.s, n, c.ev, c.part);
        chk_launch(c);
        finish(c, 1, 1, 1, C_DMAX);
        const int ab = (k + NWAVE - 1) / NWAVE;
        hipLaunchKernelGGL(k_amplitudes, dim3(ab), dim3(NT), 0, c.s, k, m, Kb, c.ev, c.part);
        chk_launch(c);
        finish(c, ab, 3, 1 | (1 << 2) | (2 << 4), C_AOV);
        if (!read_ctl(c)) return 0;
        const double tol = fmax(c.o.jacobi_tol * fmax(c.h[C_DMAX], 1.0),
                                loose > 0.0 ? loose : PD_FINAL_REL * c.o.x_atol_max);
        const double aov = c.h[C_AOV], kmax = c.h[C_KMAX], dmin = c.h[C_DMIN];
        if (aov <= tol) { ret = 1; break; }
        const int fail = pass == 0 ? 0 : -1;
        if (!(dmin > 0.0) || !(kmax <= c.o.pd_kmax)) { ret = fail; break; }
        // an exact call's last pass only measures
        if (loose == 0.0 && pass == npass - 1) { ret = fail; break; }
        double* X = c.T;
        double* Z = c.T + (size_t)k * k;
        gemm_sym<false, true>(c, k, m, Kb, m, Kb, m, X, k);                                  // K K^T
        gemm_t<false, false, false>(c, k, n, m, 1.0, Kb, m, Cv, n, 0.0, Z, n, false);         // K C_vir
        hipLaunchKernelGGL(k_sumsq, dim3(256), dim3(NT), 0, c.s, (size_t)k * k, X, c.part);
        chk_launch(c);
        finish(c, 256, 1, 0, C_XF);
        if (!read_ctl(c)) return 0;
        const double xf = sqrt(c.h[C_XF]);
        const double tolj = loose > 0.0 ? fmin(1e-13, fmax(1e-17, 1e-3 * loose)) : 1e-17;
        int J = 0;
        double term = 0.5 * xf;
        bool toolong = false;
        while (term > tolj) {
            ++J;
            term *= xf / ((2.0 * J + 1.0) * (2.0 * J + 2.0));
            if (J > 8) { toolong = true; break; }
        }
        if (toolong) { ret = fail; break; }
        double cc[9], ss[9], ph[9];
        {
            double f = 1.0;  // (2j)!
            for (int j = 0; j <= 8; ++j) {
                if (j > 0) f *= (2.0 * j - 1.0) * (2.0 * j);
                const double sg = (j & 1) ? -1.0 : 1.0;
                cc[j] = sg / f;
                ss[j] = sg / (f * (2.0 * j + 1.0));
                ph[j] = -sg / (f * (2.0 * j + 1.0) * (2.0 * j + 2.0));
            }
        }
        const size_t kn = (size_t)k * n;
        double* cur = c.F;
        double* nxt = c.F + kn;
        // R = sinc(X) C_occ + phi(X) Z
        axpby(c, kn, ss[J], c.Ct, ph[J], Z, cur);
        for (int j = J - 1; j >= 0; --j) {
            gemm_t<false, false, true>(c, k, n, k, 1.0, X, k, cur, n, 0.0, nxt, n, false, c.Ct, Z, ss[j], ph[j]);
            double* t = cur; cur = nxt; nxt = t;
        }
        gemm_t<true, false, false>(c, m, n, k, 1.0, Kb, m, cur, n, 1.0, Cv, n, false);  // C_vir += K^T R
        // C_occ <- cos(X) C_occ - sinc(X) Z
        axpby(c, kn, cc[J], c.Ct, -ss[J], Z, cur);
        for (int j = J - 1; j >= 0; --j) {
            gemm_t<false, false, true>(c, k, n, k, 1.0, X, k, cur, n, 0.0, nxt, n, false, c.Ct, Z, cc[j], -ss[j]);
            double* t = cur; cur = nxt; nxt = t;
        }
        axpby(c, kn, 1.0, cur, 0.0, nullptr, c.Ct);
        if (loose > 0.0 && kmax * aov <= tol) { ret = 1; break; }
    }
    c.t_pd += now_s() - t0;
    return ret;
}

// one SCF map vin -> vout (scf_step in scf.hip)
int step(Ctx& c, const double* vin, double* vout, int& status, double loose, bool allow_bd, bool sort_rows,
         bool final_step, int* decoupled, bool last) {
    const int n = c.n;
    const xtb_scf_opts& o = c.o;
    int cb_lo = -1, cb_hi = -1;
    if ((allow_bd || final_step) && o.bd_gap_kt > 0.0) {
        const double na = c.nel_a, nb = c.nel_b;
        const int ia = (int)rint(na), ib = (int)rint(nb);
        bool ok = fabs(na - ia) < 1e-12 && fabs(nb - ib) < 1e-12 && ib <= ia;
        const double need = o.bd_gap_kt * fmax(o.kt, 9.5e-4);
        if (ok && ia > 0 && ia < n) ok = c.gap_a >= need;
        if (ok && ib > 0 && ib < n) ok = c.gap_b >= need;
        if (ok) { cb_lo = ib; cb_hi = ia; }
    }
    fock(c, vin);
    int sw = 0;
    bool pd_done = false;
    if (final_step) {
        bool intra = false;
        if (cb_lo >= 0 && cb_lo == cb_hi && o.pd_kmax > 0.0) {
            const int r = pd(c, cb_lo, 0.0, vin);
            intra = r > 0;
            if (r != 0) fock(c, vin);  // F may have served as scratch
        }
        pd_done = intra && !o.orbitals;
        if (pd_done && decoupled) *decoupled = cb_lo;
    } else if (cb_lo >= 0 && cb_lo == cb_hi && o.pd_kmax > 0.0) {
        const int r = pd(c, cb_lo, loose, vin);
        pd_done = r > 0;
        if (r < 0) fock(c, vin);
    }
    if (pd_done) {
        ++c.pd_steps;
    } else {
        // A = C F C^T: T = F C^T, A = C T (symmetric) into F
        gemm_t<false, true, false>(c, n, n, n, 1.0, c.F, n, c.Ct, n, 0.0, c.T, n, false);
        gemm_sym<false, false>(c, n, n, c.Ct, n, c.T, n, c.F, n);
        sw = eig(c, o.jacobi_tol, loose);
    }
    if (sw < 0 && final_step) status |= XTB_MOL_JACOBI;
    // sort, Fermi occupations, frontier gaps, occupied prefix
    const size_t lds = (256 + (size_t)c.npow) * 8 + (size_t)c.npow * 4;
    hipLaunchKernelGGL(k_sort_fermi, dim3(1), dim3(NT), lds, c.s, n, c.nao, c.npow, c.ev, c.skey, c.sidx, c.ftot,
                       c.st->occ, c.nel_a, c.nel_b, c.ia, c.ib, o.kt, o.fermi_thresh, o.fermi_maxiter, c.ctl);
    chk_launch(c);
    if (!read_ctl(c)) return 0;
    if (c.h[C_FOK] == 0.0) status |= XTB_MOL_FERMI;
    c.gap_a = c.h[C_GAPA];
    c.gap_b = c.h[C_GAPB];
    const int kocc = (int)c.h[C_KOCC];
    const bool perm = c.h[C_PERM] != 0.0;
    // density P = C_occ^T diag(f) C_occ in ascending order
    if (perm) {
        const int rows = sort_rows ? n : kocc;
        if (rows > 0) {
            hipLaunchKernelGGL(k_gather, dim3(rows), dim3(256), 0, c.s, n, c.Ct, c.sidx, c.F);
            chk_launch(c);
        }
        if (kocc > 0) {
            hipLaunchKernelGGL(k_scale_rows, dim3(kocc), dim3(256), 0, c.s, n, c.F, c.sidx, c.ftot,
                               (const double*)nullptr, c.T);
            chk_launch(c);
        }
        gemm_sym<true, false>(c, n, kocc, c.F, n, c.T, n, c.P, n);
        if (sort_rows) {
            // the gathered rows become the orbitals (ascending order)
            double* t = c.Ct;
            c.Ct = c.F;
            c.F = t;
        }
    } else {
        if (kocc > 0) {
            hipLaunchKernelGGL(k_scale_rows, dim3(kocc), dim3(256), 0, c.s, n, c.Ct, (const int*)nullptr, c.ftot,
                               (const double*)nullptr, c.T);
            chk_launch(c);
        }
        gemm_sym<true, false>(c, n, kocc, c.Ct, n, c.T, n, c.P, n);
    }
    // Mulliken charges, shell/atom charges, ES2 + ES3 potential
    const xtb_batch* b = c.b;
    hipLaunchKernelGGL(k_mulliken, dim3((n + NWAVE - 1) / NWAVE), dim3(NT), 0, c.s, n, c.nao, c.P, c.S, c.H0, b->n0,
                       last ? 1 : 0, c.st->q, c.st->eorb);
    chk_launch(c);
    hipLaunchKernelGGL(k_shell_atom, dim3(1), dim3(NT), 0, c.s, c.nsh, c.nat, b->sh_ao, b->sh_us, c.par->us_l,
                       b->atom_sh0, b->atom_nsh, c.st->q, c.st->qsh, c.qat);
    chk_launch(c);
    hipLaunchKernelGGL(k_shell_pot, dim3((c.nsh + NWAVE - 1) / NWAVE), dim3(NT), 0, c.s, c.nsh, c.gam, c.st->qsh,
                       c.qat, b->sh_atom, b->atom_sp, c.par->sp_gam3, c.gq, c.vsh);
    chk_launch(c);
    hipLaunchKernelGGL(k_spread, dim3((n + 255) / 256), dim3(256), 0, c.s, n, c.nao, b->ao_sh, c.vsh, vout);
    chk_launch(c);
    return sw < 0 ? 0 : sw;
}

bool mix(Ctx& c, Mixer& mx, Broyden& bm, const double* xnew, const double* xold, double* xout, double* resid) {
    hipLaunchKernelGGL(k_mix, dim3(1), dim3(NT), 0, c.s, mx, bm, c.n, c.o, c.st->hist_x, c.st->hist_f, xnew, xold,
                       xout, c.ctl);
    chk_launch(c);
    if (!read_ctl(c)) return false;
    // the kernel's own copy of the mixer advanced the same way
    if (c.o.mixer == 1) {
        bm.step += 1;
        bm.rank = (int)c.h[C_BRANK];
        bm.dxnorm = c.h[C_BDX];
    } else {
        mx.step += 1;
        mx.head = (mx.head - 1 + mx.G) % mx.G;
    }
    *resid = c.h[C_RESID];
    return c.h[C_CONV] != 0.0;
}

}  // namespace

extern "C" int32_t xtb_scf_large_workspace(const xtb_batch* batch, int64_t* doubles) {
    if (!batch || !doubles || batch->nmol != 1) return XTB_EINVAL;
    if (batch->max_n > LARGE_MAX_N) return XTB_ETOOBIG;
    *doubles = (int64_t)make_layout(batch->max_n, batch->max_nsh, batch->max_nat).total;
    return XTB_OK;
}

extern "C" int32_t xtb_scf_large(const xtb_batch* batch, const xtb_params* par, const double* S, const double* H0,
                                 const double* gamma, const xtb_scf_opts* opts, const xtb_scf_state* state,
                                 double* work, int64_t work_doubles, void* stream) {
    if (!batch || !par || !S || !H0 || !gamma || !opts || !state || !work) return XTB_EINVAL;
    if (batch->nmol != 1 || batch->max_n <= 0 || batch->has_padding) return XTB_EINVAL;
    if (batch->max_n > LARGE_MAX_N) return XTB_ETOOBIG;
    if (opts->mixer != 0 && opts->mixer != 1) return XTB_EINVAL;
    if (opts->generations < 1 || opts->generations > (opts->mixer == 1 ? BROYDEN_MAXR : MIX_MAXG))
        return XTB_EINVAL;
    if (opts->jacobi_inner < 1 || opts->jacobi_inner > 20 || opts->eig_inexact < 0.0) return XTB_EINVAL;
    if (!state->F || !state->C || !state->T || !state->P || !state->pot || !state->vnew || !state->q ||
        !state->eorb || !state->emo || !state->occ || !state->hist_x || !state->hist_f || !state->qsh ||
        !state->out)
        return XTB_EINVAL;
    const Layout L = make_layout(batch->max_n, batch->max_nsh, batch->max_nat);
    if (work_doubles < (int64_t)L.total) return XTB_EINVAL;
    const double tstart = now_s();
    Ctx c;
    c.s = (hipStream_t)stream;
    c.b = batch;
    c.par = par;
    c.o = *opts;
    c.st = state;
    c.n = batch->max_n;
    c.nao = batch->max_n;
    c.nsh = batch->max_nsh;
    c.nat = batch->max_nat;
    c.npow = 1;
    while (c.npow < c.n) c.npow <<= 1;
    const int nb = (c.n + JB - 1) / JB;
    c.nb2 = nb + (nb & 1);
    c.npair = c.nb2 / 2;
    c.S = S;
    c.H0 = H0;
    c.gam = gamma;
    c.F = state->F;
    c.Ct = state->C;
    c.T = state->T;
    c.P = state->P;
    c.E = work + L.E;
    c.Qbuf = work + L.Q;
    c.ev = work + L.ev;
    c.skey = work + L.skey;
    c.ftot = work + L.ftot;
    c.sidx = reinterpret_cast<int*>(work + L.sidx);
    c.ident = reinterpret_cast<int*>(work + L.ident);
    c.part = work + L.part;
    c.ctl = work + L.ctl;
    c.qat = work + L.qat;
    c.gq = work + L.gq;
    c.vsh = work + L.vsh;
    static bool attr = false;
    if (!attr) {
        chk(c, hipFuncSetAttribute((const void*)k_sort_fermi, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (256 + LARGE_MAX_N) * 8 + LARGE_MAX_N * 4));
        attr = true;
    }
    double nel[2] = {0.0, 0.0};
    double prev_status = 1.0;  // a warm start needs a previous call that succeeded
    chk(c, hipMemcpyAsync(nel, batch->mol_nel, sizeof(nel), hipMemcpyDeviceToHost, c.s));
    if (opts->warm_start)
        chk(c, hipMemcpyAsync(&prev_status, state->out + XTB_OUT_STATUS, sizeof(double), hipMemcpyDeviceToHost,
                              c.s));
    chk(c, hipStreamSynchronize(c.s));
    if (c.err != hipSuccess) return XTB_EHIP;
    c.nel_a = nel[0];
    c.nel_b = nel[1];
    c.ia = (int)rint(nel[0]);
    c.ib = (int)rint(nel[1]);
    double* out = state->out;
    int status = 0;
    Rec rec;
    rec.mask = 0u;
    auto put = [&](int slot, double v) {
        rec.v[slot] = v;
        rec.mask |= 1u << slot;
    };
    // warm start from the orbitals and potential a previous call on this
    // state left (a failed re-orthonormalisation starts cold)
    bool warm = opts->warm_start != 0 && prev_status == 0.0;
    if (warm) warm = warm_orthonormalise(c, state->pot);
    const bool ok_ortho = warm || (opts->ortho == 1 ? ortho_chol(c) : ortho(c));
    const double t_ortho = now_s() - tstart;
    if (c.err != hipSuccess) return XTB_EHIP;
    if (!ok_ortho) {
        status |= XTB_MOL_OVERLAP;
        for (int k = 0; k < XTB_OUT_STRIDE; ++k)
            if (k != XTB_OUT_EREP) put(k, 0.0);
        put(XTB_OUT_ESCF, NAN);
        put(XTB_OUT_STATUS, (double)status);
        hipLaunchKernelGGL(k_put, dim3(1), dim3(XTB_OUT_STRIDE), 0, c.s, out, rec);
        chk_launch(c);
        return c.err == hipSuccess ? XTB_OK : XTB_EHIP;
    }
    const int n = c.n;
    if (!warm) chk(c, hipMemsetAsync(state->pot, 0, sizeof(double) * n, c.s));
    const xtb_scf_opts& o = c.o;
    Mixer mx{0, 0, o.generations + 1, !warm && o.soft_start != 0};
    Broyden bm{0, 0, o.generations, 0.0};
    const bool use_broyden = o.mixer == 1;
    int iters = 0, sweeps = 0;
    double resid = 1e-2;
    // first step: a cold start diagonalises to first_tol; a warm start
    // decouples to 1e-2 of that
    const double loose0 = o.eig_inexact > 0.0 ? (warm ? 1e-2 * o.first_tol : o.first_tol) : 0.0;
    const int sw_first = step(c, state->pot, state->vnew, status, loose0, warm && o.eig_inexact > 0.0, true, false,
                              nullptr, false);
    sweeps += sw_first;
    ++iters;
    if (o.mix_guess || use_broyden) {
        mix(c, mx, bm, state->vnew, state->pot, state->pot, &resid);
    } else {
        chk(c, hipMemcpyAsync(state->pot, state->vnew, sizeof(double) * n, hipMemcpyDeviceToDevice, c.s));
    }
    bool converged = false;
    for (int it = 0; it < o.maxiter && c.err == hipSuccess; ++it) {
        // fractional occupations near the Fermi level tighten the inexact
        // eigensolve by 2 kT (as in scf.hip)
        const double soft = c.h[C_FSOFT] > 1e-6 ? fmin(1.0, 2.0 * o.kt) : 1.0;
        sweeps += step(c, state->pot, state->vnew, status, o.eig_inexact * resid * soft, o.eig_inexact > 0.0,
                       true, false, nullptr, false);
        ++iters;
        if (mix(c, mx, bm, state->vnew, state->pot, state->pot, &resid)) {
            converged = true;
            break;
        }
    }
    if (use_broyden) {
        // the final density comes from the root search's result (in pot)
        chk(c, hipMemcpyAsync(state->vnew, state->pot, sizeof(double) * n, hipMemcpyDeviceToDevice, c.s));
    }
    if (!converged && o.maxiter > 0) status |= XTB_MOL_MAXITER;
    int kdec = 0;
    const int sw_last = step(c, state->vnew, state->pot, status, 0.0, false, false, o.eig_inexact > 0.0, &kdec, true);
    sweeps += sw_last;
    hipLaunchKernelGGL(k_energies, dim3(1), dim3(NT), 0, c.s, n, c.nao, c.nsh, c.nat, state->eorb, state->qsh, c.gq,
                       c.qat, batch->atom_sp, par->sp_gam3, state->occ, o.kt, out);
    chk_launch(c);
    hipLaunchKernelGGL(k_copy_emo, dim3((n + 255) / 256), dim3(256), 0, c.s, n, c.skey, kdec > 0 ? 1 : 0,
                       state->emo);
    chk_launch(c);
    if (o.want_w && kdec > 0) {
        // W = C_occ^T A' C_occ with A = C_occ F C_occ^T, A'_ij = (f_i + f_j)/2 A_ij
        // (the final step's input potential; rows not reordered)
        const int k = kdec;
        fock(c, state->vnew);
        gemm_t<false, false, false>(c, k, n, n, 1.0, c.Ct, n, c.F, n, 0.0, c.T, n, false);
        gemm_t<false, true, false>(c, k, k, n, 1.0, c.T, n, c.Ct, n, 0.0, c.F, k, false);
        hipLaunchKernelGGL(k_scale_pairs, dim3(k), dim3(256), 0, c.s, k, c.F, (const int*)nullptr, c.ftot);
        chk_launch(c);
        double* R = c.F + (size_t)k * k;
        gemm_t<false, false, false>(c, k, n, k, 1.0, c.F, k, c.Ct, n, 0.0, R, n, false);
        gemm_sym<true, false>(c, n, k, c.Ct, n, R, n, c.T, n);
    } else if (o.want_w) {
        // W = C diag(f e) C^T over the occupied orbitals in ascending order
        const int kocc = (int)c.h[C_KOCC];
        if (kocc > 0) {
            hipLaunchKernelGGL(k_gather, dim3(kocc), dim3(256), 0, c.s, n, c.Ct, c.sidx, c.F);
            chk_launch(c);
            hipLaunchKernelGGL(k_scale_rows, dim3(kocc), dim3(256), 0, c.s, n, c.F, c.sidx, c.ftot, c.ev, c.E);
            chk_launch(c);
        }
        gemm_sym<true, false>(c, n, kocc, c.F, n, c.E, n, c.T, n);
    }
    chk(c, hipStreamSynchronize(c.s));
    put(XTB_OUT_ITER, (double)iters);
    put(XTB_OUT_STATUS, (double)status);
    put(XTB_OUT_EFERMI_A, c.h[C_EFA]);
    put(XTB_OUT_EFERMI_B, c.h[C_EFB]);
    put(XTB_OUT_SWEEPS, (double)sweeps);
    put(XTB_OUT_T_ORTHO, t_ortho * 1e6);
    put(XTB_OUT_T_EIGEN, (c.t_eig + c.t_pd) * 1e6);
    put(XTB_OUT_T_PD, c.t_pd * 1e6);
    put(XTB_OUT_PD_STEPS, (double)c.pd_steps);
    put(XTB_OUT_SWEEPS_FIRST, (double)sw_first);
    put(XTB_OUT_SWEEPS_LAST, (double)sw_last);
    put(XTB_OUT_T_TOTAL, (now_s() - tstart) * 1e6);
    hipLaunchKernelGGL(k_put, dim3(1), dim3(XTB_OUT_STRIDE), 0, c.s, out, rec);
    chk_launch(c);
    if (c.err != hipSuccess) return XTB_EHIP;
    // the state's orbital buffer holds the orbitals, whichever buffer they
    // ended in
    if (c.Ct != state->C) {
        chk(c, hipMemcpyAsync(state->C, c.Ct, sizeof(double) * n * n, hipMemcpyDeviceToDevice, c.s));
    }
    return c.err == hipSuccess ? XTB_OK : XTB_EHIP;
}
