// Workgroup-level FP64 building blocks for gfx950 (one molecule per
// workgroup).  Every function here is called by ALL threads of the
// workgroup (they contain barriers) and assumes blockDim.x == XTB_NT.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef XTB_NT
#define XTB_NT 512
#endif

namespace xtb {

constexpr int NT = XTB_NT;
constexpr int WAVE = 64;
constexpr int NWAVE = NT / WAVE;

// Address-space-qualified pointers.  A function compiled out of line cannot
// tell whether a plain pointer addresses LDS or global memory and falls back
// to flat instructions (slower, and every wait covers both kinds of access);
// these types keep ds_* and global_* addressing across calls.
typedef __attribute__((address_space(3))) double ldsf;
typedef __attribute__((address_space(3))) int ldsi;
typedef __attribute__((address_space(1))) double glbf;
__device__ inline ldsf* to_lds(double* p) { return (ldsf*)p; }
__device__ inline ldsi* to_lds(int* p) { return (ldsi*)p; }
__device__ inline glbf* to_glb(double* p) { return (glbf*)p; }
__device__ inline const glbf* to_glb(const double* p) { return (const glbf*)p; }

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
    return v;
}

// Sum over the workgroup; `red` holds >= NWAVE doubles of LDS.  Result is
// returned to every thread.
__device__ inline double wg_sum(double v, ldsf* red) {
    v = wave_sum(v);
    const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NWAVE; ++i) s += red[i];
    __syncthreads();
    return s;
}

__device__ inline double wg_max(double v, ldsf* red) {
    v = wave_max(v);
    const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int i = 1; i < NWAVE; ++i) s = fmax(s, red[i]);
    __syncthreads();
    return s;
}

// Several sums at once: vals[K] per thread -> out[K] on every thread.
// `red` must hold K*NWAVE doubles.
template <int K>
__device__ inline void wg_sum_n(double (&vals)[K], ldsf* red) {
    const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
#pragma unroll
    for (int k = 0; k < K; ++k) vals[k] = wave_sum(vals[k]);
    __syncthreads();
    if (l == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) red[k * NWAVE + w] = vals[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int i = 0; i < NWAVE; ++i) s += red[k * NWAVE + i];
        vals[k] = s;
    }
    __syncthreads();
}

// Y[e] = a X[e] + b Z[e] (Z null: a X[e]) over `len` contiguous elements, by
// the workgroup, four elements per thread in flight.  Y may alias X or Z
// (element-wise).  No barrier inside.
__device__ inline void wg_axpby(size_t len, double a, const glbf* X, double b, const glbf* Z, glbf* Y) {
    size_t e = threadIdx.x;
    for (; e + 3 * NT < len; e += 4 * NT) {
        double x[4], z[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            x[u] = X[e + u * NT];
            z[u] = Z ? Z[e + u * NT] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) Y[e + u * NT] = Z ? a * x[u] + b * z[u] : a * x[u];
    }
    for (; e < len; e += NT) Y[e] = Z ? a * X[e] + b * Z[e] : a * X[e];
}

// ---------------------------------------------------------------------------
// FP64 matrix cores.  v_mfma_f64_16x16x4_f64: D(16x16) += A(16x4) B(4x16);
// lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds
// D[l/16 + 4r][l%16] in element r of its 4-vector (checked on gfx950).  So
// element r of an output fragment is exactly the B operand of k-step r of a
// following product: fragments chain without a trip through LDS.
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ inline v4d mfma16(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ inline v4d v4zero() {
    v4d z = {0.0, 0.0, 0.0, 0.0};
    return z;
}

// C = alpha op(A) op(B) + beta C on matrix cores.  Workgroup tile TM x TN =
// 32 FM x 64 FN, 8 waves as 2 (rows) x 4 (cols) of FM x FN 16x16 fragments
// each, k panels of KP staged in LDS as As[k][m], Bs[k][n] (row pitches
// TM+1, TN+1: conflict-free transposed stores), two stages: the next panel
// is fetched into registers before the current panel's MFMAs and stored to
// the other stage after them, so one barrier per panel separates the
// stages.  `lds` holds >= 2*KP*(TM+TN+2) doubles.  Each global load instruction covers contiguous memory: lanes run
// along m (or n) when the operand is stored that way, along k otherwise.
// Panel loop per fragment pattern (split) or one loop for both: measured per
// operand layout under a full chip of workgroups.
#ifndef XTB_GEMM_SPLIT
#define XTB_GEMM_SPLIT(TA, TB) (true)
#endif

// One TM x TN output tile at (m0, n0).  With EPI, C = alpha op(A) op(B) + s1 E1 + s2 E2 (E1, E2 with C's leading
// dimension; beta unused): an elementwise update fused into the epilogue.  C may
// be E1 or E2 (each element is read and then written by the same thread, once);
// A and B must not overlap C.
//
// With MULL (C not written): the tile's elements x = alpha (op(A) op(B))_ij
// E1_ij over rows and columns below `nao` are summed per row (over the
// tile's columns) and per column (over its rows); part[0..prow) gets the
// first prow row sums and, with pcol, part[TM..TM+TN) the column sums.
// Lanes reduce by xor butterflies (the same bits in every lane), waves
// through LDS in index order: the sums have the same bits on every run.
template <bool TA, bool TB, int FM, int FN, int KP, bool EPI = false, bool MULL = false>
__device__ __forceinline__ void gemm_tile(int m0, int n0, int M, int N, int K, double alpha,
                                          const glbf* __restrict__ A, int lda,
                                          const glbf* __restrict__ B, int ldb, double beta,
                                          glbf* C, int ldc, ldsf* lds,
                                          const glbf* E1, const glbf* E2,
                                          double s1, double s2, int nao = 0, glbf* part = nullptr,
                                          int prow = 0, bool pcol = false) {
    static_assert(NT == 512, "wave layout assumes 8 waves");
    constexpr int TM = 32 * FM, TN = 64 * FN;
    constexpr int LDA_S = TM + 1, LDB_S = TN + 1;
    constexpr int EA = TM * KP / NT, EB = TN * KP / NT;  // staged elements per thread
    static_assert(EA * NT == TM * KP && EB * NT == TN * KP, "panel split over the threads");
    constexpr int STAGE = KP * (LDA_S + LDB_S);
    const int t = threadIdx.x;
    // wave index made scalar, so that per-wave decisions are scalar branches
    const int w = __builtin_amdgcn_readfirstlane(t / WAVE);
    const int lane = t % WAVE, li = lane % 16, lg = lane / 16;
    const int wm = w / 4, wn = w % 4;
    // element e of this thread: (row along the contiguous direction)
    auto a_mk = [&](int e, int& m, int& k) {
        const int idx = t + e * NT;
        if (TA) { m = idx % TM; k = idx / TM; }       // A[k][m]: contiguous m
        else { k = idx % KP; m = idx / KP; }          // A[m][k]: contiguous k
    };
    auto b_nk = [&](int e, int& n, int& k) {
        const int idx = t + e * NT;
        if (TB) { k = idx % KP; n = idx / KP; }       // B[n][k]: contiguous k
        else { n = idx % TN; k = idx / TN; }          // B[k][n]: contiguous n
    };
    // 16x16 fragments of this wave that hold any output (the last tile
    // row/column is partial): the others skip their MFMAs
    const int rrem = M - (m0 + wm * (TM / 2)), crem = N - (n0 + wn * (TN / 4));
    const int nfm = rrem <= 0 ? 0 : (rrem >= FM * 16 ? FM : (rrem + 15) / 16);
    const int nfn = crem <= 0 ? 0 : (crem >= FN * 16 ? FN : (crem + 15) / 16);
    const bool full = nfm == FM && nfn == FN;
    v4d acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = v4zero();
    double pa[EA], pb[EB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
            int m, k;
            a_mk(e, m, k);
            const int gm = m0 + m, gk = k0 + k;
            pa[e] = (gm < M && gk < K) ? (TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk]) : 0.0;
        }
#pragma unroll
        for (int e = 0; e < EB; ++e) {
            int n, k;
            b_nk(e, n, k);
            const int gn = n0 + n, gk = k0 + k;
            pb[e] = (gn < N && gk < K) ? (TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn]) : 0.0;
        }
    };
    auto store = [&](ldsf* st) {
        ldsf* As = st;
        ldsf* Bs = st + KP * LDA_S;
#pragma unroll
        for (int e = 0; e < EA; ++e) {
            int m, k;
            a_mk(e, m, k);
            As[k * LDA_S + m] = pa[e];
        }
#pragma unroll
        for (int e = 0; e < EB; ++e) {
            int n, k;
            b_nk(e, n, k);
            Bs[k * LDB_S + n] = pb[e];
        }
    };
    // MFMAs of one panel when every fragment of the wave holds output
    auto panel_full = [&](const ldsf* As, const ldsf* Bs) {
#pragma unroll
        for (int kk = 0; kk < KP / 4; ++kk) {
            double a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) a[i] = As[(4 * kk + lg) * LDA_S + wm * (TM / 2) + i * 16 + li];
#pragma unroll
            for (int j = 0; j < FN; ++j) b[j] = Bs[(4 * kk + lg) * LDB_S + wn * (TN / 4) + j * 16 + li];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = mfma16(a[i], b[j], acc[i][j]);
            __builtin_amdgcn_sched_group_barrier(0x100, FM + FN, 0);  // DS reads
            __builtin_amdgcn_sched_group_barrier(0x008, FM * FN, 0);  // MFMAs
        }
    };
    // partial tiles: fragments without output skip their MFMAs
    auto panel_part = [&](const ldsf* As, const ldsf* Bs) {
#pragma unroll
        for (int kk = 0; kk < KP / 4; ++kk) {
            double a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) a[i] = As[(4 * kk + lg) * LDA_S + wm * (TM / 2) + i * 16 + li];
#pragma unroll
            for (int j = 0; j < FN; ++j) b[j] = Bs[(4 * kk + lg) * LDB_S + wn * (TN / 4) + j * 16 + li];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    if (i < nfm && j < nfn) acc[i][j] = mfma16(a[i], b[j], acc[i][j]);
        }
    };
    const int np = (K + KP - 1) / KP;
    if (np > 0) {
        fetch(0);
        store(lds);
        __syncthreads();
    }
    if (XTB_GEMM_SPLIT(TA, TB)) {
        // one panel loop per fragment pattern, so the accumulators are
        // not merged between the two code paths after every panel
        if (full) {
            for (int p = 0; p < np; ++p) {
                ldsf* cur = lds + (p & 1) * STAGE;
                if (p + 1 < np) fetch((p + 1) * KP);
                panel_full(cur, cur + KP * LDA_S);
                if (p + 1 < np) store(lds + ((p + 1) & 1) * STAGE);
                __syncthreads();
            }
        } else {
            for (int p = 0; p < np; ++p) {
                ldsf* cur = lds + (p & 1) * STAGE;
                if (p + 1 < np) fetch((p + 1) * KP);
                panel_part(cur, cur + KP * LDA_S);
                if (p + 1 < np) store(lds + ((p + 1) & 1) * STAGE);
                __syncthreads();
            }
        }
    } else {
        for (int p = 0; p < np; ++p) {
            ldsf* cur = lds + (p & 1) * STAGE;
            if (p + 1 < np) fetch((p + 1) * KP);
            if (full) panel_full(cur, cur + KP * LDA_S);
            else panel_part(cur, cur + KP * LDA_S);
            if (p + 1 < np) store(lds + ((p + 1) & 1) * STAGE);
            __syncthreads();
        }
    }
    if constexpr (MULL) {
        double rs[FM][4], cs[FN];
#pragma unroll
        for (int j = 0; j < FN; ++j) cs[j] = 0.0;
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                rs[i][r] = 0.0;
                const int gm = m0 + wm * (TM / 2) + i * 16 + lg + 4 * r;
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    const int gn = n0 + wn * (TN / 4) + j * 16 + li;
                    const double x =
                        (gm < nao && gn < nao) ? alpha * acc[i][j][r] * E1[(size_t)gm * ldc + gn] : 0.0;
                    rs[i][r] += x;
                    cs[j] += x;
                }
            }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) rs[i][r] += __shfl_xor(rs[i][r], o, WAVE);
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int o = 16; o < WAVE; o <<= 1) cs[j] += __shfl_xor(cs[j], o, WAVE);
        ldsf* rred = lds;           // [4 column waves][TM]
        ldsf* cred = lds + 4 * TM;  // [2 row waves][TN]
        if (li == 0)
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) rred[wn * TM + wm * (TM / 2) + i * 16 + lg + 4 * r] = rs[i][r];
        if (lg == 0)
#pragma unroll
            for (int j = 0; j < FN; ++j) cred[wm * TN + wn * (TN / 4) + j * 16 + li] = cs[j];
        __syncthreads();
        for (int e = t; e < prow; e += NT) part[e] = ((rred[e] + rred[TM + e]) + rred[2 * TM + e]) + rred[3 * TM + e];
        if (pcol)
            for (int e = t; e < TN; e += NT) part[TM + e] = cred[e] + cred[TN + e];
        __syncthreads();  // the LDS stages of the next tile reuse these rows
        return;
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * (TM / 2) + i * 16 + lg + 4 * r;
                const int gn = n0 + wn * (TN / 4) + j * 16 + li;
                if (gm < M && gn < N) {
                    const size_t e = (size_t)gm * ldc + gn;
                    if (EPI) C[e] = fma(alpha, acc[i][j][r], fma(s1, E1[e], s2 * E2[e]));
                    else C[e] = (beta == 0.0) ? alpha * acc[i][j][r] : fma(alpha, acc[i][j][r], beta * C[e]);
                }
            }
}

// C = alpha op(A) op(B) + beta C on matrix cores, all tiles by one workgroup
// (see gemm_tile).  With `upper` (M == N, a product known to be symmetric)
// only the tiles on and above the diagonal are formed; gemm_sym mirrors the
// rest.
template <bool TA, bool TB, int FM, int FN, int KP, bool EPI = false>
__device__ __noinline__ void wg_gemm_mfma(int M, int N, int K, double alpha, const glbf* __restrict__ A,
                                          int lda, const glbf* __restrict__ B, int ldb, double beta,
                                          glbf* C, int ldc, ldsf* lds, int upper,
                                          const glbf* E1 = nullptr,
                                          const glbf* E2 = nullptr, double s1 = 0.0,
                                          double s2 = 0.0) {
    constexpr int TM = 32 * FM, TN = 64 * FN;
    // arguments of an out-of-line call arrive in vector registers; they are
    // uniform, and saying so turns the tile and fragment tests into scalar
    // branches instead of exec-masked code
    M = __builtin_amdgcn_readfirstlane(M);
    N = __builtin_amdgcn_readfirstlane(N);
    K = __builtin_amdgcn_readfirstlane(K);
    lda = __builtin_amdgcn_readfirstlane(lda);
    ldb = __builtin_amdgcn_readfirstlane(ldb);
    ldc = __builtin_amdgcn_readfirstlane(ldc);
    upper = __builtin_amdgcn_readfirstlane(upper);
    const int tm = (M + TM - 1) / TM, tn = (N + TN - 1) / TN;
    for (int tile = 0; tile < tm * tn; ++tile) {
        if (upper && tile / tn > tile % tn) continue;
        gemm_tile<TA, TB, FM, FN, KP, EPI>((tile / tn) * TM, (tile % tn) * TN, M, N, K, alpha, A, lda, B,
                                           ldb, beta, C, ldc, lds, E1, E2, s1, s2);
    }
    __syncthreads();
}

// k-panel depth of the workgroup GEMM
#ifndef XTB_GEMM_KP
#define XTB_GEMM_KP 16
#endif
constexpr int GEMM_KP = XTB_GEMM_KP;
constexpr int NGEMM_MFMA_LDS = 2 * GEMM_KP * (128 + 128 + 2);  // largest tile, 128 x 128

// Tile shape for an M x N product: the least padded area, with one 64 x 64
// tile's worth of area added per tile for its fixed costs; ties keep the
// larger tile.  Returns FM * 4 + FN (FM in 2..4 fragments of 32 rows, FN in
// 1..2 of 64 columns).
__device__ inline int gemm_shape(int M, int N) {
    int best = 0;
    long cost = -1;
    for (int fm = 4; fm >= 2; --fm)
        for (int fn = 2; fn >= 1; --fn) {
            const long tm = (M + 32 * fm - 1) / (32 * fm), tn = (N + 64 * fn - 1) / (64 * fn);
            const long c = tm * tn * (32L * fm * 64 * fn + 4096);
            if (cost < 0 || c < cost) { cost = c; best = fm * 4 + fn; }
        }
    return best;
}

#define XTB_GEMM_SHAPES(CALL)        \
    switch (gemm_shape(M, N)) {      \
        case 2 * 4 + 1: CALL(2, 1); break; \
        case 2 * 4 + 2: CALL(2, 2); break; \
        case 3 * 4 + 1: CALL(3, 1); break; \
        case 3 * 4 + 2: CALL(3, 2); break; \
        case 4 * 4 + 1: CALL(4, 1); break; \
        default: CALL(4, 2); break;  \
    }

// dispatch: matrix-core GEMM with a tile shape matched to the size
template <bool TA, bool TB>
__device__ inline void gemm(int M, int N, int K, double alpha, const glbf* A, int lda,
                            const glbf* B, int ldb, double beta, glbf* C, int ldc, ldsf* lds) {
#define XTB_G(fm, fn) wg_gemm_mfma<TA, TB, fm, fn, GEMM_KP>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lds, 0)
    XTB_GEMM_SHAPES(XTB_G)
#undef XTB_G
}

// C = A B + s1 E1 + s2 E2 (row-major, E1 and E2 with C's leading dimension)
__device__ inline void gemm_epi(int M, int N, int K, const glbf* A, int lda, const glbf* B, int ldb,
                                glbf* C, int ldc, const glbf* E1, double s1, const glbf* E2, double s2,
                                ldsf* lds) {
#define XTB_G(fm, fn) \
    wg_gemm_mfma<false, false, fm, fn, GEMM_KP, true>(M, N, K, 1.0, A, lda, B, ldb, 0.0, C, ldc, lds, 0, E1, E2, s1, s2)
    XTB_GEMM_SHAPES(XTB_G)
#undef XTB_G
}

// C (n x n) = alpha op(A) op(B) for a product known to be symmetric: the
// tiles on and above the diagonal on matrix cores, the others mirrored
// through LDS in 64 x 64 blocks (coalesced both ways).  `lds` holds
// >= max(NGEMM_MFMA_LDS, 64 * 65) doubles.
template <bool TA, bool TB>
__device__ inline void gemm_sym(int n, int K, double alpha, const glbf* A, int lda, const glbf* B, int ldb,
                                glbf* C, int ldc, ldsf* lds) {
    // square tiles: 64 x 64 for small n, 128 x 128 otherwise
    const int tile = n <= 64 ? 64 : 128;
    if (tile == 64) wg_gemm_mfma<TA, TB, 2, 1, GEMM_KP>(n, n, K, alpha, A, lda, B, ldb, 0.0, C, ldc, lds, 1);
    else wg_gemm_mfma<TA, TB, 4, 2, GEMM_KP>(n, n, K, alpha, A, lda, B, ldb, 0.0, C, ldc, lds, 1);
    // 64-blocks (bi, bj) whose tile lies below the diagonal: C_bi,bj = C_bj,bi^T
    const int nb = (n + 63) / 64, per = tile / 64;
    for (int bi = 1; bi < nb; ++bi)
        for (int bj = 0; bj < bi; ++bj) {
            if (bi / per <= bj / per) continue;  // formed by the GEMM
            for (int e = threadIdx.x; e < 64 * 64; e += NT) {
                const int r = e / 64, c = e % 64;  // source row bj*64 + r, column bi*64 + c
                const int gr = bj * 64 + r, gc = bi * 64 + c;
                lds[r * 65 + c] = (gr < n && gc < n) ? C[(size_t)gr * ldc + gc] : 0.0;
            }
            __syncthreads();
            for (int e = threadIdx.x; e < 64 * 64; e += NT) {
                const int c = e / 64, r = e % 64;  // destination row bi*64 + c, column bj*64 + r
                const int gr = bi * 64 + c, gc = bj * 64 + r;
                if (gr < n && gc < n) C[(size_t)gr * ldc + gc] = lds[r * 65 + c];
            }
            __syncthreads();
        }
}

// Mulliken sums a_i = sum_{j < nao} P_ij S_ij (0 for i >= nao) of the
// symmetric n x n product P = op(A) op(B), without forming P: the tiles on
// and above the diagonal on matrix cores, each reducing P o S in its
// epilogue per row and per column (a tile above the diagonal stands, by its
// column sums, for the rows of its mirror image); the partials are added per
// row in a fixed order.  Square tiles as gemm_sym.  `part`: scratch of
// 2 T tm^2 doubles (T the tile edge, tm = ceil(n / T)) when n > T; one tile
// writes its row sums straight to `out`.  S with leading dimension n.
template <bool TA, bool TB, int FM>
__device__ __noinline__ void wg_gemm_sym_mull(int n, int K, const glbf* __restrict__ A, int lda,
                                              const glbf* __restrict__ B, int ldb, const glbf* S, int nao,
                                              glbf* part, glbf* out, ldsf* lds) {
    constexpr int FN = FM / 2;
    constexpr int T = 32 * FM;
    static_assert(T == 64 * FN, "square tiles");
    n = __builtin_amdgcn_readfirstlane(n);
    K = __builtin_amdgcn_readfirstlane(K);
    lda = __builtin_amdgcn_readfirstlane(lda);
    ldb = __builtin_amdgcn_readfirstlane(ldb);
    nao = __builtin_amdgcn_readfirstlane(nao);
    const int tm = (n + T - 1) / T;
    if (tm == 1) {
        gemm_tile<TA, TB, FM, FN, GEMM_KP, false, true>(0, 0, n, n, K, 1.0, A, lda, B, ldb, 0.0, nullptr, n, lds,
                                                         S, nullptr, 0.0, 0.0, nao, out, n, false);
        __syncthreads();
        return;
    }
    for (int tile = 0; tile < tm * tm; ++tile) {
        const int I = tile / tm, J = tile % tm;
        if (I > J) continue;
        gemm_tile<TA, TB, FM, FN, GEMM_KP, false, true>(I * T, J * T, n, n, K, 1.0, A, lda, B, ldb, 0.0, nullptr, n,
                                                         lds, S, nullptr, 0.0, 0.0, nao,
                                                         part + (size_t)tile * 2 * T, T, I != J);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += NT) {
        const int I = i / T, r = i - I * T;
        double a = 0.0;
        for (int J = I; J < tm; ++J) a += part[(size_t)(I * tm + J) * 2 * T + r];
        for (int J = 0; J < I; ++J) a += part[(size_t)(J * tm + I) * 2 * T + T + r];
        out[i] = a;
    }
    __syncthreads();
}

template <bool TA, bool TB>
__device__ inline void gemm_sym_mull(int n, int K, const glbf* A, int lda, const glbf* B, int ldb, const glbf* S,
                                     int nao, glbf* part, glbf* out, ldsf* lds) {
    if (n <= 64) wg_gemm_sym_mull<TA, TB, 2>(n, K, A, lda, B, ldb, S, nao, part, out, lds);
    else wg_gemm_sym_mull<TA, TB, 4>(n, K, A, lda, B, ldb, S, nao, part, out, lds);
}

// ---------------------------------------------------------------------------
// Symmetric matrix storage used by the Jacobi eigensolver.  Only the upper
// triangle (i <= j) is referenced.  Packed form lives in LDS, full form in
// global memory.
// ---------------------------------------------------------------------------
struct SymPacked {
    ldsf* a;
    int n;
    __device__ inline ldsf& at(int i, int j) const {
        if (i > j) { int t = i; i = j; j = t; }
        return a[i * n - (i * (i - 1)) / 2 + (j - i)];
    }
};

struct SymFull {
    glbf* a;
    int n;
    __device__ inline glbf& at(int i, int j) const {
        if (i > j) { int t = i; i = j; j = t; }
        return a[(size_t)i * n + j];
    }
};

// Round-robin ("circle method") pairing of n2 (even) players: in round r,
// pair k matches (n2-1, r) for k = 0 and ((r+k) mod (n2-1), (r-k) mod (n2-1))
// otherwise.  Every unordered pair meets once in n2-1 rounds.
__device__ inline void rr_pair(int r, int k, int n2, int& p, int& q) {
    const int m = n2 - 1;
    if (k == 0) { p = m; q = r; }
    else { p = (r + k) % m; q = (r - k + m) % m; }
}

// Cyclic two-sided Jacobi with parallel (round-robin) ordering.
// Diagonalises the symmetric n x n matrix `A` in place (eigenvalues end on
// the diagonal) and applies every rotation to the rows of Ct (Ct[k][mu],
// row k = vector k), i.e. C <- C J.  Ct rows have stride `ldc`.
// `cs` is LDS for 3*(n/2+1) doubles: cos, sin and the packed pair (p, q) of
// each rotation of the round.  Loops are laid out wave x lane so that no
// index needs an integer division.  Returns the number of sweeps, or -1
// when max_sweeps were spent without reaching `tol` (absolute).
//
// Occupation classes (cb_lo >= 0): index i has class (i < cb_lo) + (i < cb_hi)
// (doubly, singly, un-occupied for sorted orbitals).  Mode 1 rotates and
// tests only pairs of different classes, which decouples the classes without
// diagonalising inside them (the density of integer occupations needs no
// more); mode 2 only pairs of the same class, which diagonalises inside
// classes that are already decoupled.  Mode 0 (cb_lo < 0): every pair.
__device__ inline int occ_class(int i, int cb_lo, int cb_hi) { return (i < cb_lo) + (i < cb_hi); }
__device__ inline bool pair_in_mode(int mode, int ci, int cj) {
    return mode == 0 || ((ci != cj) == (mode == 1));
}

template <class Sym, class CAcc>
__device__ int wg_jacobi(Sym A, int n, CAcc Ct, int max_sweeps, double tol, ldsf* cs,
                         ldsf* red, int cb_lo = -1, int cb_hi = -1, bool intra = false) {
    const int mode = cb_lo < 0 ? 0 : (intra ? 2 : 1);
    const int n2 = n + (n & 1);
    const int npair = n2 / 2;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    ldsf* cc = cs;
    ldsf* ss = cs + npair;
    ldsi* pqs = reinterpret_cast<ldsi*>(cs + 2 * npair);  // p | q << 16
    int sweep = 0;
    for (;;) {
        // convergence test: largest off-diagonal element
        double off = 0.0;
        for (int i = w; i < n; i += NWAVE) {
            const int ci = occ_class(i, cb_lo, cb_hi);
            for (int j = i + 1 + lane; j < n; j += WAVE)
                if (pair_in_mode(mode, ci, occ_class(j, cb_lo, cb_hi))) off = fmax(off, fabs(A.at(i, j)));
        }
        off = wg_max(off, red);
        if (off <= tol) return sweep;
        if (sweep >= max_sweeps) return -1;
        const double skip = 1e-3 * tol;
        for (int r = 0; r < n2 - 1; ++r) {
            // 1. rotation parameters (Rutishauser's stable formulas)
            for (int k = threadIdx.x; k < npair; k += NT) {
                int p, q;
                rr_pair(r, k, n2, p, q);
                double c = 1.0, s = 0.0;
                if (p < n && q < n && pair_in_mode(mode, occ_class(p, cb_lo, cb_hi), occ_class(q, cb_lo, cb_hi))) {
                    const double apq = A.at(p, q);
                    if (fabs(apq) > skip) {
                        const double app = A.at(p, p), aqq = A.at(q, q);
                        const double theta = (aqq - app) / (2.0 * apq);
                        double t = 1.0 / (fabs(theta) + sqrt(fma(theta, theta, 1.0)));
                        if (theta < 0.0) t = -t;
                        c = 1.0 / sqrt(fma(t, t, 1.0));
                        s = t * c;
                    }
                }
                cc[k] = c;
                ss[k] = s;
                pqs[k] = p | (q << 16);
            }
            __syncthreads();
            // 2a. A <- J^T A J on 2x2 blocks (P <= R): a wave per block row P
            for (int P = w; P < npair; P += NWAVE) {
                const int pq = pqs[P];
                const int p = pq & 0xffff, q = pq >> 16;
                if (p >= n || q >= n) continue;
                const double sp = ss[P], cp = cc[P];
                for (int R = P + lane; R < npair; R += WAVE) {
                    const double sr = ss[R];
                    if (sp == 0.0 && sr == 0.0) continue;
                    const int xy = pqs[R];
                    const int x = xy & 0xffff, y = xy >> 16;
                    if (x >= n || y >= n) continue;
                    const double cr = cc[R];
                    if (P == R) {
                        const double apq = A.at(p, q);
                        const double t = sp / cp;
                        A.at(p, p) = A.at(p, p) - t * apq;
                        A.at(q, q) = A.at(q, q) + t * apq;
                        A.at(p, q) = 0.0;
                    } else {
                        const double m00 = A.at(p, x), m01 = A.at(p, y);
                        const double m10 = A.at(q, x), m11 = A.at(q, y);
                        // rows: J_P^T M
                        const double r00 = cp * m00 - sp * m10, r01 = cp * m01 - sp * m11;
                        const double r10 = sp * m00 + cp * m10, r11 = sp * m01 + cp * m11;
                        // cols: (.) J_R
                        A.at(p, x) = r00 * cr - r01 * sr;
                        A.at(p, y) = r00 * sr + r01 * cr;
                        A.at(q, x) = r10 * cr - r11 * sr;
                        A.at(q, y) = r10 * sr + r11 * cr;
                    }
                }
            }
            // 2b. vectors: rows p, q of Ct, a wave per rotation
            for (int k = w; k < npair; k += NWAVE) {
                const double s = ss[k];
                if (s == 0.0) continue;
                const int pq = pqs[k];
                const int p = pq & 0xffff, q = pq >> 16;
                if (p >= n || q >= n) continue;
                const double c = cc[k];
                for (int mu = lane; mu < n; mu += WAVE) {
                    auto& vp = Ct.at(p, mu);
                    auto& vq = Ct.at(q, mu);
                    const double a = vp, b = vq;
                    vp = c * a - s * b;
                    vq = s * a + c * b;
                }
            }
            __syncthreads();
        }
        ++sweep;
    }
}

// ---------------------------------------------------------------------------
// Block Jacobi for matrices that do not fit in LDS.
//
// Indices are grouped in blocks of JB = 16; a round of the round-robin
// ordering pairs the blocks, and every pair (I, J) defines a subproblem on
// the <= 32 indices of I u J.  One wavefront diagonalises its subproblem in
// LDS with scalar cyclic Jacobi (no workgroup barrier inside), producing an
// orthogonal Q_k.  The workgroup then applies all Q_k of the round to A
// (rows, then columns) and to the vectors Ct, streaming each element once
// per phase and reusing every loaded value 32 times.  Per sweep this moves
// ~3 n^3 bytes instead of the ~32 n^3 of the scalar method.
// ---------------------------------------------------------------------------
constexpr int JB = 16;
constexpr int JS = 2 * JB;                  // subproblem dimension
constexpr int JPK = JS * (JS + 1) / 2;      // packed subproblem
constexpr int JWAVE_LDS = JPK + JS * JS + JS;  // M (packed), Q, c/s
constexpr int NPAIR32 = JS / 2;                      // 16 pairs per round
constexpr int NBLK32 = NPAIR32 * (NPAIR32 + 1) / 2;  // 136 blocks
constexpr int SCHED_INTS = (JS - 1) * NPAIR32 + NBLK32;
constexpr int NBLOCK_LDS = NWAVE * JWAVE_LDS + 64 + (SCHED_INTS + 1) / 2;  // + flags, schedules

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline ldsf& pk(ldsf* M, int i, int j) {
    if (i > j) { int t = i; i = j; j = t; }
    return M[i * JS - (i * (i - 1)) / 2 + (j - i)];
}

// Global index of slot i (0..31) of the subproblem of block pair (I, J):
// slots 0..15 are block I, 16..31 block J; -1 where the index is beyond n
// (partial last block, or the dummy block of an odd block count).
__device__ inline int slot_index(int I, int J, int i, int n) {
    const int g = (i < JB) ? I * JB + i : J * JB + (i - JB);
    return g < n ? g : -1;
}

// Round-robin pair of a 32-slot subproblem (constant modulus).
__device__ inline void rr32(int r, int k, int& p, int& q) {
    if (k == 0) { p = 31; q = r; }
    else { p = (r + k) % 31; q = (r - k + 31) % 31; }
}

// Schedules of the 32-slot subproblem, built once per block-Jacobi call in
// LDS: rrtab[r*16+k] = p | q<<8 for pair k of round r, and blktab[e] =
// P | R<<8 for the e-th 2x2 block (P <= R) of a round.

__device__ inline void build_schedules(ldsi* tab) {
    for (int e = threadIdx.x; e < (JS - 1) * NPAIR32; e += NT) {
        int p, q;
        rr32(e / NPAIR32, e % NPAIR32, p, q);
        tab[e] = p | (q << 8);
    }
    for (int e = threadIdx.x; e < NBLK32; e += NT) {
        int P = 0, rem = e;
        while (rem >= NPAIR32 - P) { rem -= NPAIR32 - P; ++P; }
        tab[(JS - 1) * NPAIR32 + e] = P | ((P + rem) << 8);
    }
}

// Wave-level cyclic Jacobi on the packed JS x JS matrix M (slots whose
// rows/columns are zero are never rotated).  Qt (JS x JS, Qt[col][row])
// accumulates the rotations so that M_final = Q^T M_0 Q.  Returns 1 if any
// rotation was applied, 0 if M was already diagonal to `tol` (Q = I).
//
// A round is two LDS round trips per lane: the 16 rotation lanes read their
// pivots and publish (c, s); then every lane gathers the schedule entries,
// (c, s) and matrix/vector elements of all its work items at once, computes,
// and stores.  Work items: the 2x2 blocks lane, lane+64, lane+128 (of 136)
// of the two-sided update, and 8 (pair, row) elements of Qt.  Items touch
// disjoint elements, so no ordering among them is needed.
__device__ __noinline__ int wave_jacobi_sub(ldsf* M, ldsf* Qt, ldsf* cs, const ldsi* sched,
                                            double tol, int max_sweeps, int mode, unsigned lo_mask,
                                            unsigned hi_mask) {
    const int lane = threadIdx.x % WAVE;
    // occupation class of a slot (pairs are selected by mode as in wg_jacobi)
    auto scls = [&](int i) { return (int)((lo_mask >> i) & 1u) + (int)((hi_mask >> i) & 1u); };
    const ldsi* rrtab = sched;
    const ldsi* blktab = sched + (JS - 1) * NPAIR32;
    constexpr int NB = (NBLK32 + WAVE - 1) / WAVE;  // 3 block slots per lane
    constexpr int NQ = NPAIR32 / (WAVE / JS);       // 8 Qt items per lane
    int bP[NB], bR[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int e = lane + WAVE * j;
        const int PR = (e < NBLK32) ? blktab[e] : -1;
        bP[j] = (e < NBLK32) ? (PR & 255) : -1;
        bR[j] = (e < NBLK32) ? (PR >> 8) : -1;
    }
    const int qrow = lane % JS, qk0 = lane / JS;
    for (int e = lane; e < JS * JS; e += WAVE) Qt[e] = ((e / JS) == (e % JS)) ? 1.0 : 0.0;
    wave_sync();
    const double skip = 1e-3 * tol;
    int rotated = 0;
    for (int sw = 0; sw < max_sweeps; ++sw) {
        double off = 0.0;
        for (int e = lane; e < JS * JS; e += WAVE) {
            const int i = e / JS, j = e % JS;
            if (j > i && pair_in_mode(mode, scls(i), scls(j))) off = fmax(off, fabs(pk(M, i, j)));
        }
        off = wave_max(off);
        if (off <= tol) break;
        rotated = 1;
        for (int r = 0; r < JS - 1; ++r) {
            const ldsi* rr = rrtab + r * NPAIR32;
            if (lane < NPAIR32) {
                const int pq = rr[lane];
                const int p = pq & 255, q = pq >> 8;
                double c = 1.0, s = 0.0;
                const double apq = pk(M, p, q), app = pk(M, p, p), aqq = pk(M, q, q);
                if (fabs(apq) > skip && pair_in_mode(mode, scls(p), scls(q))) {
                    const double theta = (aqq - app) / (2.0 * apq);
                    double t = 1.0 / (fabs(theta) + sqrt(fma(theta, theta, 1.0)));
                    if (theta < 0.0) t = -t;
                    c = 1.0 / sqrt(fma(t, t, 1.0));
                    s = t * c;
                }
                cs[lane] = c;
                cs[JB + lane] = s;
            }
            wave_sync();
            // gather: schedule entries and rotation parameters
            int bpq[NB], bxy[NB];
            double bcp[NB], bsp[NB], bcr[NB], bsr[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int P = bP[j] < 0 ? 0 : bP[j], R = bR[j] < 0 ? 0 : bR[j];
                bpq[j] = rr[P];
                bxy[j] = rr[R];
                bcp[j] = cs[P];
                bsp[j] = cs[JB + P];
                bcr[j] = cs[R];
                bsr[j] = cs[JB + R];
            }
            int qpq[NQ];
            double qc[NQ], qs[NQ];
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                const int k = qk0 + 2 * j;
                qpq[j] = rr[k];
                qc[j] = cs[k];
                qs[j] = cs[JB + k];
            }
            // gather: matrix and vector elements
            double m[NB][4];
            int mi[NB][4];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int p = bpq[j] & 255, q = bpq[j] >> 8, x = bxy[j] & 255, y = bxy[j] >> 8;
                const bool diag = bP[j] == bR[j];
                int i0 = diag ? p : p, j0 = diag ? p : x;
                int i1 = diag ? q : p, j1 = diag ? q : y;
                int i2 = diag ? p : q, j2 = diag ? q : x;
                int i3 = diag ? p : q, j3 = diag ? q : y;
                mi[j][0] = (i0 <= j0) ? i0 * JS - (i0 * (i0 - 1)) / 2 + (j0 - i0)
                                      : j0 * JS - (j0 * (j0 - 1)) / 2 + (i0 - j0);
                mi[j][1] = (i1 <= j1) ? i1 * JS - (i1 * (i1 - 1)) / 2 + (j1 - i1)
                                      : j1 * JS - (j1 * (j1 - 1)) / 2 + (i1 - j1);
                mi[j][2] = (i2 <= j2) ? i2 * JS - (i2 * (i2 - 1)) / 2 + (j2 - i2)
                                      : j2 * JS - (j2 * (j2 - 1)) / 2 + (i2 - j2);
                mi[j][3] = (i3 <= j3) ? i3 * JS - (i3 * (i3 - 1)) / 2 + (j3 - i3)
                                      : j3 * JS - (j3 * (j3 - 1)) / 2 + (i3 - j3);
#pragma unroll
                for (int u = 0; u < 4; ++u) m[j][u] = M[mi[j][u]];
            }
            double qa[NQ], qb[NQ];
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                const int p = qpq[j] & 255, q = qpq[j] >> 8;
                qa[j] = Qt[p * JS + qrow];
                qb[j] = Qt[q * JS + qrow];
            }
            // compute and store.  Diagonal block (P == R): elements
            // (p,p), (q,q), (p,q); off-diagonal: (p,x), (p,y), (q,x), (q,y).
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const double sp = bsp[j], sr = bsr[j], cp = bcp[j], cr = bcr[j];
                if (bP[j] < 0 || (sp == 0.0 && sr == 0.0)) continue;
                if (bP[j] == bR[j]) {
                    const double apq = m[j][2];
                    const double t = sp / cp;
                    M[mi[j][0]] = m[j][0] - t * apq;
                    M[mi[j][1]] = m[j][1] + t * apq;
                    M[mi[j][2]] = 0.0;
                } else {
                    const double m00 = m[j][0], m01 = m[j][1], m10 = m[j][2], m11 = m[j][3];
                    const double r00 = cp * m00 - sp * m10, r01 = cp * m01 - sp * m11;
                    const double r10 = sp * m00 + cp * m10, r11 = sp * m01 + cp * m11;
                    M[mi[j][0]] = r00 * cr - r01 * sr;
                    M[mi[j][1]] = r00 * sr + r01 * cr;
                    M[mi[j][2]] = r10 * cr - r11 * sr;
                    M[mi[j][3]] = r10 * sr + r11 * cr;
                }
            }
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                const double s = qs[j];
                if (s == 0.0) continue;
                const double c = qc[j];
                const int p = qpq[j] & 255, q = qpq[j] >> 8;
                Qt[p * JS + qrow] = c * qa[j] - s * qb[j];
                Qt[q * JS + qrow] = s * qa[j] + c * qb[j];
            }
            wave_sync();
        }
    }
    return rotated;
}

// Two-sided tile update on matrix cores (one wavefront):
// A[P_k, P_l] <- Q_k^T A[P_k, P_l] Q_l, written to both triangles.
// A is symmetric before the round, so the tile is read as A[P_l, P_k]
// (row segments, coalesced).  Y = A_kl Q_l stays in registers and feeds the
// second product as its B operand (element r = k-step r).
__device__ __noinline__ void tile_update(glbf* A, int n, const glbf* Qk, const glbf* Ql,
                                         int Ik, int Jk, int Il, int Jl, int idk, int idl) {
    const int lane = threadIdx.x % WAVE, li = lane % 16, lg = lane / 16;
    int gk[2], gl_[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        gk[b] = slot_index(Ik, Jk, b * 16 + li, n);
        gl_[b] = slot_index(Il, Jl, b * 16 + li, n);
    }
    v4d y[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) y[i][j] = v4zero();
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int kp = kb * 16 + lg + 4 * kk;
            const int g = slot_index(Il, Jl, kp, n);
            double a[2], bq[2];
#pragma unroll
            for (int bm = 0; bm < 2; ++bm)
                a[bm] = (g >= 0 && gk[bm] >= 0) ? A[(size_t)g * n + gk[bm]] : 0.0;
#pragma unroll
            for (int bn = 0; bn < 2; ++bn)
                bq[bn] = idl ? ((kp == bn * 16 + li) ? 1.0 : 0.0) : Ql[kp * JS + bn * 16 + li];
#pragma unroll
            for (int bm = 0; bm < 2; ++bm)
#pragma unroll
                for (int bn = 0; bn < 2; ++bn) y[bm][bn] = mfma16(a[bm], bq[bn], y[bm][bn]);
        }
    v4d z[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) z[i][j] = v4zero();
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = kb * 16 + lg + 4 * r;
            double a[2];
#pragma unroll
            for (int bm = 0; bm < 2; ++bm)
                a[bm] = idk ? ((k == bm * 16 + li) ? 1.0 : 0.0) : Qk[k * JS + bm * 16 + li];
#pragma unroll
            for (int bm = 0; bm < 2; ++bm)
#pragma unroll
                for (int bn = 0; bn < 2; ++bn) z[bm][bn] = mfma16(a[bm], y[kb][bn][r], z[bm][bn]);
        }
#pragma unroll
    for (int bm = 0; bm < 2; ++bm)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gi = slot_index(Ik, Jk, bm * 16 + lg + 4 * r, n);
            if (gi < 0) continue;
#pragma unroll
            for (int bn = 0; bn < 2; ++bn) {
                const int gj = gl_[bn];
                if (gj < 0) continue;
                A[(size_t)gi * n + gj] = z[bm][bn][r];
                A[(size_t)gj * n + gi] = z[bm][bn][r];
            }
        }
}

// Vector update on matrix cores (one wavefront): rows P_k of the row-major
// matrix X (stride ld), columns c0..c0+63:  X[P_k, c] <- Q_k^T X[P_k, c].
__device__ __noinline__ void rows_update(glbf* X, int n, size_t ld, int ncol, const glbf* Qk,
                                         int Ik, int Jk, int c0) {
    const int lane = threadIdx.x % WAVE, li = lane % 16, lg = lane / 16;
    v4d acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4zero();
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int kp = kb * 16 + lg + 4 * kk;
            const int g = slot_index(Ik, Jk, kp, n);
            double a[2], b[4];
#pragma unroll
            for (int bm = 0; bm < 2; ++bm) a[bm] = Qk[kp * JS + bm * 16 + li];
#pragma unroll
            for (int bn = 0; bn < 4; ++bn) {
                const int c = c0 + bn * 16 + li;
                b[bn] = (g >= 0 && c < ncol) ? X[(size_t)g * ld + c] : 0.0;
            }
#pragma unroll
            for (int bm = 0; bm < 2; ++bm)
#pragma unroll
                for (int bn = 0; bn < 4; ++bn) acc[bm][bn] = mfma16(a[bm], b[bn], acc[bm][bn]);
        }
#pragma unroll
    for (int bm = 0; bm < 2; ++bm)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int g = slot_index(Ik, Jk, bm * 16 + lg + 4 * r, n);
            if (g < 0) continue;
#pragma unroll
            for (int bn = 0; bn < 4; ++bn) {
                const int c = c0 + bn * 16 + li;
                if (c < ncol) X[(size_t)g * ld + c] = acc[bm][bn][r];
            }
        }
}

// A: full symmetric n x n (row-major, global); Ct: vectors (rows, global);
// Qbuf: global scratch >= (nblocks/2 + 1) * JS * JS doubles.  Returns sweeps
// or -1 when max_sweeps were spent.
template <class CAcc>
__device__ __noinline__ int wg_block_jacobi(glbf* A, int n, CAcc Ct, glbf* Qbuf, int max_sweeps, double tol,
                                            ldsf* lds, ldsf* red, double* tacc, int inner,
                                            int cb_lo, int cb_hi, bool intra) {
    const int mode = cb_lo < 0 ? 0 : (intra ? 2 : 1);
    // mode 2 skips pairs of pure blocks of different classes; every block then
    // still meets a block of its own class if each class spans two blocks
    const bool skip_cross = mode == 2 && cb_lo >= 2 * JB && n - cb_hi >= 2 * JB;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    ldsf* M = lds + w * JWAVE_LDS;
    ldsf* Q = M + JPK;
    ldsf* cs = Q + JS * JS;
    const int nb = (n + JB - 1) / JB;
    const int nb2 = nb + (nb & 1);
    const int npair = nb2 / 2;
    const int ntile = npair * (npair - 1) / 2;
    const int nch = (n + WAVE - 1) / WAVE;
    ldsi* ident = reinterpret_cast<ldsi*>(lds + NWAVE * JWAVE_LDS);  // per pair: Q == I
    ldsi* sched = ident + 128;                                        // subproblem schedules
    build_schedules(sched);
    // symmetrise from the upper triangle
    for (int i = w; i < n; i += NWAVE)
        for (int j = lane; j < i; j += WAVE) A[(size_t)i * n + j] = A[(size_t)j * n + i];
    __syncthreads();
    int sweep = 0;
    for (;;) {
        // (four rows per wave: the loads of four rows in flight)
        double off = 0.0;
        for (int i0 = w * 4; i0 < n; i0 += NWAVE * 4) {
            int ii[4], ci[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                ii[u] = i0 + u < n ? i0 + u : n - 1;
                ci[u] = occ_class(ii[u], cb_lo, cb_hi);
            }
            for (int j = i0 + 1 + lane; j < n; j += WAVE) {
                double a[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) a[u] = A[(size_t)ii[u] * n + j];
                const int cj = occ_class(j, cb_lo, cb_hi);
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (j > i0 + u && i0 + u < n && pair_in_mode(mode, ci[u], cj)) off = fmax(off, fabs(a[u]));
            }
        }
        off = wg_max(off, red);
        if (off <= tol) return sweep;
        if (sweep >= max_sweeps) return -1;
        for (int r = 0; r < nb2 - 1; ++r) {
            const double t0 = (double)wall_clock64() * 0.01;
            // 1. subproblems, one wavefront each (invalid slots stay decoupled:
            //    zero rows/columns are never rotated)
            for (int k = w; k < npair; k += NWAVE) {
                int I, J;
                rr_pair(r, k, nb2, I, J);
                unsigned lo_mask = 0u, hi_mask = 0u;
                if (mode != 0) {
                    // slot classes (slots 0..15 block I, 16..31 block J)
                    const int g = slot_index(I, J, lane % JS, n);
                    const unsigned vm = (unsigned)__ballot(lane < JS && g >= 0);
                    lo_mask = (unsigned)__ballot(lane < JS && g >= 0 && g < cb_lo);
                    hi_mask = (unsigned)__ballot(lane < JS && g >= 0 && g < cb_hi);
                    bool skip;
                    if (mode == 1) {
                        // a pair whose valid slots share one class has nothing to decouple
                        skip = (lo_mask == 0u || lo_mask == vm) && (hi_mask == 0u || hi_mask == vm);
                    } else {
                        // two non-empty pure blocks of different classes: no
                        // same-class pair spans them
                        const unsigned vI = vm & 0xFFFFu, vJ = vm >> 16;
                        const unsigned lI = lo_mask & 0xFFFFu, lJ = lo_mask >> 16;
                        const unsigned hI = hi_mask & 0xFFFFu, hJ = hi_mask >> 16;
                        const bool pI = (lI == 0u || lI == vI) && (hI == 0u || hI == vI);
                        const bool pJ = (lJ == 0u || lJ == vJ) && (hJ == 0u || hJ == vJ);
                        const int cI = (lI != 0u) + (hI != 0u), cJ = (lJ != 0u) + (hJ != 0u);
                        skip = skip_cross && vI != 0u && vJ != 0u && pI && pJ && cI != cJ;
                    }
                    if (skip) {
                        if (lane == 0) ident[k] = 1;
                        continue;
                    }
                }
                for (int e = lane; e < JS * JS; e += WAVE) {
                    const int i = e / JS, j = e % JS;
                    if (j >= i) {
                        const int gi = slot_index(I, J, i, n), gj = slot_index(I, J, j, n);
                        double v = 0.0;
                        if (gi >= 0 && gj >= 0)
                            v = A[(size_t)(gi < gj ? gi : gj) * n + (gi < gj ? gj : gi)];
                        pk(M, i, j) = v;
                    }
                }
                wave_sync();
                const int rot = wave_jacobi_sub(M, Q, cs, sched, tol, inner, mode, lo_mask, hi_mask);
                glbf* qg = Qbuf + (size_t)k * JS * JS;
                if (lane == 0) ident[k] = rot ? 0 : 1;
                if (rot) {
                    for (int e = lane; e < JS * JS; e += WAVE) {
                        qg[e] = Q[(e % JS) * JS + e / JS];  // Q[i][j] = Qt[j][i]
                        const int i = e / JS, j = e % JS;
                        const int gi = slot_index(I, J, i, n), gj = slot_index(I, J, j, n);
                        if (gi >= 0 && gj >= 0) A[(size_t)gi * n + gj] = pk(M, i, j);
                    }
                }
                wave_sync();
            }
            __syncthreads();
            const double t1 = (double)wall_clock64() * 0.01;
            // 2. off-diagonal tiles of A and the vectors, matrix cores
            const int nitem = ntile + npair * nch;
            for (int it = w; it < nitem; it += NWAVE) {
                if (it < ntile) {
                    // tile index -> (k, l), k < l
                    int k = 0, rem = it;
                    while (rem >= npair - 1 - k) { rem -= npair - 1 - k; ++k; }
                    const int l = k + 1 + rem;
                    if (ident[k] && ident[l]) continue;  // both rotations are I
                    int Ik, Jk, Il, Jl;
                    rr_pair(r, k, nb2, Ik, Jk);
                    rr_pair(r, l, nb2, Il, Jl);
                    tile_update(A, n, Qbuf + (size_t)k * JS * JS, Qbuf + (size_t)l * JS * JS, Ik, Jk,
                                Il, Jl, ident[k], ident[l]);
                } else {
                    const int v = it - ntile;
                    const int k = v / nch, ch = v % nch;
                    if (ident[k]) continue;
                    int Ik, Jk;
                    rr_pair(r, k, nb2, Ik, Jk);
                    rows_update(Ct.a, n, (size_t)Ct.ld, n, Qbuf + (size_t)k * JS * JS, Ik, Jk,
                                ch * WAVE);
                }
            }
            __syncthreads();
            const double t2 = (double)wall_clock64() * 0.01;
            tacc[4] += t1 - t0;
            tacc[5] += t2 - t1;
        }
        ++sweep;
    }
}

struct RowsGlobal {
    glbf* a;
    int ld;
    __device__ inline glbf& at(int k, int mu) const { return a[(size_t)k * ld + mu]; }
};

struct RowsLds {
    ldsf* a;
    int ld;
    __device__ inline ldsf& at(int k, int mu) const { return a[k * ld + mu]; }
};

// Ascending bitonic sort of (key, idx) pairs in LDS; len is a power of two
// and entries beyond the valid range must hold +inf keys.
__device__ inline void wg_bitonic_sort(ldsf* key, ldsi* idx, int len) {
    for (int size = 2; size <= len; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < len / 2; t += NT) {
                const int lo = 2 * t - (t & (stride - 1));
                const int hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const double a = key[lo], b = key[hi];
                if ((a > b) == up) {
                    key[lo] = b; key[hi] = a;
                    const int ti = idx[lo]; idx[lo] = idx[hi]; idx[hi] = ti;
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace xtb
