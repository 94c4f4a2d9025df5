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
__device__ inline double wg_sum(double v, double* red) {
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

__device__ inline double wg_max(double v, double* red) {
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
__device__ inline void wg_sum_n(double (&vals)[K], double* red) {
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

// ---------------------------------------------------------------------------
// GEMM: C = alpha * op(A) * op(B) + beta * C, all row-major in global memory.
// op(A)[m][k] = TA ? A[k*lda+m] : A[m*lda+k];  op(B)[k][n] likewise.
// The workgroup splits into NT/256 groups; each group computes 64x64 output
// tiles with a 4x4 register micro-tile per thread, staging 64x16 / 16x64
// operand panels in LDS.  `lds` must hold NGEMM_LDS doubles.
// ---------------------------------------------------------------------------
constexpr int GBM = 64, GBN = 64, GBK = 16;
constexpr int GPAD = GBM + 1;
constexpr int NGROUP = NT / 256;
constexpr int NGEMM_LDS = NGROUP * 2 * GBK * GPAD;

template <bool TA, bool TB>
__device__ __noinline__ void wg_gemm(int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
                        const double* __restrict__ B, int ldb, double beta, double* __restrict__ C,
                        int ldc, double* lds) {
    const int g = threadIdx.x / 256;
    const int t = threadIdx.x % 256;
    double* As = lds + g * 2 * GBK * GPAD;  // [GBK][GPAD]
    double* Bs = As + GBK * GPAD;           // [GBK][GPAD]
    const int tr = t / 16, tc = t % 16;
    const int tm = (M + GBM - 1) / GBM, tn = (N + GBN - 1) / GBN;
    const int ntiles = tm * tn;
    // all groups run the same number of tile iterations (barriers inside)
    const int iters = (ntiles + NGROUP - 1) / NGROUP;
    for (int it = 0; it < iters; ++it) {
        const int tile = it * NGROUP + g;
        const bool active = tile < ntiles;
        const int m0 = active ? (tile / tn) * GBM : 0;
        const int n0 = active ? (tile % tn) * GBN : 0;
        double acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
        for (int k0 = 0; k0 < K; k0 += GBK) {
            // stage op(A)[m0:m0+64][k0:k0+16] as As[k][m]
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int m, k;
                if (TA) { m = t % 64; k = t / 64 + 4 * r; }
                else { k = t % 16; m = t / 16 + 16 * r; }
                const int gm = m0 + m, gk = k0 + k;
                double val = 0.0;
                if (active && gm < M && gk < K) val = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
                As[k * GPAD + m] = val;
            }
            // stage op(B)[k0:k0+16][n0:n0+64] as Bs[k][n]
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int n, k;
                if (TB) { k = t % 16; n = t / 16 + 16 * r; }
                else { n = t % 64; k = t / 64 + 4 * r; }
                const int gn = n0 + n, gk = k0 + k;
                double val = 0.0;
                if (active && gn < N && gk < K) val = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
                Bs[k * GPAD + n] = val;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < GBK; ++k) {
                double a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = As[k * GPAD + tr * 4 + i];
#pragma unroll
                for (int j = 0; j < 4; ++j) b[j] = Bs[k * GPAD + tc * 4 + j];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
            }
            __syncthreads();
        }
        if (active) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int gm = m0 + tr * 4 + i;
                if (gm >= M) continue;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int gn = n0 + tc * 4 + j;
                    if (gn >= N) continue;
                    double* c = C + (size_t)gm * ldc + gn;
                    *c = (beta == 0.0) ? alpha * acc[i][j] : fma(alpha, acc[i][j], beta * *c);
                }
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Symmetric matrix storage used by the Jacobi eigensolver.  Only the upper
// triangle (i <= j) is referenced.  Packed form lives in LDS, full form in
// global memory.
// ---------------------------------------------------------------------------
struct SymPacked {
    double* a;
    int n;
    __device__ inline double& at(int i, int j) const {
        if (i > j) { int t = i; i = j; j = t; }
        return a[i * n - (i * (i - 1)) / 2 + (j - i)];
    }
};

struct SymFull {
    double* a;
    int n;
    __device__ inline double& at(int i, int j) const {
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
// `cs` is LDS for 2*(n/2+1) doubles.  Returns the number of sweeps, or
// -1 when max_sweeps were spent without reaching `tol` (absolute).
template <class Sym, class CAcc>
__device__ int wg_jacobi(Sym A, int n, CAcc Ct, int max_sweeps, double tol, double* cs,
                         double* red) {
    const int n2 = n + (n & 1);
    const int npair = n2 / 2;
    double* cc = cs;
    double* ss = cs + npair;
    int sweep = 0;
    for (;;) {
        // convergence test: largest off-diagonal element
        double off = 0.0;
        for (int idx = threadIdx.x; idx < n * n; idx += NT) {
            const int i = idx / n, j = idx % n;
            if (j > i) off = fmax(off, fabs(A.at(i, j)));
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
                if (p < n && q < n) {
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
            }
            __syncthreads();
            // 2a. A <- J^T A J on 2x2 blocks (P <= R)
            const int nblk = npair * npair;
            for (int idx = threadIdx.x; idx < nblk; idx += NT) {
                const int P = idx / npair, R = idx % npair;
                if (R < P) continue;
                const double sp = ss[P], sr = ss[R];
                if (sp == 0.0 && sr == 0.0) continue;
                int p, q, x, y;
                rr_pair(r, P, n2, p, q);
                rr_pair(r, R, n2, x, y);
                if (p >= n || q >= n || x >= n || y >= n) continue;
                const double cp = cc[P], cr = cc[R];
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
            // 2b. vectors: rows p, q of Ct
            const int nvec = npair * n;
            for (int idx = threadIdx.x; idx < nvec; idx += NT) {
                const int k = idx / n, mu = idx % n;
                const double s = ss[k];
                if (s == 0.0) continue;
                int p, q;
                rr_pair(r, k, n2, p, q);
                if (p >= n || q >= n) continue;
                const double c = cc[k];
                double& vp = Ct.at(p, mu);
                double& vq = Ct.at(q, mu);
                const double a = vp, b = vq;
                vp = c * a - s * b;
                vq = s * a + c * b;
            }
            __syncthreads();
        }
        ++sweep;
    }
}

struct RowsGlobal {
    double* a;
    int ld;
    __device__ inline double& at(int k, int mu) const { return a[(size_t)k * ld + mu]; }
};

struct RowsLds {
    double* a;
    int ld;
    __device__ inline double& at(int k, int mu) const { return a[k * ld + mu]; }
};

// Ascending bitonic sort of (key, idx) pairs in LDS; len is a power of two
// and entries beyond the valid range must hold +inf keys.
__device__ inline void wg_bitonic_sort(double* key, int* idx, int len) {
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
