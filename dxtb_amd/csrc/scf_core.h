// Per-molecule SCF pieces shared by the one-workgroup-per-molecule kernel
// (scf.hip) and the whole-GPU path for one large molecule (large.hip): the
// molecule view, the LDS view, Fermi smearing and the Anderson mixer.  Every
// function here is called by all threads of one workgroup.
#pragma once

#include "common.h"
#include "wg.h"

namespace xtb {

struct Mol {
    int m, n, nao, nsh, nat, vec0, sh0, at0;
    size_t mat0, gam0;
    const glbf *S, *H0, *gam;
    glbf *F, *Ct, *T, *P;
    glbf *pot, *vnew, *q, *eorb, *emo, *occ, *qsh_out;
    double nel_a, nel_b;
    int nvec_total;
};

struct Shared {
    ldsf *red, *cs, *skey, *ev, *ftot, *qsh, *vsh, *gq, *qat, *big;
    ldsi* sidx;
    int npow, big_cap;
    double* tacc;  // per-thread stage timers (meaningful on thread 0)
#ifdef XTB_PD_TIMERS
    double* ptacc;  // phases of the perturbative step (instrumented builds only)
#endif
};

#ifdef XTB_PD_TIMERS
#define PD_MARK(sh, i, t) do { const double t_ = now_us(); (sh).ptacc[i] += t_ - (t); (t) = t_; } while (0)
#else
#define PD_MARK(sh, i, t) do { } while (0)
#endif

// device wall clock (constant 100 MHz) in microseconds
__device__ inline double now_us() { return (double)wall_clock64() * 0.01; }

// Passes of the perturbative occupied-virtual rotation.  The amplitudes
// divide by the diagonals of the occupied and virtual blocks only, and those
// blocks drift away from diagonal over an SCF run, so the passes converge
// linearly: the largest coupling falls by a factor 0.2 per pass (median;
// 0.35-0.45 at the 90th percentile, 2400 molecules of the serving stream),
// not quadratically.  An inexact SCF step stops after two passes, or once
// the second-order estimate max|K| max|A_ov| is below its tolerance.  The
// final step stops only on a measured coupling below
// max(jacobi_tol max(|e|, 1), PD_FINAL_REL x_atol_max): a coupling d turns
// the orbitals by ~d / gap (gaps >= 60 kT = 0.057 Eh here), so the charges
// and forces move by < 40 d, at most what the SCF's own convergence
// threshold on the potential leaves.  At x_atol_max = 1e-10, 79 % of the
// 2400 molecules meet 1e-11 without a rotation and all within 6
// measurements; 1e-2 x_atol_max needed up to 9 and more, and made that
// launch 17 % longer.  After PD_EXACT_PASSES measurements the step falls
// back to a full eigensolve, which costs as much as ~15 passes on a
// 500-orbital molecule.
constexpr int PD_INEXACT_PASSES = 2;
constexpr int PD_EXACT_PASSES = 7;
constexpr double PD_FINAL_REL = 1e-1;

// ---------------------------------------------------------------------------
// Diagonal blocks of the blocked Cholesky orthogonaliser (one wave, 32 x 32
// in LDS, row-major).  chol_block_factor: D <- its lower Cholesky factor in
// place (the upper triangle is left stale); returns false when a pivot is
// not positive (the pivot is then taken as 1 and the factor is meaningless).
// ---------------------------------------------------------------------------
__device__ inline bool chol_block_factor(ldsf* D, int lane) {
    constexpr int NB = 32;
    bool ok = true;
    for (int c = 0; c < NB; ++c) {
        const double d = D[c * NB + c];
        wave_sync();
        const double sd = d > 0.0 ? sqrt(d) : 1.0;
        if (!(d > 0.0)) ok = false;
        if (lane == 0) D[c * NB + c] = sd;
        for (int r = c + 1 + lane; r < NB; r += WAVE) D[r * NB + c] /= sd;
        wave_sync();
        const int m = NB - c - 1;
        for (int e = lane; e < m * m; e += WAVE) {
            const int rr = c + 1 + e / m, qq = c + 1 + e % m;
            if (qq <= rr) D[rr * NB + qq] -= D[rr * NB + c] * D[qq * NB + c];
        }
        wave_sync();
    }
    return ok;
}

__device__ inline void chol_block_inverse(ldsf* D, ldsf* Di, glbf* Dg, int nb, int lane) {
    // lower-triangular D (NB x NB, LDS) -> Di = D^-1 (LDS) and its leading
    // nb x nb block to Dg (global, leading dimension nb: it must fit in an
    // n x n buffer when n < NB); lane j forms column j by forward substitution
    constexpr int NB = 32;
    if (lane < NB) {
        const int j = lane;
        for (int i = 0; i < NB; ++i) {
            double v;
            if (i < j) v = 0.0;
            else if (i == j) v = 1.0 / D[j * NB + j];
            else {
                double acc = 0.0;
                for (int k = j; k < i; ++k) acc = fma(D[i * NB + k], Di[k * NB + j], acc);
                v = -acc / D[i * NB + i];
            }
            Di[i * NB + j] = v;
            if (i < nb && j < nb) Dg[i * nb + j] = v;
        }
    }
}

// electronic temperatures below this (0.1 K in Hartree) fill the orbitals
// by the aufbau principle instead of Fermi smearing
// (src/dxtb/_src/scf/base.py:888-889)
constexpr double KT_AUFBAU = 3e-7;

// ---------------------------------------------------------------------------
// Fermi smearing of one spin channel (src/dxtb/_src/wavefunction/
// filling.py:201-366): mid-gap start, Newton on the rounded electron count
// (scf/base.py:879), exponent cutoff 50, threshold on |N - sum f|.  Below
// KT_AUFBAU the aufbau occupations of the unrounded count instead
// (filling.py:116-198; base.py:888-889 keeps the initial filling of
// iterator.py:147-189): orbital j (from 0) is full while j + 1 < nel, the
// next one takes nel - ceil(nel) + 1.  Sorted eigenvalues in sh.skey.
// Writes the sorted occupations to occ_out and adds them to sh.ftot.
// ---------------------------------------------------------------------------
template <class SH>
__device__ __noinline__ bool fermi_channel(const Mol& M, const SH& sh, double nel, double kt, double thr,
                              int maxiter, bool padded, glbf* occ_out, double* efermi) {
    const int n = M.n;
    const double eps = 2.220446049250313e-16;
    if (fabs(nel) < eps) {
        for (int j = threadIdx.x; j < n; j += NT) occ_out[j] = 0.0;
        *efermi = 0.0;
        __syncthreads();
        return true;
    }
    if (kt < KT_AUFBAU) {
        const double frac = nel - ceil(nel) + 1.0;
        for (int j = threadIdx.x; j < n; j += NT) {
            const double idx = (double)(j + 1);
            const double f = idx < nel ? 1.0 : (idx < nel + 1.0 ? frac : 0.0);
            occ_out[j] = f;
            sh.ftot[sh.sidx[j]] += f;
        }
        *efermi = 0.0;
        __syncthreads();
        return true;
    }
    nel = rint(nel);
    int homo = (int)ceil(nel - 5e-15) - 1;
    if (homo < 0) homo = 0;
    if (homo > n - 1) homo = n - 1;
    const bool lumo_missing = (n - 1 <= homo);
    const int lumo = lumo_missing ? homo : homo + 1;
    double ef = 0.5 * (sh.skey[homo] + sh.skey[lumo]);
    if (padded && (homo >= M.nao || lumo >= M.nao)) ef *= 2.0;  // filling.py:254-264
    const double target = (double)(homo + 1);
    for (int it = 0; it < maxiter; ++it) {
        double acc[2] = {0.0, 0.0};
        for (int j = threadIdx.x; j < n; j += NT) {
            const double x = (sh.skey[j] - ef) / kt;
            if (x < 50.0) {
                const double et = exp(x);
                acc[0] += 1.0 / (et + 1.0);
                acc[1] += et / (kt * (et + 1.0) * (et + 1.0));
            } else {
                acc[1] += eps;
            }
        }
        wg_sum_n<2>(acc, sh.red);
        const double diff = target - acc[0];
        if (fabs(diff) <= thr) {
            for (int j = threadIdx.x; j < n; j += NT) {
                const double x = (sh.skey[j] - ef) / kt;
                const double f = (x < 50.0) ? 1.0 / (exp(x) + 1.0) : 0.0;
                occ_out[j] = f;
                sh.ftot[sh.sidx[j]] += f;
            }
            *efermi = ef;
            __syncthreads();
            return true;
        }
        ef += diff / acc[1];
    }
    for (int j = threadIdx.x; j < n; j += NT) occ_out[j] = 0.0;
    *efermi = ef;
    __syncthreads();
    return false;
}

// ---------------------------------------------------------------------------
// Anderson mixing with soft start (scf/mixer/anderson.py:163-317), history in
// a ring buffer of G = generations+1 slots.  Returns the convergence flag of
// the step (mixer/base.py:229-256 on the step's residual).
// ---------------------------------------------------------------------------
// largest history (generations) the mixer takes; its Gram system's 27 sums
// fit one pass of the reduction scratch, and 8 generations (44 sums) grow
// the kernel's stack by ~0.9 KB per lane for ~1 % fewer iterations
constexpr int MIX_MAXG = 6;

struct Mixer {
    int step, head, G;
    bool soft;  // upstream's soft start; a warm start, already near the fixed point, skips it
};

template <class SH>
__device__ __noinline__ bool anderson(Mixer& mx, const Mol& M, const SH& sh, const xtb_scf_opts& o,
                         glbf* hx, glbf* hf, const glbf* xnew, const glbf* xold,
                         glbf* xout, double* resid_inf) {
    const int n = M.n, G = mx.G;
    auto X = [&](int h) { return hx + (size_t)((mx.head + h) % G) * M.nvec_total + M.vec0; };
    auto Fh = [&](int h) { return hf + (size_t)((mx.head + h) % G) * M.nvec_total + M.vec0; };
    if (mx.step == 0) {
        glbf* x0 = X(0);
        for (int i = threadIdx.x; i < n; i += NT) x0[i] = xold[i];
        __syncthreads();
    }
    mx.step += 1;
    const int gen = G - 1;
    glbf* x0 = X(0);
    glbf* f0 = Fh(0);
    double l2 = 0.0, linf = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) {
        const double f = xnew[i] - x0[i];
        f0[i] = f;
        l2 = fma(f, f, l2);
        linf = fmax(linf, fabs(f));
    }
    l2 = wg_sum(l2, sh.red);
    linf = wg_max(linf, sh.red);
    const bool conv = (sqrt(l2) < o.x_atol) && (linf < o.x_atol_max);
    *resid_inf = linf;
    double theta[MIX_MAXG];
#pragma unroll
    for (int h = 0; h < MIX_MAXG; ++h) theta[h] = 0.0;
    int nn = 0;
    if (mx.step > gen || !mx.soft) nn = (mx.step - 1 < gen) ? mx.step - 1 : gen;  // soft start
    if (nn > 0) {
        // a_hj = (f0 - f_h).(f0 - f_j), b_h = (f0 - f_h).f0 ; h, j = 1..nn.
        // One pass per h over the elements with at most MIX_MAXG + 1 sums
        // live; each sum is reduced at once (wave butterfly, then the waves'
        // partials in LDS added in wave order), so no per-thread array of all
        // NACC sums stays live (it spilled to scratch on every call)
        constexpr int NA = MIX_MAXG * (MIX_MAXG + 1) / 2, NACC = NA + MIX_MAXG;
        static_assert(NACC * NWAVE + NACC <= 256, "reduction scratch");
        const int wv = threadIdx.x / WAVE, ln = threadIdx.x % WAVE;
        ldsf* part = sh.red;                 // [NACC][NWAVE]
        ldsf* tot = sh.red + NACC * NWAVE;   // [NACC]
        int k0 = 0;
#pragma unroll
        for (int h = 0; h < MIX_MAXG; ++h) {
            double sj[MIX_MAXG + 1];
#pragma unroll
            for (int j = 0; j <= MIX_MAXG; ++j) sj[j] = 0.0;
            if (h < nn) {
                for (int i = threadIdx.x; i < n; i += NT) {
                    const double fi = f0[i];
                    const double dh = fi - Fh(h + 1)[i];
#pragma unroll
                    for (int j = h; j < MIX_MAXG; ++j)
                        if (j < nn) sj[j] = fma(dh, fi - Fh(j + 1)[i], sj[j]);
                    sj[MIX_MAXG] = fma(dh, fi, sj[MIX_MAXG]);
                }
            }
#pragma unroll
            for (int j = h; j <= MIX_MAXG; ++j) {
                const double v = wave_sum(sj[j]);
                const int e = j < MIX_MAXG ? k0 + (j - h) : NA + h;
                if (ln == 0) part[e * NWAVE + wv] = v;
            }
            k0 += MIX_MAXG - h;
        }
        __syncthreads();
        if (threadIdx.x < NACC) {
            double v = 0.0;
            for (int x = 0; x < NWAVE; ++x) v += part[threadIdx.x * NWAVE + x];
            tot[threadIdx.x] = v;
        }
        __syncthreads();
        // solve (a * diag(1 + off^2)) theta = b by Gaussian elimination with
        // partial pivoting; one thread, system held in LDS (over the partials)
        constexpr int W = MIX_MAXG + 1;
        ldsf* a = sh.red;                  // [MIX_MAXG][W]
        ldsf* th = sh.red + MIX_MAXG * W;  // [MIX_MAXG]
        static_assert(MIX_MAXG * W + MIX_MAXG <= NACC * NWAVE, "solver rows below the totals");
        if (threadIdx.x == 0) {
            int k = 0;
            for (int h = 0; h < MIX_MAXG; ++h)
                for (int j = h; j < MIX_MAXG; ++j) { a[h * W + j] = tot[k]; a[j * W + h] = tot[k]; ++k; }
            for (int h = 0; h < MIX_MAXG; ++h) a[h * W + MIX_MAXG] = tot[NA + h];
            for (int h = 0; h < nn; ++h) a[h * W + h] *= 1.0 + o.diag_offset * o.diag_offset;
            for (int c = 0; c < nn; ++c) {
                int piv = c;
                for (int r = c + 1; r < nn; ++r)
                    if (fabs(a[r * W + c]) > fabs(a[piv * W + c])) piv = r;
                if (piv != c)
                    for (int x = 0; x < W; ++x) {
                        const double t = a[c * W + x];
                        a[c * W + x] = a[piv * W + x];
                        a[piv * W + x] = t;
                    }
                for (int r = c + 1; r < nn; ++r) {
                    const double fct = a[r * W + c] / a[c * W + c];
                    for (int x = c; x < W; ++x) a[r * W + x] -= fct * a[c * W + x];
                }
            }
            for (int r = nn - 1; r >= 0; --r) {
                double sacc = a[r * W + MIX_MAXG];
                for (int x = r + 1; x < nn; ++x) sacc -= a[r * W + x] * th[x];
                th[r] = sacc / a[r * W + r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < MIX_MAXG; ++h) theta[h] = (h < nn) ? th[h] : 0.0;
        __syncthreads();
    }
    // new slot for the mixed vector = oldest slot
    const int newhead = (mx.head - 1 + G) % G;
    glbf* xn = hx + (size_t)newhead * M.nvec_total + M.vec0;
    for (int i = threadIdx.x; i < n; i += NT) {
        double xm;
        if (nn > 0) {
            double xb = x0[i], fb = f0[i];
#pragma unroll
            for (int h = 0; h < MIX_MAXG; ++h) {
                if (h < nn) {
                    const double xh = X(h + 1)[i], fh = Fh(h + 1)[i];
                    xb += theta[h] * (xh - x0[i]);
                    fb -= theta[h] * (f0[i] - fh);
                }
            }
            xm = xb + o.damp * fb;
        } else {
            xm = x0[i] + (mx.soft ? o.damp_init : o.damp) * f0[i];
        }
        xout[i] = xm;
        xn[i] = xm;  // slot newhead may alias X(G-1): element-wise read-before-write
    }
    __syncthreads();
    mx.head = newhead;
    return conv;
}

// ---------------------------------------------------------------------------
// Limited-memory Broyden root search ("broyden1", good Broyden) on
// g(x) = x - f(x), as upstream's implicit SCF modes run it
// (scf/implicit/default.py, scf/pure/iterator.py:125-147 -> xitorch
// equilibrium, optimize/root/rootsolver.py:42-256 with BroydenFirst,
// _jacobian.py:50-112): inverse Jacobian H = h0 I + sum_k c_k d_k^T with
// h0 = -alpha = 0.5, no line search, x <- x + damp dx with dx = -H g.  After
// each map evaluation (x_prev, g_prev -> x, g): dx = x - x_prev,
// dy = g - g_prev, v = H^T dx, c = dx - H dy, d = v / (dy.v), H += c d^T.
// Converged (TerminationCondition, rootsolver.py:471-492) when the last
// step's |dx| < x_atol and |g| < f_atol.  Upstream keeps every update; here
// at most R = generations (<= BROYDEN_MAXR) are kept and the memory restarts
// from h0 I when full (upstream's "restart" reduction at max_rank = R).
// Storage: c_k in hx slot k, d_k in hf slot k (k < R); x_prev, g_prev in
// slot R.  The first call only records x_prev and g_prev.
// ---------------------------------------------------------------------------
constexpr int BROYDEN_MAXR = 24;

struct Broyden {
    int step, rank, R;
    double dxnorm;  // |dx| of the last step (before damping)
};

// Projections sum_i U_k[i] w_i for k < r, reduced in chunks of 8 and left
// in LDS at red[128 + k] (red holds 256 doubles; wg_sum_n<8> uses the first
// 64).  U_k(i) and w(i) are callables.
template <class SH, class UK, class W>
__device__ inline void wg_projections(int r, int n, const SH& sh, UK U, W w, int out0) {
    for (int k0 = 0; k0 < r; k0 += 8) {
        double t[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int i = threadIdx.x; i < n; i += NT) {
            const double wi = w(i);
#pragma unroll
            for (int kk = 0; kk < 8; ++kk)
                if (k0 + kk < r) t[kk] = fma(U(k0 + kk, i), wi, t[kk]);
        }
        wg_sum_n<8>(t, sh.red);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int kk = 0; kk < 8; ++kk)
                if (k0 + kk < r) sh.red[128 + out0 + k0 + kk] = t[kk];
        }
        __syncthreads();
    }
}

template <class SH>
__device__ __noinline__ bool broyden(Broyden& bm, const Mol& M, const SH& sh, const xtb_scf_opts& o,
                                     glbf* hx, glbf* hf, const glbf* fx, const glbf* x, glbf* xout,
                                     double* resid_inf) {
    static_assert(128 + 2 * BROYDEN_MAXR <= 256, "projections fit the reduction scratch");
    const int n = M.n, R = bm.R;
    constexpr double h0 = 0.5;  // -alpha, alpha = -0.5 (scf/pure/iterator.py:132)
    auto C = [&](int k) { return hx + (size_t)k * M.nvec_total + M.vec0; };
    auto D = [&](int k) { return hf + (size_t)k * M.nvec_total + M.vec0; };
    glbf* xp = C(R);
    glbf* gp = D(R);
    ldsf* proj = sh.red + 128;  // [0, R): first set, [R, 2R): second set
    double g2 = 0.0, ginf = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) {
        const double g = x[i] - fx[i];
        g2 = fma(g, g, g2);
        ginf = fmax(ginf, fabs(g));
    }
    g2 = wg_sum(g2, sh.red);
    ginf = wg_max(ginf, sh.red);
    *resid_inf = ginf;
    if (bm.step > 0) {
        if (bm.dxnorm < o.x_atol && sqrt(g2) < o.f_atol) {
            // the root search stops before it takes x: its result is the
            // previous input (rootsolver.py:200-215 breaks before x = xnew)
            for (int i = threadIdx.x; i < n; i += NT) xout[i] = xp[i];
            __syncthreads();
            bm.step += 1;
            return true;
        }
        // c_k.dx and d_k.dy over the kept pairs; the new pair goes to slot rr
        // (the memory restarts when full)
        const int r = bm.rank;
        const int rr = r >= R ? 0 : r;
        wg_projections(rr, n, sh, [&](int k, int i) { return C(k)[i]; },
                       [&](int i) { return x[i] - xp[i]; }, 0);
        wg_projections(rr, n, sh, [&](int k, int i) { return D(k)[i]; },
                       [&](int i) { return (x[i] - fx[i]) - gp[i]; }, R);
        glbf* cn = C(rr);
        glbf* dn = D(rr);
        double dyv = 0.0;
        for (int i = threadIdx.x; i < n; i += NT) {
            const double dx = x[i] - xp[i], dy = (x[i] - fx[i]) - gp[i];
            double v = h0 * dx, hdy = h0 * dy;
            for (int k = 0; k < rr; ++k) {
                v = fma(D(k)[i], proj[k], v);
                hdy = fma(C(k)[i], proj[R + k], hdy);
            }
            cn[i] = dx - hdy;
            dn[i] = v;
            dyv = fma(dy, v, dyv);
        }
        dyv = wg_sum(dyv, sh.red);
        const double sc = 1.0 / dyv;
        for (int i = threadIdx.x; i < n; i += NT) dn[i] *= sc;
        __syncthreads();
        bm.rank = rr + 1;
    } else {
        bm.rank = 0;
    }
    // dx = -H g with the updated memory
    const int rk = bm.rank;
    wg_projections(rk, n, sh, [&](int k, int i) { return D(k)[i]; }, [&](int i) { return x[i] - fx[i]; }, 0);
    double dn2 = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) {
        const double xi = x[i], g = xi - fx[i];
        double hg = h0 * g;
        for (int k = 0; k < rk; ++k) hg = fma(C(k)[i], proj[k], hg);
        // x_prev, g_prev <- x, g
        xp[i] = xi;
        gp[i] = g;
        dn2 = fma(hg, hg, dn2);
        xout[i] = xi - o.damp * hg;  // xout may alias x: element-wise read-before-write
    }
    bm.dxnorm = sqrt(wg_sum(dn2, sh.red));
    bm.step += 1;
    return false;
}
}  // namespace xtb
