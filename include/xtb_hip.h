/*
 * xtb_hip.h -- C ABI of the MI355X (gfx950) GFN1-xTB single-point library.
 *
 * The library computes, for a batch of molecules stored in a ragged layout,
 * the quantities upstream dxtb computes on its GFN1 single-point path:
 *
 *   xtb_build_s_h0   overlap S, core Hamiltonian H0, coordination numbers
 *                    replaces Integrals.build_overlap / build_hcore
 *                    (src/dxtb/_src/integral/container.py:178-246,
 *                     src/dxtb/_src/xtb/base.py:251-360) and cn_d3
 *                    (called at src/dxtb/_src/xtb/gfn1.py:57-64)
 *   xtb_build_gamma  shell-resolved ES2 Coulomb matrix and ES3 hardness
 *                    derivatives; replaces ES2.get_cache / ES3.get_cache
 *                    (src/dxtb/_src/components/interactions/coulomb/
 *                     secondorder.py:216-276,799-870; thirdorder.py:182-250)
 *   xtb_build_s_h0_gamma  both of the above in one pass (one launch fewer)
 *   xtb_scf          persistent per-molecule SCF (Fock build, generalized
 *                    eigensolve, Fermi smearing, density, Mulliken charges,
 *                    ES2/ES3 potential, Anderson mixing, energies); replaces
 *                    scf.solve -> SelfConsistentFieldFull
 *                    (src/dxtb/_src/scf/iterator.py:51-144,
 *                     src/dxtb/_src/scf/unrolling/default.py:71-401)
 *   xtb_repulsion    classical repulsion energy per molecule; replaces
 *                    Repulsion.get_energy
 *                    (src/dxtb/_src/components/classicals/repulsion/
 *                     base.py:269-334, rep.py:65-98)
 *   xtb_forces       analytic nuclear gradient dE/dR; replaces
 *                    AnalyticalCalculator.forces_analytical
 *                    (src/dxtb/_src/calculators/types/analytical.py:61-222)
 *                    with GFN1Hamiltonian.get_gradient
 *                    (src/dxtb/_src/xtb/gfn1.py:185-408)
 *   xtb_param_grad   dE/d(Hamiltonian, Coulomb, repulsion parameters);
 *                    replaces autograd with differentiable ParamModule
 *                    entries (src/dxtb/_src/param/module/utils.py:123-240)
 *
 * Conventions
 *  - All array pointers are DEVICE pointers owned by the caller; the library
 *    allocates no device memory.  Sizes of the work buffers follow from the
 *    batch descriptor (see xtb_workspace_doubles).
 *  - Every call is stream ordered on `stream` (a hipStream_t passed as
 *    void*), returns an int32 status (XTB_OK or an XTB_E* code) and never
 *    throws.  Per-molecule outcomes (SCF convergence, Fermi failure,
 *    non-finite values) are written to device memory (xtb_scf_state.out).
 *  - Units: bohr, Hartree.  Arithmetic is IEEE FP64 throughout.
 *  - Matrices are row-major, one n x n block per molecule at offset
 *    mol_mat0[b] (in doubles) of a packed buffer; n = mol_n[b] is the
 *    matrix dimension (the molecule's own orbital count in the ragged
 *    layout, the batch maximum in the padded layout), of which the first
 *    mol_nao[b] orbitals are real.
 */
#ifndef XTB_HIP_H
#define XTB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XTB_ABI_VERSION 1

/* call status */
#define XTB_OK 0
#define XTB_EINVAL 1      /* bad argument or descriptor */
#define XTB_EHIP 2        /* a HIP runtime call failed */
#define XTB_ETOOBIG 3     /* a size exceeds a kernel limit */

/* per-molecule status word (xtb_scf_state.out[b*XTB_OUT_STRIDE+XTB_OUT_STATUS]) */
#define XTB_MOL_CONVERGED 0
#define XTB_MOL_MAXITER 1   /* SCF not converged within maxiter */
#define XTB_MOL_FERMI 2     /* Fermi level failed to converge */
#define XTB_MOL_NONFINITE 4 /* NaN/Inf met */
#define XTB_MOL_JACOBI 8    /* the final step's eigensolve hit its sweep limit */
#define XTB_MOL_OVERLAP 16  /* overlap matrix not positive definite: no SCF was run
                               (upstream's Cholesky reduction raises there) */

/* layout of the per-molecule result record */
#define XTB_OUT_STRIDE 32
#define XTB_OUT_ESCF 0      /* E_h0 + E_es2 + E_es3 */
#define XTB_OUT_EH0 1
#define XTB_OUT_EES2 2
#define XTB_OUT_EES3 3
#define XTB_OUT_FENERGY 4   /* electronic free energy -TS */
#define XTB_OUT_ITER 5      /* SCF iterations (as counted upstream) */
#define XTB_OUT_STATUS 6
#define XTB_OUT_EFERMI_A 7
#define XTB_OUT_EFERMI_B 8
#define XTB_OUT_EREP 9      /* written by xtb_repulsion */
#define XTB_OUT_SWEEPS 10   /* total Jacobi sweeps spent */
/* stage timers of the SCF kernel, microseconds (device wall clock) */
#define XTB_OUT_T_ORTHO 11  /* orthogonaliser (Cholesky L^-1 or Loewdin S^-1/2) */
#define XTB_OUT_T_GEMM 12   /* Fock transform, density and W products */
#define XTB_OUT_T_EIGEN 13  /* eigensolves / decoupling of the SCF steps */
#define XTB_OUT_T_TOTAL 14  /* whole molecule */
#define XTB_OUT_T_BSUB 15   /* block Jacobi: subproblem solves */
#define XTB_OUT_T_BROW 16   /* block Jacobi: row-panel updates (A and vectors) */
#define XTB_OUT_PD_STEPS 17 /* SCF steps decoupled by the perturbative rotation */
#define XTB_OUT_SWEEPS_FIRST 18  /* Jacobi sweeps of the first SCF step */
#define XTB_OUT_SWEEPS_LAST 19   /* Jacobi sweeps of the final (exact) step */
#define XTB_OUT_T_START 20  /* device wall clock at the molecule's start, microseconds */
#define XTB_OUT_T_PD 21     /* perturbative decoupling steps (part of T_EIGEN) */
#define XTB_OUT_XCD 22      /* accelerator complex die (XCD) the molecule's workgroup ran on */
#define XTB_OUT_HW_ID 23    /* the workgroup's HW_ID register (CU, shader array, engine) */
#define XTB_OUT_T_FOCK 24   /* Fock builds F = H0 - S o (v_i + v_j)/2, microseconds */
#define XTB_OUT_N_FOCK 25   /* Fock builds done (each reads H0 and S, writes F) */
#define XTB_OUT_T_MULL 26   /* Mulliken charges + ES2/ES3 potential, microseconds */
#define XTB_OUT_N_MULL 27   /* P, S (and H0) reads of those passes in units of
                               16 n^2 bytes: 1.5 for the last step (orbital
                               energies too); the other steps take the charges
                               from the density product and count 0 */
#define XTB_OUT_T_INTEGRALS 28  /* xtb_singlepoint: the molecule's integral pass, microseconds */
#define XTB_OUT_T_FORCES 29     /* xtb_singlepoint: the molecule's gradient, microseconds */

/* Ragged batch descriptor.  All pointers are device pointers. */
typedef struct xtb_batch {
    int32_t nmol;
    int32_t nat_total;   /* atoms over the batch */
    int32_t nsh_total;   /* shells over the batch */
    int32_t nvec_total;  /* sum of mol_n */
    int32_t max_n;       /* max of mol_n */
    int32_t max_nsh;     /* max of mol_nsh */
    int32_t max_nat;     /* max of mol_nat */
    int32_t ntile_sh;    /* entries of the shell-pair tile list */
    int32_t ntile_at;    /* entries of the atom-pair tile list */
    int32_t has_padding; /* 1 if some molecule has mol_n > mol_nao (padded layout) */
    int64_t nmat_total;  /* doubles of one packed n x n buffer */
    int64_t ngam_total;  /* doubles of the packed nsh x nsh buffer */
    /* per molecule [nmol] */
    const int32_t* mol_nat;
    const int32_t* mol_nsh;
    const int32_t* mol_nao;
    const int32_t* mol_n;
    const int32_t* mol_atom0;
    const int32_t* mol_sh0;
    const int32_t* mol_vec0;
    const int64_t* mol_mat0;
    const int64_t* mol_gam0;
    const double* mol_nel;   /* [2*nmol] alpha, beta electrons, not rounded: Fermi
                                smearing rounds them, the aufbau filling used
                                below kT = 3e-7 Hartree takes them as they are */
    /* per atom [nat_total] */
    const int32_t* atom_sp;  /* species index into xtb_params */
    const int32_t* atom_mol;
    const int32_t* atom_sh0; /* first shell of the atom, local to the molecule */
    const int32_t* atom_nsh; /* shells on the atom */
    /* per shell [nsh_total] */
    const int32_t* sh_atom;  /* atom index local to the molecule */
    const int32_t* sh_us;    /* unique-shell index into xtb_params */
    const int32_t* sh_ao;    /* first orbital, local to the molecule */
    /* per orbital slot [nvec_total] */
    const int32_t* ao_sh;    /* shell local to the molecule, -1 = padding */
    const double* n0;        /* reference occupation */
    /* 16x16 tiles of the shell-pair and atom-pair spaces (upper triangle) */
    const int32_t* tile_sh;  /* [3*ntile_sh]: molecule, i0, j0 */
    const int32_t* tile_at;  /* [3*ntile_at]: molecule, a0, b0 */
} xtb_batch;

/* GFN1 parameters, reduced to the species/shells present in the batch. */
typedef struct xtb_params {
    int32_t nsp;             /* species */
    int32_t nus;             /* unique shells */
    const double* sp_rad;    /* [nsp] atomic radius of the H0 distance polynomial, bohr */
    const double* sp_rcov;   /* [nsp] D3 covalent radius, bohr */
    const double* sp_gam3;   /* [nsp] third-order Hubbard derivative */
    const double* sp_zeff;   /* [nsp] repulsion effective charge */
    const double* sp_arep;   /* [nsp] repulsion exponent */
    const double* us_level;  /* [nus] self-energy, Hartree */
    const double* us_kcn;    /* [nus] CN shift of the self-energy, Hartree */
    const double* us_shpoly; /* [nus] distance-polynomial coefficient */
    const double* us_hard;   /* [nus] shell hardness lgam*gam */
    const int32_t* us_l;     /* [nus] angular momentum (0 or 1) */
    const int32_t* us_nprim; /* [nus] primitives (<= 8) */
    const double* us_alpha;  /* [nus*8] primitive exponents */
    const double* us_coeff;  /* [nus*8] normalised contraction coefficients */
    const double* us_K;      /* [nus*nus] K_ab = hscale*kpair*X(EN) or hscale */
    double kexp;             /* repulsion distance exponent */
    double rep_cutoff;       /* bohr */
    double cn_cutoff;        /* bohr */
    double cn_steep;         /* steepness of the counting function */
    double int_cutoff;       /* bohr, overlap cutoff */
} xtb_params;

typedef struct xtb_scf_opts {
    int32_t maxiter;
    int32_t mix_guess;
    int32_t generations;     /* Anderson history length, 1..6 */
    int32_t fermi_maxiter;
    int32_t want_w;          /* also form the energy-weighted density */
    int32_t jacobi_max_sweeps;
    int32_t padded;          /* 1: padded layout semantics */
    int32_t jacobi_inner;    /* sweeps per 32x32 block-Jacobi subproblem */
    double x_atol;
    double x_atol_max;
    double kt;               /* electronic temperature, Hartree; below 3e-7 (0.1 K)
                                the orbitals are filled by the aufbau principle */
    double damp;
    double damp_init;
    double diag_offset;
    double fermi_thresh;
    double jacobi_tol;
    double eig_inexact;      /* eigensolve tolerance = this x last |dv|_inf (0: exact) */
    double bd_gap_kt;        /* intermediate SCF steps only decouple occupied from
                                virtual orbitals when the previous step's frontier
                                gaps are >= this many kT (0: always full eigensolve) */
    double pd_kmax;          /* closed shells: decouple by one perturbative
                                occupied-virtual rotation when its largest
                                amplitude is <= this (0: Jacobi only) */
    double first_tol;        /* eigensolve tolerance of the first SCF step
                                (used when eig_inexact > 0) */
    int32_t ortho;           /* orthogonaliser: 0 Loewdin S^-1/2, 1 Cholesky L^-1 */
    int32_t orbitals;        /* 1: the final step forms the orbital energies (emo)
                                by a full eigensolve; 0: for gapped closed shells
                                it stops once occupied and virtual orbitals are
                                decoupled (density, energies and W are the same),
                                emo is then NaN and the Fermi levels are mid-gap
                                values of the decoupled blocks' diagonals */
    int32_t keep_state;      /* 1: leave the orbitals (C, rows in ascending energy
                                order) and the final potential (pot) in the state
                                buffers for a following warm start */
    int32_t warm_start;      /* 1: start from what a previous keep_state call on
                                the same batch and state buffers left there (new
                                positions allowed): the kept orbitals are
                                re-orthonormalised against the new overlap, the
                                SCF starts from the kept potential with a fresh
                                mixer.  Molecules whose previous status was not
                                0 start cold.  Converges to the same fixed point */
    int32_t soft_start;      /* 1: the mixer's first `generations` steps are simple
                                mixing with damp_init (upstream's soft start);
                                0: Anderson acceleration from the second step */
    int32_t wg_skip;         /* 0..7: the launch starts with this many workgroups
                                that exit at once.  Workgroup i of a launch runs
                                on XCD (i + first) mod 8 where the dispatcher
                                mostly starts each launch at XCD 0, so launches
                                of a few molecules pile up on the first XCDs;
                                a serving loop rotates this to spread them */
    int32_t mixer;           /* 0: Anderson (upstream's full mode, mixer/anderson.py);
                                1: limited-memory Broyden root search on
                                x - f(x) (upstream's implicit modes: xitorch
                                broyden1, alpha -0.5, no line search); its
                                memory holds `generations` (1..24) updates */
    double f_atol;           /* Broyden: converged when the last step's |dx| <
                                x_atol and |x - f(x)| < f_atol */
} xtb_scf_opts;

/* Work and result buffers of xtb_scf (device, caller-owned). */
typedef struct xtb_scf_state {
    double* F;      /* packed n x n: Fock / transformed Fock */
    double* C;      /* packed n x n: MO coefficients stored by vector,
                       C[k*n + mu] = coefficient of orbital mu in MO k;
                       scratch on exit when want_w */
    double* T;      /* packed n x n: scratch; energy-weighted density W on exit if want_w */
    double* P;      /* packed n x n: density */
    double* pot;    /* [nvec] orbital potential of the final charges */
    double* vnew;   /* [nvec] scratch: potential produced by an SCF step */
    double* q;      /* [nvec] orbital Mulliken charges */
    double* eorb;   /* [nvec] diag(P H0) */
    double* emo;    /* [nvec] orbital energies, ascending */
    double* occ;    /* [2*nvec] alpha/beta occupations, ascending-energy order */
    double* hist_x; /* [(generations+1)*nvec] */
    double* hist_f; /* [(generations+1)*nvec] */
    double* qsh;    /* [nsh_total] shell charges of the final density */
    double* out;    /* [nmol*XTB_OUT_STRIDE] result records */
    const int32_t* order; /* [nmol] processing order (largest first) */
} xtb_scf_state;

int32_t xtb_abi_version(void);

/* sizeof(xtb_batch), sizeof(xtb_params), sizeof(xtb_scf_opts),
 * sizeof(xtb_scf_state): lets bindings verify their struct mirrors. */
int32_t xtb_abi_struct_sizes(int64_t* out4);

/* Doubles each buffer class needs for this batch:
 * out[0] = packed n x n buffer, out[1] = packed nsh x nsh, out[2] = nvec,
 * out[3] = history (generations+1)*nvec.  Host-only, no device access. */
int32_t xtb_workspace_doubles(const xtb_batch* batch, int32_t generations,
                              int64_t* out4);

int32_t xtb_build_s_h0(const xtb_batch* batch, const xtb_params* par,
                       const double* positions, double* S, double* H0,
                       double* cn, void* stream);

int32_t xtb_build_gamma(const xtb_batch* batch, const xtb_params* par,
                        const double* positions, double* gamma, void* stream);

/* xtb_build_s_h0 and xtb_build_gamma in one pass over the shell-pair tiles
 * (the same results; one launch fewer per batch).  Replaces the same two
 * upstream call sites as the pair: GFN1Hamiltonian.build
 * (_src/xtb/base.py:251-360) and ES2.get_cache (coulomb/secondorder.py:216-276). */
int32_t xtb_build_s_h0_gamma(const xtb_batch* batch, const xtb_params* par,
                             const double* positions, double* S, double* H0,
                             double* cn, double* gamma, void* stream);

/* Dynamic LDS the SCF kernel will request for this batch, and the part of
 * it available to the eigensolver (doubles).  Host-only. */
int32_t xtb_scf_lds_bytes(const xtb_batch* batch, int32_t* bytes, int32_t* big_cap);

int32_t xtb_scf(const xtb_batch* batch, const xtb_params* par,
                const double* S, const double* H0, const double* gamma,
                const xtb_scf_opts* opts, const xtb_scf_state* state,
                void* stream);

/* Scratch (doubles) of xtb_scf_large for this one-molecule batch: one more
 * n x n matrix, the block-Jacobi rotations, orderings, reduction partials
 * and control values.  Host-only.  XTB_EINVAL unless nmol == 1;
 * XTB_ETOOBIG above 8192 orbitals. */
int32_t xtb_scf_large_workspace(const xtb_batch* batch, int64_t* doubles);

/* xtb_scf for ONE large molecule (thousands of orbitals) on the whole GPU:
 * the same SCF (orthogonaliser, Fock build, perturbative occupied-virtual
 * rotation for gapped closed shells, block-Jacobi eigensolve otherwise,
 * Fermi smearing, density, Mulliken charges, ES2/ES3 potential, Anderson
 * mixing, energies, W) with every stage a grid kernel over all CUs and the
 * SCF loop on the host (a few control values read back per step, so the
 * call returns after the SCF has finished on `stream`).  Same state buffers
 * (F, C, T, P: one n x n each, stored at offset 0) and result record as
 * xtb_scf, so xtb_forces and xtb_param_grad follow unchanged.  The orbitals
 * always stay in C (keep_state has no further effect); with warm_start and
 * a previous call on this state that succeeded (its record's status 0), the
 * SCF starts from them, re-orthonormalised against the new S, and from the
 * kept potential.  `work` holds xtb_scf_large_workspace doubles.
 * Replaces the same upstream call sites as xtb_scf
 * (src/dxtb/_src/scf/iterator.py:51-144, scf/unrolling/default.py:71-401,
 * with scf/unrolling/base.py:142-175 for the eigensolve). */
int32_t xtb_scf_large(const xtb_batch* batch, const xtb_params* par,
                      const double* S, const double* H0, const double* gamma,
                      const xtb_scf_opts* opts, const xtb_scf_state* state,
                      double* work, int64_t work_doubles, void* stream);

int32_t xtb_repulsion(const xtb_batch* batch, const xtb_params* par,
                      const double* positions, double* out, void* stream);

/* The whole integral pass in one launch: CN, S, H0 (xtb_build_s_h0), gamma
 * (xtb_build_gamma) and the repulsion energy (xtb_repulsion), with the same
 * results bit for bit.  Replaces the same upstream call sites as those three:
 * GFN1Hamiltonian.build (_src/xtb/base.py:251-360) with cn_d3, ES2.get_cache
 * (coulomb/secondorder.py:216-276) and Repulsion.get_energy
 * (classicals/repulsion/base.py:269-334). */
int32_t xtb_build_integrals(const xtb_batch* batch, const xtb_params* par,
                            const double* positions, double* S, double* H0, double* cn,
                            double* gamma, double* out, void* stream);

/* Gradient of the repulsion energy alone, per atom [3*nat_total] (written,
 * not added); replaces the autograd of Repulsion.get_energy
 * (src/dxtb/_src/components/classicals/repulsion/base.py:337-406).  With
 * xtb_forces it gives the gradient when the repulsion or the electronic
 * energy is excluded from the total. */
int32_t xtb_repulsion_gradient(const xtb_batch* batch, const xtb_params* par,
                               const double* positions, double* grad, void* stream);

/* Nuclear gradient dE/dR of E = E_scf + (-TS) + E_rep, per atom [3*nat_total].
 * Needs the converged P, W (state.T when want_w), the orbital potential,
 * shell charges and CN from the same batch.  `dedcn` is [nat_total] scratch. */
int32_t xtb_forces(const xtb_batch* batch, const xtb_params* par,
                   const double* positions, const double* cn,
                   const double* gamma, const xtb_scf_state* state,
                   double* dedcn, double* grad, void* stream);

/* The single point of every molecule in ONE launch: each molecule's
 * workgroup computes its CN, S, H0, gamma and repulsion energy (as
 * xtb_build_integrals, the same results bit for bit), runs its SCF (as
 * xtb_scf with the same opts and state) and, with opts->want_w, writes its
 * dE/dCN to `dedcn` [nat_total] and its gradient dE/dR [3*nat_total] to
 * `grad` (written, not added; per-atom sums in a fixed order, so the same
 * bits on every run; xtb_forces agrees to rounding).  Replaces the chain
 * upstream's singlepoint + forces_analytical run (calculators/types/
 * energy.py:78-402, analytical.py:61-222) with one kernel per batch, so
 * that under a busy serving loop no small integral or gradient launch
 * waits for compute units between a batch's stages.  The record fields
 * XTB_OUT_T_INTEGRALS and XTB_OUT_T_FORCES time those parts, T_TOTAL the
 * whole molecule.  state.F is scratch after the SCF (the pair terms of the
 * gradient).  dedcn and grad may be NULL without want_w. */
int32_t xtb_singlepoint(const xtb_batch* batch, const xtb_params* par,
                        const double* positions, double* S, double* H0, double* cn,
                        double* gamma, const xtb_scf_opts* opts, const xtb_scf_state* state,
                        double* dedcn, double* grad, void* stream);

/* Atoms a batch may hold in xtb_forces_noalloc (its scratch: 8 bytes per
 * atom, 2 MiB per device, held until the process ends). */
#define XTB_NOALLOC_MAX_ATOMS 262144

/* xtb_forces without the `dedcn` scratch argument: allocates nothing and
 * takes no workspace; the scratch lives in a fixed device array of
 * XTB_NOALLOC_MAX_ATOMS doubles inside the library, one per device.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, results and status codes as
 * xtb_forces; XTB_ETOOBIG when nat_total > XTB_NOALLOC_MAX_ATOMS.  All calls
 * on one device share the scratch array: calls that overlap in time on one
 * device (different streams or threads) race. */
int32_t xtb_forces_noalloc(const xtb_batch* batch, const xtb_params* par,
                           const double* positions, const double* cn,
                           const double* gamma, const xtb_scf_state* state,
                           double* grad, void* stream);

/* Parameter gradient dE/dtheta of E = E_scf + (-TS) + E_rep at the converged
 * density (the energy is stationary in it, so no response term); replaces
 * autograd through the SCF with differentiable parameters
 * (src/dxtb/_src/param/module/utils.py:123-240, examples/optim/param.py).
 * Needs S and CN of the geometry and state.P / state.qsh of a finished
 * xtb_scf.  Per molecule, 4*nus + 3*nsp + nus*nus doubles: dE/d level[nus],
 * d kcn[nus], d shpoly[nus], d eta[nus] (ES2 shell hardness), d gam3[nsp],
 * d zeff[nsp], d arep[nsp], d us_K[nus][nus] (the xtb_params tables,
 * atomic units; us_K as each off-site shell pair reads it).
 * nus <= 32, nsp <= 16. */
int32_t xtb_param_grad(const xtb_batch* batch, const xtb_params* par,
                       const double* positions, const double* cn, const double* S,
                       const xtb_scf_state* state, double* grad, void* stream);

/* Parameter gradient for one Slater exponent: adds dE/dzeta to column `col`
 * of grad [nmol x ncol] (caller zeroes it).  The shells' primitives rebuilt
 * at zeta + h and zeta - h (alpha/coeff tables as in xtb_params, nus x 8);
 * bit u of `mask` marks the unique shells they change.  Needs P, W (state.T
 * of an xtb_scf with want_w) and the potential of the same batch.  dS/dzeta
 * is the central difference of the contracted overlaps; the contraction with
 * the energy's derivative in S is that of xtb_forces.  Replaces autograd
 * with differentiable "slater" entries (examples/optim/param.py). */
int32_t xtb_param_grad_basis(const xtb_batch* batch, const xtb_params* par,
                             const double* alpha_plus, const double* coeff_plus,
                             const double* alpha_minus, const double* coeff_minus,
                             uint32_t mask, double h, const double* positions,
                             const double* cn, const xtb_scf_state* state, double* grad,
                             int32_t col, int32_t ncol, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* XTB_HIP_H */
