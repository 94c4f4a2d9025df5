"""HIP path vs the CPU oracle on the same seeded inputs (FP64).

Tolerances: integrals 1e-12 (absolute, O(1) entries), energies 1e-9 Eh,
orbital Mulliken charges 1e-8 e, forces 1e-8 Eh/bohr -- tighter than the
1e-8 Eh / 1e-7 e / 1e-7 Eh/bohr parity targets.
"""

import numpy as np
import pytest
import torch

import oracle.gfn1 as og
from conftest import mol
from dxtb_amd import workloads as W
from dxtb_amd.engine import ScfOptions
from helpers import TIGHT_ORACLE, block, run_native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _mols(fx):
    out = [mol(fx, k) for k in ("h2o", "ch4", "lys_xao")]
    out.append(W.config0())
    return out


def test_integrals_vs_oracle(fx):
    mols = _mols(fx)
    print("case: integrals h2o ch4 lys_xao c8h18", flush=True)
    pipe = run_native(mols, forces=False)
    for b, (z, p) in enumerate(mols):
        pos = torch.tensor(p)
        bas = og.build_basis(z)
        S = og.overlap(bas, pos).numpy()
        cn = og.coordination_number(list(z), pos).numpy()
        H = og.core_hamiltonian(bas, pos, torch.tensor(S), torch.tensor(cn)).numpy()
        G = og.gamma_matrix(bas, pos).numpy()
        a0 = int(pipe.L.atom0[b])
        assert np.abs(block(pipe, pipe.S, b) - S).max() < 1e-13
        assert np.abs(block(pipe, pipe.H0, b) - H).max() < 1e-12
        assert np.abs(pipe.cn[a0:a0 + len(z)].cpu().numpy() - cn).max() < 1e-12
        g0, m = int(pipe.L.gam0[b]), int(pipe.L.nsh[b])
        Gn = pipe.gamma[g0:g0 + m * m].view(m, m).cpu().numpy()
        assert np.abs(Gn - G).max() < 1e-13
        erep = og.repulsion(list(z), pos).sum().item()
        assert abs(pipe.records()[b, 9].item() - erep) < 1e-12


def test_energy_charges_forces_vs_oracle(fx):
    mols = _mols(fx)
    print("case: scf+forces h2o ch4 lys_xao c8h18", flush=True)
    pipe = run_native(mols, forces=True)
    rec = pipe.records().cpu().numpy()
    E = pipe.energies().cpu().numpy()
    G = pipe.gradient().cpu().numpy()
    for b, (z, p) in enumerate(mols):
        r = og.singlepoint(z, p, opts=TIGHT_ORACLE, with_forces=True)
        assert int(rec[b, 6]) == 0, f"status {rec[b, 6]}"
        assert abs(E[b] - r["energy"]) < 1e-9
        assert abs(rec[b, 0] - r["scf"]) < 1e-9
        q = pipe.orbital_charges(b).cpu().numpy()
        assert np.abs(q - r["charges_orbital"].numpy()).max() < 1e-8
        a0 = int(pipe.L.atom0[b])
        assert np.abs(G[a0:a0 + len(z)] - r["gradient"].numpy()).max() < 1e-8


def test_scf_energy_vs_tblite(fx):
    keys = ("h2", "h2o", "ch4", "lys_xao", "c60")
    mols = [mol(fx, k) for k in keys]
    print("case: tblite energies", flush=True)
    pipe = run_native(mols)
    rec = pipe.records().cpu().numpy()
    for b, k in enumerate(keys):
        assert abs(rec[b, 0] - float(fx[f"{k}_escf"])) < 1e-9, k


def test_padded_layout_matches_ragged(fx):
    mols = _mols(fx)
    print("case: padded vs ragged", flush=True)
    a = run_native(mols, padded=False, forces=True)
    b = run_native(mols, padded=True, forces=True)
    assert np.abs(a.energies().cpu().numpy() - b.energies().cpu().numpy()).max() < 1e-10
    assert np.abs(a.gradient().cpu().numpy() - b.gradient().cpu().numpy()).max() < 1e-10


def test_large_molecules_block_jacobi_vs_oracle():
    """n > LDS capacity: exercises the block-Jacobi eigensolver."""
    rng = np.random.default_rng(5)
    mols = [W.chain(rng, 40), W.water_cluster(rng, 50), W.chain(rng, 30)]
    print("case: block jacobi n=" + ",".join(str(len(z)) for z, _ in mols) + " atoms", flush=True)
    pipe = run_native(mols, forces=True)
    rec = pipe.records().cpu().numpy()
    E = pipe.energies().cpu().numpy()
    G = pipe.gradient().cpu().numpy()
    assert int(pipe.L.n.max()) > 200
    for b, (z, p) in enumerate(mols):
        r = og.singlepoint(z, p, opts=TIGHT_ORACLE, with_forces=True)
        assert int(rec[b, 6]) == 0, f"status {rec[b, 6]}"
        assert abs(E[b] - r["energy"]) < 1e-9
        q = pipe.orbital_charges(b).cpu().numpy()
        assert np.abs(q - r["charges_orbital"].numpy()).max() < 1e-8
        a0 = int(pipe.L.atom0[b])
        assert np.abs(G[a0:a0 + len(z)] - r["gradient"].numpy()).max() < 1e-8


def test_inexact_eigensolves_reach_the_same_fixed_point(fx):
    """The SCF steps' eigensolves are converged only as far as the SCF
    residual warrants; the final step is exact.  Results must equal the fully
    exact path (every step to 1e-13) to the parity tolerances: the two SCF
    trajectories stop at different points within the convergence threshold
    (|dv| < 1e-11), and the energy of a non-self-consistent density moves to
    first order in that residual (~1e-10 Eh for a 30-water cluster)."""
    rng = np.random.default_rng(6)
    mols = [mol(fx, "lys_xao"), W.chain(rng, 35), W.water_cluster(rng, 30)]
    print("case: inexact vs exact eigensolves", flush=True)
    a = run_native(mols, forces=True)
    b = run_native(mols, forces=True, eig_inexact=0.0)
    assert np.abs(a.energies().cpu().numpy() - b.energies().cpu().numpy()).max() < 1e-9
    assert np.abs(a.gradient().cpu().numpy() - b.gradient().cpu().numpy()).max() < 1e-8
    for k in range(len(mols)):
        qa, qb = a.orbital_charges(k).cpu().numpy(), b.orbital_charges(k).cpu().numpy()
        assert np.abs(qa - qb).max() < 1e-8


def test_block_diagonal_steps_reach_the_same_fixed_point(fx):
    """Intermediate SCF steps of gapped molecules only decouple occupied from
    virtual orbitals (the density of 0/1 occupations needs nothing more), by a
    perturbative rotation or by Jacobi on the inter-class pairs; the final
    step is a full eigensolve.  Results must equal those of full
    eigensolves in every step, to the parity tolerances.  Covers the packed
    scalar Jacobi (n < 190) and block Jacobi (n > 190) paths."""
    rng = np.random.default_rng(8)
    mols = [mol(fx, "lys_xao"), W.chain(rng, 12), W.chain(rng, 35), W.water_cluster(rng, 30)]
    print("case: occupied/virtual decoupling vs full eigensolves", flush=True)
    b = run_native(mols, forces=True, bd_gap_kt=0.0)
    for opt in ({"orbital_energies": True}, {"pd_kmax": 0.0}):
        # default: perturbative rotations where they apply, else Jacobi on the
        # inter-class pairs; pd_kmax=0: Jacobi on the inter-class pairs only
        a = run_native(mols, forces=True, **opt)
        rec = a.records().cpu().numpy()
        assert np.abs(rec[:, 6]).max() == 0
        if "pd_kmax" not in opt:
            assert rec[:, 17].max() > 0  # the perturbative path ran
        assert np.abs(a.energies().cpu().numpy() - b.energies().cpu().numpy()).max() < 1e-9
        assert np.abs(a.gradient().cpu().numpy() - b.gradient().cpu().numpy()).max() < 1e-8
        for k in range(len(mols)):
            qa, qb = a.orbital_charges(k).cpu().numpy(), b.orbital_charges(k).cpu().numpy()
            assert np.abs(qa - qb).max() < 1e-8
            ea, eb = a.emo[int(a.L.vec0[k]):int(a.L.vec0[k]) + int(a.L.n[k])], \
                b.emo[int(b.L.vec0[k]):int(b.L.vec0[k]) + int(b.L.n[k])]
            assert (ea - eb).abs().max().item() < 1e-8


@pytest.mark.parametrize("key", ["h2", "h2o", "ch4"])
def test_small_molecule_alone_vs_oracle(fx, key):
    """A molecule alone in its batch: every n x n work buffer is exactly
    n^2 long (n = 2..12 < 32, the width of the Cholesky and Jacobi blocks),
    so a block routine writing past n^2 would corrupt the other buffers."""
    z, p = mol(fx, key)
    print(f"case: {key} alone", flush=True)
    pipe = run_native([(z, p)], forces=True)
    r = og.singlepoint(z, p, opts=TIGHT_ORACLE, with_forces=True)
    assert abs(pipe.energies().item() - r["energy"]) < 1e-9
    assert np.abs(pipe.gradient().cpu().numpy() - r["gradient"].numpy()).max() < 1e-8
    assert np.abs(pipe.orbital_charges(0).cpu().numpy() - r["charges_orbital"].numpy()).max() < 1e-8


def test_loewdin_and_cholesky_orthogonalisers_agree(fx):
    """The first-step basis (Loewdin S^-1/2 or Cholesky L^-1) changes the SCF
    path, not the converged eigenpairs: results agree to the parity bar."""
    rng = np.random.default_rng(10)
    mols = [mol(fx, "lys_xao"), mol(fx, "ch4"), W.chain(rng, 30), W.water_cluster(rng, 12)]
    print("case: Loewdin vs Cholesky orthogonaliser", flush=True)
    a = run_native(mols, forces=True, ortho=0)
    b = run_native(mols, forces=True, ortho=1)
    assert np.abs(a.energies().cpu().numpy() - b.energies().cpu().numpy()).max() < 1e-9
    assert np.abs(a.gradient().cpu().numpy() - b.gradient().cpu().numpy()).max() < 1e-8


def test_orbital_energies_only_on_request(fx):
    """Without orbital energies the final step of a gapped closed shell stops
    once occupied and virtual orbitals are decoupled to full tolerance and W
    comes from the occupied block: the energy, charges and forces equal those
    of the full final eigensolve (the SCF paths are identical up to that
    step) up to rounding (the energy sums over n^2 products carry ~1e-13
    relative error: 3e-11 Eh for 66 waters), and the orbital energies are
    NaN.  With them, they match the oracle's eigenvalues; so do the
    calculator's get_mo_energies."""
    from dxtb_amd import GFN1Calculator

    rng = np.random.default_rng(11)
    mols = [mol(fx, "lys_xao"), W.chain(rng, 12), W.chain(rng, 35), W.water_cluster(rng, 30)]
    print("case: final step with and without orbital energies", flush=True)
    a = run_native(mols, forces=True)
    b = run_native(mols, forces=True, orbital_energies=True)
    assert np.abs(a.energies().cpu().numpy() - b.energies().cpu().numpy()).max() < 1e-10
    assert np.abs(a.gradient().cpu().numpy() - b.gradient().cpu().numpy()).max() < 1e-10
    for k in range(len(mols)):
        qa, qb = a.orbital_charges(k).cpu().numpy(), b.orbital_charges(k).cpu().numpy()
        assert np.abs(qa - qb).max() < 1e-10
        v0, n = int(a.L.vec0[k]), int(a.L.n[k])
        assert torch.isnan(a.emo[v0:v0 + n]).all()
        assert torch.isfinite(b.emo[v0:v0 + n]).all()
    z, p = mols[0]
    r = og.singlepoint(z, p, opts=TIGHT_ORACLE)
    v0, n = int(b.L.vec0[0]), int(b.L.n[0])
    assert np.abs(b.emo[v0:v0 + n].cpu().numpy() - r["emo"].numpy()).max() < 1e-8
    calc = GFN1Calculator(torch.tensor(z, device="cuda:0"), dtype=torch.float64,
                          opts={**TIGHT_ORACLE, "exclude": ["disp"]})
    pos = torch.tensor(p, dtype=torch.float64, device="cuda:0")
    e = calc.get_energy(pos)
    emo = calc.get_mo_energies(pos)
    assert abs(float(e) - r["energy"]) < 1e-9
    assert np.abs(emo.cpu().numpy() - r["emo"].numpy()).max() < 1e-8


def test_warm_start_along_a_trajectory(fx):
    """Molecular-dynamics-style calls on one batch: each geometry starts from
    the orbitals and potential the previous call kept (the orbitals
    re-orthonormalised against the new overlap).  Every geometry's results
    equal a cold start's to the parity tolerances, without the first step's
    eigensolve.  A molecule whose previous call did not converge starts
    cold, exactly as a fresh run does."""
    from helpers import TIGHT, native_pipeline

    rng = np.random.default_rng(12)
    geo = [[mol(fx, "lys_xao"), W.chain(rng, 30), W.water_cluster(rng, 25), mol(fx, "h2o")]]
    for _ in range(2):
        geo.append([(z, p + 0.03 * rng.standard_normal(p.shape)) for z, p in geo[-1]])
    print("case: warm start along a 3-geometry trajectory", flush=True)
    pipe, run = native_pipeline(geo[0])
    warm = ScfOptions(**TIGHT, warm_start=True)
    for g, mols in enumerate(geo):
        run(mols, warm, forces=True)
        c = run_native(mols, forces=True)
        rw, rc = pipe.records().cpu().numpy(), c.records().cpu().numpy()
        assert np.abs(rw[:, 6]).max() == 0
        assert np.abs(pipe.energies().cpu().numpy() - c.energies().cpu().numpy()).max() < 1e-9
        assert np.abs(pipe.gradient().cpu().numpy() - c.gradient().cpu().numpy()).max() < 1e-8
        for k in range(len(mols)):
            qa, qb = pipe.orbital_charges(k).cpu().numpy(), c.orbital_charges(k).cpu().numpy()
            assert np.abs(qa - qb).max() < 1e-8
        print(f"geometry {g}: iterations warm {rw[:, 5]} cold {rc[:, 5]}, first-step sweeps "
              f"warm {rw[:, 18]} cold {rc[:, 18]}", flush=True)
        # and against the oracle at every geometry (energies, charges, forces)
        Ew, Gw = pipe.energies().cpu().numpy(), pipe.gradient().cpu().numpy()
        for k, (z, p) in enumerate(mols):
            r = og.singlepoint(z, p, opts=TIGHT_ORACLE, with_forces=True)
            assert abs(Ew[k] - r["energy"]) < 1e-9, (g, k)
            qw = pipe.orbital_charges(k).cpu().numpy()
            assert np.abs(qw - r["charges_orbital"].numpy()).max() < 1e-8, (g, k)
            a0 = int(pipe.L.atom0[k])
            assert np.abs(Gw[a0:a0 + len(z)] - r["gradient"].numpy()).max() < 1e-8, (g, k)
        if g > 0:
            # the saving is the first step's eigensolve (the kept orbitals
            # replace it); with random 0.03 bohr displacements the iteration
            # counts stay about the same as a cold start's
            assert rw[:, 18].sum() < rc[:, 18].sum(), (rw[:, 18], rc[:, 18])
            assert rw[:, 5].sum() <= rc[:, 5].sum() + len(mols), (rw[:, 5], rc[:, 5])
    # an unconverged call (maxiter 2) leaves status != 0: the next call starts cold
    run(geo[1], ScfOptions(**{**TIGHT, "maxiter": 2}, warm_start=True), forces=False)
    assert (pipe.records()[:, 6] != 0).all()
    run(geo[2], warm, forces=True)
    c = run_native(geo[2], forces=True)
    assert np.array_equal(pipe.records()[:, 5].cpu().numpy(), c.records()[:, 5].cpu().numpy())
    assert np.abs(pipe.energies().cpu().numpy() - c.energies().cpu().numpy()).max() < 1e-12


def test_kept_orbitals_stay_orthonormal(fx):
    """The intermediate steps' rotations truncate their series at 1e-13 (the
    final step at 1e-17) and are never re-orthonormalised: after a full SCF
    the orbitals kept for a warm start must still satisfy C S C^T = I to
    1e-11, and the density must hold the electron count to 1e-10."""
    from helpers import TIGHT, native_pipeline

    rng = np.random.default_rng(13)
    mols = [W.water_cluster(rng, 60), W.chain(rng, 60), mol(fx, "lys_xao")]
    print("case: orthonormality of the kept orbitals", flush=True)
    pipe, run = native_pipeline(mols)
    run(mols, ScfOptions(**TIGHT, warm_start=True), forces=False)
    assert int(pipe.records()[:, 6].abs().max().item()) == 0
    for b in range(len(mols)):
        C = pipe.matrix(pipe.C, b)
        S = pipe.matrix(pipe.S, b)
        G = C @ S @ C.T
        err = (G - torch.eye(G.shape[0], dtype=G.dtype, device=G.device)).abs().max().item()
        assert err < 1e-11, (b, err)
        P = pipe.matrix(pipe.P, b)
        nel = float(pipe.L.nel_a[b] + pipe.L.nel_b[b])
        assert abs((P * S).sum().item() - nel) < 1e-10


def test_mixer_settings_reach_the_same_fixed_point(fx):
    """Upstream's mixer (soft start, 5 generations), this path's default (no
    soft start, 6) and a short history (3) converge to the same energies,
    charges and forces: the mixer changes the route, not the fixed point."""
    rng = np.random.default_rng(21)
    mols = [mol(fx, "lys_xao"), W.chain(rng, 40), W.water_cluster(rng, 25), mol(fx, "h2o")]
    print("case: mixer settings", flush=True)
    runs = [run_native(mols, forces=True, **kw) for kw in (
        dict(soft_start=True, generations=5), dict(), dict(generations=3), dict(soft_start=True, generations=1))]
    e0, g0 = runs[0].energies().cpu().numpy(), runs[0].gradient().cpu().numpy()
    for r in runs[1:]:
        assert np.abs(r.energies().cpu().numpy() - e0).max() < 1e-10
        assert np.abs(r.gradient().cpu().numpy() - g0).max() < 1e-9
        assert int(r.records()[:, 6].abs().max()) == 0
    it = [float(r.records()[:, 5].mean()) for r in runs]
    assert it[1] < it[0]


def test_xcd_shift_changes_placement_not_results(fx):
    """`wg_skip` leading workgroups exit at once and move every molecule's
    workgroup to later XCDs: the results are those of the unshifted launch
    (the forces kernels accumulate with atomics, hence the 1e-12)."""
    from helpers import TIGHT, native_pipeline

    mols = _mols(fx)
    print("case: XCD shift 0 vs 5", flush=True)
    a = run_native(mols, forces=True)
    pipe, _ = native_pipeline(mols)
    pos = torch.tensor(np.concatenate([p for _, p in mols]), dtype=torch.float64, device="cuda:0")
    pipe.run(pos, ScfOptions(**TIGHT), forces=True, wg_skip=5)
    torch.cuda.synchronize()
    assert np.abs(a.energies().cpu().numpy() - pipe.energies().cpu().numpy()).max() < 1e-12
    assert np.abs(a.gradient().cpu().numpy() - pipe.gradient().cpu().numpy()).max() < 1e-12
    ra, rb = a.records().cpu().numpy(), pipe.records().cpu().numpy()
    assert np.array_equal(ra[:, 5], rb[:, 5]) and (rb[:, 6] == 0).all()
    assert ((rb[:, 22] >= 0) & (rb[:, 22] < 8)).all()


def test_fused_integrals_equal_the_separate_entry_points(fx):
    """xtb_build_integrals (what the pipeline calls: CN, S, H0, gamma and the
    repulsion in one launch) writes exactly what xtb_build_s_h0,
    xtb_build_gamma, xtb_build_s_h0_gamma and xtb_repulsion write."""
    import ctypes as ct

    from dxtb_amd import _lib

    mols = _mols(fx)
    print("case: fused S/H0/gamma vs separate entry points", flush=True)
    p = run_native(mols, forces=False)
    S, H0, gam, cn = (torch.full_like(t, float("nan")) for t in (p.S, p.H0, p.gamma, p.cn))
    lib = _lib.load()
    s = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    pos = p.positions.data_ptr()
    _lib.check(lib.xtb_build_s_h0(ct.byref(p.batch), ct.byref(p.dt.s), pos, S.data_ptr(), H0.data_ptr(),
                                  cn.data_ptr(), s), "xtb_build_s_h0")
    _lib.check(lib.xtb_build_gamma(ct.byref(p.batch), ct.byref(p.dt.s), pos, gam.data_ptr(), s),
               "xtb_build_gamma")
    out = torch.full_like(p.out, float("nan"))
    _lib.check(lib.xtb_repulsion(ct.byref(p.batch), ct.byref(p.dt.s), pos, out.data_ptr(), s),
               "xtb_repulsion")
    S2, H2, g2, c2 = (torch.full_like(t, float("nan")) for t in (p.S, p.H0, p.gamma, p.cn))
    _lib.check(lib.xtb_build_s_h0_gamma(ct.byref(p.batch), ct.byref(p.dt.s), pos, S2.data_ptr(),
                                        H2.data_ptr(), c2.data_ptr(), g2.data_ptr(), s),
               "xtb_build_s_h0_gamma")
    torch.cuda.synchronize()
    assert torch.equal(cn, p.cn) and torch.equal(c2, p.cn)
    assert torch.equal(out.view(-1, 32)[:, 9], p.records()[:, 9])
    for b in range(len(mols)):
        # each molecule's blocks (the packed buffers keep alignment gaps
        # between molecules that no kernel writes)
        assert torch.equal(p.matrix(S, b), p.matrix(p.S, b))
        assert torch.equal(p.matrix(H0, b), p.matrix(p.H0, b))
        assert torch.equal(p.matrix(S2, b), p.matrix(p.S, b))
        assert torch.equal(p.matrix(H2, b), p.matrix(p.H0, b))
        g0, m = int(p.L.gam0[b]), int(p.L.nsh[b])
        assert torch.equal(gam[g0:g0 + m * m], p.gamma[g0:g0 + m * m])
        assert torch.equal(g2[g0:g0 + m * m], p.gamma[g0:g0 + m * m])


def test_extrapolated_warm_start_reaches_the_same_fixed_point(fx):
    """warm_extrapolate starts each call from 2 v_(k-1) - v_(k-2) of the kept
    potentials: along a smooth trajectory (constant drift) every geometry's
    results equal a cold start's to the parity tolerances."""
    from helpers import TIGHT, native_pipeline

    rng = np.random.default_rng(13)
    m0 = [W.chain(rng, 30), W.water_cluster(rng, 20), mol(fx, "lys_xao")]
    drift = [0.01 * rng.standard_normal(p.shape) for _, p in m0]
    geo = [[(z, p + k * d) for (z, p), d in zip(m0, drift)] for k in range(4)]
    print("case: extrapolated warm start along a 4-geometry drift", flush=True)
    pipe, run = native_pipeline(geo[0])
    o = ScfOptions(**TIGHT, warm_start=True, warm_extrapolate=True)
    for g, mols in enumerate(geo):
        run(mols, o, forces=True)
        c = run_native(mols, forces=True)
        rw, rc = pipe.records().cpu().numpy(), c.records().cpu().numpy()
        assert np.abs(rw[:, 6]).max() == 0
        assert np.abs(pipe.energies().cpu().numpy() - c.energies().cpu().numpy()).max() < 1e-9
        assert np.abs(pipe.gradient().cpu().numpy() - c.gradient().cpu().numpy()).max() < 1e-8
        print(f"geometry {g}: iterations extrapolated warm {rw[:, 5]} cold {rc[:, 5]}", flush=True)


def test_final_step_decoupling_meets_its_tolerance():
    """The final step's occupied-virtual coupling, measured on the host from
    the orbitals it leaves, C_occ F(v) C_vir^T with F built from S, H0 and its
    input potential, is within max(1e-13 max(|e|, 1), 1e-1 x_atol_max): the
    rotation's passes converge linearly (the amplitudes divide by the blocks'
    diagonals only), so its exit is a measured coupling, not the
    second-order estimate.  At the default threshold (x_atol_max 1e-5) the
    three 440-520-orbital molecules of the serving stream get there without
    a fallback to Jacobi; at x_atol_max 1e-11 the bound is ~1e-12 Eh."""
    stream = W.config4(seed=4, count=1000)
    mols = [(stream[i][1], stream[i][2]) for i in (900, 992, 831)]
    for xmax in (1e-5, 1e-11):
        print(f"case: final-step coupling, x_atol_max {xmax}", flush=True)
        # warm_start keeps the final orbitals in C (occupied rows first)
        a = run_native(mols, forces=True, x_atol=max(10 * xmax, 1e-10), x_atol_max=xmax, warm_start=True)
        rec = a.records().cpu().numpy()
        assert np.all(rec[:, 6] == 0)
        assert np.all(rec[:, 17] > 0)  # perturbative steps ran
        if xmax >= 1e-5:
            assert np.all(rec[:, 19] == 0)  # no Jacobi sweep in the final step
        for b in range(len(mols)):
            n, v0, k = int(a.L.n[b]), int(a.L.vec0[b]), int(round(float(a.L.nel_a[b])))
            S, H0, C = a.matrix(a.S, b), a.matrix(a.H0, b), a.matrix(a.C, b)
            v = a.vnew[v0:v0 + n]
            F = H0 - 0.5 * S * (v[:, None] + v[None, :])
            A = C @ F @ C.T
            tol = max(1e-13 * max(A.diagonal().abs().max().item(), 1.0), 1e-1 * xmax)
            aov = A[:k, k:].abs().max().item()
            assert aov <= 1.01 * tol, (b, aov, tol)
            # the orbitals stay S-orthonormal
            assert (C @ S @ C.T - torch.eye(n, dtype=C.dtype, device=C.device)).abs().max().item() < 1e-10
