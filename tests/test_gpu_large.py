"""Whole-GPU large-system path (xtb_scf_large: grid GEMMs, grid block
Jacobi, device-side Fermi smearing, density, Mulliken and mixing) vs the
oracle and vs the per-molecule kernel on the same molecules: gapped closed
shells (perturbative steps), hot molecules with fractional occupations
(block Jacobi in every step), charged and open-shell molecules, orbital
energies, and the 600-water cluster of config 3."""

import numpy as np
import pytest
import torch

import oracle.gfn1 as og
from conftest import mol
from dxtb_amd import _lib
from dxtb_amd import workloads as W
from dxtb_amd.engine import ScfOptions
from dxtb_amd.large import LargeSystem
from helpers import TIGHT, TIGHT_ORACLE, run_native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _check(r, o, e_tol=1e-9, q_tol=1e-8, f_tol=1e-8):
    assert r["converged"] and r["status"] == 0, r["status"]
    assert abs(r["energy"] - o["energy"]) < e_tol, (r["energy"], o["energy"])
    assert abs(r["free_energy"] - o["fenergy"]) < e_tol
    assert np.abs(r["charges_orbital"].cpu().numpy() - o["charges_orbital"].numpy()).max() < q_tol
    assert np.abs(r["charges_atom"].cpu().numpy() - o["charges_atom"].numpy()).max() < q_tol
    assert np.abs(r["gradient"].cpu().numpy() - o["gradient"].numpy()).max() < f_tol


@pytest.mark.parametrize("case", ["lys_xao", "water40", "chain30", "water40-loewdin"])
def test_large_path_matches_oracle(fx, case):
    """Cholesky orthogonaliser (default), and Loewdin's by Newton-Schulz."""
    rng = np.random.default_rng(8)
    if case == "lys_xao":
        z, p = mol(fx, "lys_xao")
    elif case.startswith("water40"):
        z, p = W.water_cluster(rng, 40)
    else:
        z, p = W.chain(rng, 30)
    print(f"case: large path {case}", flush=True)
    ls = LargeSystem(z, "cuda:0", opts=ScfOptions(**TIGHT, ortho=0 if case.endswith("loewdin") else 1))
    r = ls.run(p, forces=True)
    o = og.singlepoint(z, p, opts=TIGHT_ORACLE, with_forces=True)
    _check(r, o)
    # gapped closed shells: intermediate steps by the perturbative rotation
    assert r["records"][_lib.OUT_PD_STEPS] > 0
    k = run_native([(z, p)], forces=True)
    assert abs(k.energies().item() - r["energy"]) < 1e-9
    assert np.abs(k.gradient().cpu().numpy() - r["gradient"].cpu().numpy()).max() < 1e-9


@pytest.mark.parametrize("etemp", [5000.0, 20000.0])
def test_large_path_fractional_occupations(etemp):
    """Hot water cluster: Fermi occupations away from 0/1 around the Fermi
    level and a sizeable -TS; every step is a block-Jacobi eigensolve on the
    whole GPU."""
    z, p = W.water_cluster(np.random.default_rng(31), 30)
    print(f"case: large path, water30 at {etemp} K", flush=True)
    ls = LargeSystem(z, "cuda:0", opts=ScfOptions(**TIGHT, fermi_etemp=etemp))
    r = ls.run(p, forces=True)
    o = og.singlepoint(z, p, opts={**TIGHT_ORACLE, "fermi_etemp": etemp}, with_forces=True)
    _check(r, o)
    assert -o["fenergy"] > (1e-6 if etemp < 1e4 else 1e-4)
    assert r["records"][_lib.OUT_PD_STEPS] == 0


def test_large_path_charged_and_open_shell():
    """A chain cation (odd electron count: doublet, as upstream picks it) and
    the same chain as a triplet: alpha and beta channels of different size."""
    z, p = W.chain(np.random.default_rng(32), 20, p_n=0.0, p_o=0.0)
    for chrg, spin in ((1.0, None), (0.0, 2)):
        print(f"case: large path, chain20 charge {chrg} spin {spin}", flush=True)
        ls = LargeSystem(z, "cuda:0", opts=ScfOptions(**TIGHT), chrg=chrg, spin=spin)
        r = ls.run(p, forces=True)
        o = og.singlepoint(z, p, chrg=chrg, spin=spin, opts=TIGHT_ORACLE, with_forces=True)
        _check(r, o)


def test_large_path_orbital_energies():
    """Orbital energies asked for: the final step diagonalises fully."""
    z, p = W.water_cluster(np.random.default_rng(33), 20)
    print("case: large path, orbital energies", flush=True)
    ls = LargeSystem(z, "cuda:0", opts=ScfOptions(**TIGHT, orbital_energies=True))
    r = ls.run(p, forces=True)
    o = og.singlepoint(z, p, opts=TIGHT_ORACLE, with_forces=True)
    _check(r, o)
    emo = ls.pipe.emo[: len(o["emo"])].cpu().numpy()
    assert np.abs(emo - o["emo"].numpy()).max() < 1e-8


def test_config3_water600_vs_oracle_fixture():
    """600 H2O (1800 atoms, 4800 orbitals) through the calculator, which sends
    a single molecule of this size to the whole-GPU path; the reference
    values come from the oracle (tests/golden/make_config3.py, CPU FP64,
    x_atol 1e-10).  Targets: |dE| <= 1e-8 Eh, atomic charges <= 1e-7 e,
    forces <= 1e-7 Eh/bohr."""
    import os

    from dxtb_amd import GFN1Calculator

    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "config3_oracle.npz"))
    z = torch.as_tensor(ref["numbers"], device="cuda:0")
    pos = torch.as_tensor(ref["positions"], dtype=torch.float64, device="cuda:0")
    print("case: config 3 (600 H2O) vs oracle fixture", flush=True)
    calc = GFN1Calculator(z, dtype=torch.float64, device="cuda:0",
                          opts={"x_atol": 1e-10, "x_atol_max": 1e-10, "maxiter": 300,
                                "exclude": ["disp"]})
    e = calc.get_energy(pos)
    assert calc._pipe.large
    assert abs(float(e) - float(ref["energy"])) < 1e-8
    f = calc.forces_analytical(pos)
    assert np.abs(-f.cpu().numpy() - ref["gradient"]).max() < 1e-7
    q = calc.get_charges(pos)
    assert q.shape == (4800,) and abs(float(q.sum())) < 1e-8
    L = calc._pipe.L
    ao_atom = L.arrays["sh_atom"][L.arrays["ao_sh"]].astype(np.int64)
    qat = np.zeros(1800)
    np.add.at(qat, ao_atom, q.cpu().numpy())
    assert np.abs(qat - ref["charges_atom"]).max() < 1e-7


@pytest.mark.parametrize("padded", [False, True])
def test_batch_with_molecule_above_kernel_limit(fx, padded):
    """A batch through the calculator whose second molecule (129 H2O, 1032
    orbitals) exceeds the per-molecule kernel's limit (1024): it runs on the
    whole-GPU path and the other two as a batch of their own.  Energies,
    orbital charges, forces and orbital energies of all three, in the
    batch's order, against the oracle.  Upstream pads the small molecules to
    1032 orbitals; the oracle with that padding (npad) is the reference, and
    at 300 K its padding levels take no occupation to double precision."""
    from dxtb_amd import GFN1Calculator
    from dxtb_amd.engine import SplitPipeline

    rng = np.random.default_rng(41)
    mols = [mol(fx, "h2o"), W.water_cluster(rng, 129), W.chain(rng, 8)]
    nums, pos = W.pad(mols)
    dev = torch.device("cuda:0")
    print(f"case: batch with a 1032-orbital molecule, padded={padded}", flush=True)
    calc = GFN1Calculator(torch.tensor(nums, device=dev), dtype=torch.float64, device=dev,
                          opts={**TIGHT_ORACLE, "exclude": ["disp"]}, padded=padded)
    p = torch.tensor(pos, dtype=torch.float64, device=dev)
    e = calc.get_energy(p).cpu().numpy()
    assert isinstance(calc._pipe, SplitPipeline)
    f = calc.forces_analytical(p).cpu().numpy()
    q = calc.get_charges(p).cpu().numpy()
    emo = calc.get_mo_energies(p)
    nmax = 1032
    for b, (z, x) in enumerate(mols):
        o = og.singlepoint(z, x, opts=TIGHT_ORACLE, with_forces=True)
        nao = o["charges_orbital"].shape[0]
        if b == 1:
            assert nao == nmax
        assert abs(e[b] - o["energy"]) < 1e-9, (b, e[b], o["energy"])
        assert np.abs(q[b, :nao] - o["charges_orbital"].numpy()).max() < 1e-8
        assert np.all(q[b, nao:] == 0.0)
        assert np.abs(-f[b, : len(z)] - o["gradient"].numpy()).max() < 1e-8
        assert np.all(f[b, len(z):] == 0.0)
        eb = emo[b].cpu().numpy()
        if padded:
            # upstream's padded spectrum: the molecule's levels and
            # nmax - nao padding levels at zero, sorted
            assert eb.shape == (nmax,)
            ref = np.sort(np.concatenate([o["emo"].numpy(), np.zeros(nmax - nao)]))
            if b != 1:
                op = og.singlepoint(z, x, opts=TIGHT_ORACLE, npad=nmax - nao)
                assert abs(op["energy"] - o["energy"]) < 1e-12
        else:
            ref = o["emo"].numpy()
        assert np.abs(eb - ref).max() < 1e-8


def test_calculator_routes_a_single_molecule_by_size():
    """GFN1Calculator sends one molecule above LARGE_N orbitals to the
    whole-GPU path and a smaller one to the per-molecule kernel; both give
    what the per-molecule kernel gives for the same molecule (energy,
    forces, charges), and the batch of the two goes to the kernel."""
    from dxtb_amd import GFN1Calculator
    from dxtb_amd.calculator import LARGE_N

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(11)
    mols = [W.chain(rng, 30), W.chain(rng, 12)]  # n = 236 and 100 orbitals
    ref = run_native(mols, forces=True, **TIGHT)
    opts = {"maxiter": 300, "x_atol": 1e-10, "x_atol_max": 1e-10, "exclude": ["disp"]}
    for k, (z, p) in enumerate(mols):
        print(f"case: calculator, one molecule of {len(z)} atoms", flush=True)
        calc = GFN1Calculator(torch.tensor(z, device=dev), opts=opts, dtype=torch.float64, device=dev)
        pos = torch.tensor(p, dtype=torch.float64, device=dev, requires_grad=True)
        e = calc.get_energy(pos)
        f = calc.get_forces(pos)
        q = calc.get_charges(pos.detach())
        n = int(ref.L.n[k])
        assert calc._pipe.large == (n > LARGE_N)
        assert abs(e.item() - ref.energies()[k].item()) < 1e-9
        a0, na = int(ref.L.atom0[k]), len(z)
        assert (f + ref.gradient()[a0:a0 + na]).abs().max().item() < 1e-8
        assert (q - ref.orbital_charges(k)).abs().max().item() < 1e-8
    assert int(ref.L.n[0]) > LARGE_N >= int(ref.L.n[1])


def test_large_path_warm_start_along_a_trajectory():
    """One molecule on the whole-GPU path along a 3-geometry trajectory with
    warm_start: each geometry starts from the orbitals the previous call
    left (re-orthonormalised against the new overlap) and its potential.
    Results equal a cold start's to the parity tolerances, and the first
    step needs no eigensolve."""
    rng = np.random.default_rng(14)
    z, p0 = W.water_cluster(rng, 40)
    geo = [p0]
    for _ in range(2):
        geo.append(geo[-1] + 0.03 * rng.standard_normal(p0.shape))
    warm = LargeSystem(z, device="cuda:0", opts=ScfOptions(**TIGHT, warm_start=True))
    cold = LargeSystem(z, device="cuda:0", opts=ScfOptions(**TIGHT))
    for g, p in enumerate(geo):
        print(f"case: whole-GPU warm start, geometry {g}", flush=True)
        rw, rc = warm.run(p, forces=True), cold.run(p, forces=True)
        assert rw["status"] == 0 and rc["status"] == 0
        assert abs(rw["energy"] - rc["energy"]) < 1e-9
        assert np.abs(rw["gradient"].cpu().numpy() - rc["gradient"].cpu().numpy()).max() < 1e-8
        assert np.abs(rw["charges_orbital"].cpu().numpy() - rc["charges_orbital"].cpu().numpy()).max() < 1e-8
        sw_w, sw_c = rw["records"][_lib.OUT_SWEEPS_FIRST], rc["records"][_lib.OUT_SWEEPS_FIRST]
        print(f"geometry {g}: iterations warm {rw['iterations']} cold {rc['iterations']}, "
              f"first-step sweeps warm {sw_w} cold {sw_c}", flush=True)
        if g > 0:
            # the saving is the first step's eigensolve; with random 0.03
            # bohr displacements the iteration counts stay about the same
            assert sw_w == 0 < sw_c
            assert rw["iterations"] <= rc["iterations"] + 3
    # a large jump (0.4 bohr per coordinate): whether the kept orbitals are
    # kept or dropped for a cold start, the result is a cold start's
    p = geo[-1] + 0.4 * rng.standard_normal(p0.shape)
    print("case: whole-GPU warm start after a 0.4 bohr jump", flush=True)
    rw, rc = warm.run(p, forces=True), cold.run(p, forces=True)
    assert rw["status"] == 0 and rc["status"] == 0
    assert abs(rw["energy"] - rc["energy"]) < 1e-9
    assert np.abs(rw["gradient"].cpu().numpy() - rc["gradient"].cpu().numpy()).max() < 1e-8
