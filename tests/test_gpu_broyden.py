"""Broyden mixer (upstream's implicit SCF modes: broyden1 root search on
v - f(v)) in the per-molecule kernel and on the whole-GPU path, against the
oracle's restatement of the same search and against the Anderson (full
mode) fixed point."""

import numpy as np
import pytest
import torch

import oracle.gfn1 as og
from dxtb_amd import GFN1Calculator
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


def _mols():
    rng = np.random.default_rng(41)
    return [W.chain(rng, 6), W.water_cluster(rng, 6), W.chain(rng, 14), W.water_cluster(rng, 12)]


def test_broyden_kernel_vs_oracle_implicit_mode():
    """Exact eigensolves in every step: the kernel's map equals the oracle's
    to rounding, so the root search takes the oracle's steps (iteration
    counts equal) to the same result; energies, charges and forces also
    equal the full-mode (Anderson) fixed point."""
    mols = _mols()
    print("case: broyden, exact eigensolves", flush=True)
    pipe = run_native(mols, forces=True, mixer="broyden", f_atol=1e-10, eig_inexact=0.0,
                      bd_gap_kt=0.0, pd_kmax=0.0)
    rec = pipe.records().cpu().numpy()
    E = pipe.energies().cpu().numpy()
    G = pipe.gradient().cpu().numpy()
    for b, (z, p) in enumerate(mols):
        r = og.singlepoint(z, p, opts={**TIGHT_ORACLE, "scf_mode": "implicit", "f_atol": 1e-10,
                                       "broyden_memory": 24}, with_forces=True)
        a = og.singlepoint(z, p, opts=TIGHT_ORACLE)
        assert r["converged"] and int(rec[b, _lib.OUT_STATUS]) == 0
        assert int(rec[b, _lib.OUT_ITER]) == r["iterations"], (b, rec[b, _lib.OUT_ITER], r["iterations"])
        assert abs(E[b] - r["energy"]) < 1e-9
        assert abs(E[b] - a["energy"]) < 1e-9
        assert np.abs(pipe.orbital_charges(b).cpu().numpy() - r["charges_orbital"].numpy()).max() < 1e-8
        a0 = int(pipe.L.atom0[b])
        assert np.abs(G[a0:a0 + len(z)] - r["gradient"].numpy()).max() < 1e-8


def test_broyden_default_settings_and_whole_gpu_path():
    """Default (inexact, perturbative) eigensolves with the Broyden mixer
    reach the same fixed point as Anderson; the whole-GPU path runs the same
    mixer."""
    mols = _mols()
    print("case: broyden, default settings", flush=True)
    a = run_native(mols, forces=True)
    b = run_native(mols, forces=True, mixer="broyden", f_atol=1e-11)
    assert np.all(b.records().cpu().numpy()[:, _lib.OUT_STATUS] == 0)
    assert np.abs(a.energies().cpu().numpy() - b.energies().cpu().numpy()).max() < 1e-9
    assert np.abs(a.gradient().cpu().numpy() - b.gradient().cpu().numpy()).max() < 1e-8
    z, p = mols[3]
    print("case: broyden, whole-GPU path", flush=True)
    ls = LargeSystem(z, "cuda:0", opts=ScfOptions(**TIGHT, mixer="broyden", f_atol=1e-11))
    r = ls.run(p, forces=True)
    assert r["status"] == 0
    assert abs(r["energy"] - float(a.energies()[3])) < 1e-9
    assert np.abs(r["gradient"].cpu().numpy() - a.gradient().cpu().numpy()[int(a.L.atom0[3]):]).max() < 1e-8


def test_calculator_scf_modes():
    """scf_mode as upstream: implicit modes run broyden1, the full mode
    Anderson (a 'broyden' mixer falls back to it); same energies."""
    z, p = W.chain(np.random.default_rng(42), 9)
    numbers = torch.as_tensor(z, device="cuda:0")
    pos = torch.as_tensor(p, dtype=torch.float64, device="cuda:0")
    base = {"x_atol": 1e-10, "x_atol_max": 1e-10, "f_atol": 1e-10, "maxiter": 300, "exclude": ["disp"]}
    print("case: calculator scf modes", flush=True)
    res = {}
    for mode in ("full", "implicit", "nonpure"):
        calc = GFN1Calculator(numbers, opts={**base, "scf_mode": mode, "mixer": "broyden"},
                              dtype=torch.float64, device="cuda:0")
        assert calc.scf_opts.mixer == ("anderson" if mode == "full" else "broyden")
        res[mode] = (float(calc.get_energy(pos)), int(calc.get_iterations(pos)))
    assert abs(res["full"][0] - res["implicit"][0]) < 1e-9
    assert res["implicit"] == res["nonpure"]
    with pytest.raises(NotImplementedError):
        GFN1Calculator(numbers, opts={**base, "mixer": "simple"}, dtype=torch.float64, device="cuda:0")
