"""Whole-GPU large-system path (purification on matrix-core GEMMs) vs the
oracle and vs the per-molecule kernel on the same molecules."""

import numpy as np
import pytest
import torch

import oracle.gfn1 as og
from conftest import mol
from dxtb_amd import workloads as W
from dxtb_amd.engine import ScfOptions
from dxtb_amd.large import LargeSystem
from helpers import TIGHT, TIGHT_ORACLE, run_native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


@pytest.mark.parametrize("case", ["lys_xao", "water40", "chain30"])
def test_large_path_matches_oracle(fx, case):
    rng = np.random.default_rng(8)
    if case == "lys_xao":
        z, p = mol(fx, "lys_xao")
    elif case == "water40":
        z, p = W.water_cluster(rng, 40)
    else:
        z, p = W.chain(rng, 30)
    print(f"case: large path {case}", flush=True)
    ls = LargeSystem(z, "cuda:0", opts=ScfOptions(**TIGHT))
    r = ls.run(p, forces=True)
    o = og.singlepoint(z, p, opts=TIGHT_ORACLE, with_forces=True)
    assert r["converged"]
    assert abs(r["energy"] - o["energy"]) < 1e-9
    assert np.abs(r["charges_orbital"].cpu().numpy() - o["charges_orbital"].numpy()).max() < 1e-8
    assert np.abs(r["gradient"].cpu().numpy() - o["gradient"].numpy()).max() < 1e-8
    k = run_native([(z, p)], forces=True)
    assert abs(k.energies().item() - r["energy"]) < 1e-9


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
    assert abs(float(e) - float(ref["energy"])) < 1e-8
    f = calc.forces_analytical(pos)
    assert np.abs(-f.cpu().numpy() - ref["gradient"]).max() < 1e-7
    q = calc.get_charges(pos)
    assert q.shape == (4800,) and abs(float(q.sum())) < 1e-8
    qat = calc._run(pos, forces=False)["large"]["charges_atom"]
    assert np.abs(qat.cpu().numpy() - ref["charges_atom"]).max() < 1e-7
