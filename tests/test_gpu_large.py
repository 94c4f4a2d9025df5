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
