"""Serving loop: every request's result is the same alone and in the mix,
and the calculator drop-in keeps upstream's API semantics."""

import numpy as np
import pytest
import torch

from dxtb_amd import GFN1Calculator, workloads as W
from dxtb_amd.engine import ScfOptions
from dxtb_amd.serving import RaggedServer, Request, padded_dropin
from helpers import run_native

pytestmark = pytest.mark.gpu

OPTS = dict(x_atol=1e-10, x_atol_max=1e-10, maxiter=300)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def test_alone_equals_in_mix():
    stream = W.config4(seed=11, count=40)
    print("case: serving alone vs mix, 40 requests seed 11", flush=True)
    reqs = [Request(i, z, p, arrival=0.0) for i, (_, z, p) in enumerate(stream)]
    srv = RaggedServer("cuda:0", opts=ScfOptions(**OPTS), streams=4)
    srv.run_stream(reqs, realtime=False)
    for r in reqs[:12]:
        alone = run_native([(r.numbers, r.positions)], forces=True, **OPTS)
        e = alone.energies().item()
        f = -alone.gradient().cpu().numpy()
        got = srv.results[r.rid]
        assert got.status == 0
        assert abs(got.energy - e) < 1e-11
        assert np.abs(got.forces - f).max() < 1e-11


def test_padded_dropin_matches_ragged():
    stream = W.config4(seed=12, count=24)
    reqs = [Request(i, z, p) for i, (_, z, p) in enumerate(stream)]
    print("case: padded drop-in vs ragged", flush=True)
    srv = RaggedServer("cuda:0", opts=ScfOptions(**OPTS))
    srv.run_stream([Request(r.rid, r.numbers, r.positions) for r in reqs], realtime=False)
    padded_dropin(reqs, 8, "cuda:0", opts=ScfOptions(**OPTS))
    for r in reqs:
        assert abs(r.energy - srv.results[r.rid].energy) < 1e-10


def test_calculator_api_semantics():
    mols = [W.config0(), W.chain(np.random.default_rng(3), 4)]
    nums, pos = W.pad(mols)
    dev = torch.device("cuda:0")
    numbers = torch.tensor(nums, device=dev)
    positions = torch.tensor(pos, dtype=torch.float64, device=dev)
    calc = GFN1Calculator(numbers, opts={**OPTS, "exclude": ["disp"]}, dtype=torch.float64)
    e = calc.get_energy(positions)
    assert e.shape == (2,)
    with pytest.raises(RuntimeError):
        calc.get_forces(positions)  # upstream requires requires_grad (decorators.py:63-82)
    with pytest.raises(ValueError):
        calc.get_energy(positions, chrg=1)  # scalar non-zero charge in a batch
    p = positions.clone().requires_grad_(True)
    f = calc.get_forces(p)
    assert f.shape == positions.shape
    assert torch.all(f[1, len(mols[1][0]):] == 0)  # padding atoms
    # single molecule == its row of the batch
    z0 = torch.tensor(mols[0][0], device=dev)
    c0 = GFN1Calculator(z0, opts={**OPTS, "exclude": ["disp"]}, dtype=torch.float64)
    p0 = torch.tensor(mols[0][1], dtype=torch.float64, device=dev, requires_grad=True)
    assert abs(c0.get_energy(p0).item() - e[0].item()) < 1e-10
    assert (c0.get_forces(p0) - f[0, : len(mols[0][0])]).abs().max().item() < 1e-9
    q = calc.get_charges(positions)
    assert q.shape[0] == 2
    # energy via autograd backward equals -forces
    p2 = positions.clone().requires_grad_(True)
    calc.reset()
    calc.get_energy(p2).sum().backward()
    assert (p2.grad + f).abs().max().item() < 1e-12


def test_staged_results_equal_plain_copies():
    """Results copied to pinned memory behind each launch (stage_results)
    equal those copied at retirement."""
    stream = W.config4(seed=12, count=30)
    print("case: serving, staged vs plain result copies, 30 requests seed 12", flush=True)
    out = []
    for staged in (False, True):
        reqs = [Request(i, z, p, arrival=0.0) for i, (_, z, p) in enumerate(stream)]
        srv = RaggedServer("cuda:0", opts=ScfOptions(**OPTS), streams=3, stage_results=staged)
        srv.run_stream(reqs, realtime=False)
        out.append(srv.results)
    for rid, r in out[0].items():
        s = out[1][rid]
        assert r.status == 0 and s.status == 0
        assert abs(r.energy - s.energy) < 1e-12
        assert np.abs(r.forces - s.forces).max() < 1e-12
