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


def test_overlap_not_positive_definite_raises():
    """Twelve H atoms 0.102 bohr apart on a line: the overlap matrix has an
    eigenvalue of -7e-10 in FP64, so upstream's Cholesky reduction raises.
    The SCF kernel skips such a molecule (status bit OVERLAP) and the
    calculator raises torch.linalg.LinAlgError; a serving request is marked
    failed with NaN results.  A normal molecule in the same batch is
    unaffected."""
    dev = torch.device("cuda:0")
    line = np.zeros((12, 3))
    line[:, 2] = 0.102 * np.arange(12)
    good = W.chain(np.random.default_rng(4), 3)
    print("case: overlap not positive definite", flush=True)
    pipe = run_native([(np.ones(12, dtype=np.int64), line), good], forces=True, **OPTS)
    st = pipe.records().cpu().numpy()[:, 6].astype(int)
    assert st[0] & 16 and st[1] == 0
    calc = GFN1Calculator(torch.ones(12, dtype=torch.int64, device=dev),
                          opts={**OPTS, "exclude": ["disp"]}, dtype=torch.float64)
    with pytest.raises(torch.linalg.LinAlgError):
        calc.get_energy(torch.tensor(line, device=dev))
    srv = RaggedServer("cuda:0", opts=ScfOptions(**OPTS), streams=2)
    reqs = [Request(0, np.ones(12, dtype=np.int64), line), Request(1, *good)]
    srv.run_stream(reqs, realtime=False)
    assert srv.results[0].failed and np.isnan(srv.results[0].energy)
    assert np.all(np.isnan(srv.results[0].forces))
    assert not srv.results[1].failed and np.isfinite(srv.results[1].energy)


def test_excluded_components_with_autograd():
    """opts["exclude"] with positions that require grad (upstream's autograd
    covers any component selection): without the repulsion, without the
    electronic energy, and without both, against the oracle's pieces."""
    import oracle.gfn1 as og

    dev = torch.device("cuda:0")
    z, p = W.chain(np.random.default_rng(13), 6)
    r = og.singlepoint(z, p, opts={"x_atol": 1e-10, "x_atol_max": 1e-10, "maxiter": 300},
                       with_forces=True)
    _, grep = og.repulsion(list(z), torch.tensor(p), grad=True)
    want = {
        "rep": (r["energy"] - r["repulsion"], r["gradient"] - grep),
        "scf": (r["repulsion"], grep),
        "all": (0.0, torch.zeros_like(grep)),
    }
    print("case: excluded components with autograd", flush=True)
    for ex, (e_ref, g_ref) in want.items():
        calc = GFN1Calculator(torch.tensor(z, device=dev), opts={**OPTS, "exclude": ["disp", ex]},
                              dtype=torch.float64)
        pos = torch.tensor(p, dtype=torch.float64, device=dev, requires_grad=True)
        e = calc.get_energy(pos)
        assert abs(e.item() - e_ref) < 1e-9, ex
        f = calc.get_forces(pos)
        assert (f.cpu() + g_ref).abs().max().item() < 1e-8, ex


def test_result_cache_checks_positions_by_value():
    """With cache_enabled a run is reused only for equal positions, charges
    and spin: a new tensor that happens to reuse a freed tensor's storage
    gets its own result.  Without it (upstream's default) nothing is
    reused."""
    dev = torch.device("cuda:0")
    z, p = W.chain(np.random.default_rng(14), 5)
    numbers = torch.tensor(z, device=dev)
    print("case: result cache", flush=True)
    for cache in (True, False):
        calc = GFN1Calculator(numbers, opts={**OPTS, "exclude": ["disp"], "cache_enabled": cache},
                              dtype=torch.float64)
        fresh = GFN1Calculator(numbers, opts={**OPTS, "exclude": ["disp"]}, dtype=torch.float64)
        for k in range(6):
            q = p.copy()
            q[0, 0] += 0.01 * k  # a distortion (a translation would not change E)
            pos = torch.tensor(q, dtype=torch.float64, device=dev)
            e = calc.get_energy(pos).item()
            assert abs(e - fresh.get_energy(pos.clone()).item()) < 1e-12
            assert abs(calc.get_energy(pos).item() - e) < 1e-12
            del pos


def test_request_above_the_largest_class_takes_the_whole_gpu_lane():
    """A request above the largest size class (1024 orbitals) is served by
    the whole-GPU path from a worker thread, while the loop keeps serving
    the others: its result equals that path's alone, theirs equal their
    own alone."""
    from dxtb_amd.large import LargeSystem

    rng = np.random.default_rng(21)
    big = W.water_cluster(rng, 130)  # 1040 orbitals
    stream = W.config4(seed=13, count=16)
    reqs = [Request(0, big[0], big[1], arrival=0.0)]
    reqs += [Request(i + 1, z, p, arrival=0.0) for i, (_, z, p) in enumerate(stream)]
    print("case: serving with a 1040-orbital request", flush=True)
    srv = RaggedServer("cuda:0", opts=ScfOptions(**OPTS), streams=4)
    srv.run_stream(reqs, realtime=False)
    assert srv.stats.large_runs == 1 and srv.stats.completed == len(reqs)
    ref = LargeSystem(big[0], device="cuda:0", opts=ScfOptions(**OPTS)).run(big[1], forces=True)
    got = srv.results[0]
    assert got.status == 0 and ref["status"] == 0
    assert abs(got.energy - ref["energy"]) < 1e-10
    assert np.abs(got.forces + ref["gradient"].cpu().numpy()).max() < 1e-10
    for r in reqs[1:7]:
        alone = run_native([(r.numbers, r.positions)], forces=True, **OPTS)
        assert abs(srv.results[r.rid].energy - alone.energies().item()) < 1e-11
