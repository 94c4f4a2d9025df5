"""The benchmark configurations at their full sizes: the padded drop-in on
256 molecules (configuration 1) and forces on 512 molecules (configuration
2), checked against the oracle on seeded subsets and, at full size, through
properties that do not depend on the oracle: padded == ragged, and the
analytic forces against central finite differences of the GPU energies."""

import numpy as np
import pytest
import torch

import oracle.gfn1 as og
from dxtb_amd import GFN1Calculator
from dxtb_amd import workloads as W
from helpers import TIGHT_ORACLE, run_native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def test_config1_padded_dropin_vs_ragged_and_oracle():
    """256 CHNO chains (2-20 C) as one padded batch through the calculator
    (atomic number 0 pads), against the ragged pipeline on all 256 and the
    oracle on 6 seeded molecules: energies 1e-9 Eh, orbital charges 1e-8 e."""
    mols = W.config1()
    nums, pos = W.pad(mols)
    dev = torch.device("cuda:0")
    print(f"case: config 1, padded batch {nums.shape}", flush=True)
    calc = GFN1Calculator(torch.tensor(nums, device=dev), dtype=torch.float64, device=dev,
                          opts={**TIGHT_ORACLE, "exclude": ["disp"]})
    p = torch.tensor(pos, dtype=torch.float64, device=dev)
    e = calc.get_energy(p).cpu().numpy()
    q = calc.get_charges(p).cpu().numpy()
    assert e.shape == (256,)
    rag = run_native(mols)
    assert np.abs(e - rag.energies().cpu().numpy()).max() < 1e-10
    rng = np.random.default_rng(21)
    for b in rng.choice(len(mols), 6, replace=False):
        z, x = mols[b]
        r = og.singlepoint(z, x, opts=TIGHT_ORACLE)
        assert abs(e[b] - r["energy"]) < 1e-9
        nao = r["charges_orbital"].shape[0]
        assert np.abs(q[b, :nao] - r["charges_orbital"].numpy()).max() < 1e-8
        assert np.all(q[b, nao:] == 0.0)


def test_config2_forces_finite_differences_and_oracle():
    """512 chains (5-40 C): analytic forces of the ragged batch against
    central differences (h = 1e-4 bohr) of the GPU energies on 16 seeded
    molecules (4 seeded coordinates each), and against the oracle's forces
    on 4 seeded molecules; all within 1e-7 Eh/bohr."""
    mols = W.config2()
    print(f"case: config 2, {len(mols)} molecules, forces", flush=True)
    pipe = run_native(mols, forces=True)
    g = pipe.gradient().cpu().numpy()
    assert int(pipe.records()[:, 6].abs().max().item()) == 0
    a0 = pipe.L.atom0
    rng = np.random.default_rng(22)
    pick = rng.choice(len(mols), 16, replace=False)
    h = 1e-4
    disp, where = [], []
    for b in pick:
        z, x = mols[b]
        for _ in range(4):
            a, c = int(rng.integers(len(z))), int(rng.integers(3))
            for s in (1.0, -1.0):
                y = x.copy()
                y[a, c] += s * h
                disp.append((z, y))
            where.append((b, a, c))
    print(f"case: config 2, {len(disp)} displaced geometries", flush=True)
    fd = run_native(disp).energies().cpu().numpy()
    err = 0.0
    for i, (b, a, c) in enumerate(where):
        d = (fd[2 * i] - fd[2 * i + 1]) / (2 * h)
        err = max(err, abs(d - g[int(a0[b]) + a, c]))
    assert err < 1e-7, err
    small = [b for b in rng.permutation(len(mols)) if len(mols[b][0]) <= 40][:4]
    for b in small:
        z, x = mols[b]
        r = og.singlepoint(z, x, opts=TIGHT_ORACLE, with_forces=True)
        assert np.abs(g[int(a0[b]):int(a0[b]) + len(z)] - r["gradient"].numpy()).max() < 1e-7


def test_command_line_matches_the_calculator(capsys):
    """`python -m dxtb_amd` on upstream's H2O coord and NO2 xyz files (one
    padded batch, --grad, JSON output) prints the calculator's energies and
    gradient; the batch holds an open shell (NO2, 23 electrons)."""
    import json
    from pathlib import Path

    from dxtb_amd import cli, io

    g = Path(__file__).resolve().parent / "golden" / "mols"
    print("case: command line, H2O coord", flush=True)
    assert cli.main([str(g / "H2O.coord"), "--grad", "--json"]) == 0
    out = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    z, p = io.read(g / "H2O.coord")
    calc = GFN1Calculator(torch.as_tensor(z, device="cuda:0"), device="cuda:0", dtype=torch.float64)
    pos = torch.as_tensor(p, dtype=torch.float64, device="cuda:0").requires_grad_(True)
    f = calc.get_forces(pos, 0)
    e = calc.get_energy(pos.detach(), 0)
    assert abs(out["energy"] - float(e)) < 1e-12
    assert np.abs(np.array(out["gradient"]) + f.detach().cpu().numpy()).max() < 1e-12
    r = og.singlepoint(z, p, opts=TIGHT_ORACLE)
    assert abs(out["energy"] - float(r["energy"])) < 1e-8
    print("case: command line, two files as a batch", flush=True)
    assert cli.main([str(g / "H2O.coord"), str(g / "NO2.xyz"), "--forces"]) == 0
    text = capsys.readouterr().out
    assert text.count("total energy") == 2 and text.count("Forces") == 2


def test_config4_stream_slice_vs_oracle():
    """Configuration 4 (serving stream, seed 4): the first 300 requests plus
    the largest chain and the largest water cluster of the first 2000 (about
    200 atoms, n ~ 530), served through the ragged continuous-batching loop;
    energies and forces of 8 seeded requests (the two largest included)
    against the oracle: 1e-9 Eh, 1e-8 Eh/bohr."""
    from dxtb_amd.engine import ScfOptions
    from dxtb_amd.serving import RaggedServer, Request

    stream = W.config4(seed=4, count=2000)
    big_chain = max((i for i in range(len(stream)) if 8 not in stream[i][1]),
                    key=lambda i: len(stream[i][1]))
    big_water = max((i for i in range(len(stream)) if 8 in stream[i][1] and 6 not in stream[i][1]),
                    key=lambda i: len(stream[i][1]))
    ids = list(range(300)) + [big_chain, big_water]
    print(f"case: config 4 slice, {len(ids)} requests, largest {len(stream[big_chain][1])} "
          f"and {len(stream[big_water][1])} atoms", flush=True)
    reqs = [Request(i, stream[i][1], stream[i][2]) for i in ids]
    srv = RaggedServer("cuda:0", opts=ScfOptions(**{k: v for k, v in TIGHT_ORACLE.items()}))
    srv.run_stream(reqs, realtime=False)
    assert all(srv.results[i].status == 0 for i in ids)
    rng = np.random.default_rng(41)
    check = [big_chain, big_water] + [int(i) for i in rng.choice(300, 6, replace=False)]
    for i in check:
        z, x = stream[i][1], stream[i][2]
        print(f"case: config 4 request {i}, {len(z)} atoms, oracle", flush=True)
        r = og.singlepoint(z, x, opts=TIGHT_ORACLE, with_forces=True)
        got = srv.results[i]
        assert abs(got.energy - r["energy"]) < 1e-9, i
        assert np.abs(got.forces + r["gradient"].numpy()).max() < 1e-8, i
