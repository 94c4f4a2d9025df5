"""The fused single point (xtb_singlepoint: each molecule's integral pass,
SCF and gradient by its own workgroup, one launch per batch) against the
three-launch path (xtb_build_integrals, xtb_scf, xtb_forces) and the
oracle.  The integral pass and the SCF run the same per-element code, so
energies, charges and orbital energies agree bit for bit; the gradient's
per-atom sums run in a fixed order instead of atomics, so it agrees to
rounding and is the same bits on every run."""

import numpy as np
import pytest
import torch

import oracle.gfn1 as og
from conftest import mol
from dxtb_amd import _lib
from dxtb_amd import workloads as W
from helpers import TIGHT_ORACLE, run_native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _batch(fx):
    rng = np.random.default_rng(61)
    return [mol(fx, "h2o"), mol(fx, "ch4"), mol(fx, "lys_xao"), W.water_cluster(rng, 25),
            W.chain(rng, 12), W.chain(rng, 40), W.water_cluster(rng, 3), mol(fx, "h2")]


@pytest.mark.parametrize("padded", [False, True])
def test_fused_equals_separate_launches(fx, padded):
    mols = _batch(fx)
    print(f"case: fused vs separate launches, padded={padded}", flush=True)
    a = run_native(mols, padded=padded, forces=True)
    b = run_native(mols, padded=padded, forces=True, fused=True)
    ra, rb = a.records().cpu().numpy(), b.records().cpu().numpy()
    assert np.all(rb[:, _lib.OUT_STATUS] == 0)
    for k in (_lib.OUT_ESCF, _lib.OUT_FENERGY, _lib.OUT_EREP, _lib.OUT_ITER):
        assert np.array_equal(ra[:, k], rb[:, k]), k
    assert torch.equal(a.cn, b.cn)
    assert torch.equal(a.q, b.q)
    ga, gb = a.gradient().cpu().numpy(), b.gradient().cpu().numpy()
    assert np.abs(ga - gb).max() < 1e-12
    assert np.all(rb[:, _lib.OUT_T_INTEGRALS] > 0) and np.all(rb[:, _lib.OUT_T_FORCES] > 0)


def test_fused_gradient_vs_oracle_and_reproducible(fx):
    """Energies and forces of the fused path against the oracle (1e-9 Eh,
    1e-8 Eh/bohr); a second run gives the same gradient bits."""
    mols = _batch(fx)
    print("case: fused vs oracle, two runs", flush=True)
    p1 = run_native(mols, forces=True, fused=True)
    g1 = p1.gradient().cpu().numpy()
    e1 = p1.energies().cpu().numpy()
    p2 = run_native(mols, forces=True, fused=True)
    assert np.array_equal(g1, p2.gradient().cpu().numpy())
    a0 = p1.L.atom0
    for b in (0, 2, 3, 5):
        z, x = mols[b]
        o = og.singlepoint(z, x, opts=TIGHT_ORACLE, with_forces=True)
        assert abs(e1[b] - o["energy"]) < 1e-9
        assert np.abs(g1[a0[b]:a0[b + 1]] - o["gradient"].numpy()).max() < 1e-8


def test_fused_charged_open_shell_and_energy_only():
    """A cation, a triplet and a neutral molecule in one fused launch against
    the oracle, then the same batch without forces (no gradient buffers
    touched: energies unchanged)."""
    rng = np.random.default_rng(62)
    mols = [W.chain(rng, 10, p_n=0.0, p_o=0.0), W.chain(rng, 8, p_n=0.0, p_o=0.0), W.water_cluster(rng, 4)]
    chrg, spin = [1.0, 0.0, 0.0], [1.0, 2.0, 0.0]
    print("case: fused, charged and open shell", flush=True)
    p = run_native(mols, forces=True, fused=True, chrg=chrg, spin=spin)
    e, g = p.energies().cpu().numpy(), p.gradient().cpu().numpy()
    a0 = p.L.atom0
    for b, (z, x) in enumerate(mols):
        o = og.singlepoint(z, x, chrg=chrg[b], spin=spin[b], opts=TIGHT_ORACLE, with_forces=True)
        assert abs(e[b] - o["energy"]) < 1e-9
        assert np.abs(g[a0[b]:a0[b + 1]] - o["gradient"].numpy()).max() < 1e-8
    q = run_native(mols, forces=False, fused=True, chrg=chrg, spin=spin)
    assert np.array_equal(q.energies().cpu().numpy(), e)


def test_fused_overlap_failure_gives_nan_gradient():
    """A molecule whose overlap is not positive definite (H atoms 0.102 bohr
    apart) gets status OVERLAP and a NaN gradient; its batch neighbour is
    unaffected."""
    line = np.zeros((12, 3))
    line[:, 2] = 0.102 * np.arange(12)
    good = W.chain(np.random.default_rng(4), 3)
    print("case: fused, overlap failure", flush=True)
    p = run_native([(np.ones(12, dtype=np.int64), line), good], forces=True, fused=True)
    st = p.records().cpu().numpy()[:, _lib.OUT_STATUS].astype(int)
    assert st[0] & _lib.MOL_OVERLAP and st[1] == 0
    g = p.gradient().cpu().numpy()
    assert np.all(np.isnan(g[:12])) and np.all(np.isfinite(g[12:]))
    ref = run_native([good], forces=True)
    assert np.abs(g[12:] - ref.gradient().cpu().numpy()).max() < 1e-12
