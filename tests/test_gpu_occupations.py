"""Charged, open-shell and fractionally occupied molecules: the HIP path
against upstream's tblite values and the CPU oracle (FP64).

The neutral closed shells of the other parity tests leave the kernel's
occupation logic untested where it matters: alpha/beta channels of
different size (singly occupied class, spin-resolved Fermi levels), charged
electron counts, fractional Fermi-Dirac occupations with a non-zero -TS, and
the aufbau filling at zero temperature.  Tolerances: |dE| <= 1e-9 Eh
against the oracle (tblite: 5e-9, its own convergence), |dq| <= 1e-8 e,
|dF| <= 1e-8 Eh/bohr -- all below the 1e-8 / 1e-7 / 1e-7 targets.
"""

import numpy as np
import pytest
import torch

import oracle.gfn1 as og
from conftest import mol
from dxtb_amd import workloads as W
from helpers import TIGHT_ORACLE, run_native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _compare(pipe, mols, chrg, spin, opts, e_tol=1e-9, q_tol=1e-8, f_tol=1e-8, forces=True):
    """Every molecule of a finished run against the oracle; returns the
    oracle results."""
    rec = pipe.records().cpu().numpy()
    E = pipe.energies().cpu().numpy()
    G = pipe.gradient().cpu().numpy() if forces else None
    out = []
    for b, (z, p) in enumerate(mols):
        r = og.singlepoint(z, p, chrg=float(chrg[b]), spin=None if spin is None else spin[b],
                           opts=opts, with_forces=forces)
        assert r["converged"]
        assert int(rec[b, 6]) == 0, f"molecule {b}: status {rec[b, 6]}"
        assert abs(E[b] - r["energy"]) < e_tol, (b, E[b], r["energy"])
        assert abs(rec[b, 4] - r["fenergy"]) < e_tol, (b, rec[b, 4], r["fenergy"])
        q = pipe.orbital_charges(b).cpu().numpy()
        assert np.abs(q - r["charges_orbital"].numpy()).max() < q_tol, b
        if forces:
            a0 = int(pipe.L.atom0[b])
            assert np.abs(G[a0:a0 + len(z)] - r["gradient"].numpy()).max() < f_tol, b
        out.append(r)
    return out


def test_atoms_and_ions_vs_tblite(fx):
    """H, C, N, O atoms, cations and anions in one ragged batch, spins from
    upstream's uhf tables: partly filled degenerate p shells give fractional
    occupations (1/3, 2/3) and -TS ~ 3.6e-3 Eh; tblite's free energies
    (test_scf/test_elements_gfn1.py) pin E_scf - TS."""
    zs, ch, sp, ref = fx["atoms_z"], fx["atoms_chrg"], fx["atoms_spin"], fx["atoms_escf"]
    mols = [(np.array([int(z)]), np.zeros((1, 3))) for z in zs]
    # automatic spin (-1) is the parity of the electron count, as upstream
    spin = [int(s) if s >= 0 else None for s in sp]
    nel = [sum(og._n0(og.build_basis([int(z)])).tolist()) - c for z, c in zip(zs, ch)]
    spin_arr = [s if s is not None else int(round(n)) % 2 for s, n in zip(spin, nel)]
    opts = dict(fermi_thresh=1e-10)
    print("case: atoms and ions", flush=True)
    pipe = run_native(mols, forces=True, chrg=ch, spin=spin_arr, **opts)
    rec = pipe.records().cpu().numpy()
    for b in range(len(mols)):
        assert abs(rec[b, 0] + rec[b, 4] - float(ref[b])) < 5e-9, (int(zs[b]), float(ch[b]))
    r = _compare(pipe, mols, ch, spin, {**TIGHT_ORACLE, "fermi_thresh": 1e-10})
    assert min(x["fenergy"] for x in r) < -1e-3  # fractional occupations were exercised


def test_charged_molecule_vs_tblite_and_oracle(fx):
    """AD7en+ (C14H25+, charge +1): tblite E_scf (test_scf/samples_charged.py)
    and the oracle's energy, charges and forces; with a neutral and a
    charged chain in the same ragged batch."""
    rng = np.random.default_rng(21)
    mols = [mol(fx, "ad7en"), W.chain(rng, 9), W.chain(rng, 14)]
    chrg = [float(fx["ad7en_chrg"]), 0.0, -2.0]
    print("case: AD7en+ and charged chains", flush=True)
    pipe = run_native(mols, forces=True, chrg=chrg)
    rec = pipe.records().cpu().numpy()
    assert abs(rec[0, 0] - float(fx["ad7en_escf"])) < 1e-9
    _compare(pipe, mols, chrg, None, TIGHT_ORACLE)


def _radical_chain(rng, nc):
    """A chain with its last hydrogen removed: an odd electron count."""
    z, p = W.chain(rng, nc, p_n=0.0, p_o=0.0)
    return z[:-1], p[:-1]


def test_open_shell_molecules_vs_oracle(fx):
    """Alpha and beta channels of different size: NO2 doublet and quartet
    (upstream's coord file), triplet O2, a doublet chain radical, a doublet
    chain cation, a triplet chain and a triplet water trimer."""
    rng = np.random.default_rng(22)
    no2 = mol(fx, "no2")
    o2 = (np.array([8, 8]), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.28]]))
    mols = [no2, no2, o2, _radical_chain(rng, 7), W.chain(rng, 10, p_n=0.0, p_o=0.0),
            W.chain(np.random.default_rng(7), 6, p_n=0.0, p_o=0.0),
            W.water_cluster(np.random.default_rng(5), 3)]
    chrg = [0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0]
    spin = [1, 3, 2, 1, 1, 2, 2]
    print("case: open shells", flush=True)
    pipe = run_native(mols, forces=True, chrg=chrg, spin=spin)
    _compare(pipe, mols, chrg, spin, TIGHT_ORACLE)


@pytest.mark.parametrize("etemp", [5000.0, 20000.0])
def test_fractional_occupations_at_high_temperature(etemp):
    """Fermi smearing at 5000 and 20000 K: occupations well away from 0/1
    around the Fermi level and a sizeable -TS, through the kernel's Fermi
    Newton iteration, the density over fractional weights and the
    energy-weighted density of the forces."""
    rng = np.random.default_rng(23)
    mols = [W.chain(rng, 8), W.water_cluster(rng, 8), W.chain(rng, 24), _radical_chain(rng, 5)]
    chrg = [0.0] * len(mols)
    print(f"case: fractional occupations at {etemp} K", flush=True)
    pipe = run_native(mols, forces=True, fermi_etemp=etemp)
    r = _compare(pipe, mols, chrg, None, {**TIGHT_ORACLE, "fermi_etemp": etemp})
    assert max(-x["fenergy"] for x in r) > (1e-5 if etemp < 1e4 else 1e-3)


def test_zero_temperature_aufbau(fx):
    """fermi_etemp = 0 (kT below 3e-7): aufbau filling of the unrounded
    alpha/beta counts (scf/base.py:888-889) -- a chain, CH4 (its triply
    degenerate HOMO filled), a water cluster and a doublet radical; -TS is 0."""
    rng = np.random.default_rng(24)
    mols = [W.chain(rng, 10), mol(fx, "ch4"), W.water_cluster(rng, 5), _radical_chain(rng, 6)]
    chrg = [0.0] * len(mols)
    print("case: zero temperature", flush=True)
    pipe = run_native(mols, forces=True, fermi_etemp=0.0)
    r = _compare(pipe, mols, chrg, None, {**TIGHT_ORACLE, "fermi_etemp": 0.0})
    assert all(x["fenergy"] == 0.0 for x in r)
    assert np.all(pipe.records().cpu().numpy()[:, 4] == 0.0)


def test_padded_open_shell_and_charged(fx):
    """The padded layout (upstream's batch semantics: padding orbitals are
    levels at energy 0 that share the Fermi filling and -TS, and the mask
    that doubles the Fermi level's start value, filling.py:201-366) for
    charged and open-shell molecules.  At 300 K the padding levels hold no
    electrons and padded == ragged; at 5000 K they hold ~1e-6 each and the
    padded results move by ~1e-7 Eh, which the oracle's padded-batch
    emulation (npad zero levels) reproduces."""
    rng = np.random.default_rng(25)
    mols = [mol(fx, "no2"), mol(fx, "ad7en"), _radical_chain(rng, 4), W.water_cluster(rng, 4)]
    chrg = [0.0, 1.0, 0.0, 0.0]
    spin = [1, 0, 1, 0]
    print("case: padded vs ragged, open shells, 300 K", flush=True)
    a = run_native(mols, padded=False, forces=True, chrg=chrg, spin=spin, fermi_etemp=300.0)
    b = run_native(mols, padded=True, forces=True, chrg=chrg, spin=spin, fermi_etemp=300.0)
    assert np.abs(a.energies().cpu().numpy() - b.energies().cpu().numpy()).max() < 1e-10
    assert np.abs(a.gradient().cpu().numpy() - b.gradient().cpu().numpy()).max() < 1e-9
    print("case: padded vs oracle with padding levels, open shells, 5000 K", flush=True)
    b = run_native(mols, padded=True, forces=True, chrg=chrg, spin=spin, fermi_etemp=5000.0)
    nmax = int(b.L.n.max())
    E = b.energies().cpu().numpy()
    rec = b.records().cpu().numpy()
    G = b.gradient().cpu().numpy()
    moved = 0.0
    for k, (z, p) in enumerate(mols):
        npad = nmax - int(b.L.nao[k])
        r = og.singlepoint(z, p, chrg=chrg[k], spin=spin[k], with_forces=True, npad=npad,
                           opts={**TIGHT_ORACLE, "fermi_etemp": 5000.0})
        u = og.singlepoint(z, p, chrg=chrg[k], spin=spin[k],
                           opts={**TIGHT_ORACLE, "fermi_etemp": 5000.0})
        assert int(rec[k, 6]) == 0
        assert abs(E[k] - r["energy"]) < 1e-9, (k, E[k], r["energy"])
        a0 = int(b.L.atom0[k])
        assert np.abs(G[a0:a0 + len(z)] - r["gradient"].numpy()).max() < 1e-8
        assert np.abs(b.orbital_charges(k).cpu().numpy() - r["charges_orbital"].numpy()).max() < 1e-8
        moved = max(moved, abs(r["energy"] - u["energy"]))
    assert moved > 1e-9  # the padding levels do take part
