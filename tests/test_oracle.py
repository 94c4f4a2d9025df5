"""The CPU oracle against upstream's golden vectors and against mathematics.

Each check is chosen so a plausible mistake fails it: a wrong primitive
normalisation or p-orbital order fails the tblite pair values and the
overlap fixture; a wrong radius, CN or unit constant fails the H0 fixture or
the tblite SCF energies (which discriminate CODATA 2010/2018 Hartree-eV at
1e-8 relative); a sign or factor slip in the gradient fails both the tblite
SCF gradients and central finite differences.
"""

import math

import numpy as np
import pytest
import torch

import oracle.gfn1 as og
from conftest import mol


def test_pair_overlaps_match_tblite(fx):
    for k in ("1s1s", "2s1s", "3s4s", "2p1s", "2p2p"):
        (ngi, ni, li), (ngj, nj, lj) = fx[f"pair_{k}_cgto"]
        ai, ci = og.slater_to_gauss(int(ngi), int(ni), int(li), 1.0)
        aj, cj = og.slater_to_gauss(int(ngj), int(nj), int(lj), 1.0)
        # upstream's vec enters md as (R_j - R_i); our rvec = R_A - R_B
        rvec = -torch.tensor(fx[f"pair_{k}_vec"])
        s, ds = og.shell_pair_overlap(int(li), int(lj), ai, ci, aj, cj, rvec, grad=True)
        assert np.allclose(s.numpy(), fx[f"pair_{k}_s"], atol=1e-5), k
        # fixture rows: flattened (mi, mj) pairs; columns: derivative with
        # respect to the x, y, z components of R_A - R_B
        g = ds.permute(1, 2, 0).reshape(-1, 3).numpy()
        assert np.allclose(g, fx[f"pair_{k}_g"], atol=1e-5), k


@pytest.mark.parametrize("key", ["h2", "h2o", "ch4", "lys_xao"])
def test_overlap_and_h0_fixtures(fx, key):
    z, p = mol(fx, key)
    bas = og.build_basis(z)
    pos = torch.tensor(p)
    S = og.overlap(bas, pos)
    cn = og.coordination_number(list(z), pos)
    H = og.core_hamiltonian(bas, pos, S, cn)
    # fixtures are float32: 2e-8 relative on O(1) entries
    assert np.abs(S.numpy() - fx[f"{key}_overlap"]).max() < 5e-8
    assert np.abs(H.numpy() - fx[f"{key}_h0"]).max() < 1e-7


def test_normalisation_and_orthogonality():
    bas = og.build_basis([1, 6])
    pos = torch.tensor([[0.0, 0.0, 0.0], [0.0, 0.0, 40.0]])
    # the STO-nG tables are normalised to ~1e-9 (as upstream uses them)
    for a, c, l in zip(bas.shell_alpha, bas.shell_coeff, bas.shell_l):
        s = og.shell_pair_overlap(l, l, a, c, a, c, torch.zeros(3))
        assert (torch.diagonal(s) - 1.0).abs().max() < 1e-8
    # H 2s is explicitly renormalised after Gram-Schmidt: exact to rounding
    s22 = og.shell_pair_overlap(0, 0, bas.shell_alpha[1], bas.shell_coeff[1],
                                bas.shell_alpha[1], bas.shell_coeff[1], torch.zeros(3))
    assert abs(float(s22) - 1.0) < 1e-14
    s12 = og.shell_pair_overlap(0, 0, bas.shell_alpha[0], bas.shell_coeff[0],
                                bas.shell_alpha[1], bas.shell_coeff[1], torch.zeros(3))
    assert abs(float(s12)) < 1e-8


def test_overlap_gradient_finite_difference(fx):
    z, p = mol(fx, "h2o")
    bas = og.build_basis(z)
    pos = torch.tensor(p)
    dS = og.overlap_grad(bas, pos)
    ao_atom = [bas.shell_atom[s] for s in bas.ao_shell]
    h = 1e-5
    for a in range(len(z)):
        for x in range(3):
            pp, pm = pos.clone(), pos.clone()
            pp[a, x] += h
            pm[a, x] -= h
            fd = (og.overlap(bas, pp) - og.overlap(bas, pm)) / (2 * h)
            for i in range(bas.nao):
                if ao_atom[i] != a:
                    continue
                for j in range(bas.nao):
                    if ao_atom[j] == a:
                        continue
                    assert abs(float(fd[i, j] - dS[i, j, x])) < 1e-8


TIGHT = {"x_atol": 1e-11, "x_atol_max": 1e-11, "maxiter": 300}


@pytest.mark.parametrize("key", ["h2", "h2o", "ch4", "lys_xao"])
def test_scf_energy_matches_tblite(fx, key):
    z, p = mol(fx, key)
    r = og.singlepoint(z, p, opts=TIGHT)
    assert r["converged"]
    assert abs(r["scf"] - float(fx[f"{key}_escf"])) < 1e-9


@pytest.mark.parametrize("key", ["h2", "h2o", "lys_xao"])
def test_repulsion_matches_tblite(fx, key):
    z, p = mol(fx, key)
    e = og.repulsion(list(z), torch.tensor(p))
    assert abs(float(e.sum()) - float(fx[f"{key}_erep"])) < 1e-12
    if key == "h2o":
        _, g = og.repulsion(list(z), torch.tensor(p), grad=True)
        assert np.abs(g.numpy() - fx["h2o_grep"]).max() < 1e-12


@pytest.mark.parametrize("key", ["h2", "h2o", "ch4"])
def test_scf_gradient_matches_tblite(fx, key):
    z, p = mol(fx, key)
    r = og.singlepoint(z, p, opts=TIGHT, with_forces=True)
    _, grep = og.repulsion(list(z), torch.tensor(p), grad=True)
    g = (r["gradient"] - grep).numpy()
    # tblite gradients are float32 and SCF-converged loosely: 1e-6
    assert np.abs(g - fx[f"{key}_gscf"]).max() < 1e-6


def test_forces_central_difference(fx):
    z, p = mol(fx, "h2o")
    r = og.singlepoint(z, p, opts=TIGHT, with_forces=True)
    h = 1e-4
    for a in range(len(z)):
        for x in range(3):
            pp, pm = p.copy(), p.copy()
            pp[a, x] += h
            pm[a, x] -= h
            fd = (og.singlepoint(z, pp, opts=TIGHT)["energy"]
                  - og.singlepoint(z, pm, opts=TIGHT)["energy"]) / (2 * h)
            assert abs(fd - float(r["gradient"][a, x])) < 1e-8


def _free_energy(f, kt):
    """-TS of occupations f (base.py:580-623 restated)."""
    eps = torch.finfo(torch.float64).eps
    o, o1 = f.clamp(min=eps), (1 - f).clamp(min=eps)
    return float(torch.log(o ** o * o1 ** o1).sum() * kt)


def test_fermi_matches_upstream_sih4_vectors(fx):
    """Fermi smearing at 5000 K and aufbau filling at 0 K against the SiH4
    vectors of test_wavefunction/test_filling.py:186-270.  The 5000 K
    occupations are fractional down to 1e-19, so a wrong exponent, cutoff,
    derivative or Boltzmann constant moves them by more than 1e-8."""
    emo = torch.tensor(fx["sih4_emo"])
    nel = float(fx["sih4_nel"])
    na, nb = og.alpha_beta(nel, 0)
    kt = 5000 * og.KELVIN2AU
    f = torch.stack([og.fermi_occupation(round(na), emo, kt), og.fermi_occupation(round(nb), emo, kt)])
    assert np.abs(f.sum(0).numpy() - fx["sih4_focc_5000"]).max() < 5e-9
    assert abs(_free_energy(f, kt) - float(fx["sih4_fenergy_5000"])) < 1e-8
    f0 = torch.stack([og.aufbau_occupation(emo.numel(), na), og.aufbau_occupation(emo.numel(), nb)])
    assert np.array_equal(f0.sum(0).numpy(), fx["sih4_focc_0"])
    assert _free_energy(f0, 0.0) == 0.0


def test_aufbau_matches_upstream_doctest():
    """filling.py:135-147: fractional electron counts fill the next orbital
    with the remainder."""
    for norb, nel, ref in ((5, 1.0, [1, 0, 0, 0, 0]), (8, 3.0, [1, 1, 1, 0, 0, 0, 0, 0]),
                           (4, 3.5, [1, 1, 1, 0.5]), (2, 1.5, [1, 0.5])):
        assert og.aufbau_occupation(norb, nel).tolist() == ref


def test_atoms_and_ions_match_tblite(fx):
    """Open shells, charges and fractional occupations: tblite free energies
    (E_scf - TS at 300 K) of the H, C, N, O atoms and ions
    (test_scf/test_elements_gfn1.py).  The degenerate p shells are partly
    filled, so -TS is ~3.6e-3 Eh and the Boltzmann constant matters at 1e-9."""
    opts = {"x_atol": 1e-12, "x_atol_max": 1e-12, "maxiter": 300, "fermi_thresh": 1e-10}
    for z, c, s, ref in zip(fx["atoms_z"], fx["atoms_chrg"], fx["atoms_spin"], fx["atoms_escf"]):
        r = og.singlepoint([int(z)], np.zeros((1, 3)), chrg=float(c), spin=None if s < 0 else int(s), opts=opts)
        assert r["converged"]
        assert abs(r["scf"] + r["fenergy"] - float(ref)) < 5e-9, (z, c)


def test_charged_molecule_matches_tblite(fx):
    """AD7en+ (C14H25+, charge +1): tblite SCF energy
    (test_scf/samples_charged.py:57-61)."""
    r = og.singlepoint(fx["ad7en_numbers"], fx["ad7en_positions"], chrg=float(fx["ad7en_chrg"]), opts=TIGHT)
    assert r["converged"]
    assert abs(r["scf"] - float(fx["ad7en_escf"])) < 1e-9


def test_spin_that_does_not_fit_raises():
    """filling.py:80-93: more unpaired electrons than electrons, or the wrong
    parity, raise ValueError."""
    with pytest.raises(ValueError):
        og.alpha_beta(8.0, 1)
    with pytest.raises(ValueError):
        og.alpha_beta(3.0, 5)
    assert og.alpha_beta(17.0, 1) == (9.0, 8.0)
    assert og.alpha_beta(17.0, None) == (9.0, 8.0)
    assert og.alpha_beta(8.0, 2) == (5.0, 3.0)
    with pytest.raises(ValueError):
        og.singlepoint([8, 1, 1], np.array([[0.0, 0, 0], [0, 0, 1.8], [1.8, 0, 0]]), opts={"fermi_etemp": -1.0})


def test_fermi_closed_form():
    # two-level system with a tiny gap: occupations follow 1/(1+exp) exactly
    kt = 300 * og.KELVIN2AU
    e = torch.tensor([-0.5, -0.1, -0.1 + 1e-4, 0.3], dtype=torch.float64)
    f = og.fermi_occupation(2.0, e, kt)
    assert abs(float(f.sum()) - 2.0) < 1e-8
    assert float(f[0]) == pytest.approx(1.0, abs=1e-15)
    ef = -0.1 + 0.5e-4  # symmetric pair around the midpoint
    assert float(f[1]) == pytest.approx(1.0 / (1.0 + math.exp((-0.1 - ef) / kt)), rel=1e-6)


def test_batched_overlap_equals_pair_loop(fx):
    z, p = mol(fx, "lys_xao")
    bas = og.build_basis(z)
    pos = torch.tensor(p)
    S1, d1 = og._overlap_impl(bas, pos, True)
    S2, d2 = og._overlap_impl_loop(bas, pos, True)
    assert (S1 - S2).abs().max() < 1e-15
    assert (d1 - d2).abs().max() < 1e-15


def test_broyden_updates_are_good_broyden():
    """The oracle's broyden1 (upstream's implicit-mode root search) keeps the
    inverse of Broyden's FIRST ("good") Jacobian update: with B = H^-1,
    B_new = B + (dy - B dx) dx^T / (dx^T dx) after every step, and H_new
    satisfies the secant condition H_new dy = dx.  The second ("bad")
    method would satisfy the secant condition too but fails the first
    identity.  On an affine contraction the search returns the fixed point."""
    rng = np.random.default_rng(3)
    n = 6
    A = torch.as_tensor(0.3 * rng.standard_normal((n, n)), dtype=torch.float64)
    b = torch.as_tensor(rng.standard_normal(n), dtype=torch.float64)
    func = lambda x: A @ x + b
    br = og.Broyden(damp=0.5)
    dense = lambda: br.h0 * torch.eye(n, dtype=torch.float64) + sum(
        (torch.outer(c, d) for c, d in zip(br.cs, br.ds)), torch.zeros(n, n, dtype=torch.float64))
    # replay the search step by step, checking each update
    br.cs, br.ds = [], []
    x = torch.zeros(n, dtype=torch.float64)
    y = x - func(x)
    for _ in range(4):
        H = dense()
        dx_step = -(H @ y)
        xn = x + 0.5 * dx_step
        yn = xn - func(xn)
        ddx, ddy = xn - x, yn - y
        v = H.T @ ddx
        br.cs.append(ddx - H @ ddy)
        br.ds.append(v / torch.dot(ddy, v))
        Hn = dense()
        assert torch.allclose(Hn @ ddy, ddx, atol=1e-12)
        B, Bn = torch.linalg.inv(H), torch.linalg.inv(Hn)
        good = B + torch.outer(ddy - B @ ddx, ddx) / torch.dot(ddx, ddx)
        assert torch.allclose(Bn, good, atol=1e-10)
        x, y = xn, yn
    xs, ok = og.Broyden(damp=0.5).solve(func, torch.zeros(n, dtype=torch.float64), 200, 1e-12, 1e-12)
    exact = torch.linalg.solve(torch.eye(n, dtype=torch.float64) - A, b)
    assert ok and float((xs - exact).abs().max()) < 1e-11


def test_implicit_mode_reaches_the_full_mode_fixed_point():
    """Upstream's implicit mode (broyden1 on v - f(v)) and full mode
    (Anderson) share the SCF fixed point; the root search takes more
    steps.  A restart of the Broyden memory changes only the path."""
    from dxtb_amd import workloads as W

    z, p = W.chain(np.random.default_rng(3), 12)
    o = {"x_atol": 1e-10, "x_atol_max": 1e-10, "maxiter": 300}
    a = og.singlepoint(z, p, opts=o)
    b = og.singlepoint(z, p, opts={**o, "scf_mode": "implicit", "f_atol": 1e-10})
    c = og.singlepoint(z, p, opts={**o, "scf_mode": "implicit", "f_atol": 1e-10, "broyden_memory": 8})
    assert a["converged"] and b["converged"] and c["converged"]
    assert abs(a["energy"] - b["energy"]) < 1e-10 and abs(b["energy"] - c["energy"]) < 1e-10
    assert float((a["charges_orbital"] - b["charges_orbital"]).abs().max()) < 1e-9
    assert b["iterations"] != a["iterations"]
