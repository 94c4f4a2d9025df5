"""Parameter gradients dE/dtheta against central finite differences of the
GPU energies in the same parameters (h = 1e-5 x max(1, |theta|)): every
level, kcn, shpoly, ES2 hardness and Slater exponent of the unique shells
and every gam3, zeff and arep of the species, on four molecules; and the
calculator's mapping to the parameter file's names and units."""

import numpy as np
import pytest
import torch

from conftest import mol
from dxtb_amd import constants as C
from dxtb_amd import workloads as W
from dxtb_amd.engine import BatchLayout, DeviceTables, Pipeline, ScfOptions
from dxtb_amd.param import SpeciesTables
from helpers import TIGHT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _energies(mols, tb, dev, grad=False):
    lay = BatchLayout([np.asarray(z) for z, _ in mols], tb, padded=False, chrg=np.zeros(len(mols)))
    pipe = Pipeline(lay, DeviceTables(tb, dev), dev)
    pos = torch.tensor(np.concatenate([p for _, p in mols]), dtype=torch.float64, device=dev)
    pipe.run(pos, ScfOptions(**TIGHT), forces=grad)
    assert int(pipe.records()[:, 6].abs().max().item()) == 0
    e = pipe.energies().cpu().numpy()
    if not grad:
        return e
    return e, pipe.param_grad().cpu().numpy(), pipe.param_grad_basis(tb).cpu().numpy()


def test_parameter_gradient_vs_finite_differences(fx):
    rng = np.random.default_rng(15)  # the chain holds N and O
    mols = [mol(fx, "h2o"), mol(fx, "ch4"), W.chain(rng, 8, p_n=0.3, p_o=0.3), W.water_cluster(rng, 4)]
    dev = torch.device("cuda:0")
    zs = np.unique(np.concatenate([z for z, _ in mols]))
    tb = SpeciesTables(zs)
    print(f"case: parameter gradients, species {list(tb.zs)}, {tb.nus} unique shells", flush=True)
    assert list(tb.zs) == [1, 6, 7, 8]
    _, g, gz = _energies(mols, tb, dev, grad=True)
    nus, nsp = tb.nus, len(tb.zs)
    cols = [("us_level", u, u) for u in range(nus)] + [("us_kcn", u, nus + u) for u in range(nus)]
    cols += [("us_shpoly", u, 2 * nus + u) for u in range(nus)] + [("us_hard", u, 3 * nus + u) for u in range(nus)]
    cols += [("sp_gam3", i, 4 * nus + i) for i in range(nsp)] + [("sp_zeff", i, 4 * nus + nsp + i) for i in range(nsp)]
    cols += [("sp_arep", i, 4 * nus + 2 * nsp + i) for i in range(nsp)]
    worst = 0.0
    for name, idx, col in cols:
        base = getattr(tb, name)[idx]
        h = 1e-5 * max(1.0, abs(base))
        ep = []
        for s in (1.0, -1.0):
            t = SpeciesTables(zs)
            getattr(t, name)[idx] = base + s * h
            ep.append(_energies(mols, t, dev))
        fd = (ep[0] - ep[1]) / (2 * h)
        err = np.abs(fd - g[:, col]).max() / max(1.0, np.abs(fd).max())
        worst = max(worst, err)
        assert err < 1e-7, (name, idx, fd, g[:, col])
    print(f"worst relative error {worst:.1e}", flush=True)
    # Slater exponents (the basis): central differences of the energy
    from dxtb_amd.param import elements_by_z

    els = elements_by_z()
    worst = 0.0
    for u in range(nus):
        z = tb.zs[int(tb.us_sp[u])]
        zeta = list(els[z][1]["slater"])
        k = int(tb.us_rec[u])
        h = 1e-5 * zeta[k]
        ep = []
        for s in (1.0, -1.0):
            zz = list(zeta)
            zz[k] += s * h
            ep.append(_energies(mols, SpeciesTables(zs, slater={z: zz}), dev))
        fd = (ep[0] - ep[1]) / (2 * h)
        err = np.abs(fd - gz[:, u]).max() / max(1.0, np.abs(fd).max())
        worst = max(worst, err)
        assert err < 1e-6, (z, k, fd, gz[:, u])
    print(f"slater: worst relative error {worst:.1e}", flush=True)
    # the pair scaling K_ab as each shell pair reads it (ordered entries)
    worst = 0.0
    kcols = 4 * nus + 3 * nsp
    for a_ in range(nus):
        for b_ in range(nus):
            h = 1e-5 * max(1.0, abs(tb.us_K[a_, b_]))
            ep = []
            for s in (1.0, -1.0):
                t = SpeciesTables(zs)
                t.us_K[a_, b_] += s * h
                ep.append(_energies(mols, t, dev))
            fd = (ep[0] - ep[1]) / (2 * h)
            err = np.abs(fd - g[:, kcols + a_ * nus + b_]).max() / max(1.0, np.abs(fd).max())
            worst = max(worst, err)
            assert err < 1e-7, (a_, b_, fd, g[:, kcols + a_ * nus + b_])
    print(f"K: worst relative error {worst:.1e}", flush=True)


def test_calculator_parameter_gradient_names_and_units(fx):
    from dxtb_amd import GFN1Calculator
    from dxtb_amd.param import elements_by_z

    z, p = mol(fx, "h2o")
    dev = torch.device("cuda:0")
    calc = GFN1Calculator(torch.tensor(z, device=dev), dtype=torch.float64,
                          opts={"x_atol": 1e-11, "x_atol_max": 1e-11, "maxiter": 300, "exclude": ["disp"]})
    pos = torch.tensor(p, dtype=torch.float64, device=dev)
    d = calc.get_param_gradient(pos)
    assert set(d) == {"H", "O", "hamiltonian"}
    tb = SpeciesTables(np.unique(z))
    _, g, gz = _energies([(z, p)], tb, dev, grad=True)
    els = elements_by_z()
    for i, zz in enumerate(tb.zs):
        sym, e = els[zz]
        us = np.nonzero(tb.us_sp == i)[0]
        nus = tb.nus
        lv = d[sym]["levels"].cpu().numpy()
        assert lv.shape == (len(e["levels"]),)
        assert np.allclose(lv[tb.us_rec[us]], g[0, us] * C.EV2AU, rtol=1e-12, atol=1e-14)
        eta = g[0, 3 * nus + us]
        assert abs(d[sym]["gam"].item() - float((eta * np.asarray(e["lgam"])[tb.us_rec[us]]).sum())) < 1e-12
        assert abs(d[sym]["arep"].item() - g[0, 4 * nus + 2 * len(tb.zs) + i]) < 1e-14
        assert np.allclose(d[sym]["slater"].cpu().numpy()[tb.us_rec[us]], gz[0, us], rtol=1e-12, atol=1e-14)
    # Hamiltonian-section parameters: central differences through K
    import copy

    from dxtb_amd.param import load_gfn1, shell_pair_k

    ham = load_gfn1()["hamiltonian"]
    en = {zz: e["en"] for zz, (_, e) in elements_by_z().items()}
    checks = [("shell", "ss"), ("shell", "sp"), ("shell", "pp"), ("kpol",), ("enscale",), ("kpair", "H-H"),
              ("en", 8)]
    for name in checks:
        vals = []
        for s in (1.0, -1.0):
            hh, ee = copy.deepcopy(ham), dict(en)
            h = 1e-5
            if name[0] in ("shell", "kpair"):
                hh[name[0]][name[1]] += s * h
            elif name[0] == "en":
                ee[name[1]] += s * h
            else:
                hh[name[0]] += s * h
            t = SpeciesTables(np.unique(z))
            t.us_K = shell_pair_k(t, hh, ee)
            vals.append(_energies([(z, p)], t, dev)[0])
        fd = (vals[0] - vals[1]) / 2e-5
        if name[0] in ("shell", "kpair"):
            an = d["hamiltonian"][name[0]].get(name[1], torch.zeros(())).item()
        elif name[0] == "en":
            an = d["O"]["en"].item()
        else:
            an = d["hamiltonian"][name[0]].item()
        assert abs(fd - an) < 1e-7 * max(1.0, abs(fd)), (name, fd, an)


def test_parameter_gradient_vs_oracle_finite_differences():
    """The calculator's dE/dtheta (names and units of the parameter file)
    against central differences of the CPU oracle's energy in the same
    parameter, one parameter of each class: a self-energy and its CN shift,
    the distance polynomial, the shell hardness and the element's Hubbard
    parameter, the third-order derivative, both repulsion parameters, a
    Slater exponent, an electronegativity, and the Hamiltonian section's
    shell, kpol, enscale and kpair scalings.  The oracle is independent of
    the kernel (its own integrals and SCF), so a wrong chain rule or unit
    in either fails here.  Relative tolerance 1e-6 (the differences carry
    the SCF convergence noise, ~1e-12 Eh over 2h)."""
    import copy

    import oracle.gfn1 as og
    from dxtb_amd import GFN1Calculator

    rng = np.random.default_rng(31)
    z, p = W.chain(rng, 5, p_n=0.4, p_o=0.4)
    assert {7, 8} <= set(z.tolist())
    dev = torch.device("cuda:0")
    calc = GFN1Calculator(torch.tensor(z, device=dev), dtype=torch.float64,
                          opts={"x_atol": 1e-11, "x_atol_max": 1e-11, "maxiter": 300, "exclude": ["disp"]})
    print("case: parameter gradients vs oracle differences", flush=True)
    d = calc.get_param_gradient(torch.tensor(p, dtype=torch.float64, device=dev))
    par0 = og.load_params()
    opts = {"x_atol": 1e-12, "x_atol_max": 1e-12, "maxiter": 300}
    checks = [
        (("elements", "C", "levels", 1), lambda: d["C"]["levels"][1]),
        (("elements", "N", "kcn", 0), lambda: d["N"]["kcn"][0]),
        (("elements", "O", "shpoly", 1), lambda: d["O"]["shpoly"][1]),
        (("elements", "C", "lgam", 1), lambda: d["C"]["lgam"][1]),
        (("elements", "H", "gam"), lambda: d["H"]["gam"]),
        (("elements", "O", "gam3"), lambda: d["O"]["gam3"]),
        (("elements", "N", "zeff"), lambda: d["N"]["zeff"]),
        (("elements", "C", "arep"), lambda: d["C"]["arep"]),
        (("elements", "H", "slater", 0), lambda: d["H"]["slater"][0]),
        (("elements", "O", "en"), lambda: d["O"]["en"]),
        (("hamiltonian", "shell", "sp"), lambda: d["hamiltonian"]["shell"]["sp"]),
        (("hamiltonian", "kpol"), lambda: d["hamiltonian"]["kpol"]),
        (("hamiltonian", "enscale"), lambda: d["hamiltonian"]["enscale"]),
        (("hamiltonian", "kpair", "N-H"), lambda: d["hamiltonian"]["kpair"]["N-H"]),
    ]
    for path, get in checks:
        node = par0
        for k in path[:-1]:
            node = node[k]
        base = node[path[-1]]
        h = 1e-4 * max(1.0, abs(base))
        e = []
        for s in (1.0, -1.0):
            par = copy.deepcopy(par0)
            node = par
            for k in path[:-1]:
                node = node[k]
            node[path[-1]] = base + s * h
            e.append(og.singlepoint(z, p, opts=opts, par=par)["energy"])
        fd = (e[0] - e[1]) / (2 * h)
        an = float(get())
        assert abs(fd - an) < 1e-6 * max(abs(fd), 1e-2), (path, fd, an)
