"""GFN1-xTB parametrisation and per-species basis tables.

The data files next to this module (``gfn1.json``, ``sto_ng.npz``) hold the
upstream GFN1 parametrisation (``src/dxtb/_src/param/gfn1/gfn1-xtb.toml``)
and the STO-nG expansion tables (``src/dxtb/_src/basis/sto-ng/*.npy``);
``extract_gfn1.py`` regenerates them.

``SpeciesTables`` turns the elements present in a batch into the flat
parameter arrays the HIP kernels read: per species (radii, third-order and
repulsion parameters) and per unique shell (self-energy, CN shift, distance
polynomial, hardness, angular momentum, primitives, and the shell-pair
scaling K = hscale * kpair * X(EN)).
"""

from __future__ import annotations

import json
import math
from functools import lru_cache
from pathlib import Path

import numpy as np

from .. import constants as C

HERE = Path(__file__).parent
MAXPRIM = 8


@lru_cache(maxsize=1)
def load_gfn1() -> dict:
    with open(HERE / "gfn1.json") as fh:
        return json.load(fh)


@lru_cache(maxsize=1)
def _sto() -> dict:
    with np.load(HERE / "sto_ng.npz", allow_pickle=False) as d:
        return {k: np.array(d[k], dtype=np.float64) for k in d.files}


GFN1_XTB = "GFN1-xTB"


def elements_by_z() -> dict:
    return {e["z"]: (sym, e) for sym, e in load_gfn1()["elements"].items()}


_DFACT = (1.0, 1.0, 3.0, 15.0, 105.0)


def slater_to_gauss(ng: int, n: int, l: int, zeta: float):
    """STO-nG contraction with normalised primitives.

    Follows src/dxtb/_src/basis/slater.py:69-135 (table row selection, the
    n = 6 special case and the primitive normalisation).
    """
    itype = n + (0, 4, 7, 9, 10)[l] - 1
    if n == 6 and ng == 6:
        itype = 15 + l
    tab = _sto()[f"sto{ng}"]
    coeff = tab[0, itype, :].copy()
    alpha = tab[1, itype, :] * zeta * zeta
    coeff *= (2.0 / math.pi * alpha) ** 0.75 * np.sqrt(4.0 * alpha) ** l / math.sqrt(_DFACT[l])
    return alpha, coeff


def _ss_overlap(ai, ci, aj, cj) -> float:
    oij = 1.0 / (ai[:, None] + aj[None, :])
    return float((np.sqrt(math.pi * oij) ** 3 * ci[:, None] * cj[None, :]).sum())


def orthogonalize(a1, c1, a2, c2):
    """Gram-Schmidt of shell 2 against shell 1 (src/dxtb/_src/basis/ortho.py:74-110)."""
    ovl = _ss_overlap(a1, c1, a2, c2)
    a = np.concatenate([a2, a1])
    c = np.concatenate([c2, -ovl * c1])
    c = c / math.sqrt(_ss_overlap(a, c, a, c))
    return a, c


@lru_cache(maxsize=None)
def species_shells(z: int, slater: tuple | None = None):
    """Contracted shells of element z: list of (l, alpha, coeff, record index).
    `slater` replaces the element's Slater exponents (one per shell)."""
    els = elements_by_z()
    if z not in els or z not in C.SUPPORTED_ELEMENTS:
        raise NotImplementedError(
            f"Element Z={z} is not supported: this build covers GFN1-xTB for "
            f"Z in {C.SUPPORTED_ELEMENTS}."
        )
    _, e = els[z]
    zeta = e["slater"] if slater is None else slater
    out = []
    for k, (n, l) in enumerate(e["shells"]):
        a, c = slater_to_gauss(e["ngauss"][k], n, l, zeta[k])
        if not e["valence"][k]:
            # src/dxtb/_src/basis/bas.py:174-184
            pa, pc = out[k - 1][1], out[k - 1][2]
            a, c = orthogonalize(pa, pc, a, c)
        out.append((l, a, c, k))
    return out


class SpeciesTables:
    """Flat parameter arrays for the species of a batch (host numpy)."""

    def __init__(self, zs, slater: dict | None = None):
        """`slater` maps an atomic number to replacement Slater exponents."""
        par = load_gfn1()
        els = elements_by_z()
        self.zs = sorted(set(int(z) for z in zs))
        self.sp_of_z = np.full(119, -1, dtype=np.int32)
        for i, z in enumerate(self.zs):
            self.sp_of_z[z] = i
        nsp = len(self.zs)
        self.sp_rad = np.array([C.ATOMIC_RAD[z] * C.AA2AU for z in self.zs])
        self.sp_rcov = np.array([4.0 / 3.0 * C.COV_2009[z] * C.AA2AU for z in self.zs])
        self.sp_gam3 = np.array([els[z][1]["gam3"] for z in self.zs], dtype=np.float64)
        self.sp_zeff = np.array([els[z][1]["zeff"] for z in self.zs], dtype=np.float64)
        self.sp_arep = np.array([els[z][1]["arep"] for z in self.zs], dtype=np.float64)
        us_sp, us_l, us_val, us_level, us_kcn, us_shpoly, us_hard, us_ref = ([] for _ in range(8))
        us_rec = []  # index of the unique shell in its element's parameter record
        us_alpha, us_coeff, us_nprim = [], [], []
        self.us_base = np.zeros(nsp, dtype=np.int32)
        self.nsh_of_sp = np.zeros(nsp, dtype=np.int32)
        for i, z in enumerate(self.zs):
            _, e = els[z]
            self.us_base[i] = len(us_l)
            shells = species_shells(z, None if not slater or z not in slater else tuple(slater[z]))
            self.nsh_of_sp[i] = len(shells)
            for l, a, c, k in shells:
                us_sp.append(i)
                us_rec.append(k)
                us_l.append(l)
                us_val.append(bool(e["valence"][k]))
                us_level.append(e["levels"][k] * C.EV2AU)
                us_kcn.append(e["kcn"][k] * C.EV2AU)
                us_shpoly.append(e["shpoly"][k])
                us_hard.append(e["lgam"][k] * e["gam"])
                us_ref.append(e["refocc"][k])
                al = np.zeros(MAXPRIM)
                co = np.zeros(MAXPRIM)
                al[: len(a)] = a
                co[: len(c)] = c
                us_alpha.append(al)
                us_coeff.append(co)
                us_nprim.append(len(a))
        self.nus = len(us_l)
        self.us_sp = np.array(us_sp, dtype=np.int32)
        self.us_rec = np.array(us_rec, dtype=np.int32)
        self.us_l = np.array(us_l, dtype=np.int32)
        self.us_valence = np.array(us_val, dtype=bool)
        self.us_level = np.array(us_level)
        self.us_kcn = np.array(us_kcn)
        self.us_shpoly = np.array(us_shpoly)
        self.us_hard = np.array(us_hard)
        self.us_refocc = np.array(us_ref)
        self.us_alpha = np.concatenate(us_alpha) if us_alpha else np.zeros(0)
        self.us_coeff = np.concatenate(us_coeff) if us_coeff else np.zeros(0)
        self.us_nprim = np.array(us_nprim, dtype=np.int32)
        self.us_K = self._shell_pair_k(par)

    def _shell_pair_k(self, par) -> np.ndarray:
        """K_ab of src/dxtb/_src/xtb/gfn1.py:66-165 and xtb/base.py:316-330."""
        return shell_pair_k(self, par["hamiltonian"], {z: e["en"] for z, (_, e) in elements_by_z().items()})


def shell_pair_k(tb, h: dict, en: dict, grad: bool = False):
    """K_ab = hscale(l_a, l_b) kpair(A, B) (1 + enscale (EN_A - EN_B)^2) for
    two valence shells, (k_a + k_b)/2 otherwise (k = shell scaling, or kpol
    for a polarisation shell), over the unique shells of `tb`, from the
    Hamiltonian section `h` and electronegativities `en` (by atomic number).
    With `grad`, also the derivatives: a list of (name, dK) with dK (nus, nus)
    for every parameter K depends on (("shell", key), ("kpol",),
    ("enscale",), ("kpair", key), ("en", z))."""
    lab = "spdf"
    shell = h["shell"]
    kpair = h["kpair"]
    els = elements_by_z()
    nus = tb.nus
    K = np.zeros((nus, nus))
    dK: dict = {}

    def add(name, i, j, v):
        if grad:
            dK.setdefault(name, np.zeros((nus, nus)))[i, j] += v

    for i in range(nus):
        for j in range(nus):
            li, lj = int(tb.us_l[i]), int(tb.us_l[j])
            vi, vj = bool(tb.us_valence[i]), bool(tb.us_valence[j])
            kii = shell.get(lab[li] * 2, 1.0) if vi else h["kpol"]
            kjj = shell.get(lab[lj] * 2, 1.0) if vj else h["kpol"]
            if vi and vj:
                za, zb = tb.zs[tb.us_sp[i]], tb.zs[tb.us_sp[j]]
                sa, sb = els[za][0], els[zb][0]
                kkey = f"{sa}-{sb}" if f"{sa}-{sb}" in kpair else (f"{sb}-{sa}" if f"{sb}-{sa}" in kpair else None)
                kp = kpair[kkey] if kkey else 1.0
                de = en[za] - en[zb]
                x = 1.0 + h["enscale"] * de * de
                hkey = lab[li] + lab[lj] if lab[li] + lab[lj] in shell else (
                    lab[lj] + lab[li] if lab[lj] + lab[li] in shell else None)
                hs = shell[hkey] if hkey else 0.5 * (kii + kjj)
                K[i, j] = hs * kp * x
                if hkey:
                    add(("shell", hkey), i, j, kp * x)
                else:
                    for key in (lab[li] * 2, lab[lj] * 2):
                        if key in shell:
                            add(("shell", key), i, j, 0.5 * kp * x)
                if kkey:
                    add(("kpair", kkey), i, j, hs * x)
                add(("enscale",), i, j, hs * kp * de * de)
                add(("en", za), i, j, hs * kp * h["enscale"] * 2.0 * de)
                add(("en", zb), i, j, -hs * kp * h["enscale"] * 2.0 * de)
            else:
                K[i, j] = 0.5 * (kii + kjj)
                for v, l in ((vi, li), (vj, lj)):
                    if not v:
                        add(("kpol",), i, j, 0.5)
                    elif lab[l] * 2 in shell:
                        add(("shell", lab[l] * 2), i, j, 0.5)
    return (K, list(dK.items())) if grad else K
