"""Regenerate the GFN1-xTB parameter data shipped with this package.

Run once, in a checkout that has the upstream dxtb sources next to it::

    python -m dxtb_amd.param.extract_gfn1 /root/reference

It reads the upstream parametrisation file
``src/dxtb/_src/param/gfn1/gfn1-xtb.toml`` and the STO-nG expansion tables
``src/dxtb/_src/basis/sto-ng/sto-{1..6}g.npy`` as *data* and writes

* ``gfn1.json``  -- global model constants plus the per-element records of
  every element whose GFN1 basis is made of s and p shells only;
* ``sto_ng.npz`` -- the STO-nG contraction tables, keys ``sto1`` .. ``sto6``,
  each of shape ``(2, ntype, ng)`` = (coefficients, exponents).

Nothing here is imported at run time; the package only reads the two data
files.  The output format is this package's own: shells are stored as
``(principal quantum number, angular momentum)`` pairs, and the valence flag
of GFN1's second hydrogen s-shell is precomputed (a shell is non-valence when
an earlier shell of the same atom has the same angular momentum, following
``src/dxtb/_src/param/module/utils.py:488-542``).
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

try:  # Python 3.11+
    import tomllib as _toml
except ModuleNotFoundError:  # pragma: no cover - Python 3.10 in this image
    import tomli as _toml

HERE = Path(__file__).parent

ELEMENT_FIELDS = (
    "levels", "slater", "ngauss", "refocc", "shpoly", "kcn", "lgam",
    "gam", "gam3", "zeff", "arep", "en", "xbond",
)

SYMBOLS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co "
    "Ni Cu Zn Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb "
    "Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re "
    "Os Ir Pt Au Hg Tl Pb Bi Po At Rn"
).split()


def extract(reference_root: Path) -> None:
    toml_path = (
        reference_root / "src/dxtb/_src/param/gfn1/gfn1-xtb.toml"
    )
    with open(toml_path, "rb") as fh:
        raw = _toml.load(fh)

    xtb = raw["hamiltonian"]["xtb"]
    out: dict = {
        "meta": {"name": raw["meta"]["name"], "source": "gfn1-xtb.toml"},
        "hamiltonian": {
            "kpol": xtb["kpol"],
            "enscale": xtb["enscale"],
            "wexp": xtb["wexp"],
            "shell": dict(xtb["shell"]),
            "kpair": dict(xtb["kpair"]),
        },
        "dispersion_d3": dict(raw["dispersion"]["d3"]),
        "repulsion": {"kexp": raw["repulsion"]["effective"]["kexp"]},
        "halogen": dict(raw["halogen"]["classical"]),
        "charge": {
            "gexp": raw["charge"]["effective"]["gexp"],
            "average": raw["charge"]["effective"]["average"],
        },
        "thirdorder": {"shell": raw["thirdorder"]["shell"]},
        "elements": {},
    }

    lmap = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4}
    for z, sym in enumerate(SYMBOLS, start=1):
        rec = raw["element"].get(sym)
        if rec is None:
            continue
        shells = [(int(s[:-1]), lmap[s[-1]]) for s in rec["shells"]]
        if any(l > 1 for _, l in shells):
            continue  # d/f shells: not supported by this package
        valence = []
        seen: set[int] = set()
        for _, l in shells:
            valence.append(l not in seen)
            seen.add(l)
        entry = {"z": z, "shells": shells, "valence": valence}
        for key in ELEMENT_FIELDS:
            entry[key] = rec[key]
        out["elements"][sym] = entry

    (HERE / "gfn1.json").write_text(json.dumps(out, indent=1, sort_keys=True))

    base = reference_root / "src/dxtb/_src/basis/sto-ng"
    tables = {
        f"sto{n}": np.load(base / f"sto-{n}g.npy", allow_pickle=False)
        for n in range(1, 7)
    }
    np.savez(HERE / "sto_ng.npz", **tables)


if __name__ == "__main__":
    extract(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
