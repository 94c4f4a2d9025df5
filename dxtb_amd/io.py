"""Coordinate and charge files: Turbomole ``coord``, ``xyz``, ``.CHRG``/``.UHF``.

Upstream reads them with tad-mctc's ``read`` / ``read_chrg`` (used in
``test/test_singlepoint/test_energy.py:55-56`` and by the CLI,
``_src/cli/driver.py:94-117,173-181``); tad-mctc is not vendored in the
reference, so this module restates the formats themselves:

* Turbomole ``coord``: a ``$coord`` block of ``x y z symbol`` lines in bohr
  (``$coord angs`` in Angstrom), ended by the next ``$`` line;
* ``xyz``: the atom count, a comment line, then ``symbol x y z`` in Angstrom;
* ``.CHRG`` / ``.UHF``: one integer (total charge / unpaired electrons),
  looked up next to the coordinate file; absent means 0.

Readers return host arrays (atomic numbers int64, positions float64 in
bohr); a batch of files is padded with atomic number 0, the padded-batch
convention of the calculator.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from . import constants as C

_SYMBOLS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se "
    "Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy "
    "Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf "
    "Es Fm Md No Lr Rf Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og"
).split()
S2Z = {s.lower(): z for z, s in enumerate(_SYMBOLS, start=1)}
Z2S = {z: s for s, z in ((s, S2Z[s.lower()]) for s in _SYMBOLS)}

FILES = {"chrg": ".CHRG", "spin": ".UHF"}


def _number(token: str, path: Path) -> int:
    t = token.strip()
    if t.isdigit():
        return int(t)
    z = S2Z.get(t.lower())
    if z is None:
        raise ValueError(f"{path}: unknown element '{token}'")
    return z


def filetype(path: str | Path) -> str:
    """'tm' for Turbomole coord files, 'xyz' for xyz files (by name)."""
    p = Path(path)
    if p.suffix.lower() == ".xyz":
        return "xyz"
    if p.name.lower() in ("coord", "control") or p.suffix.lower() in (".coord", ".tmol", ".tm"):
        return "tm"
    raise ValueError(f"{p}: cannot tell the file type from the name (use ftype='xyz' or 'tm')")


def read_coord(path: str | Path) -> tuple[np.ndarray, np.ndarray]:
    """Turbomole ``$coord`` block -> (numbers, positions in bohr)."""
    p = Path(path)
    numbers, pos = [], []
    scale, inside, seen = 1.0, False, False
    for line in p.read_text().splitlines():
        s = line.strip()
        if s.startswith("$"):
            if inside:
                break
            head = s.split()
            if head[0] == "$coord":
                inside, seen = True, True
                opt = head[1].lower() if len(head) > 1 else ""
                if opt.startswith("frac"):
                    raise NotImplementedError(f"{p}: fractional coordinates (periodic systems)")
                scale = C.AA2AU if opt.startswith("ang") else 1.0
            continue
        if inside and s:
            f = s.split()
            if len(f) < 4:
                raise ValueError(f"{p}: malformed $coord line '{line}'")
            pos.append([float(x) for x in f[:3]])
            numbers.append(_number(f[3], p))
    if not seen:
        raise ValueError(f"{p}: no $coord block")
    if not numbers:
        raise ValueError(f"{p}: empty $coord block")
    return np.asarray(numbers, dtype=np.int64), np.asarray(pos, dtype=np.float64) * scale


def read_xyz(path: str | Path) -> tuple[np.ndarray, np.ndarray]:
    """xyz file (first frame) -> (numbers, positions in bohr)."""
    p = Path(path)
    lines = p.read_text().splitlines()
    try:
        nat = int(lines[0].split()[0])
    except (IndexError, ValueError):
        raise ValueError(f"{p}: the first line must hold the number of atoms") from None
    body = lines[2:2 + nat]
    if len(body) < nat:
        raise ValueError(f"{p}: {nat} atoms announced, {len(body)} lines found")
    numbers, pos = [], []
    for line in body:
        f = line.split()
        if len(f) < 4:
            raise ValueError(f"{p}: malformed atom line '{line}'")
        numbers.append(_number(f[0], p))
        pos.append([float(x) for x in f[1:4]])
    return np.asarray(numbers, dtype=np.int64), np.asarray(pos, dtype=np.float64) * C.AA2AU


def read(path: str | Path, ftype: str | None = None) -> tuple[np.ndarray, np.ndarray]:
    """Read one structure; `ftype` 'xyz' or 'tm' ('coord', 'turbomole'), else by name."""
    t = (ftype or filetype(path)).lower()
    if t in ("tm", "coord", "turbomole", "tmol"):
        return read_coord(path)
    if t == "xyz":
        return read_xyz(path)
    raise ValueError(f"unknown file type '{ftype}'")


def read_chrg(path: str | Path) -> int:
    """The integer in a ``.CHRG`` (or ``.UHF``) file."""
    p = Path(path)
    tok = p.read_text().split()
    if not tok:
        raise ValueError(f"{p}: empty")
    v = float(tok[0])
    if v != int(v):
        raise ValueError(f"{p}: '{tok[0]}' is not an integer")
    return int(v)


def attr_next_to(path: str | Path, attr: str) -> int:
    """Total charge ('chrg') or unpaired electrons ('spin') from the file
    next to `path`, 0 when there is none."""
    f = Path(path).resolve().parent / FILES[attr]
    return read_chrg(f) if f.is_file() else 0


def pack(structures: list[tuple[np.ndarray, np.ndarray]]) -> tuple[np.ndarray, np.ndarray]:
    """Pad several structures to a batch: numbers (B, N) with 0, positions (B, N, 3)."""
    n = max(len(z) for z, _ in structures)
    num = np.zeros((len(structures), n), dtype=np.int64)
    pos = np.zeros((len(structures), n, 3), dtype=np.float64)
    for b, (z, p) in enumerate(structures):
        num[b, :len(z)] = z
        pos[b, :len(z)] = p
    return num, pos
