"""Write tests/golden/fixtures.npz from the upstream test data (data only).

    python tests/golden/make_fixtures.py /root/reference

Collected (all paths relative to the upstream repository):

* geometries (bohr) of H2, H2O, CH4, LYS_xao, C60 from
  test/test_singlepoint/mols/<name>/coord (Turbomole format);
* overlap and core-Hamiltonian matrices (tblite, float32) from
  test/test_overlap/overlap.npz and test/test_hamiltonian/h0.npz;
* SCF-only nuclear gradients (tblite 0.2.1, float32) from
  test/test_scf/grad.npz;
* SCF electronic energies (tblite) typed in from test/test_scf/samples.py
  (egfn1 of H2 :67-70, H2O :161-165, CH4 :166-170, LYS_xao :570-574,
  C60 :580-584);
* total energies (tblite, including D3) from
  test/test_singlepoint/samples.py:45-90;
* repulsion energies from test/test_classical/test_repulsion/samples.py
  (H2 :49-54, H2O :116-117, LYS_xao :409-411) and the H2O repulsion
  gradient (:118-137);
* contracted pair overlaps and their gradients for zeta = 1 Slater
  functions (tblite) from test/test_overlap/test_grad.py:225-380.

No upstream code is imported or executed.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).parent / "fixtures.npz"
SYM = {"h": 1, "c": 6, "n": 7, "o": 8}
MOLS = {"H2": "h2", "H2O": "h2o", "CH4": "ch4", "LYS_xao": "lys_xao", "C60": "c60"}

ESCF = {  # test/test_scf/samples.py, egfn1
    "h2": -1.0585984032484, "h2o": -5.8052489623704, "ch4": -4.3393059719255,
    "lys_xao": -4.8850798066902e01, "c60": -1.2879148324775e02,
}
ETOT = {  # test/test_singlepoint/samples.py, egfn1 (includes D3)
    "h2": -1.0362714373390, "h2o": -5.7686218257620, "ch4": -4.2741992424931,
    "lys_xao": -4.8324739766346e01, "c60": -1.2673081838911e02,
}
EREP = {  # test/test_classical/test_repulsion/samples.py, gfn1
    "h2": 2.2362490584964e-02, "h2o": 3.6764721202060e-02, "lys_xao": 0.54175667737478617,
}
GREP_H2O = np.array([
    [0.0, 0.0, 0.10833611039049376],
    [6.9798837064846950e-02, 0.0, -5.4168055195246881e-02],
    [-6.9798837064846950e-02, 0.0, -5.4168055195246881e-02],
])

# test/test_overlap/test_grad.py: (ng, n, l) of both CGTOs, vec, overlap, gradient
PAIRS = {
    "1s1s": ((6, 1, 0), (6, 1, 0), [0.0, 0.0, -1.405], [[0.75156098243428548]],
             [[0.0, 0.0, -0.27637559271358614]]),
    "2s1s": ((6, 2, 0), (6, 1, 0), [0.0, 0.0, -1.405], [[0.76296162997970662]],
             [[0.0, 0.0, -0.14526133860726220]]),
    "3s4s": ((6, 3, 0), (6, 4, 0), [0.0, 0.0, -1.405], [[0.893829]],
             [[0.0, 0.0, -0.056111]]),
    "2p1s": ((6, 2, 1), (6, 1, 0), [1.245, -6.789, 0.123], [[-0.081147], [0.001470], [0.014881]],
             [[-0.011176, 0.048990, -0.001104], [0.000202, -0.001104, -0.011933],
              [-0.009903, -0.011176, 0.000202]]),
}
S2P2P = np.diag([8.2922059116650637e-01, 5.3250035739126811e-01, 8.2922059116650637e-01])
G2P2P = np.array([
    [0, 0, 2.1118877848771400e-01], [0, 2.1118877848771400e-01, 0], [0, 0, 0],
    [0, 2.1118877848771400e-01, 0], [0, 0, 5.2445032996407726e-01], [2.1118877848771400e-01, 0, 0],
    [0, 0, 0], [2.1118877848771400e-01, 0, 0], [0, 0, 2.1118877848771400e-01],
])


def read_coord(path: Path):
    nums, pos = [], []
    on = False
    for line in path.read_text().splitlines():
        if line.startswith("$coord"):
            on = True
            continue
        if line.startswith("$"):
            on = False
            continue
        if on and line.strip():
            p = line.split()
            pos.append([float(x) for x in p[:3]])
            nums.append(SYM[p[3].lower()])
    return np.array(nums, dtype=np.int64), np.array(pos, dtype=np.float64)


def main(ref: Path) -> None:
    t = ref / "test"
    out = {}
    ov = np.load(t / "test_overlap/overlap.npz", allow_pickle=False)
    h0 = np.load(t / "test_hamiltonian/h0.npz", allow_pickle=False)
    gs = np.load(t / "test_scf/grad.npz", allow_pickle=False)
    for name, key in MOLS.items():
        z, p = read_coord(t / "test_singlepoint/mols" / name / "coord")
        out[f"{key}_numbers"] = z
        out[f"{key}_positions"] = p
        out[f"{key}_escf"] = np.array(ESCF[key])
        out[f"{key}_etot"] = np.array(ETOT[key])
        out[f"{key}_gscf"] = gs[key].astype(np.float64)
        if key in ov.files:
            out[f"{key}_overlap"] = ov[key].astype(np.float64)
            out[f"{key}_h0"] = h0[key].astype(np.float64)
        if key in EREP:
            out[f"{key}_erep"] = np.array(EREP[key])
    out["h2o_grep"] = GREP_H2O
    for k, (ci, cj, vec, s, g) in PAIRS.items():
        out[f"pair_{k}_cgto"] = np.array([ci, cj], dtype=np.int64)
        out[f"pair_{k}_vec"] = np.array(vec)
        out[f"pair_{k}_s"] = np.array(s)
        out[f"pair_{k}_g"] = np.array(g)
    out["pair_2p2p_cgto"] = np.array([(6, 2, 1), (6, 2, 1)], dtype=np.int64)
    out["pair_2p2p_vec"] = np.array([0.0, 0.0, 1.405])
    out["pair_2p2p_s"] = S2P2P
    out["pair_2p2p_g"] = G2P2P
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
