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
  functions (tblite) from test/test_overlap/test_grad.py:225-380;
* the charged molecule AD7en+ (C14H25+, charge +1): geometry from
  test/test_singlepoint/mols/AD7en+/coord, tblite SCF energy typed in from
  test/test_scf/samples_charged.py:57-61 and total energy from
  test/test_singlepoint/samples.py:86-89;
* the open-shell NO2 geometry from test/test_singlepoint/mols/NO2/coord;
* tblite SCF free energies of the H, C, N, O atoms, their cations and
  anions (spins from test/test_scf/uhf_table.py), typed in from
  test/test_scf/test_elements_gfn1.py (ref :45-53, ref_cation :137-145,
  ref_anion :229-237; H+ has no electron and is left out);
* the SiH4 Fermi-smearing vectors of test/test_wavefunction: orbital
  energies (samples.py:965-983), occupations and electronic free energy at
  5000 K and 0 K (test_filling.py:186-238).

No upstream code is imported or executed.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).parent / "fixtures.npz"
SYM = {"h": 1, "c": 6, "n": 7, "o": 8}

# charged and open-shell molecules (coord files)
EXTRA_MOLS = {"AD7en+": "ad7en", "NO2": "no2"}
AD7EN_ESCF = -4.3226840214360e01  # test/test_scf/samples_charged.py:57-61 (charge +1)
AD7EN_ETOT = -4.2547841532513e01  # test/test_singlepoint/samples.py:86-89 (includes D3)

# atoms: (Z, charge, unpaired electrons or -1 for automatic, tblite energy)
ATOMS = [
    (1, 0.0, -1, -0.4014294744618301), (6, 0.0, -1, -1.7411359557542052),
    (7, 0.0, -1, -2.8988862104065958), (8, 0.0, -1, -4.3526523408648030),
    (6, 1.0, 1, -1.1060742863488979), (7, 1.0, 0, -2.1782414865973068),
    (8, 1.0, 1, -3.3929027848641802),
    (1, -1.0, 0, -0.5678094489236601), (6, -1.0, 1, -1.9170116002810331),
    (7, -1.0, 0, -3.1284233130873655), (8, -1.0, 1, -4.7053860329689856),
]

# test/test_wavefunction/samples.py:965-983 (SiH4, 8 electrons, one channel)
SIH4_EMO = [
    -0.6158187985420227, -0.4695263206958771, -0.4695262610912323, -0.4695262312889099,
    -0.1725806295871735, -0.1725805550813675, -0.1725803762674332, -0.0903786197304726,
    -0.0903783515095711, -0.0675993412733078, -0.0675990283489227, -0.0675989910960197,
    -0.0500356517732143, 0.3943647742271423, 1.2486594915390015, 1.2486598491668701,
    1.2486605644226074,
]
# test/test_wavefunction/test_filling.py:193-238: alpha + beta occupations and
# the electronic free energy at 5000 K; 0 K is the aufbau filling
SIH4_FOCC_5000 = [
    1.9999999835050197, 1.9998302497800045, 1.9998302497800045, 1.9998302497800045,
    1.6888179157220813e-004, 1.6888179157220572e-004, 1.6888179157220273e-004,
    9.3970980586158547e-007, 9.3970980586158028e-007, 2.2294515200028562e-007,
    2.2294515200028403e-007, 2.2294515200027612e-007, 7.3530914097230426e-008,
    4.7590701921523143e-020, 0.0, 0.0, 0.0,
]
SIH4_FE_5000 = -1.6758385176418445e-004
SIH4_FOCC_0 = [2.0] * 4 + [0.0] * 13
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
    for name, key in EXTRA_MOLS.items():
        z, p = read_coord(t / "test_singlepoint/mols" / name / "coord")
        out[f"{key}_numbers"] = z
        out[f"{key}_positions"] = p
    out["ad7en_chrg"] = np.array(1.0)
    out["ad7en_escf"] = np.array(AD7EN_ESCF)
    out["ad7en_etot"] = np.array(AD7EN_ETOT)
    a = np.array(ATOMS)
    out["atoms_z"] = a[:, 0].astype(np.int64)
    out["atoms_chrg"] = a[:, 1]
    out["atoms_spin"] = a[:, 2].astype(np.int64)
    out["atoms_escf"] = a[:, 3]
    out["sih4_emo"] = np.array(SIH4_EMO)
    out["sih4_nel"] = np.array(8.0)
    out["sih4_focc_5000"] = np.array(SIH4_FOCC_5000)
    out["sih4_fenergy_5000"] = np.array(SIH4_FE_5000)
    out["sih4_focc_0"] = np.array(SIH4_FOCC_0)
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
