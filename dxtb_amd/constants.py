"""Physical constants and model defaults of the GFN1-xTB path.

Upstream reads the unit constants from tad-mctc 0.7.0; the values here are
the ones that reproduce the tblite reference energies stored in the upstream
test-suite (``test/test_scf/samples.py``) to 1e-11 Eh.
"""

AA2AU = 1.0 / 0.52917721090380  # Angstrom -> bohr
EV2AU = 1.0 / 27.21138505  # eV -> Hartree (xTB/tblite value, CODATA 2010)
# k_B in Hartree/K, the xTB/tblite value: the tblite free energies of the
# H, C, N, O atoms and ions in test/test_scf/test_elements_gfn1.py (with
# fractional p occupations) agree with it to 2e-12 Eh, not with CODATA 2018
KELVIN2AU = 3.166808578545117e-06
# below this kT (0.1 K) upstream fills orbitals by the aufbau principle
# instead of Fermi smearing (src/dxtb/_src/scf/base.py:888-889)
KT_AUFBAU = 3e-7

# Covalent radii for the D3 coordination number (Pyykkoe & Atsumi 2009,
# Angstrom; scaled by 4/3 as DFT-D3 does) and atomic radii of the xTB
# distance polynomial (Angstrom).  Only elements whose values are pinned by
# the upstream fixtures are listed; other elements are rejected.
COV_2009 = {1: 0.32, 6: 0.75, 7: 0.71, 8: 0.63}
ATOMIC_RAD = {1: 0.32, 6: 0.75, 7: 0.71, 8: 0.64}
SUPPORTED_ELEMENTS = tuple(sorted(COV_2009))

CN_STEEPNESS = 16.0  # exponential counting function
CN_CUTOFF = 25.0  # bohr
INTCUTOFF = 50.0  # bohr (src/dxtb/_src/constants/defaults.py:93)
REPULSION_CUTOFF = 25.0  # bohr (src/dxtb/_src/constants/xtb.py:36)

PAD = 0  # atomic number of padding atoms
