"""CPU oracle for the GFN1-xTB single-point path (TEST INFRASTRUCTURE ONLY).

This package restates, in plain PyTorch FP64 on the CPU, the algorithm of the
upstream dxtb GFN1-xTB single point: basis construction, overlap and its
nuclear derivative, coordination number, core Hamiltonian H0, shell-resolved
second-order and atom-resolved third-order electrostatics, classical
repulsion, the full-mode SCF (potential as self-consistent quantity, Anderson
mixing), Fermi smearing, energies, and the analytic (variational) nuclear
gradient.  Every function cites the upstream file and line range it follows.

Only the test-suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker or the timed
CPU reference.  The product package ``dxtb_amd`` never imports it.

Pinning status (what checks each function; tests/test_oracle.py):

* STO-nG expansion, orthogonalisation, overlap, overlap derivative:
  the tblite pair overlaps and gradients held in the upstream tests
  (test/test_overlap/test_grad.py), the upstream ``overlap.npz`` fixture,
  exact normalisation of the orthogonalised H 2s shell and central finite
  differences of S.
* H0 and coordination number: the upstream ``h0.npz`` fixture (float32;
  sensitive to the atomic and covalent radii and the CN).  The CN gradient
  is checked only through the forces (finite differences, tblite SCF
  gradients).
* ES2/ES3, SCF, Mulliken, energies, the eV unit: tblite SCF energies of H2,
  H2O, CH4, LYS_xao (test/test_scf/samples.py) and of the charged AD7en+
  (test/test_scf/samples_charged.py), <= 1e-9 Eh.
* Fermi smearing, -TS and the Boltzmann constant: the SiH4 5000 K
  occupations and free energy of test/test_wavefunction/test_filling.py
  (5e-9 / 1e-8), tblite free energies of the H, C, N, O atoms, cations and
  anions (test/test_scf/test_elements_gfn1.py: open shells, fractional
  occupations, <= 5e-9 Eh), and a two-level closed form.
* Aufbau filling at zero temperature: upstream's doctest values
  (wavefunction/filling.py:135-147) and the SiH4 0 K occupations.
* alpha/beta split and its errors: upstream's rules (filling.py:41-113).
* Implicit-mode root search (Broyden, xitorch broyden1): each update is
  checked to be the inverse of Broyden's first ("good") Jacobian update and
  to satisfy the secant condition; it returns the exact fixed point of an
  affine contraction and the full-mode (Anderson) SCF fixed point.  Its
  step-by-step path (iteration counts) has no upstream vector: parity of
  the path is unpinned beyond those identities.
* Padded-batch filling (npad padding levels at energy 0 sharing the Fermi
  filling and -TS): follows filling.py:201-366, where only the start value
  is masked; upstream's batched filling tests
  (test/test_wavefunction/test_filling.py:125-185) pack zero-energy padding
  into the same filling.  No upstream vector holds a padded result at high
  temperature: parity unpinned beyond the unpadded limit (npad = 0).
* Repulsion: test/test_classical/test_repulsion/samples.py (energies and
  the H2O gradient).
* Analytic forces: tblite SCF-only gradients (test/test_scf/grad.npz,
  float32, 1e-6) and central finite differences of the oracle energy.
* D3(BJ) dispersion: parity unpinned -- its reference C6 table lives in the
  third-party package tad-dftd3 0.6.0, which is absent; the oracle does not
  implement it and every parity comparison excludes dispersion.
* The CN cutoff (25 bohr): parity unpinned -- no fixture geometry has an
  atom pair near it.
"""

from .gfn1 import (  # noqa: F401
    Basis,
    build_basis,
    coordination_number,
    coordination_number_grad,
    core_hamiltonian,
    forces,
    gamma_matrix,
    overlap,
    overlap_grad,
    repulsion,
    singlepoint,
)
