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

Pinning status (what checks each function; see ``tests/test_oracle_*.py``):

* STO-nG expansion, orthogonalisation, overlap, overlap derivative:
  pinned by the tblite pair values held in the upstream tests, the upstream
  ``overlap.npz`` fixture, invariants and finite differences.
* H0 and coordination number: pinned by the upstream ``h0.npz`` fixture and
  ``grad_no_overlap.npz`` (``*_dcn``), plus finite differences.
* ES2/ES3, SCF, Fermi, Mulliken, energies: pinned by the tblite SCF energies
  held in the upstream ``test_scf/samples.py`` and the Fermi/density samples
  of ``test_wavefunction/samples.py``.
* Repulsion: pinned by ``test_classical/test_repulsion/samples.py``.
* Analytic forces: pinned by ``test_scf/grad.npz`` (SCF-only tblite
  gradients) and central finite differences of the oracle energy.
* D3(BJ) dispersion: parity unpinned -- its reference C6 table lives in the
  third-party package tad-dftd3 0.6.0, which is absent; the oracle does not
  implement it and every parity comparison excludes dispersion.
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
