"""Whole-GPU path for one large system (thousands of orbitals).

The per-molecule SCF kernel gives one compute unit to each molecule, which is
right for streams of molecules but not for a single system of n ~ 5000
orbitals.  For such a molecule the integrals, forces and parameter
gradients come from the same kernels as for a batch, and the SCF from
``xtb_scf_large`` (dxtb_amd/csrc/large.hip): the same SCF map, eigensolver
and mixer as the per-molecule kernel, with every stage a grid kernel that
spreads over all CUs and the SCF loop on the host:

* Loewdin orthogonaliser X = S^-1/2 by coupled Newton-Schulz on matrix-core
  GEMMs (128 x 128 tiles per workgroup, XCD-aware tile order);
* per SCF step, for gapped closed shells, the perturbative
  occupied-virtual rotation of the previous step's orbitals (GEMMs);
  otherwise A = C F C^T and parallel-ordered block Jacobi (a wave per 32 x 32
  subproblem, the rotations applied to A and the orbitals on matrix cores),
  which gives the eigenpairs upstream's eigh does
  (src/dxtb/_src/scf/unrolling/base.py:142-175);
* sort and Fermi smearing (src/dxtb/_src/wavefunction/filling.py:201-366),
  density, Mulliken charges, ES2/ES3 potential and Anderson mixing
  (src/dxtb/_src/scf/mixer/anderson.py:163-317) on the device; the host
  reads a few control values back per step.

The SCF follows src/dxtb/_src/scf/unrolling/default.py:71-140.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .engine import BatchLayout, DeviceTables, Pipeline, ScfOptions
from .param import SpeciesTables


class LargeSystem:
    """Energy and forces of one large molecule on the whole GPU."""

    def __init__(self, numbers, device="cuda:0", opts: ScfOptions | None = None, chrg: float = 0.0,
                 spin: float | None = None):
        self.device = torch.device(device)
        self.opts = opts or ScfOptions()
        z = np.asarray(numbers, dtype=np.int64)
        self.numbers = z[z > 0]
        self.tables = SpeciesTables(np.unique(self.numbers))
        self.layout = BatchLayout([self.numbers], self.tables, padded=False, chrg=np.array([chrg]),
                                  spin=None if spin is None else np.array([spin], dtype=np.float64))
        self.dt = DeviceTables(self.tables, self.device)
        self.pipe = Pipeline(self.layout, self.dt, self.device, large=True)

    def run(self, positions, forces: bool = True) -> dict:
        pipe, dev = self.pipe, self.device
        if isinstance(positions, torch.Tensor):
            pos = positions.detach().to(device=dev, dtype=torch.float64).contiguous()
        else:
            pos = torch.as_tensor(np.asarray(positions), dtype=torch.float64, device=dev).contiguous()
        pipe.run(pos, self.opts, forces=forces)
        rec = pipe.records()[0].cpu().numpy()
        nsh = int(self.layout.nsh[0])
        sh_atom = torch.as_tensor(self.layout.arrays["sh_atom"].astype(np.int64), device=dev)
        qat = torch.zeros(int(self.layout.nat[0]), dtype=torch.float64, device=dev)
        qat.index_add_(0, sh_atom, pipe.qsh[:nsh])
        status = int(rec[_lib.OUT_STATUS])
        res = {"energy": float(rec[_lib.OUT_ESCF] + rec[_lib.OUT_FENERGY] + rec[_lib.OUT_EREP]),
               "scf": float(rec[_lib.OUT_ESCF]), "free_energy": float(rec[_lib.OUT_FENERGY]),
               "repulsion": float(rec[_lib.OUT_EREP]), "iterations": int(rec[_lib.OUT_ITER]),
               "status": status, "converged": (status & _lib.MOL_MAXITER) == 0,
               "charges_orbital": pipe.orbital_charges(0).clone(), "charges_atom": qat,
               "records": rec}
        if forces:
            res["gradient"] = pipe.gradient().clone()
        return res
