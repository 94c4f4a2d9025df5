"""Shared helpers for the GPU tests (run the native pipeline on molecules)."""

import numpy as np
import torch

from dxtb_amd.engine import BatchLayout, DeviceTables, Pipeline, ScfOptions
from dxtb_amd.param import SpeciesTables

TIGHT = dict(x_atol=1e-11, x_atol_max=1e-11, maxiter=300)
TIGHT_ORACLE = {"x_atol": 1e-11, "x_atol_max": 1e-11, "maxiter": 300}


def run_native(mols, padded=False, forces=False, chrg=None, spin=None, fused=False, **opt):
    """mols: list of (numbers, positions[bohr]); chrg, spin: per molecule
    (spin None: automatic, as upstream); fused: one launch (xtb_singlepoint)
    instead of integrals, SCF and forces; returns the Pipeline."""
    dev = torch.device("cuda:0")
    zs = np.unique(np.concatenate([z for z, _ in mols]))
    tb = SpeciesTables(zs)
    lay = BatchLayout([np.asarray(z) for z, _ in mols], tb, padded=padded,
                      chrg=np.zeros(len(mols)) if chrg is None else np.asarray(chrg, dtype=float),
                      spin=None if spin is None else np.asarray(spin, dtype=float))
    pipe = Pipeline(lay, DeviceTables(tb, dev), dev)
    pos = torch.tensor(np.concatenate([p for _, p in mols]), dtype=torch.float64, device=dev)
    o = ScfOptions(**{**TIGHT, **opt})
    pipe.run(pos, o, forces=forces, fused=fused)
    torch.cuda.synchronize()
    return pipe


def native_pipeline(mols, padded=False):
    """One pipeline for a fixed set of molecules, and a function that runs it
    at new positions: (pipe, run(mols, opts, forces))."""
    dev = torch.device("cuda:0")
    zs = np.unique(np.concatenate([z for z, _ in mols]))
    tb = SpeciesTables(zs)
    lay = BatchLayout([np.asarray(z) for z, _ in mols], tb, padded=padded, chrg=np.zeros(len(mols)))
    pipe = Pipeline(lay, DeviceTables(tb, dev), dev)

    def run(ms, opts, forces=False):
        pos = torch.tensor(np.concatenate([p for _, p in ms]), dtype=torch.float64, device=dev)
        pipe.run(pos, opts, forces=forces)
        torch.cuda.synchronize()

    return pipe, run


def block(pipe, buf, b):
    """Real-orbital block of a packed matrix buffer for molecule b (numpy)."""
    m = pipe.matrix(buf, b).cpu().numpy()
    k = int(pipe.L.nao[b])
    return m[:k, :k]
