"""Oracle values for the config-3 system (600 H2O, 1800 atoms, 4800
orbitals) written to config3_oracle.npz.

The oracle (CPU, FP64) needs a few minutes and ~20 GB of memory here; the
GPU test compares the whole-GPU path against these numbers.

    python tests/golden/make_config3.py
"""

import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle.gfn1 as og  # noqa: E402
from dxtb_amd import workloads as W  # noqa: E402


def main():
    torch.set_num_threads(os.cpu_count() or 1)
    z, p = W.config3()
    t = time.time()
    r = og.singlepoint(z, p, opts={"x_atol": 1e-10, "x_atol_max": 1e-10, "maxiter": 300},
                       with_forces="--forces" in sys.argv)
    out = {
        "numbers": np.asarray(z), "positions": np.asarray(p),
        "energy": np.array(r["energy"]), "scf": np.array(r["scf"]),
        "repulsion": np.array(r["repulsion"]), "fenergy": np.array(r["fenergy"]),
        "iterations": np.array(r["iterations"]),
        "charges_atom": r["charges_atom"].numpy(),
    }
    if "gradient" in r:
        out["gradient"] = r["gradient"].numpy()
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "config3_oracle.npz"), **out)
    print(f"energy {r['energy']:.12f} iterations {r['iterations']} ({time.time() - t:.0f} s)")


if __name__ == "__main__":
    main()
