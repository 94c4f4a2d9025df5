"""Command line single point: ``python -m dxtb_amd FILE [FILE ...] [--grad | --forces]``.

The options follow upstream's CLI (``_src/cli/argparser.py``, driver
``_src/cli/driver.py:118-370``) for the GFN1 energy/gradient path: total
charge and unpaired electrons from ``--chrg``/``--spin`` or from the
``.CHRG``/``.UHF`` files next to each structure, several files as one padded
batch, the SCF options by upstream's names, and the gradient/forces table in
upstream's layout (``driver.py:45-67``).  Everything runs through
`GFN1Calculator` on the GPU; there is no CPU path.
"""

from __future__ import annotations

import argparse
import json

import numpy as np

from . import io


def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="dxtb_amd", description="GFN1-xTB single point on the GPU")
    ap.add_argument("file", nargs="+", help="coordinate file(s): Turbomole coord or xyz")
    ap.add_argument("--filetype", default=None, choices=("xyz", "tm", "coord", "turbomole"),
                    help="file type (default: from the file name)")
    ap.add_argument("-c", "--chrg", type=int, default=None,
                    help="total charge (default: .CHRG next to the file, else 0)")
    ap.add_argument("--spin", "--uhf", dest="spin", type=int, default=None,
                    help="unpaired electrons (default: .UHF next to the file, else 0)")
    ap.add_argument("--method", default="gfn1", choices=("gfn1",), help="only GFN1-xTB")
    ap.add_argument("--dtype", default="float64", choices=("float64", "double"),
                    help="float64 only")
    ap.add_argument("--device", default="cuda:0")
    ap.add_argument("--maxiter", type=int, default=None)
    ap.add_argument("--etemp", "--fermi-etemp", "--fermi_etemp", dest="etemp", type=float, default=None)
    ap.add_argument("--fermi-maxiter", "--fermi_maxiter", dest="fermi_maxiter", type=int, default=None)
    ap.add_argument("--xtol", "--x-atol", dest="xtol", type=float, default=None,
                    help="SCF convergence threshold on the potential")
    ap.add_argument("--damp", type=float, default=None)
    ap.add_argument("--force-convergence", "--force_convergence", dest="force_convergence",
                    action="store_true", help="raise instead of warn on SCF non-convergence")
    g = ap.add_mutually_exclusive_group()
    g.add_argument("--grad", action="store_true", help="print the nuclear gradient dE/dR")
    g.add_argument("--forces", action="store_true", help="print the forces -dE/dR")
    ap.add_argument("--json", action="store_true", help="print one JSON object instead of text")
    return ap


def options(args: argparse.Namespace) -> dict:
    opts: dict = {}
    for key, val in (("maxiter", args.maxiter), ("fermi_etemp", args.etemp),
                     ("fermi_maxiter", args.fermi_maxiter), ("x_atol", args.xtol), ("damp", args.damp)):
        if val is not None:
            opts[key] = val
    if args.force_convergence:
        opts["force_convergence"] = True
    return opts


def load(args: argparse.Namespace):
    """(numbers, positions, chrg, spin) for the files of `args`: one
    structure as (N,), (N, 3); several as a padded batch."""
    structs = [io.read(f, args.filetype) for f in args.file]

    def attr(name: str, given):
        if given is not None:
            return [given] * len(args.file)
        return [io.attr_next_to(f, name) for f in args.file]

    chrg, spin = attr("chrg", args.chrg), attr("spin", args.spin)
    if len(structs) == 1:
        z, p = structs[0]
        return z, p, chrg[0], spin[0]
    z, p = io.pack(structs)
    return z, p, chrg, spin


def format_table(values: np.ndarray, numbers: np.ndarray, title: str = "Forces") -> str:
    """Per-atom vectors in upstream's CLI layout (``driver.py:45-67``)."""
    rule = "-" * 56
    out = ["", "", title, "-" * len(title), "", "  #  Z            dX              dY              dZ", rule]
    for i, row in enumerate(values):
        sym = io.Z2S.get(int(numbers[i]), "?")
        out.append(f"{i + 1:>3}  {sym:<2}  {row[0]:>15.9f} {row[1]:>15.9f} {row[2]:>15.9f}")
    out.append(rule)
    return "\n".join(out)


def main(argv: list[str] | None = None) -> int:
    args = build_parser().parse_args(argv)
    import torch

    from .calculator import GFN1Calculator

    numbers, positions, chrg, spin = load(args)
    dev = torch.device(args.device)
    num = torch.as_tensor(numbers, device=dev)
    pos = torch.as_tensor(positions, dtype=torch.float64, device=dev)
    ch = torch.as_tensor(chrg, dtype=torch.float64, device=dev)
    sp = None if not np.any(spin) else torch.as_tensor(spin, device=dev)
    calc = GFN1Calculator(num, opts=options(args), device=dev, dtype=torch.float64)
    want = args.grad or args.forces
    vec = None
    if want:
        # one SCF with the analytic gradient behind autograd (upstream: the
        # autograd gradient of result.total, driver.py:236-240)
        pos.requires_grad_(True)
        energy = calc.energy(pos, ch, sp)
        (g,) = torch.autograd.grad(energy.sum(), pos)
        vec = (g if args.grad else -g).cpu().numpy()
    else:
        energy = calc.energy(pos, ch, sp)
    e = energy.detach().cpu().numpy()
    if args.json:
        rec = {"files": args.file, "energy": e.tolist(), "unit": "Eh"}
        if vec is not None:
            rec["gradient" if args.grad else "forces"] = vec.tolist()
            rec["vector_unit"] = "Eh/bohr"
        print(json.dumps(rec))
        return 0
    es = np.atleast_1d(e)
    for k, f in enumerate(args.file):
        print(f"{f}: total energy {es[k]: .12f} Eh")
    if vec is not None:
        title = "Gradient" if args.grad else "Forces"
        if vec.ndim == 2:
            print(format_table(vec, numbers, title))
        else:
            for k in range(vec.shape[0]):
                n = int((numbers[k] > 0).sum())
                print(f"\n{args.file[k]}")
                print(format_table(vec[k, :n], numbers[k, :n], title))
    return 0
