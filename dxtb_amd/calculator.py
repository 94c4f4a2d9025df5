"""Drop-in ``GFN1Calculator`` on PyTorch-ROCm tensors.

Mirrors upstream's calculator surface for the GFN1 single point
(src/dxtb/_src/calculators/gfn1.py:38-70, types/base.py:451-734,
types/abc.py:88-170,429-486, types/energy.py:78-441, types/autograd.py:80-201,
types/analytical.py:61-222): same constructor arguments, padded-batch
convention (atomic number 0 = padding), ``get_energy`` / ``get_forces`` /
``get_charges`` / ``reset``, the same option names, exception and warning
types.  All arithmetic runs in the HIP library (FP64); the Python code only
lays out the batch and moves tensors.

Differences from upstream, by design:
* dtype must be float64 (the kernels are FP64 only); ``dtype=None`` means
  float64.
* D3 dispersion needs the reference C6 tables of tad-dftd3, which are not
  part of this package; the dispersion term is left out (with a warning
  unless "disp" is in ``opts["exclude"]``).
* ``scf_mode`` selects the mixer as upstream does: the full (unrolling)
  mode mixes the potential with Anderson (a "broyden" mixer falls back to
  Anderson there, src/dxtb/_src/scf/unrolling/base.py:76-99); the implicit
  modes ("implicit", "default", "nonpure", ...) run xitorch's broyden1 root
  search (src/dxtb/_src/scf/pure/iterator.py:125-147).  Both reach the same
  fixed point; forces are analytic in either mode.  A limited Broyden
  memory (24 updates, then a restart) stands in for upstream's unlimited one.
"""

from __future__ import annotations

import dataclasses
import warnings
from typing import Any

import numpy as np
import torch

from . import _lib
from . import constants as C
from .engine import WG_MAX_N, BatchLayout, DeviceTables, Pipeline, ScfOptions, SplitPipeline
from .exceptions import (
    DeviceError,
    DtypeError,
    PropertyNotImplementedError,
    SCFConvergenceError,
    SCFConvergenceWarning,
)
from .param import SpeciesTables

_DISP_WARNED = False
# a single molecule with more orbitals than this runs its SCF on the whole
# GPU (xtb_scf_large); batches use the per-molecule kernel, except for their
# molecules above its limit (WG_MAX_N orbitals), which take the whole GPU
# one after the other.  Energy + forces of one molecule alone at a 1e-10
# threshold, one CU / whole GPU: n = 100: 5.6 / 6.8 ms, 154: 15.0 / 9.0,
# 236: 19.9 / 11.5, 350: 53 / 16, 460: 99 / 18 (chains; water clusters alike)
LARGE_N = 128


def _opts_to_scf(opts: dict) -> tuple[ScfOptions, dict]:
    o = ScfOptions()
    rest = dict(opts)
    for key, attr in (
        ("maxiter", "maxiter"), ("mix_guess", "mix_guess"), ("damp", "damp"),
        ("damp_init", "damp_init"), ("damp_generations", "generations"),
        ("damp_diagonal_offset", "diag_offset"), ("damp_soft_start", "soft_start"), ("x_atol", "x_atol"),
        ("x_atol_max", "x_atol_max"), ("fermi_etemp", "fermi_etemp"),
        ("fermi_maxiter", "fermi_maxiter"), ("fermi_thresh", "fermi_thresh"),
        # not an upstream option: successive calls on the same molecules
        # start from the previous call's orbitals and potential
        ("warm_start", "warm_start"), ("warm_extrapolate", "warm_extrapolate"),
    ):
        if key in rest:
            setattr(o, attr, rest.pop(key))
    scp = str(rest.pop("scp_mode", "potential")).lower()
    if scp not in ("potential", "1"):
        raise NotImplementedError(f"scp_mode '{scp}': only 'potential' is implemented")
    if rest.pop("damp_dynamic", False):
        raise NotImplementedError("damp_dynamic is not implemented")
    if not 1 <= int(o.generations) <= 6:
        raise ValueError("damp_generations must be in 1..6")
    if "broyden_memory" in rest:
        o.broyden_memory = int(rest.pop("broyden_memory"))
        if not 1 <= o.broyden_memory <= 24:
            raise ValueError("broyden_memory must be in 1..24")
    return o, rest


# upstream's SCF mode names and codes (src/dxtb/_src/constants/labels/scf.py)
_FULL = ("full", "full_tracking", "unrolling", "0")
_IMPLICIT = ("default", "implicit", "1", "implicit_old", "implicit_nonpure", "nonpure", "non-pure",
             "old", "2")
_MIXERS = {"anderson": "anderson", "a": "anderson", "1": "anderson",
           "broyden": "broyden", "b": "broyden", "2": "broyden",
           "linear": "linear", "l": "linear", "simple": "linear", "s": "linear", "0": "linear"}


def _scf_mode(scf_mode, mixer) -> str:
    """The mixer of the device SCF for upstream's scf_mode / mixer options:
    implicit modes always run broyden1 (scf/pure/iterator.py:125-147,
    scf/implicit/default.py: the mixer option is ignored); the full mode
    runs the configured mixer, with "broyden" falling back to Anderson
    (scf/unrolling/base.py:76-99)."""
    m = str(scf_mode).lower()
    mix = _MIXERS.get(str(mixer).lower())
    if mix is None:
        raise ValueError(f"Unknown mixer '{mixer}'.")
    if m in _IMPLICIT:
        return "broyden"
    if m in _FULL:
        if mix == "linear":
            raise NotImplementedError("the simple (linear) mixer is not implemented; use 'anderson'")
        return "anderson"
    if m in ("experimental", "perfect", "single-shot", "single_shot", "3"):
        raise NotImplementedError(f"scf_mode '{scf_mode}' is not implemented")
    raise ValueError(f"Unknown SCF mode '{scf_mode}'.")


class _EnergyFn(torch.autograd.Function):
    """Energy with a hand-written backward: the analytic gradient kernel."""

    @staticmethod
    def forward(ctx, positions, calc, chrg, spin):
        res = calc._run(positions.detach(), chrg, spin, forces=True)
        ctx.save_for_backward(res["gradient_full"])
        return res["energy"].clone()

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        if g.ndim == 2:
            return grad_out * g, None, None, None
        return grad_out[:, None, None] * g, None, None, None


class GFN1Calculator:
    """GFN1-xTB calculator (drop-in for upstream ``dxtb.GFN1Calculator``)."""

    def __init__(self, numbers: torch.Tensor, par: Any = None, *, classical=None,
                 interaction=None, opts: dict | None = None, timer: bool = False,
                 device: torch.device | str | None = None, dtype: torch.dtype | None = None,
                 **kwargs: Any) -> None:
        if not isinstance(numbers, torch.Tensor):
            numbers = torch.as_tensor(numbers)
        if numbers.dtype not in (torch.int16, torch.int32, torch.int64):
            raise DtypeError(f"Dtype '{numbers.dtype}' of 'numbers' must be an integer type.")
        if classical is not None or interaction is not None:
            raise NotImplementedError("custom classical/interaction components are not supported")
        self.dtype = torch.float64 if dtype is None else dtype
        if self.dtype != torch.float64:
            raise DtypeError("this calculator computes in float64 only")
        self.device = torch.device(device) if device is not None else numbers.device
        self.numbers = numbers.to(self.device)
        self.batched = numbers.ndim == 2
        opts = dict(opts or {})
        opts.pop("verbosity", None)
        self.force_convergence = bool(opts.pop("force_convergence", False))
        exclude = opts.pop("exclude", [])
        self.exclude = {exclude} if isinstance(exclude, str) else set(exclude)
        # upstream's result cache is off by default (constants/defaults.py:266)
        self.cache_enabled = bool(opts.pop("cache_enabled", False))
        mode = _scf_mode(opts.pop("scf_mode", "full"), opts.pop("mixer", "broyden"))
        f_atol = opts.pop("f_atol", None)
        for k in ("int_driver", "int_level", "int_uplo", "batch_mode",
                  "method", "strict", "guess", "fermi_partition",
                  "max_element", "anomaly", "grad", "skip_compat_checks", "int_cutoff"):
            opts.pop(k, None)
        self.scf_opts, rest = _opts_to_scf(opts)
        self.scf_opts.mixer = mode
        if f_atol is not None:
            self.scf_opts.f_atol = float(f_atol)
        self.scf_opts.jacobi_max_sweeps = int(kwargs.pop("jacobi_max_sweeps", 40))
        self.layout_padded = bool(kwargs.pop("padded", True))
        if rest:
            raise TypeError(f"unknown options: {sorted(rest)}")
        nums = self.numbers.detach().cpu().numpy().astype(np.int64)
        self._mols = [r[r > 0] for r in nums] if self.batched else [nums[nums > 0]]
        self._mask = (self.numbers > 0)
        zs = np.unique(np.concatenate(self._mols))
        for z in zs:
            if int(z) not in C.SUPPORTED_ELEMENTS:
                raise NotImplementedError(
                    f"Element Z={int(z)} not supported (supported: {C.SUPPORTED_ELEMENTS}).")
        self.tables = SpeciesTables(zs)
        self._dtables = None
        self._pipe = None
        self._pipe_key = None
        self._cache: dict = {}
        global _DISP_WARNED
        if not ({"disp", "all", "scf"} & self.exclude) and not _DISP_WARNED:
            warnings.warn("D3 dispersion is not available in this build (its reference C6 data "
                          "ship with tad-dftd3); energies and forces exclude it. Pass "
                          "opts={'exclude': ['disp']} to silence this warning.", UserWarning)
            _DISP_WARNED = True

    # ------------------------------------------------------------------ core
    def reset(self) -> None:
        self._cache = {}

    def _check(self, positions: torch.Tensor) -> None:
        if positions.dtype != self.dtype:
            raise DtypeError(f"positions dtype {positions.dtype} != calculator dtype {self.dtype}")
        if positions.device != self.device:
            raise DeviceError(f"positions on {positions.device}, calculator on {self.device}")
        if positions.shape[:-1] != self.numbers.shape or positions.shape[-1] != 3:
            raise ValueError("positions must have shape (*numbers.shape, 3)")

    def _charges_array(self, chrg) -> np.ndarray:
        B = len(self._mols)
        if isinstance(chrg, torch.Tensor):
            c = chrg.detach().cpu().numpy().astype(np.float64).reshape(-1)
        else:
            c = np.asarray([float(chrg)])
        if self.batched and c.size == 1:
            if c[0] != 0.0:
                raise ValueError("Cannot set charge for batched calculations with a single "
                                 "float or integer. Please provide a 1D tensor of charges.")
            c = np.zeros(B)
        if c.size != B:
            raise ValueError("charge tensor does not match the batch")
        return c

    def _pipeline(self, chrg, spin) -> Pipeline:
        c = self._charges_array(chrg)
        sp = None if spin is None else np.asarray(
            spin.detach().cpu().numpy() if isinstance(spin, torch.Tensor) else spin, dtype=np.float64)
        key = (c.tobytes(), None if sp is None else sp.tobytes())
        if self._pipe is None or self._pipe_key != key:
            if self._dtables is None:
                self._dtables = DeviceTables(self.tables, self.device)
            layout = BatchLayout(self._mols, self.tables, padded=self.batched and self.layout_padded,
                                 chrg=c, spin=sp)
            # one molecule above LARGE_N orbitals: its SCF on the whole GPU;
            # in a batch, the molecules above the per-molecule kernel's limit
            large = layout.B == 1 and not layout.padded and int(layout.n[0]) > LARGE_N
            if not large and np.any(layout.nao > WG_MAX_N):
                self._pipe = SplitPipeline(layout, self._dtables, self.device)
            else:
                self._pipe = Pipeline(layout, self._dtables, self.device, large=large)
            self._pipe_key = key
        return self._pipe

    def _cached(self, positions: torch.Tensor, ckey, forces: bool, orbitals: bool):
        """The cached run for these positions (compared by value with the
        copy the cache holds), charges and spin, if it computed at least what
        is asked for; only with cache_enabled, as upstream."""
        if not self.cache_enabled:
            return None
        hit = self._cache.get("run")
        if hit is None:
            return None
        (key, pos, had_forces, had_orb), res = hit
        if key != ckey or (forces and not had_forces) or (orbitals and not had_orb):
            return None
        p = positions.detach()
        if pos.shape != p.shape or not torch.equal(pos, p):
            return None
        return res

    def _store(self, positions: torch.Tensor, ckey, forces: bool, orbitals: bool, res: dict) -> None:
        if self.cache_enabled:
            self._cache["run"] = ((ckey, positions.detach().clone(), forces, orbitals), res)

    def _run(self, positions: torch.Tensor, chrg=0, spin=None, forces: bool = False,
             orbitals: bool = False) -> dict:
        self._check(positions)
        ckey = (self._charges_array(chrg).tobytes(), str(spin))
        hit = self._cached(positions, ckey, forces, orbitals)
        if hit is not None:
            return hit
        pipe = self._pipeline(chrg, spin)
        pos = positions.detach()
        flat = pos[self._mask].contiguous() if self.batched else pos[self._mask].contiguous()
        opts = self.scf_opts
        orbitals = orbitals or opts.orbital_energies
        if orbitals != opts.orbital_energies:
            opts = dataclasses.replace(opts, orbital_energies=True)
        # one launch per batch (integral pass, SCF and forces per molecule;
        # the gradient's sums in a fixed order, so the same bits every run)
        pipe.run(flat, opts, forces=forces, fused=True)
        rec = pipe.records().clone()
        status = rec[:, _lib.OUT_STATUS].to(torch.int64).cpu().numpy()
        self._raise_on_status(status)
        use_scf, use_rep = self._components()
        e = torch.zeros(len(self._mols), dtype=torch.float64, device=self.device)
        if use_scf:
            e = e + rec[:, _lib.OUT_ESCF] + rec[:, _lib.OUT_FENERGY]
        if use_rep:
            e = e + rec[:, _lib.OUT_EREP]
        res = {"energy": e if self.batched else e[0], "records": rec, "pipeline": pipe}
        if forces:
            # the forces kernel differentiates E_scf - TS + E_rep; excluded
            # terms are taken out with the repulsion gradient
            if use_scf and use_rep:
                gm = pipe.gradient().clone()
            elif use_scf:
                gm = pipe.gradient() - pipe.repulsion_gradient()
            elif use_rep:
                gm = pipe.repulsion_gradient()
            else:
                gm = torch.zeros_like(pipe.gradient())
            g = torch.zeros(positions.shape, dtype=torch.float64, device=self.device)
            g[self._mask] = gm
            res["gradient_full"] = g
        self._store(positions, ckey, forces, orbitals, res)
        return res

    def _components(self) -> tuple[bool, bool]:
        """(electronic energy, repulsion) included in the total; "disp" and
        "hal" change nothing here (dispersion is not built, halogen bonding
        is zero for H/C/N/O)."""
        ex = self.exclude
        return not ({"scf", "all"} & ex), not ({"rep", "all"} & ex)

    def _raise_on_status(self, status: np.ndarray) -> None:
        if np.any(status & _lib.MOL_OVERLAP):
            # upstream's Cholesky reduction of the generalized eigenproblem
            # (tad-mctc eighb) raises here
            raise torch.linalg.LinAlgError(
                "linalg.cholesky: The factorization could not be completed because the overlap "
                "matrix is not positive-definite (molecules "
                f"{np.nonzero(status & _lib.MOL_OVERLAP)[0].tolist()}).")
        if np.any(status & _lib.MOL_JACOBI):
            msg = ("The eigensolver of the final SCF step reached its sweep limit "
                   f"(molecules {np.nonzero(status & _lib.MOL_JACOBI)[0].tolist()}); "
                   "orbitals and energies may be inaccurate.")
            if self.force_convergence:
                raise SCFConvergenceError(msg)
            warnings.warn(msg, SCFConvergenceWarning)
        if np.any(status & _lib.MOL_FERMI):
            raise RuntimeError("Fermi energy failed to converge.")
        if np.any(status & _lib.MOL_NONFINITE):
            raise RuntimeError("Non-finite values met in the SCF.")
        if np.any(status & _lib.MOL_MAXITER):
            msg = (f"SCF does not converge after {self.scf_opts.maxiter} cycles using Anderson "
                   f"mixing with a damping factor of {self.scf_opts.damp}.")
            if self.force_convergence:
                raise SCFConvergenceError(msg)
            warnings.warn(msg, SCFConvergenceWarning)

    # ----------------------------------------------------------- properties
    def singlepoint(self, positions: torch.Tensor, chrg=0, spin=None, **kwargs) -> dict:
        res = self._run(positions, chrg, spin, forces=False)
        rec = res["records"]
        out = {
            "energy": res["energy"],
            "scf": rec[:, _lib.OUT_ESCF], "fenergy": rec[:, _lib.OUT_FENERGY],
            "repulsion": rec[:, _lib.OUT_EREP],
            "iterations": rec[:, _lib.OUT_ITER].to(torch.int64),
        }
        if not self.batched:
            out = {k: (v[0] if k != "energy" else v) for k, v in out.items()}
        return out

    def energy(self, positions: torch.Tensor, chrg=0, spin=None, **kwargs) -> torch.Tensor:
        if positions.requires_grad and torch.is_grad_enabled():
            return _EnergyFn.apply(positions, self, chrg, spin)
        return self._run(positions, chrg, spin, forces=False)["energy"]

    get_energy = energy

    def forces(self, positions: torch.Tensor, chrg=0, spin=None, **kwargs) -> torch.Tensor:
        if not positions.requires_grad:
            raise RuntimeError("Position tensor needs ``requires_grad=True``.")
        e = self.energy(positions, chrg, spin)
        (g,) = torch.autograd.grad(e.sum(), positions)
        return -g

    def get_forces(self, positions: torch.Tensor, chrg=0, spin=None, grad_mode: str = "autograd",
                   **kwargs) -> torch.Tensor:
        if grad_mode not in ("autograd", "backward", "functorch", "row", "analytical"):
            raise PropertyNotImplementedError(f"grad_mode '{grad_mode}'")
        if grad_mode == "analytical":
            return self.forces_analytical(positions, chrg, spin)
        return self.forces(positions, chrg, spin)

    def forces_analytical(self, positions: torch.Tensor, chrg=0, spin=None, **kwargs) -> torch.Tensor:
        return -self._run(positions, chrg, spin, forces=True)["gradient_full"]

    def get_charges(self, positions: torch.Tensor, chrg=0, spin=None, **kwargs) -> torch.Tensor:
        """Orbital-resolved Mulliken charges, shape (..., nao) (types/abc.py:429-486)."""
        res = self._run(positions, chrg, spin, forces=False)
        pipe = res["pipeline"]
        L = pipe.L
        nmax = int(L.nao.max())
        out = torch.zeros((L.B, nmax), dtype=torch.float64, device=self.device)
        for b in range(L.B):
            out[b, : int(L.nao[b])] = pipe.orbital_charges(b)
        return out if self.batched else out[0]

    def get_param_gradient(self, positions: torch.Tensor, chrg=0, spin=None) -> dict:
        """dE/dtheta for the GFN1 parameters of the Hamiltonian, the Coulomb
        terms and the repulsion, by element and under the parameter file's
        names and units (src/dxtb/_src/param/gfn1/gfn1-xtb.toml): per shell
        ``levels`` and ``kcn`` (eV), ``shpoly``, ``lgam``, ``slater``; per
        element ``gam``, ``gam3``, ``zeff``, ``arep``, ``en``; and under
        ``"hamiltonian"`` the pair scaling's ``shell`` (ss, sp, pp),
        ``kpol``, ``enscale`` and ``kpair`` entries.  Each value has a leading
        batch dimension in batched mode.  Upstream obtains these by autograd
        with ParamModule.set_differentiable (param/module/utils.py:123-240);
        here the energy's stationarity in the density gives them from the
        converged density and charges without a response term."""
        from .param import elements_by_z

        res = self._run(positions, chrg, spin, forces=True)
        g = res["pipeline"].param_grad()
        tb = self.tables
        d_zeta = res["pipeline"].param_grad_basis(tb)
        nus, nsp = tb.nus, len(tb.zs)
        # the pair scaling K_ab: chain rule to the Hamiltonian section's
        # parameters and the electronegativities
        from .param import load_gfn1, shell_pair_k

        d_K = g[:, 4 * nus + 3 * nsp:].reshape(-1, nus, nus)
        en = {z: e["en"] for z, (_, e) in elements_by_z().items()}
        _, dks = shell_pair_k(tb, load_gfn1()["hamiltonian"], en, grad=True)
        ham: dict = {"shell": {}, "kpair": {}}
        d_en = {}
        for name, dk in dks:
            v = (d_K * torch.as_tensor(dk, device=g.device)).sum((-1, -2))
            if name[0] in ("shell", "kpair"):
                ham[name[0]][name[1]] = v if self.batched else v[0]
            elif name[0] == "en":
                d_en[name[1]] = v
            else:
                ham[name[0]] = v if self.batched else v[0]
        d_level, d_kcn, d_poly, d_eta = (g[:, k * nus:(k + 1) * nus] for k in range(4))
        d_gam3, d_zeff, d_arep = (g[:, 4 * nus + k * nsp:4 * nus + (k + 1) * nsp] for k in range(3))
        els = elements_by_z()
        out = {}
        for i, z in enumerate(tb.zs):
            sym, e = els[z]
            us = np.nonzero(tb.us_sp == i)[0]
            rec = tb.us_rec[us]
            nrec = len(e["levels"])

            def per_shell(v, scale=1.0):
                t = torch.zeros((g.shape[0], nrec), dtype=g.dtype, device=g.device)
                t[:, torch.as_tensor(rec, device=g.device)] = v[:, torch.as_tensor(us, device=g.device)] * scale
                return t

            lgam = torch.as_tensor(np.asarray(e["lgam"], dtype=np.float64)[rec], device=g.device)
            eta = d_eta[:, torch.as_tensor(us, device=g.device)]
            d = {
                "levels": per_shell(d_level, C.EV2AU), "kcn": per_shell(d_kcn, C.EV2AU),
                "shpoly": per_shell(d_poly), "lgam": per_shell(d_eta, float(e["gam"])),
                "gam": (eta * lgam).sum(-1), "gam3": d_gam3[:, i], "zeff": d_zeff[:, i],
                "arep": d_arep[:, i], "slater": per_shell(d_zeta),
                "en": d_en.get(z, torch.zeros_like(d_arep[:, i])),
            }
            out[sym] = d if self.batched else {k: v[0] for k, v in d.items()}
        out["hamiltonian"] = ham
        return out

    def get_iterations(self, positions, chrg=0, spin=None):
        return self.singlepoint(positions, chrg, spin)["iterations"]

    def get_mo_energies(self, positions, chrg=0, spin=None):
        res = self._run(positions, chrg, spin, forces=False, orbitals=True)
        pipe = res["pipeline"]
        L = pipe.L
        out = [pipe.emo[int(L.vec0[b]): int(L.vec0[b]) + int(L.n[b])].clone() for b in range(L.B)]
        return torch.stack(out) if self.batched and L.padded else (out if self.batched else out[0])
