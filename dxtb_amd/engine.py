"""Ragged batch descriptor and the native GFN1 pipeline.

A *batch* is a list of molecules, each given by its real atoms (atomic
numbers > 0) and positions in bohr.  ``BatchLayout`` builds, with vectorised
numpy on the host, the index maps the kernels need (atom -> shell ->
orbital, offsets of every molecule's packed matrices, tile lists of the
shell-pair and atom-pair spaces).  In the *ragged* layout each molecule's
matrices are n_b x n_b with n_b its own orbital count; in the *padded*
layout every molecule gets the batch-maximum dimension and its padding
orbitals carry a unit overlap diagonal and zero Hamiltonian, which is what
upstream's padded batch computes (src/dxtb/_src/scf/unrolling/base.py:122-139).

``Pipeline.run`` then calls, on one HIP stream, the entry points of
include/xtb_hip.h: S/H0/CN, gamma, repulsion, the persistent SCF and
optionally the analytic gradient.
"""

from __future__ import annotations

import ctypes as ct
import functools
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from . import constants as C
from .param import SpeciesTables

MAT_ALIGN = 32  # doubles (256 bytes) between molecules' matrices


def _cumsum0(a: np.ndarray) -> np.ndarray:
    out = np.zeros(len(a) + 1, dtype=np.int64)
    np.cumsum(a, out=out[1:])
    return out


def alpha_beta(nel: np.ndarray, spin) -> tuple[np.ndarray, np.ndarray]:
    """Alpha/beta electrons (src/dxtb/_src/wavefunction/filling.py:41-113).

    Not rounded: the SCF kernel rounds them for Fermi smearing as upstream
    rounds the occupation sum (scf/base.py:879) and fills the unrounded
    counts by the aufbau principle at zero temperature.  A spin that does not
    fit the electron count raises ValueError, as upstream (:80-93)."""
    nel = np.asarray(nel, dtype=np.float64)
    nr = np.round(nel)
    if spin is None:
        uhf = np.zeros_like(nel)
    else:
        uhf = np.broadcast_to(np.asarray(spin, dtype=np.float64), nel.shape)
        if np.any(uhf > nr):
            raise ValueError(f"Number of unpaired electrons ({uhf}) larger than "
                             f"number of electrons ({nel}).")
        if np.any(np.remainder(uhf, 2) != np.remainder(nr, 2)):
            raise ValueError(f"Odd (even) number of unpaired electrons ({uhf}) but even "
                             f"(odd) number of electrons ({nel}) given.")
    nuhf = np.where(np.remainder(uhf, 2) == np.remainder(nr, 2), uhf, np.remainder(nr, 2))
    diff = np.minimum(nuhf, nel)
    nb = (nel - diff) / 2.0
    na = nb + diff
    return na, nb


@functools.lru_cache(maxsize=None)
def _triu16(t: int) -> np.ndarray:
    """(i0, j0) of the upper-triangle 16x16 tiles of a t x t tile grid."""
    iu, ju = np.triu_indices(t)
    return np.stack([iu * 16, ju * 16], axis=1).astype(np.int32)


def _tiles(counts: np.ndarray) -> np.ndarray:
    """Upper-triangle 16x16 tiles of every molecule's pair space: rows (mol, i0, j0)."""
    t = (np.asarray(counts, dtype=np.int64) + 15) // 16
    ntile = t * (t + 1) // 2
    if ntile.sum() == 0:
        return np.zeros((0, 3), dtype=np.int32)
    out = np.empty((int(ntile.sum()), 3), dtype=np.int32)
    out[:, 0] = np.repeat(np.arange(len(t), dtype=np.int32), ntile)
    out[:, 1:] = np.concatenate([_triu16(int(x)) for x in t if x > 0])
    return out


@dataclass
class BatchLayout:
    numbers: list  # list of int arrays (real atoms)
    tables: SpeciesTables
    padded: bool
    chrg: np.ndarray
    spin: object = None
    arrays: dict = field(default_factory=dict)

    def __post_init__(self):
        tb = self.tables
        mols = [np.asarray(z, dtype=np.int64) for z in self.numbers]
        B = len(mols)
        nat = np.array([len(z) for z in mols], dtype=np.int64)
        if B == 0 or np.any(nat == 0):
            raise ValueError("every molecule needs at least one atom")
        Z = np.concatenate(mols)
        sp = tb.sp_of_z[Z]
        if np.any(sp < 0):
            raise ValueError("element outside the species tables")
        atom0 = _cumsum0(nat)
        atom_mol = np.repeat(np.arange(B), nat)
        atom_nsh = tb.nsh_of_sp[sp].astype(np.int64)
        nsh = np.add.reduceat(atom_nsh, atom0[:-1])
        sh0 = _cumsum0(nsh)
        first_sh = _cumsum0(atom_nsh)[:-1]
        sh_atom_g = np.repeat(np.arange(len(Z)), atom_nsh)
        k_in_atom = np.arange(sh0[-1]) - first_sh[sh_atom_g]
        sh_us = tb.us_base[sp[sh_atom_g]] + k_in_atom
        sh_mol = atom_mol[sh_atom_g]
        sh_size = 2 * tb.us_l[sh_us].astype(np.int64) + 1
        nao = np.add.reduceat(sh_size, sh0[:-1])
        sh_start_g = _cumsum0(sh_size)[:-1]
        ao0_real = _cumsum0(nao)
        sh_ao = sh_start_g - ao0_real[sh_mol]
        n = np.full(B, nao.max()) if self.padded else nao.copy()
        vec0 = _cumsum0(n)
        ao_sh = np.full(vec0[-1], -1, dtype=np.int64)
        n0 = np.zeros(vec0[-1])
        ao_shell_g = np.repeat(np.arange(sh0[-1]), sh_size)
        ao_mol = sh_mol[ao_shell_g]
        slot = vec0[ao_mol] + (np.arange(ao0_real[-1]) - ao0_real[ao_mol])
        ao_sh[slot] = ao_shell_g - sh0[ao_mol]
        n0[slot] = tb.us_refocc[sh_us[ao_shell_g]] / sh_size[ao_shell_g]
        nel = np.add.reduceat(n0, vec0[:-1]) - self.chrg
        na, nb = alpha_beta(nel, self.spin)
        msz = ((n * n + MAT_ALIGN - 1) // MAT_ALIGN) * MAT_ALIGN
        mat0 = _cumsum0(msz)
        gam0 = _cumsum0(nsh * nsh)
        self.B, self.nat, self.nsh, self.nao, self.n = B, nat, nsh, nao, n
        self.atom0, self.sh0, self.vec0, self.mat0, self.gam0 = atom0, sh0, vec0, mat0, gam0
        self.Z = Z
        self.nel_a, self.nel_b = na, nb
        self.sh_us = sh_us
        self.ao_sh = ao_sh
        i32 = np.int32
        self.arrays = {
            "mol_nat": nat.astype(i32), "mol_nsh": nsh.astype(i32), "mol_nao": nao.astype(i32),
            "mol_n": n.astype(i32), "mol_atom0": atom0[:-1].astype(i32),
            "mol_sh0": sh0[:-1].astype(i32), "mol_vec0": vec0[:-1].astype(i32),
            "mol_mat0": mat0[:-1].astype(np.int64), "mol_gam0": gam0[:-1].astype(np.int64),
            "mol_nel": np.stack([na, nb], axis=1).reshape(-1).astype(np.float64),
            "atom_sp": sp.astype(i32), "atom_mol": atom_mol.astype(i32),
            "atom_sh0": (first_sh - sh0[atom_mol]).astype(i32), "atom_nsh": atom_nsh.astype(i32),
            "sh_atom": (sh_atom_g - atom0[sh_mol]).astype(i32), "sh_us": sh_us.astype(i32),
            "sh_ao": sh_ao.astype(i32), "ao_sh": ao_sh.astype(i32), "n0": n0,
            "tile_sh": _tiles(nsh).reshape(-1), "tile_at": _tiles(nat).reshape(-1),
        }
        self.ntile_sh = len(self.arrays["tile_sh"]) // 3
        self.ntile_at = len(self.arrays["tile_at"]) // 3
        self.nmat_total = int(mat0[-1])
        self.ngam_total = int(gam0[-1])

    @property
    def nvec_total(self) -> int:
        return int(self.vec0[-1])

    def work_doubles(self) -> int:
        """Device doubles held by one pipeline run (for admission control)."""
        gen = 6
        return int(7 * self.nmat_total + self.ngam_total + (9 + 2 * gen) * self.nvec_total
                   + 8 * int(self.nat.sum()) + self.B * _lib.OUT_STRIDE)


class DeviceTables:
    """Parameter arrays of a SpeciesTables on the device + the ctypes struct."""

    def __init__(self, tb: SpeciesTables, device):
        f = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)
        i = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)
        self.t = {
            "sp_rad": f(tb.sp_rad), "sp_rcov": f(tb.sp_rcov), "sp_gam3": f(tb.sp_gam3),
            "sp_zeff": f(tb.sp_zeff), "sp_arep": f(tb.sp_arep), "us_level": f(tb.us_level),
            "us_kcn": f(tb.us_kcn), "us_shpoly": f(tb.us_shpoly), "us_hard": f(tb.us_hard),
            "us_l": i(tb.us_l), "us_nprim": i(tb.us_nprim), "us_alpha": f(tb.us_alpha),
            "us_coeff": f(tb.us_coeff), "us_K": f(tb.us_K),
        }
        from .param import load_gfn1

        kexp = load_gfn1()["repulsion"]["kexp"]
        self.s = _lib.XtbParams(
            nsp=len(tb.zs), nus=tb.nus, kexp=kexp, rep_cutoff=C.REPULSION_CUTOFF,
            cn_cutoff=C.CN_CUTOFF, cn_steep=C.CN_STEEPNESS, int_cutoff=C.INTCUTOFF,
            **{k: v.data_ptr() for k, v in self.t.items()},
        )


@dataclass
class ScfOptions:
    maxiter: int = 100
    mix_guess: bool = True
    # Anderson history (upstream: 5); 6 without the soft start below takes
    # ~20 % fewer SCF iterations to the same fixed point
    generations: int = 6
    fermi_maxiter: int = 200
    fermi_etemp: float = 300.0
    fermi_thresh: float | None = None
    x_atol: float = 1e-4
    x_atol_max: float = 1e-5
    damp: float = 0.5
    damp_init: float = 0.1
    diag_offset: float = 0.01
    jacobi_max_sweeps: int = 40
    jacobi_tol: float = 1e-13
    jacobi_inner: int = 1
    eig_inexact: float = 1e-2
    # intermediate SCF steps decouple occupied from virtual orbitals only
    # (all that the density needs) when the frontier gaps exceed this many kT
    bd_gap_kt: float = 60.0
    # closed shells: decouple by one perturbative occupied-virtual rotation
    # (matrix-core GEMMs) when its largest amplitude is at most this
    pd_kmax: float = 0.2
    # eigensolve tolerance of the first SCF step (inexact SCF only); with the
    # default mixer 3e-2 takes ~12 % less SCF time than 1e-2 (one or two more
    # iterations, fewer first-step Jacobi sweeps), 1e-1 diverges into more
    # iterations
    first_tol: float = 3e-2
    # orthogonaliser of the first step: 0 Loewdin S^-1/2 (Newton-Schulz),
    # 1 Cholesky L^-1 (same eigenpairs, ~10x fewer flops)
    ortho: int = 1
    # form the orbital energies in the final step (a full eigensolve); off,
    # gapped closed shells stop at the occupied-virtual decoupling, which
    # fixes the density, energies and W but not the orbital energies
    orbital_energies: bool = False
    # molecular dynamics / optimisation: keep each molecule's orbitals and
    # potential in the pipeline and start the next call on the same batch
    # from them (new positions); same fixed point, fewer iterations
    warm_start: bool = False
    # warm starts from the second on: start from the potential extrapolated
    # linearly from the last two calls (for smooth trajectories: MD with a
    # small time step, optimisation paths); the fixed point is the same
    warm_extrapolate: bool = False
    # upstream's soft start: simple mixing with damp_init for the first
    # `generations` steps before Anderson acceleration (warm starts always
    # skip it); off, Anderson starts at the second step
    soft_start: bool = False
    # "anderson" (upstream's full mode) or "broyden" (upstream's implicit
    # modes: xitorch's broyden1 root search on v - f(v), converged when the
    # last step's |dv| < x_atol and |v - f(v)| < f_atol)
    mixer: str = "anderson"
    f_atol: float = 1e-4
    # Broyden memory (updates kept before it restarts; upstream keeps all)
    broyden_memory: int = 24

    def history_slots(self) -> int:
        """Vectors per molecule in each of the two mixer history buffers."""
        return (self.broyden_memory if self.mixer == "broyden" else self.generations) + 1

    def to_struct(self, want_w: bool, padded: bool, warm: bool = False, wg_skip: int = 0) -> _lib.XtbScfOpts:
        if self.mixer not in ("anderson", "broyden"):
            raise ValueError(f"Unknown mixer '{self.mixer}'.")
        if not self.fermi_etemp >= 0.0:
            # filling.py:320-324
            raise ValueError(f"Electronic Temperature must be positive or None ({self.fermi_etemp}).")
        thr = self.fermi_thresh if self.fermi_thresh is not None else float(np.finfo(np.float64).eps) ** 0.5
        return _lib.XtbScfOpts(
            maxiter=self.maxiter, mix_guess=int(self.mix_guess),
            generations=self.history_slots() - 1,
            fermi_maxiter=self.fermi_maxiter, want_w=int(want_w),
            jacobi_max_sweeps=self.jacobi_max_sweeps, padded=int(padded),
            jacobi_inner=self.jacobi_inner, eig_inexact=self.eig_inexact,
            x_atol=self.x_atol, x_atol_max=self.x_atol_max, kt=self.fermi_etemp * C.KELVIN2AU,
            damp=self.damp, damp_init=self.damp_init, diag_offset=self.diag_offset,
            fermi_thresh=thr, jacobi_tol=self.jacobi_tol, bd_gap_kt=self.bd_gap_kt,
            pd_kmax=self.pd_kmax, first_tol=self.first_tol, ortho=self.ortho,
            orbitals=int(self.orbital_energies), keep_state=int(self.warm_start),
            warm_start=int(self.warm_start and warm), soft_start=int(self.soft_start),
            wg_skip=int(wg_skip), mixer=int(self.mixer == "broyden"), f_atol=self.f_atol,
        )


class HostArena:
    """A pinned host buffer that successive launches on one stream reuse:
    the layout upload, the positions and the staged results of a launch are
    carved from it.  New pinned allocations synchronise the whole device in
    the HIP runtime, so a serving loop allocates them once, not per launch.
    Reuse is safe once the stream's previous launch has been retired (its
    copies are complete)."""

    def __init__(self, nbytes: int = 16 << 20):
        self.buf = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
        self.off = 0

    @staticmethod
    def need(layout: "BatchLayout") -> int:
        """Upper bound of the bytes one launch on `layout` takes."""
        nat = int(layout.nat.sum())
        # the layout arrays, the launch order and the zeroed result records
        lay = (sum(a.nbytes + 256 for a in layout.arrays.values()) + 4 * layout.B + 256
               + 8 * layout.B * _lib.OUT_STRIDE + 256)
        return lay + 2 * (24 * nat + 256) + 8 * layout.B * _lib.OUT_STRIDE + 1024

    def reserve(self, nbytes: int) -> None:
        """Start a launch: grow (one new pinned allocation) if it would not fit."""
        if nbytes > self.buf.numel():
            self.buf = torch.empty(max(nbytes, 2 * self.buf.numel()), dtype=torch.uint8, pin_memory=True)
        self.off = 0

    def take(self, nbytes: int) -> torch.Tensor:
        n = (nbytes + 255) // 256 * 256
        if self.off + n > self.buf.numel():
            raise RuntimeError("host arena exhausted: reserve() the launch's need first")
        v = self.buf[self.off:self.off + nbytes]
        self.off += n
        return v


class Pipeline:
    """Device buffers for one batch and the calls into the HIP library."""

    def __init__(self, layout: BatchLayout, dtables: DeviceTables, device, arena: HostArena | None = None,
                 large: bool = False):
        """``large``: one molecule, whose SCF runs on the whole GPU
        (xtb_scf_large: grid kernels, a host-side driver) instead of one
        workgroup (xtb_scf)."""
        if large and (layout.B != 1 or layout.padded):
            raise ValueError("the whole-GPU path takes one molecule in the ragged layout")
        self.large = large
        self.work = None
        self.L = layout
        self.arena = arena
        self.dt = dtables
        self.device = torch.device(device)
        self.lib = _lib.load()
        dev = self.device
        self.d = {}
        # every layout array (and the launch order) packed into one pinned
        # staging buffer and one asynchronous copy: a pageable copy would
        # block the host until the stream (or a busy stream sharing its
        # hardware queue) reaches it, and one copy costs less host time than
        # twenty.  The staging buffer lives as long as the pipeline.
        order = np.argsort(-layout.n, kind="stable").astype(np.int32)  # largest first
        arrays = dict(layout.arrays)
        arrays["order"] = order
        # the result records start at zero (a warm start reads the previous
        # status): zeros in the same upload instead of a fill kernel
        arrays["out"] = np.zeros(layout.B * _lib.OUT_STRIDE)
        offs, total = {}, 0
        for k, a in arrays.items():
            a = np.ascontiguousarray(a)
            arrays[k] = a
            offs[k] = total
            total += (a.nbytes + 255) // 256 * 256
        total = max(total, 256)
        if arena is not None:
            self._staging = arena.take(total)
            host = self._staging.numpy()
        else:
            host = np.zeros(total, dtype=np.uint8)
        for k, a in arrays.items():
            host[offs[k]:offs[k] + a.nbytes] = a.view(np.uint8).reshape(-1)
        if arena is None:
            self._staging = torch.from_numpy(host).pin_memory()
        buf = self._staging.to(dev, non_blocking=True)
        tdt = {np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
               np.dtype(np.float64): torch.float64}
        for k, a in arrays.items():
            self.d[k] = buf[offs[k]:offs[k] + max(a.nbytes, a.itemsize)].view(tdt[a.dtype])[: a.size]
        z = lambda k: torch.empty(k, dtype=torch.float64, device=dev)
        L = layout
        nm, nv = L.nmat_total, L.nvec_total
        natt = int(L.nat.sum())
        self.S, self.H0 = z(nm), z(nm)
        self.F, self.C, self.T, self.P = z(nm), z(nm), z(nm), z(nm)
        self.gamma = z(max(L.ngam_total, 1))
        self.cn = z(natt)
        self.pot, self.vnew, self.q, self.eorb, self.emo = z(nv), z(nv), z(nv), z(nv), z(nv)
        self.occ = z(2 * nv)
        self.qsh = z(int(L.sh0[-1]))
        self.out = self.d["out"]
        self._kept = False  # the state buffers hold orbitals and potential to warm-start from
        self._pot_prev = None  # kept potential of the call before the last (extrapolation)
        self._have_w = False  # the last run formed the energy-weighted density
        self.grad = z(3 * natt)
        self.dedcn = z(natt)
        self.hist_x = self.hist_f = None
        self.order = self.d["order"]
        a = self.d
        self.batch = _lib.XtbBatch(
            nmol=L.B, nat_total=natt, nsh_total=int(L.sh0[-1]), nvec_total=nv,
            max_n=int(L.n.max()), max_nsh=int(L.nsh.max()), max_nat=int(L.nat.max()),
            ntile_sh=L.ntile_sh, ntile_at=L.ntile_at, has_padding=int((L.n > L.nao).any()), nmat_total=nm,
            ngam_total=L.ngam_total,
            **{k: a[k].data_ptr() for k in (
                "mol_nat", "mol_nsh", "mol_nao", "mol_n", "mol_atom0", "mol_sh0", "mol_vec0",
                "mol_mat0", "mol_gam0", "mol_nel", "atom_sp", "atom_mol", "atom_sh0", "atom_nsh",
                "sh_atom", "sh_us", "sh_ao", "ao_sh", "n0", "tile_sh", "tile_at")},
        )

    def _state(self) -> _lib.XtbScfState:
        return _lib.XtbScfState(
            F=self.F.data_ptr(), C=self.C.data_ptr(), T=self.T.data_ptr(), P=self.P.data_ptr(),
            pot=self.pot.data_ptr(), vnew=self.vnew.data_ptr(), q=self.q.data_ptr(),
            eorb=self.eorb.data_ptr(), emo=self.emo.data_ptr(), occ=self.occ.data_ptr(),
            hist_x=self.hist_x.data_ptr(), hist_f=self.hist_f.data_ptr(), qsh=self.qsh.data_ptr(),
            out=self.out.data_ptr(), order=self.order.data_ptr(),
        )

    def run(self, positions: torch.Tensor, opts: ScfOptions, forces: bool, stream=None,
            timing: bool = False, wg_skip: int = 0, fused: bool = False) -> None:
        """positions: (nat_total, 3) float64 on the device, bohr.

        With ``timing`` HIP events bracket the SCF launch on its stream
        (``self.scf_events``) and the integral pass (CN, S, H0, gamma:
        ``self.int_events``, which includes the repulsion energy) for
        kernel-duration measurements.  ``wg_skip``
        (0..7) shifts the SCF launch's workgroups to later XCDs.  ``fused``:
        integrals, SCF and forces in one launch (xtb_singlepoint; the
        gradient's per-atom sums in a fixed order); ``scf_events`` then
        bracket that launch and there are no ``int_events``."""
        if positions.dtype != torch.float64 or not positions.is_contiguous():
            raise ValueError("positions must be contiguous float64")
        if positions.shape != (int(self.L.nat.sum()), 3):
            raise ValueError("positions do not match the batch layout")
        self.positions = positions
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        s = ct.c_void_p(st.cuda_stream)
        lib, b, p = self.lib, ct.byref(self.batch), ct.byref(self.dt.s)
        g = opts.history_slots()
        if self.hist_x is None or self.hist_x.numel() != g * self.L.nvec_total:
            self.hist_x = torch.empty(g * self.L.nvec_total, dtype=torch.float64, device=self.device)
            self.hist_f = torch.empty_like(self.hist_x)
        pos = positions.data_ptr()
        fused = fused and not self.large
        self.int_events = None
        if not fused:
            if timing:
                ei0, ei1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                ei0.record(st)
            _lib.check(lib.xtb_build_integrals(b, p, pos, self.S.data_ptr(), self.H0.data_ptr(),
                                               self.cn.data_ptr(), self.gamma.data_ptr(),
                                               self.out.data_ptr(), s), "xtb_build_integrals")
            if timing:
                ei1.record(st)
                self.int_events = (ei0, ei1)
        # a warm start needs the state a previous keep_state launch on these
        # buffers left (in stream order, so the launches need no other sync)
        if self._kept and opts.warm_extrapolate:
            # start from 2 v_(k-1) - v_(k-2): along a smooth trajectory the
            # potential moves about linearly from step to step
            with torch.cuda.stream(st):
                last = self.pot.clone()
                if self._pot_prev is not None:
                    self.pot.mul_(2.0).sub_(self._pot_prev)
                self._pot_prev = last
        elif not self._kept:
            self._pot_prev = None
        o = opts.to_struct(want_w=forces, padded=self.L.padded, warm=self._kept, wg_skip=wg_skip)
        self._kept = False
        self.state = self._state()
        if timing:
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(st)
        if self.large:
            if self.work is None:
                nw = ct.c_int64(0)
                _lib.check(lib.xtb_scf_large_workspace(b, ct.byref(nw)), "xtb_scf_large_workspace")
                self.work = torch.empty(int(nw.value), dtype=torch.float64, device=self.device)
            _lib.check(lib.xtb_scf_large(b, p, self.S.data_ptr(), self.H0.data_ptr(),
                                         self.gamma.data_ptr(), ct.byref(o), ct.byref(self.state),
                                         self.work.data_ptr(), self.work.numel(), s),
                       "xtb_scf_large")
        elif fused:
            _lib.check(lib.xtb_singlepoint(b, p, pos, self.S.data_ptr(), self.H0.data_ptr(), self.cn.data_ptr(),
                                           self.gamma.data_ptr(), ct.byref(o), ct.byref(self.state),
                                           self.dedcn.data_ptr() if forces else None,
                                           self.grad.data_ptr() if forces else None, s),
                       "xtb_singlepoint")
        else:
            _lib.check(lib.xtb_scf(b, p, self.S.data_ptr(), self.H0.data_ptr(),
                                   self.gamma.data_ptr(), ct.byref(o), ct.byref(self.state), s),
                       "xtb_scf")
        if timing:
            ev1.record(st)
            self.scf_events = (ev0, ev1)
        self._kept = bool(opts.warm_start)
        self._have_w = bool(forces)
        if forces and not fused:
            self.run_forces(stream)

    def stage_results(self, stream=None) -> None:
        """Queue copies of the result records and (after a run with forces)
        the gradient into pinned host buffers behind the launches on
        `stream`: once the stream passes them, ``host_records`` and
        ``host_gradient`` hold the results with no further device access."""
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        def host(n):
            if self.arena is None:
                return torch.empty(n, dtype=torch.float64, pin_memory=True)
            return self.arena.take(8 * n).view(torch.float64)

        with torch.cuda.stream(st):
            self.h_out = host(self.out.numel())
            self.h_out.copy_(self.out, non_blocking=True)
            self.h_grad = None
            if self._have_w:
                self.h_grad = host(self.grad.numel())
                self.h_grad.copy_(self.grad, non_blocking=True)

    def host_records(self) -> np.ndarray:
        return self.h_out.numpy().reshape(self.L.B, _lib.OUT_STRIDE)

    def host_gradient(self) -> np.ndarray | None:
        return None if self.h_grad is None else self.h_grad.numpy().reshape(-1, 3)

    def run_forces(self, stream=None) -> None:
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        s = ct.c_void_p(st.cuda_stream)
        _lib.check(self.lib.xtb_forces(ct.byref(self.batch), ct.byref(self.dt.s),
                                       self.positions.data_ptr(), self.cn.data_ptr(),
                                       self.gamma.data_ptr(), ct.byref(self.state),
                                       self.dedcn.data_ptr(), self.grad.data_ptr(), s),
                   "xtb_forces")

    def repulsion_gradient(self, stream=None) -> torch.Tensor:
        """dE_rep/dR of the last run's positions, (nat_total, 3)."""
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        g = torch.empty_like(self.grad)
        _lib.check(self.lib.xtb_repulsion_gradient(ct.byref(self.batch), ct.byref(self.dt.s),
                                                   self.positions.data_ptr(), g.data_ptr(),
                                                   ct.c_void_p(st.cuda_stream)),
                   "xtb_repulsion_gradient")
        return g.view(-1, 3)

    def param_grad(self, stream=None) -> torch.Tensor:
        """dE/dtheta of the last run's energies for the species tables'
        parameters: (B, 4 nus + 3 nsp + nus^2) = level, kcn, shpoly, eta (per
        unique shell), gam3, zeff, arep (per species), us_K (per ordered
        unique-shell pair), atomic units."""
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        tb = self.dt
        nus, nsp = int(tb.s.nus), int(tb.s.nsp)
        g = torch.empty((self.L.B, 4 * nus + 3 * nsp + nus * nus), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.xtb_param_grad(ct.byref(self.batch), ct.byref(tb.s), self.positions.data_ptr(),
                                           self.cn.data_ptr(), self.S.data_ptr(), ct.byref(self.state),
                                           g.data_ptr(), ct.c_void_p(st.cuda_stream)),
                   "xtb_param_grad")
        return g

    def param_grad_basis(self, tables, rel_step: float = 1e-4, stream=None) -> torch.Tensor:
        """dE/dzeta of the last run's energies for the Slater exponent of
        every unique shell of `tables` (the SpeciesTables this pipeline's
        DeviceTables were made from): (B, nus).  Needs a run with forces
        (the energy-weighted density).  dS/dzeta is a central difference of
        the contracted overlaps at zeta (1 +- rel_step)."""
        from .param import SpeciesTables, elements_by_z

        if not self._have_w:
            raise RuntimeError("parameter gradients of the basis need a run with forces")
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        s = ct.c_void_p(st.cuda_stream)
        els = elements_by_z()
        nus = tables.nus
        g = torch.zeros((self.L.B, nus), dtype=torch.float64, device=self.device)
        keep = []
        for u in range(nus):
            z = tables.zs[int(tables.us_sp[u])]
            zeta = list(els[z][1]["slater"])
            k = int(tables.us_rec[u])
            h = rel_step * zeta[k]
            side = []
            for sg in (1.0, -1.0):
                zz = list(zeta)
                zz[k] = zeta[k] + sg * h
                t = SpeciesTables(tables.zs, slater={z: zz})
                side.append((torch.as_tensor(t.us_alpha, device=self.device),
                             torch.as_tensor(t.us_coeff, device=self.device), t))
            tp, tm = side[0][2], side[1][2]
            diff = (np.abs(tp.us_alpha - tm.us_alpha) + np.abs(tp.us_coeff - tm.us_coeff))
            changed = diff.reshape(nus, -1).max(axis=1) > 0
            mask = int(sum(1 << v for v in np.nonzero(changed)[0]))
            keep.append(side)
            _lib.check(self.lib.xtb_param_grad_basis(
                ct.byref(self.batch), ct.byref(self.dt.s), side[0][0].data_ptr(), side[0][1].data_ptr(),
                side[1][0].data_ptr(), side[1][1].data_ptr(), ct.c_uint32(mask), ct.c_double(h),
                self.positions.data_ptr(), self.cn.data_ptr(), ct.byref(self.state), g.data_ptr(),
                u, nus, s), "xtb_param_grad_basis")
        st.synchronize()  # the perturbed tables in `keep` are read by the launches
        return g

    def scf_flops(self) -> float:
        """Algorithmic FP64 flops of the SCF launch (fixed convention): per
        molecule (iterations + 1) x (9 1/3 n^3 eigen + 2 n^2 k density)
        + 5 1/3 n^3 orthogonaliser + 2 n^2 k energy-weighted density, with
        k = occupied orbitals (alpha count) and the measured iterations."""
        rec = self.records().cpu().numpy() if getattr(self, "h_out", None) is None else self.host_records()
        n = self.L.n.astype(np.float64)
        k = self.L.nel_a.astype(np.float64)
        it = rec[:, _lib.OUT_ITER] + 1.0
        return float(np.sum(it * (28.0 / 3.0 * n**3 + 2.0 * n * n * k) + 16.0 / 3.0 * n**3
                            + 2.0 * n * n * k))

    # ---- result views -------------------------------------------------------
    def records(self) -> torch.Tensor:
        return self.out.view(self.L.B, _lib.OUT_STRIDE)

    def energies(self) -> torch.Tensor:
        """Total energy per molecule: E_scf + (-TS) + E_rep."""
        r = self.records()
        return r[:, _lib.OUT_ESCF] + r[:, _lib.OUT_FENERGY] + r[:, _lib.OUT_EREP]

    def gradient(self) -> torch.Tensor:
        return self.grad.view(-1, 3)

    def orbital_charges(self, b: int) -> torch.Tensor:
        v0 = int(self.L.vec0[b])
        return self.q[v0: v0 + int(self.L.nao[b])]

    def matrix(self, buf: torch.Tensor, b: int) -> torch.Tensor:
        n = int(self.L.n[b])
        m0 = int(self.L.mat0[b])
        return buf[m0: m0 + n * n].view(n, n)


# the per-molecule SCF kernel's size limit (orbitals of one molecule)
WG_MAX_N = 1024


class SplitPipeline:
    """A batch holding molecules above the per-molecule kernel's limit
    (WG_MAX_N orbitals): each of those runs its SCF on the whole GPU
    (xtb_scf_large), one after the other, and the remaining molecules run
    together as a batch of their own (xtb_scf), on the same stream.  The
    accessors return the results in the order and shapes of the original
    layout.  Upstream runs any batch through one padded SCF
    (src/dxtb/_src/scf/unrolling/base.py:122-139); in a padded batch the
    molecules here are padded to the largest of their own part, which drops
    only the Fermi occupations of the zero-energy padding levels that
    upstream's larger padding adds, exp(eF / kT) of a level each."""

    large = False

    def __init__(self, layout: BatchLayout, dtables: DeviceTables, device):
        self.L = L = layout
        self.device = torch.device(device)
        big = np.nonzero(L.nao > WG_MAX_N)[0]
        small = np.nonzero(L.nao <= WG_MAX_N)[0]
        groups = ([(small, False)] if len(small) else []) + [(np.array([b]), True) for b in big]
        spin = L.spin
        self.parts = []
        for idx, large in groups:
            sp = None if spin is None else np.broadcast_to(np.asarray(spin, dtype=np.float64), (L.B,))[idx]
            sub = BatchLayout([L.numbers[b] for b in idx], L.tables, padded=L.padded and not large,
                              chrg=np.asarray(L.chrg, dtype=np.float64)[idx], spin=sp)
            atoms = np.concatenate([np.arange(L.atom0[b], L.atom0[b + 1]) for b in idx])
            self.parts.append((torch.as_tensor(idx, device=self.device),
                               torch.as_tensor(atoms, device=self.device),
                               Pipeline(sub, dtables, device, large=large)))
        # molecule -> (part, index in the part)
        self._where = {}
        for k, (idx, _, _) in enumerate(self.parts):
            for j, b in enumerate(idx.tolist()):
                self._where[b] = (k, j)
        self.out = torch.zeros((L.B, _lib.OUT_STRIDE), dtype=torch.float64, device=self.device)
        self.grad = torch.zeros((int(L.nat.sum()), 3), dtype=torch.float64, device=self.device)

    def run(self, positions: torch.Tensor, opts: ScfOptions, forces: bool, stream=None, fused: bool = False,
            **kw) -> None:
        if positions.shape != (int(self.L.nat.sum()), 3):
            raise ValueError("positions do not match the batch layout")
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        with torch.cuda.stream(st):
            for idx, atoms, p in self.parts:
                p.run(positions.index_select(0, atoms).contiguous(), opts, forces=forces, stream=st, fused=fused)
                self.out.index_copy_(0, idx, p.records())
                if forces:
                    self.grad.index_copy_(0, atoms, p.gradient())

    def records(self) -> torch.Tensor:
        return self.out

    def energies(self) -> torch.Tensor:
        r = self.out
        return r[:, _lib.OUT_ESCF] + r[:, _lib.OUT_FENERGY] + r[:, _lib.OUT_EREP]

    def gradient(self) -> torch.Tensor:
        return self.grad

    def repulsion_gradient(self, stream=None) -> torch.Tensor:
        g = torch.zeros_like(self.grad)
        for _, atoms, p in self.parts:
            g.index_copy_(0, atoms, p.repulsion_gradient(stream))
        return g

    def orbital_charges(self, b: int) -> torch.Tensor:
        k, j = self._where[b]
        return self.parts[k][2].orbital_charges(j)

    @property
    def emo(self) -> torch.Tensor:
        """Orbital energies in the original layout's vectors (L.vec0, L.n):
        in a padded batch a molecule of a part padded less than the batch
        gets the missing padding levels' energy, zero, sorted in."""
        L = self.L
        out = torch.zeros(L.nvec_total, dtype=torch.float64, device=self.device)
        for b in range(L.B):
            k, j = self._where[b]
            p = self.parts[k][2]
            v0, n = int(p.L.vec0[j]), int(p.L.n[j])
            e = p.emo[v0:v0 + n]
            m = int(L.n[b])
            if m > n:
                e = torch.sort(torch.cat([e, e.new_zeros(m - n)])).values
            out[int(L.vec0[b]):int(L.vec0[b]) + m] = e
        return out

    def param_grad(self, stream=None) -> torch.Tensor:
        rows = [(idx, p.param_grad(stream)) for idx, _, p in self.parts]
        g = torch.zeros((self.L.B, rows[0][1].shape[1]), dtype=torch.float64, device=self.device)
        for idx, r in rows:
            g.index_copy_(0, idx, r)
        return g

    def param_grad_basis(self, tables, rel_step: float = 1e-4, stream=None) -> torch.Tensor:
        rows = [(idx, p.param_grad_basis(tables, rel_step, stream)) for idx, _, p in self.parts]
        g = torch.zeros((self.L.B, rows[0][1].shape[1]), dtype=torch.float64, device=self.device)
        for idx, r in rows:
            g.index_copy_(0, idx, r)
        return g
