"""ctypes binding of the HIP library ``libxtbhip.so`` (C ABI: include/xtb_hip.h).

The structs below mirror the C structs field by field.  Loading fails
loudly: there is no CPU fallback for any entry point.
"""

from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path

from .exceptions import NativeLibraryError

LIB_NAME = "libxtbhip.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

XTB_OK = 0
XTB_EINVAL, XTB_EHIP, XTB_ETOOBIG = 1, 2, 3
# atoms a batch may hold in xtb_forces_noalloc (XTB_NOALLOC_MAX_ATOMS)
NOALLOC_MAX_ATOMS = 262144
STATUS_NAMES = {1: "invalid argument", 2: "HIP runtime error", 3: "size exceeds kernel limit"}

MOL_MAXITER = 1
MOL_FERMI = 2
MOL_NONFINITE = 4
MOL_JACOBI = 8
MOL_OVERLAP = 16

OUT_STRIDE = 32
OUT_ESCF, OUT_EH0, OUT_EES2, OUT_EES3, OUT_FENERGY = 0, 1, 2, 3, 4
OUT_ITER, OUT_STATUS, OUT_EF_A, OUT_EF_B, OUT_EREP, OUT_SWEEPS = 5, 6, 7, 8, 9, 10
OUT_T_ORTHO, OUT_T_EIGEN, OUT_T_TOTAL = 11, 13, 14  # stage and SCF times of the molecule, microseconds
OUT_PD_STEPS, OUT_SWEEPS_FIRST, OUT_SWEEPS_LAST, OUT_T_PD = 17, 18, 19, 21
OUT_T_START = 20  # device wall clock at the molecule's start, microseconds
OUT_XCD = 22  # XCD the molecule's workgroup ran on
OUT_T_FOCK, OUT_N_FOCK, OUT_T_MULL, OUT_N_MULL = 24, 25, 26, 27  # stage timers (us) and counts
OUT_T_INTEGRALS, OUT_T_FORCES = 28, 29  # xtb_singlepoint: the molecule's integral pass and gradient (us)

P = ct.c_void_p
i32 = ct.c_int32
i64 = ct.c_int64
f64 = ct.c_double


class XtbBatch(ct.Structure):
    _fields_ = [
        ("nmol", i32), ("nat_total", i32), ("nsh_total", i32), ("nvec_total", i32),
        ("max_n", i32), ("max_nsh", i32), ("max_nat", i32), ("ntile_sh", i32),
        ("ntile_at", i32), ("has_padding", i32), ("nmat_total", i64), ("ngam_total", i64),
        ("mol_nat", P), ("mol_nsh", P), ("mol_nao", P), ("mol_n", P), ("mol_atom0", P),
        ("mol_sh0", P), ("mol_vec0", P), ("mol_mat0", P), ("mol_gam0", P), ("mol_nel", P),
        ("atom_sp", P), ("atom_mol", P), ("atom_sh0", P), ("atom_nsh", P),
        ("sh_atom", P), ("sh_us", P), ("sh_ao", P),
        ("ao_sh", P), ("n0", P),
        ("tile_sh", P), ("tile_at", P),
    ]


class XtbParams(ct.Structure):
    _fields_ = [
        ("nsp", i32), ("nus", i32),
        ("sp_rad", P), ("sp_rcov", P), ("sp_gam3", P), ("sp_zeff", P), ("sp_arep", P),
        ("us_level", P), ("us_kcn", P), ("us_shpoly", P), ("us_hard", P),
        ("us_l", P), ("us_nprim", P), ("us_alpha", P), ("us_coeff", P), ("us_K", P),
        ("kexp", f64), ("rep_cutoff", f64), ("cn_cutoff", f64), ("cn_steep", f64),
        ("int_cutoff", f64),
    ]


class XtbScfOpts(ct.Structure):
    _fields_ = [
        ("maxiter", i32), ("mix_guess", i32), ("generations", i32), ("fermi_maxiter", i32),
        ("want_w", i32), ("jacobi_max_sweeps", i32), ("padded", i32), ("jacobi_inner", i32),
        ("x_atol", f64), ("x_atol_max", f64), ("kt", f64), ("damp", f64), ("damp_init", f64),
        ("diag_offset", f64), ("fermi_thresh", f64), ("jacobi_tol", f64), ("eig_inexact", f64),
        ("bd_gap_kt", f64), ("pd_kmax", f64), ("first_tol", f64), ("ortho", i32), ("orbitals", i32),
        ("keep_state", i32), ("warm_start", i32), ("soft_start", i32),
        ("wg_skip", i32), ("mixer", i32), ("f_atol", f64),
    ]


class XtbScfState(ct.Structure):
    _fields_ = [
        ("F", P), ("C", P), ("T", P), ("P", P),
        ("pot", P), ("vnew", P), ("q", P), ("eorb", P), ("emo", P), ("occ", P),
        ("hist_x", P), ("hist_f", P), ("qsh", P), ("out", P), ("order", P),
    ]


_lib = None


def load(path: str | os.PathLike | None = None):
    """Load the library once; raise NativeLibraryError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise NativeLibraryError(
            f"{p} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m dxtb_amd.build`."
        )
    lib = ct.CDLL(str(p))
    bp, pp = ct.POINTER(XtbBatch), ct.POINTER(XtbParams)
    lib.xtb_abi_version.restype = i32
    lib.xtb_abi_version.argtypes = []
    lib.xtb_abi_struct_sizes.restype = i32
    lib.xtb_abi_struct_sizes.argtypes = [ct.POINTER(i64)]
    lib.xtb_workspace_doubles.restype = i32
    lib.xtb_workspace_doubles.argtypes = [bp, i32, ct.POINTER(i64)]
    lib.xtb_build_s_h0.restype = i32
    lib.xtb_build_s_h0.argtypes = [bp, pp, P, P, P, P, P]
    lib.xtb_build_gamma.restype = i32
    lib.xtb_build_gamma.argtypes = [bp, pp, P, P, P]
    lib.xtb_build_s_h0_gamma.restype = i32
    lib.xtb_build_s_h0_gamma.argtypes = [bp, pp, P, P, P, P, P, P]
    lib.xtb_scf.restype = i32
    lib.xtb_scf.argtypes = [bp, pp, P, P, P, ct.POINTER(XtbScfOpts), ct.POINTER(XtbScfState), P]
    lib.xtb_scf_lds_bytes.restype = i32
    lib.xtb_scf_lds_bytes.argtypes = [bp, ct.POINTER(i32), ct.POINTER(i32)]
    lib.xtb_repulsion.restype = i32
    lib.xtb_repulsion.argtypes = [bp, pp, P, P, P]
    lib.xtb_build_integrals.restype = i32
    lib.xtb_build_integrals.argtypes = [bp, pp, P, P, P, P, P, P, P]
    lib.xtb_repulsion_gradient.restype = i32
    lib.xtb_repulsion_gradient.argtypes = [bp, pp, P, P, P]
    lib.xtb_forces.restype = i32
    lib.xtb_forces.argtypes = [bp, pp, P, P, P, ct.POINTER(XtbScfState), P, P, P]
    lib.xtb_forces_noalloc.restype = i32
    lib.xtb_forces_noalloc.argtypes = [bp, pp, P, P, P, ct.POINTER(XtbScfState), P, P]
    lib.xtb_param_grad.restype = i32
    lib.xtb_param_grad.argtypes = [bp, pp, P, P, P, ct.POINTER(XtbScfState), P, P]
    lib.xtb_param_grad_basis.restype = i32
    lib.xtb_param_grad_basis.argtypes = [bp, pp, P, P, P, P, ct.c_uint32, ct.c_double, P, P,
                                         ct.POINTER(XtbScfState), P, i32, i32, P]
    lib.xtb_scf_large_workspace.restype = i32
    lib.xtb_scf_large_workspace.argtypes = [bp, ct.POINTER(i64)]
    lib.xtb_scf_large.restype = i32
    lib.xtb_scf_large.argtypes = [bp, pp, P, P, P, ct.POINTER(XtbScfOpts), ct.POINTER(XtbScfState), P, i64, P]
    lib.xtb_singlepoint.restype = i32
    lib.xtb_singlepoint.argtypes = [bp, pp, P, P, P, P, P, ct.POINTER(XtbScfOpts), ct.POINTER(XtbScfState),
                                    P, P, P]
    if lib.xtb_abi_version() != 1:
        raise NativeLibraryError("libxtbhip.so ABI version mismatch")
    sizes = (i64 * 4)()
    lib.xtb_abi_struct_sizes(sizes)
    mine = [ct.sizeof(s) for s in (XtbBatch, XtbParams, XtbScfOpts, XtbScfState)]
    if list(sizes) != mine:
        raise NativeLibraryError(f"struct layout mismatch: C {list(sizes)} vs ctypes {mine}")
    _lib = lib
    return lib


EXPORTED = (
    "xtb_abi_version", "xtb_abi_struct_sizes", "xtb_workspace_doubles", "xtb_build_s_h0", "xtb_build_gamma",
    "xtb_scf", "xtb_scf_lds_bytes", "xtb_repulsion", "xtb_forces", "xtb_param_grad",
    "xtb_param_grad_basis", "xtb_build_s_h0_gamma", "xtb_repulsion_gradient",
    "xtb_forces_noalloc", "xtb_build_integrals", "xtb_scf_large_workspace", "xtb_scf_large",
    "xtb_singlepoint",
)


def check(rc: int, what: str) -> None:
    if rc != XTB_OK:
        raise NativeLibraryError(f"{what} failed: {STATUS_NAMES.get(rc, rc)}")
