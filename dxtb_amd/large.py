"""Whole-GPU path for one large system (thousands of orbitals).

The per-molecule SCF kernel gives one compute unit to each molecule, which is
right for streams of molecules but not for a single system of n ~ 5000
orbitals.  Here the whole GPU works on one molecule, and every O(n^3) step is
a dense FP64 GEMM on the matrix cores (rocBLAS through torch.matmul):

* Loewdin orthogonaliser X = S^-1/2 by the coupled Newton-Schulz iteration;
* density by trace-correcting purification (TC2) of the orthogonalised Fock
  matrix F' = X F X, per spin channel: D <- D^2 or 2D - D^2, ~30 products;
  P = X (D_alpha + D_beta) X.  For a gapped system the Fermi-Dirac
  occupations at the electronic temperature are exactly 0 or 1 in FP64, so
  this is the same density upstream builds from its eigenvectors
  (src/dxtb/_src/scf/base.py:861-907).  A Lanczos estimate of the
  HOMO-LUMO gap verifies that: with the Fermi level mid-gap, a gap of
  60 kT leaves occupation tails below 1/(1 + e^30) ~ 1e-13; below that the
  system is refused;
* energy-weighted density W = P F P / 2 per closed shell, for the analytic
  gradient kernel.

Integrals (S, H0, CN, gamma, repulsion) and forces use the same HIP kernels
as the batched path; the SCF map, mixing (upstream's Anderson) and
convergence test follow src/dxtb/_src/scf/unrolling/default.py:71-140.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from .constants import KELVIN2AU
from .engine import BatchLayout, DeviceTables, Pipeline, ScfOptions
from .param import SpeciesTables

# smallest HOMO-LUMO gap (in kT) for which 0/1 occupations equal the
# Fermi-Dirac ones to ~1e-13
GAP_MIN_KT = 60.0


def newton_schulz_inv_sqrt(S: torch.Tensor, tol: float = 1e-14, maxiter: int = 60) -> torch.Tensor:
    n = S.shape[0]
    c = float(S.abs().sum(1).max()) * 1.0001
    eye = torch.eye(n, dtype=S.dtype, device=S.device)
    Y = S / c
    Z = eye.clone()
    extra = -1
    for _ in range(maxiter):
        T = 1.5 * eye - 0.5 * (Z @ Y)
        err = float((T - eye).abs().max())
        if extra == 0 or err < tol:
            break
        if extra > 0:
            extra -= 1
        elif err < 1e-9:
            extra = 1
        Y, Z = Y @ T, T @ Z
    X = Z / math.sqrt(c)
    return 0.5 * (X + X.T)


def tc2(Fp: torch.Tensor, nocc: int, maxiter: int = 120) -> torch.Tensor:
    """Trace-correcting purification: projector onto the nocc lowest
    eigenvectors of the symmetric matrix Fp."""
    n = Fp.shape[0]
    if nocc <= 0:
        return torch.zeros_like(Fp)
    if nocc >= n:
        return torch.eye(n, dtype=Fp.dtype, device=Fp.device)
    d = torch.diagonal(Fp)
    r = Fp.abs().sum(1) - d.abs()
    emin, emax = float((d - r).min()), float((d + r).max())
    D = (emax * torch.eye(n, dtype=Fp.dtype, device=Fp.device) - Fp) / (emax - emin)
    last = None
    for _ in range(maxiter):
        D2 = D @ D
        tr, tr2 = float(torch.trace(D)), float(torch.trace(D2))
        idem = abs(tr - tr2)
        if tr > nocc:
            D = D2
        else:
            D = 2.0 * D - D2
        # quadratic convergence: stop when idempotency stagnates at rounding level
        if idem < 1e-9 * n and last is not None and idem >= 0.5 * last:
            break
        last = idem
    return 0.5 * (D + D.T)


def lanczos_extreme(A: torch.Tensor, proj: torch.Tensor, largest: bool, steps: int = 60) -> float:
    """Extreme eigenvalue of A on range(proj) (proj an orthogonal projector
    commuting with A): Lanczos with full reorthogonalisation on
    B = proj A proj + c (I - proj), where c puts the complement below (largest)
    or above (smallest) the whole spectrum of A (Gershgorin bounds), so the
    wanted value is the extreme eigenvalue of B on the full space.  Ritz values
    bound it from inside."""
    n = A.shape[0]
    d = torch.diagonal(A)
    r = A.abs().sum(1) - d.abs()
    lo, hi = float((d - r).min()), float((d + r).max())
    c = lo - 1.0 if largest else hi + 1.0
    tiny = 1e-10 * (hi - lo + 1.0)

    def op(x):
        px = proj @ x
        return proj @ (A @ px) + c * (x - px)

    g = torch.Generator(device="cpu").manual_seed(1234)
    v = torch.randn(n, generator=g, dtype=A.dtype).to(A.device)
    v = v / v.norm()
    V = [v]
    alphas, betas = [], []
    for k in range(min(steps, n)):
        w = op(V[-1])
        a = float(w @ V[-1])
        alphas.append(a)
        # two Gram-Schmidt passes ("twice is enough") keep V orthonormal when
        # the Krylov space runs out (B has few distinct eigenvalues)
        for _ in range(2):
            for u in V:
                w = w - (w @ u) * u
        b = float(w.norm())
        if b < tiny or k == min(steps, n) - 1:
            break
        betas.append(b)
        V.append(w / b)
    m = len(alphas)
    T = np.diag(alphas) + np.diag(betas[: m - 1], 1) + np.diag(betas[: m - 1], -1)
    ev = np.linalg.eigvalsh(T)
    return float(ev[-1] if largest else ev[0])


class LargeSystem:
    """Energy and forces of one large molecule on the whole GPU."""

    def __init__(self, numbers, device="cuda:0", opts: ScfOptions | None = None, chrg: float = 0.0):
        self.device = torch.device(device)
        self.opts = opts or ScfOptions()
        z = np.asarray(numbers, dtype=np.int64)
        self.numbers = z[z > 0]
        self.tables = SpeciesTables(np.unique(self.numbers))
        self.layout = BatchLayout([self.numbers], self.tables, padded=False, chrg=np.array([chrg]))
        self.dt = DeviceTables(self.tables, self.device)
        self.pipe = Pipeline(self.layout, self.dt, self.device)

    def run(self, positions, forces: bool = True) -> dict:
        import ctypes as ct

        L, pipe, dev = self.layout, self.pipe, self.device
        if isinstance(positions, torch.Tensor):
            pos = positions.detach().to(device=dev, dtype=torch.float64).contiguous()
        else:
            pos = torch.as_tensor(np.asarray(positions), dtype=torch.float64, device=dev).contiguous()
        pipe.positions = pos
        s = ct.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        lib, b, p = pipe.lib, ct.byref(pipe.batch), ct.byref(self.dt.s)
        _lib.check(lib.xtb_build_s_h0(b, p, pos.data_ptr(), pipe.S.data_ptr(), pipe.H0.data_ptr(),
                                      pipe.cn.data_ptr(), s), "xtb_build_s_h0")
        _lib.check(lib.xtb_build_gamma(b, p, pos.data_ptr(), pipe.gamma.data_ptr(), s), "xtb_build_gamma")
        _lib.check(lib.xtb_repulsion(b, p, pos.data_ptr(), pipe.out.data_ptr(), s), "xtb_repulsion")
        n = int(L.n[0])
        nsh, nat = int(L.nsh[0]), int(L.nat[0])
        S = pipe.matrix(pipe.S, 0)
        H0 = pipe.matrix(pipe.H0, 0)
        G = pipe.gamma[: nsh * nsh].view(nsh, nsh)
        o = self.opts
        kt = o.fermi_etemp * KELVIN2AU
        na, nb = int(L.nel_a[0]), int(L.nel_b[0])
        n0 = torch.as_tensor(L.arrays["n0"], device=dev)
        ao_sh = torch.as_tensor(L.arrays["ao_sh"].astype(np.int64), device=dev)
        sh_at = torch.as_tensor(L.arrays["sh_atom"].astype(np.int64), device=dev)
        hd = self.dt.t["sp_gam3"][torch.as_tensor(L.arrays["atom_sp"].astype(np.int64), device=dev)]

        X = newton_schulz_inv_sqrt(S)

        def density(v):
            F = H0 - 0.5 * S * (v[:, None] + v[None, :])
            Fp = X @ F @ X
            Da = tc2(Fp, na)
            Db = Da if nb == na else tc2(Fp, nb)
            return F, Fp, Da, Db, X @ (Da + Db) @ X

        def charges(P):
            q = n0 - (P * S).sum(1)
            eorb = (P * H0).sum(1)
            qsh = torch.zeros(nsh, dtype=q.dtype, device=dev).index_add_(0, ao_sh, q)
            qat = torch.zeros(nat, dtype=q.dtype, device=dev).index_add_(0, sh_at, qsh)
            gq = G @ qsh
            vsh = gq + (hd * qat * qat)[sh_at]
            return q, eorb, qsh, qat, gq, vsh[ao_sh]

        # Anderson mixing as upstream (scf/mixer/anderson.py:163-317)
        gen = o.generations
        xh, fh = [], []
        step = 0

        def mix(xnew, xold):
            nonlocal step, xh, fh
            if step == 0:
                xh = [xold.clone()]
            step += 1
            f0 = xnew - xh[0]
            fh = [f0] + fh
            if step > gen:
                nn = min(step - 1, gen)
                df = torch.stack([f0 - fh[h] for h in range(1, nn + 1)])
                a = df @ df.T
                a = a * (torch.eye(nn, dtype=a.dtype, device=dev) * (o.diag_offset ** 2)
                         + torch.ones_like(a))
                th = torch.linalg.solve(a.cpu(), (df @ f0).cpu()).to(dev)
                xs = torch.stack(xh[1:nn + 1])
                xbar = xh[0] + th @ (xs - xh[0])
                fbar = f0 - th @ df
                xm = xbar + o.damp * fbar
            else:
                xm = xh[0] + o.damp_init * f0
            xh = [xm] + xh[:gen]
            fh = fh[:gen + 1]
            conv = float(f0.norm()) < o.x_atol and float(f0.abs().max()) < o.x_atol_max
            return xm, conv

        v = torch.zeros(n, dtype=torch.float64, device=dev)
        iters = 0
        _, _, _, _, P = density(v)
        vnew = charges(P)[5]
        iters += 1
        v, _ = mix(vnew, v) if o.mix_guess else (vnew, False)
        converged = False
        for _ in range(o.maxiter):
            _, _, _, _, P = density(v)
            vnew = charges(P)[5]
            iters += 1
            v, c = mix(vnew, v)
            if c:
                converged = True
                break
        F, Fp, Da, Db, P = density(vnew)
        q, eorb, qsh, qat, gq, vfin = charges(P)
        # gap check: occupations are exactly 0/1 in FP64 only if the gap >> kT
        eye = torch.eye(n, dtype=torch.float64, device=dev)
        homo = lanczos_extreme(Fp, Da, largest=True)
        lumo = lanczos_extreme(Fp, eye - Da, largest=False)
        gap = lumo - homo
        if gap < GAP_MIN_KT * kt:
            raise RuntimeError(f"HOMO-LUMO gap {gap:.3e} Eh too small for the purification path")
        e_h0 = float(eorb.sum())
        e_es2 = float(0.5 * (qsh * gq).sum())
        e_es3 = float((hd * qat ** 3 / 3.0).sum())
        torch.cuda.synchronize(dev)
        erep = float(pipe.out[_lib.OUT_EREP])
        res = {"energy": e_h0 + e_es2 + e_es3 + erep, "scf": e_h0 + e_es2 + e_es3,
               "repulsion": erep, "iterations": iters, "converged": converged,
               "charges_orbital": q, "charges_atom": qat, "gap": gap}
        if forces:
            W = 0.5 * (P @ F @ P) if na == nb else None
            if W is None:
                Wa = X @ Da @ Fp @ Da @ X
                Wb = X @ Db @ Fp @ Db @ X
                W = Wa + Wb
            pipe.matrix(pipe.P, 0).copy_(P)
            pipe.matrix(pipe.T, 0).copy_(W)
            pipe.pot[:n].copy_(vfin)
            pipe.qsh[:nsh].copy_(qsh)
            pipe.hist_x = pipe.hist_f = torch.empty(1, dtype=torch.float64, device=dev)
            pipe.state = pipe._state()
            pipe.run_forces()
            res["gradient"] = pipe.gradient().clone()
        return res
