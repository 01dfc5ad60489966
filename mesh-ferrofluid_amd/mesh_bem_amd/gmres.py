"""Restarted GMRES with FP32 matrix and Krylov basis on the HIP building blocks.

* FP32: the matrix (complex64 or float32), the Arnoldi basis and every inner
  matrix-vector product -- the HBM traffic of the method.
* FP64: the iterate x, the restart residual r = b - A x (FP64 accumulation
  over the FP32 matrix), and the Hessenberg least-squares problem (Givens
  rotations on the host, from one device-to-host copy per iteration).
Orthogonalisation is classical Gram-Schmidt applied twice (CGS2) on device.

Convergence is declared on the true relative residual ||b - A x|| / ||b|| <=
tol, recomputed at every restart; the Arnoldi estimate ends an inner cycle
early.  Three restarts without a 10% drop of the true residual end the
iteration unconverged.

Fallbacks (recorded in ``GmresInfo.fallback``): a warm start x0 holding NaN or
Inf is replaced by zero ("x0-nonfinite"); with ``fallback="lu"`` an
unconverged solve of a dense matrix finishes with a complex64 LU of the
matrix and FP64-residual refinement ("lu").
"""

from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np
import torch

from . import _lib

F32, F64 = 0, 1


@dataclass
class GmresInfo:
    iterations: int
    restarts: int
    converged: bool
    residual: float  # final true relative residual
    history: List[float] = field(default_factory=list)  # Arnoldi estimates per iteration
    fallback: str = ""  # "", "x0-nonfinite", "lu" or "x0-nonfinite+lu"


def _vprec(dt: torch.dtype) -> int:
    return F64 if dt in (torch.complex128, torch.float64) else F32


class DenseOperator:
    """y = A x (or b - A x) for a resident row-major FP32 matrix; x, b, y FP32 or FP64."""

    def __init__(self, A: torch.Tensor):
        if A.dim() != 2 or A.shape[0] != A.shape[1] or A.stride(1) != 1:
            raise ValueError("DenseOperator needs a square row-major matrix")
        if A.dtype not in (torch.complex64, torch.float32):
            raise TypeError("DenseOperator takes a complex64 or float32 matrix")
        self.A = A
        self.n = A.shape[0]
        self.field = _lib.MBA_COMPLEX if A.dtype == torch.complex64 else _lib.MBA_REAL
        self.dtype = A.dtype

    def apply(self, x: torch.Tensor, y: torch.Tensor, b: Optional[torch.Tensor] = None):
        lib = _lib.load()
        vp = _vprec(x.dtype)
        if _vprec(y.dtype) != vp or (b is not None and _vprec(b.dtype) != vp):
            raise TypeError("x, y and b must share one precision")
        st = lib.mba_mv32(self.field, _lib.ptr(self.A), self.n, self.n, self.A.stride(0), vp, _lib.ptr(x),
                          _lib.ptr(b) if b is not None else None, _lib.ptr(y), _lib.stream())
        _lib.check(st, "mba_mv32")


def matvec(A: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """y = A x with an FP32 matrix; y has the dtype of x (FP32 or FP64 vector)."""
    y = torch.empty(A.shape[0], dtype=x.dtype, device=A.device)
    DenseOperator(A).apply(x.contiguous(), y)
    return y


class _Blocks:
    def __init__(self, n, m, fld, dev):
        self.lib = _lib.load()
        self.field = fld
        self.n = n
        self.wsb = self.lib.mba_gmres_workspace_bytes(n, m + 1)
        self.ws = torch.empty(self.wsb, dtype=torch.uint8, device=dev)

    def dots(self, basis, nvec, w, out):
        st = self.lib.mba_gmres_dots(self.field, _lib.ptr(basis), basis.stride(0), nvec, _vprec(w.dtype),
                                     _lib.ptr(w), self.n, _lib.ptr(out), _lib.ptr(self.ws), self.wsb,
                                     _lib.stream())
        _lib.check(st, "mba_gmres_dots")

    def update(self, basis, nvec, coef, sign, w):
        st = self.lib.mba_gmres_update(self.field, _lib.ptr(basis), basis.stride(0), nvec, _lib.ptr(coef),
                                       float(sign), _vprec(w.dtype), _lib.ptr(w), self.n, _lib.stream())
        _lib.check(st, "mba_gmres_update")

    def scale(self, src, nrm2, dst):
        st = self.lib.mba_gmres_scale(self.field, _vprec(src.dtype), _lib.ptr(src), self.n, _lib.ptr(nrm2),
                                      _lib.ptr(dst), _lib.stream())
        _lib.check(st, "mba_gmres_scale")


def _givens(a, c):
    """(cs real, sn complex) with [cs sn; -conj(sn) cs] [a; c] = [rho; 0]."""
    if abs(a) == 0.0:
        return 0.0, 1.0 + 0j
    r = np.hypot(abs(a), abs(c))
    return abs(a) / r, (a / abs(a)) * np.conj(c) / r


def _lu_finish(op: DenseOperator, b: torch.Tensor, x: torch.Tensor, tol: float, maxit: int = 10):
    """x refined with a complex64 LU of the operator's matrix: r = b - A x in
    FP64 (FP32 matrix, FP64 accumulation), x += LU \\ r, until ||r|| <= tol ||b||."""
    from .dense import LUFactor

    A = op.A
    lu = LUFactor(A.to(torch.complex64) if A.dtype == torch.float32 else A.clone(), overwrite=True)
    r = torch.empty_like(x)
    bn = float(torch.linalg.vector_norm(b))
    res = float("inf")
    for _ in range(maxit + 1):
        op.apply(x, r, b)
        res = float(torch.linalg.vector_norm(r)) / bn
        if res <= tol:
            break
        d = lu.solve(r.to(torch.complex64))
        x += d.real.to(x.dtype) if not x.is_complex() else d.to(x.dtype)
    return x, res


def gmres(A, b: torch.Tensor, x0: Optional[torch.Tensor] = None, restart: int = 50, tol: float = 1e-6,
          maxiter: int = 2000, callback: Optional[Callable] = None, fallback: str = "none"):
    """Solve A x = b.  A: square FP32 tensor (complex64 / float32) or an operator
    with ``apply(x, y, b=None)``, ``n`` and ``dtype``.  Returns (x FP64, GmresInfo).

    fallback="lu" (dense A only): an unconverged iteration is finished by a
    complex64 LU with FP64-residual refinement."""
    if fallback not in ("none", "lu"):
        raise ValueError("fallback: 'none' or 'lu'")
    op = A if hasattr(A, "apply") else DenseOperator(A)
    if fallback == "lu" and not isinstance(op, DenseOperator):
        raise ValueError("fallback='lu' needs a dense matrix")
    n = op.n
    dev = b.device
    if restart < 1 or restart > 500:
        raise ValueError("restart must be in [1, 500]")
    cplx = op.dtype == torch.complex64
    dt32 = op.dtype
    dt64 = torch.complex128 if cplx else torch.float64
    fld = _lib.MBA_COMPLEX if cplx else _lib.MBA_REAL
    blk = _Blocks(n, restart, fld, dev)
    b = b.reshape(-1).to(dt64).contiguous()
    x = torch.zeros(n, dtype=dt64, device=dev) if x0 is None else x0.reshape(-1).to(dt64).clone().contiguous()
    fb = []
    if x0 is not None and not bool(torch.isfinite(torch.view_as_real(x) if cplx else x).all()):
        x.zero_()
        fb.append("x0-nonfinite")
    V = torch.empty((restart + 1, n), dtype=dt32, device=dev)
    w = torch.empty(n, dtype=dt32, device=dev)
    r = torch.empty(n, dtype=dt64, device=dev)
    # device scalars: rows [0, m+1) first pass, [m+2, 2m+3) second pass, last row ||w||^2
    hbuf = torch.empty((2 * (restart + 2) + 1, 2), dtype=torch.float64, device=dev)
    yb = torch.empty((restart, 2), dtype=torch.float64, device=dev)
    nb = torch.empty((1, 2), dtype=torch.float64, device=dev)
    blk.dots(V, 0, b, nb)
    bnorm = float(np.sqrt(nb[0, 0].item()))
    if bnorm == 0.0:
        x.zero_()
        return x, GmresInfo(0, 0, True, 0.0, fallback="+".join(fb))
    it = 0
    restarts = 0
    hist = []
    best, stalled = np.inf, 0
    o2 = restart + 2
    while True:
        op.apply(x, r, b)  # r = b - A x  (FP64)
        blk.dots(V, 0, r, nb)
        beta = float(np.sqrt(nb[0, 0].item()))
        res = beta / bnorm
        if res <= tol or it >= maxiter:
            break
        if res > 0.9 * best:
            stalled += 1
            if stalled >= 3:
                break
        else:
            stalled = 0
        best = min(best, res)
        blk.scale(r, nb, V[0])
        H = np.zeros((restart + 1, restart), dtype=np.complex128)
        cs = np.zeros(restart)
        sn = np.zeros(restart, dtype=np.complex128)
        g = np.zeros(restart + 1, dtype=np.complex128)
        g[0] = beta
        k = 0
        for j in range(restart):
            op.apply(V[j], w)
            h1, h2, nr = hbuf[: j + 2], hbuf[o2: o2 + j + 2], hbuf[-1:]
            blk.dots(V, j + 1, w, h1)
            blk.update(V, j + 1, h1, -1.0, w)
            blk.dots(V, j + 1, w, h2)
            blk.update(V, j + 1, h2, -1.0, w)
            blk.dots(V, 0, w, nr)
            blk.scale(w, nr, V[j + 1])
            hh = hbuf.cpu().numpy()
            H[: j + 1, j] = (hh[: j + 1, 0] + hh[o2: o2 + j + 1, 0]) + 1j * (hh[: j + 1, 1] + hh[o2: o2 + j + 1, 1])
            hnext = np.sqrt(max(hh[-1, 0], 0.0))
            H[j + 1, j] = hnext
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -np.conj(sn[i]) * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            cs[j], sn[j] = _givens(H[j, j], H[j + 1, j])
            H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
            H[j + 1, j] = 0.0
            g[j + 1] = -np.conj(sn[j]) * g[j]
            g[j] = cs[j] * g[j]
            it += 1
            k = j + 1
            est = abs(g[j + 1]) / bnorm
            hist.append(float(est))
            if callback is not None:
                callback(it, est)
            if est <= tol or it >= maxiter or hnext == 0.0:
                break
        y = np.zeros(k, dtype=np.complex128)
        for i in range(k - 1, -1, -1):
            y[i] = (g[i] - H[i, i + 1:k] @ y[i + 1:k]) / H[i, i]
        if not cplx:
            y = y.real.astype(np.complex128)
        yb[:k] = torch.from_numpy(np.stack([y.real, y.imag], 1)).to(dev)
        blk.update(V, k, yb, 1.0, x)
        restarts += 1
    if res > tol and fallback == "lu":
        x, res = _lu_finish(op, b, x, tol)
        fb.append("lu")
    return x, GmresInfo(it, restarts, res <= tol, float(res), hist, "+".join(fb))
