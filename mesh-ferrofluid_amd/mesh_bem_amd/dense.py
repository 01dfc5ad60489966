"""Dense complex solves on the HIP library.

* ``LUFactor``: blocked row-pivoted LU (MFMA trailing update) of a complex64 or
  complex128 matrix, then triangular solves -- the getrf/getrs pair behind the
  reference's ``torch.linalg.solve`` (src/mesh_BEM/mesh_BEM_solver_3d.py:174).
* ``refine_solve``: mixed precision -- factor in FP32, then iterative
  refinement whose residual r = V - Z J is formed in FP64 against the
  complex128 matrix, until the correction falls below ``tol`` relative to J.
"""

import ctypes
from dataclasses import dataclass
from typing import Optional

import torch

from . import _lib


def gemm(A: torch.Tensor, B: torch.Tensor, C: Optional[torch.Tensor] = None, alpha: float = 1.0,
         beta: float = 0.0) -> torch.Tensor:
    """C = alpha A B + beta C for row-major complex tensors (MFMA kernel)."""
    lib = _lib.load()
    if A.dim() != 2 or B.dim() != 2:
        raise ValueError("gemm: 2-D operands")
    M, K = A.shape
    K2, N = B.shape
    if K != K2 or A.dtype != B.dtype or A.dtype not in (torch.complex64, torch.complex128):
        raise ValueError("gemm: complex A [M, K], B [K, N] of one dtype")
    if C is None:
        C = torch.zeros((M, N), dtype=A.dtype, device=A.device)
        beta = 0.0
    if C.dtype != A.dtype or tuple(C.shape) != (M, N):
        raise ValueError("gemm: C [M, N] of the operands' dtype")
    for t in (A, B, C):
        if t.stride(1) != 1:
            raise ValueError("gemm: row-major operands with unit column stride")
    st = lib.mba_gemm(_lib.prec_of(A.dtype), M, N, K, float(alpha), _lib.ptr(A), A.stride(0),
                      _lib.ptr(B), B.stride(0), float(beta), _lib.ptr(C), C.stride(0), _lib.stream())
    _lib.check(st, "mba_gemm")
    return C


def gemm_split_f16(A: torch.Tensor, B: torch.Tensor, C: Optional[torch.Tensor] = None, alpha: float = 1.0,
                   beta: float = 0.0) -> torch.Tensor:
    """C = alpha A B + beta C for row-major complex64 tensors (K <= 512) on the
    FP16 matrix cores with FP32 accuracy (power-of-two scaled hi/lo FP16
    pieces, three piece products, FP32 accumulation)."""
    lib = _lib.load()
    M, K = A.shape
    K2, N = B.shape
    if K != K2 or A.dtype != torch.complex64 or B.dtype != torch.complex64:
        raise ValueError("gemm_split_f16: complex64 A [M, K], B [K, N]")
    if C is None:
        C = torch.zeros((M, N), dtype=torch.complex64, device=A.device)
        beta = 0.0
    for t in (A, B, C):
        if t.stride(1) != 1:
            raise ValueError("gemm_split_f16: row-major operands with unit column stride")
    wsb = lib.mba_gemm_split_f16_workspace_bytes(M, N, K)
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=A.device)
    st = lib.mba_gemm_split_f16(M, N, K, float(alpha), _lib.ptr(A), A.stride(0), _lib.ptr(B), B.stride(0),
                                float(beta), _lib.ptr(C), C.stride(0), _lib.ptr(ws), wsb, _lib.stream())
    _lib.check(st, "mba_gemm_split_f16")
    return C


def set_split_tile(tile: int) -> None:
    """Workgroup tile of the split-FP16 product (process-wide): 0 = 256 x 128
    complex (8 waves, one workgroup per CU; default), 1 = 128 x 128 (4 waves,
    two workgroups per CU)."""
    _lib.check(_lib.load().mba_gemm_split_f16_set_tile(int(tile)), "mba_gemm_split_f16_set_tile")


LU_UPDATES = {"fp32": 0, "split_f16": 1}


def set_lu_update(kind: str) -> None:
    """Trailing-update product of complex64 factorisations (process-wide):
    "fp32" (FP32 MFMA) or "split_f16" (FP32-accurate split-FP16 MFMA)."""
    if kind not in LU_UPDATES:
        raise ValueError(f"set_lu_update: one of {sorted(LU_UPDATES)}")
    _lib.check(_lib.load().mba_lu_set_update(LU_UPDATES[kind]), "mba_lu_set_update")


LU_PANELS = {"partial": 0, "tournament": 1}


def set_lu_panel(kind: str) -> None:
    """Panels of complex64 factorisations (process-wide): "partial" (partial
    pivoting, LAPACK's row order, default) or "tournament" (tournament
    pivoting of 64-column panels with >= 2048 rows: a tree of partial-pivoting
    eliminations over rows nominated per workgroup; pair it with refinement)."""
    if kind not in LU_PANELS:
        raise ValueError(f"set_lu_panel: one of {sorted(LU_PANELS)}")
    _lib.check(_lib.load().mba_lu_set_panel(LU_PANELS[kind]), "mba_lu_set_panel")


LU_SOLVES = {"per_block": 0, "persistent": 1}


def set_lu_solve(kind: str) -> None:
    """Triangular solves of complex64 ``LUFactor.solve`` with <= 4 right-hand
    sides (process-wide): "persistent" (one launch per triangle, default) or
    "per_block" (one launch per 64-row block)."""
    if kind not in LU_SOLVES:
        raise ValueError(f"set_lu_solve: one of {sorted(LU_SOLVES)}")
    _lib.check(_lib.load().mba_lu_set_solve(LU_SOLVES[kind]), "mba_lu_set_solve")


class LUFactor:
    """In-place LU with partial pivoting of a square complex matrix."""

    def __init__(self, A: torch.Tensor, overwrite: bool = False, profile: bool = False, asynchronous: bool = False,
                 max_wg: int = 0):
        """asynchronous: enqueue on the current stream without any host
        synchronisation (call ``check()`` after synchronising); max_wg caps the
        workgroups of each panel so factorisations on several streams fit side
        by side."""
        lib = _lib.load()
        if A.dim() != 2 or A.shape[0] != A.shape[1]:
            raise ValueError("LUFactor needs a square matrix")
        if A.dtype not in (torch.complex64, torch.complex128):
            raise TypeError("LUFactor takes complex64 or complex128")
        self.lu = A if overwrite else A.clone()
        if self.lu.stride(1) != 1:
            self.lu = self.lu.contiguous()
        n = A.shape[0]
        self.n = n
        self.prec = _lib.prec_of(A.dtype)
        dev = A.device
        self.ipiv = torch.empty(n, dtype=torch.int32, device=dev)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        wsb = lib.mba_lu_workspace_bytes(self.prec, n)
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        self.profile = None
        if profile:
            import ctypes

            ms = (ctypes.c_double * 5)()
            nl = (ctypes.c_int64 * 5)()
            fl = (ctypes.c_double * 2)()
            st = lib.mba_lu_factor_profiled(self.prec, _lib.ptr(self.lu), n, self.lu.stride(0),
                                            _lib.ptr(self.ipiv), _lib.ptr(info), _lib.ptr(ws), wsb,
                                            ms, nl, fl, _lib.stream())
            _lib.check(st, "mba_lu_factor_profiled")
            phases = ("panel", "laswp", "trsm", "gemm", "split")
            self.profile = {
                "phase_ms": dict(zip(phases, list(ms))),
                "phase_launches": dict(zip(phases, list(nl))),
                "gemm_flops": fl[0],
                "split_flops": fl[1],
            }
        elif asynchronous:
            self._err = torch.zeros(1, dtype=torch.int32, device=dev)
            self._ws = ws
            st = lib.mba_lu_factor_async(self.prec, _lib.ptr(self.lu), n, self.lu.stride(0), _lib.ptr(self.ipiv),
                                         _lib.ptr(info), _lib.ptr(ws), wsb, int(max_wg), _lib.ptr(self._err),
                                         _lib.stream())
            _lib.check(st, "mba_lu_factor_async")
        else:
            st = lib.mba_lu_factor(self.prec, _lib.ptr(self.lu), n, self.lu.stride(0),
                                   _lib.ptr(self.ipiv), _lib.ptr(info), _lib.ptr(ws), wsb,
                                   _lib.stream())
            _lib.check(st, "mba_lu_factor")
        auxb = lib.mba_lu_solve_aux_bytes(self.prec, n)
        self.aux = torch.empty(auxb, dtype=torch.uint8, device=dev)
        st = lib.mba_lu_solve_prepare(self.prec, _lib.ptr(self.lu), n, self.lu.stride(0),
                                      _lib.ptr(self.ipiv), _lib.ptr(self.aux), auxb, _lib.stream())
        _lib.check(st, "mba_lu_solve_prepare")
        self.perm = self.aux[: 4 * n].view(torch.int32)
        self._info_t = info
        if asynchronous:
            self.info = None
            return
        self.info = int(info.item())
        if self.info != 0:
            raise _lib.MBAError(f"mba_lu_factor: exactly zero pivot in column {self.info - 1}")

    @classmethod
    def _from_factored(cls, lu: torch.Tensor, ipiv: torch.Tensor, info: torch.Tensor, err: torch.Tensor):
        """Wrap a factorisation done elsewhere (batched); the solve tables in
        ``aux`` are filled by the caller (mba_lu_solve_prepare_batched)."""
        lib = _lib.load()
        self = cls.__new__(cls)
        self.lu, self.ipiv, self.n = lu, ipiv, lu.shape[0]
        self.prec = _lib.prec_of(lu.dtype)
        self.profile = None
        auxb = lib.mba_lu_solve_aux_bytes(self.prec, self.n)
        self.aux = torch.empty(auxb, dtype=torch.uint8, device=lu.device)
        self.perm = self.aux[: 4 * self.n].view(torch.int32)
        self._info_t, self._err, self.info = info, err, None
        return self

    def check(self):
        """Status of an asynchronous factorisation (after the stream has been synchronised)."""
        if getattr(self, "_err", None) is not None and int(self._err.item()) != 0:
            raise _lib.MBAError("mba_lu_factor_async: a cooperative panel wait timed out")
        self.info = int(self._info_t.item())
        if self.info != 0:
            raise _lib.MBAError(f"mba_lu_factor: exactly zero pivot in column {self.info - 1}")

    def solve(self, B: torch.Tensor) -> torch.Tensor:
        """X = A^{-1} B for B [n] or [n, nrhs] (returns a new tensor of the factor dtype)."""
        lib = _lib.load()
        vec = B.dim() == 1
        X = B.reshape(self.n, -1).to(self.lu.dtype).clone().contiguous()
        nrhs = X.shape[1]
        wsb = lib.mba_lu_solve_workspace_bytes(self.prec, self.n, nrhs)
        ws = torch.empty(wsb, dtype=torch.uint8, device=X.device)
        st = lib.mba_lu_solve(self.prec, _lib.ptr(self.lu), self.n, self.lu.stride(0),
                              _lib.ptr(self.aux), _lib.ptr(X), nrhs, X.stride(0), _lib.ptr(ws), wsb,
                              _lib.stream())
        _lib.check(st, "mba_lu_solve")
        return X.reshape(-1) if vec else X


def lu_factor_batched(mats, overwrite: bool = True, persistent: bool = True, options=None):
    """LU of several square complex matrices (same dtype); returns one
    LUFactor per matrix.  ``persistent`` (default): the whole batch in one
    persistent cooperative launch (``mba_lu_factor_batched_persistent``:
    32-column panels of every matrix together, phases separated by grid
    barriers; complex128 takes the lock-step path); otherwise the lock-step
    multi-launch form (``mba_lu_factor_batched``).  ``options`` (persistent,
    complex64): (super_panels, trailing) -- 32-column panels per super panel
    and the trailing product (0: FP32 3M, 1: split FP16); None: (8, 0).
    Nothing is synchronised: call ``check()`` on each after synchronising."""
    lib = _lib.load()
    if not mats:
        return []
    dt = mats[0].dtype
    prec = _lib.prec_of(dt)
    dev = mats[0].device
    n_mat = len(mats)
    lus, keep = [], []
    desc = (_lib.LuMatrix * n_mat)()
    # status words of the batch: [0, n_mat) wait errors, [n_mat, 2 n_mat) info
    status = torch.zeros(2 * n_mat, dtype=torch.int32, device=dev)
    err = status[:n_mat]
    for i, A in enumerate(mats):
        if A.dim() != 2 or A.shape[0] != A.shape[1] or A.dtype != dt:
            raise ValueError("lu_factor_batched needs square matrices of one dtype")
        lu = A if overwrite else A.clone()
        if lu.stride(1) != 1:
            lu = lu.contiguous()
        n = lu.shape[0]
        ipiv = torch.empty(n, dtype=torch.int32, device=dev)
        info = status[n_mat + i:n_mat + i + 1]
        wsb = lib.mba_lu_workspace_bytes(prec, n)
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        desc[i].a, desc[i].n, desc[i].lda = lu.data_ptr(), n, lu.stride(0)
        desc[i].ipiv, desc[i].info = ipiv.data_ptr(), info.data_ptr()
        desc[i].workspace, desc[i].workspace_bytes = ws.data_ptr(), wsb
        lus.append((lu, ipiv, info))
        keep.append(ws)
    nmax = max(A.shape[0] for A in mats)
    if persistent:
        sb = lib.mba_lu_batched_persistent_scratch_bytes(prec, n_mat, nmax)
        scratch = torch.empty(sb, dtype=torch.uint8, device=dev)
        if options is not None and prec == _lib.MBA_F32:
            opts = _lib.BatchedLuOptions(int(options[0]), int(options[1]))
            st = lib.mba_lu_factor_batched_persistent_opts(prec, desc, n_mat, ctypes.byref(opts), _lib.ptr(scratch),
                                                           sb, _lib.ptr(err), _lib.stream())
            _lib.check(st, "mba_lu_factor_batched_persistent_opts")
        else:
            st = lib.mba_lu_factor_batched_persistent(prec, desc, n_mat, _lib.ptr(scratch), sb, _lib.ptr(err),
                                                      _lib.stream())
            _lib.check(st, "mba_lu_factor_batched_persistent")
    else:
        sb = lib.mba_lu_batched_scratch_bytes(prec, n_mat, nmax)
        scratch = torch.empty(sb, dtype=torch.uint8, device=dev)
        st = lib.mba_lu_factor_batched(prec, desc, n_mat, _lib.ptr(scratch), sb, _lib.ptr(err), _lib.stream())
        _lib.check(st, "mba_lu_factor_batched")
    out = [LUFactor._from_factored(lu, ipiv, info, err[i:i + 1]) for i, (lu, ipiv, info) in enumerate(lus)]
    for i, f in enumerate(out):
        f._batch = (status, i, n_mat)
    jobs = (_lib.LuPrepareJob * n_mat)()
    for i, f in enumerate(out):
        jobs[i].lu, jobs[i].n, jobs[i].lda = f.lu.data_ptr(), f.n, f.lu.stride(0)
        jobs[i].ipiv, jobs[i].aux, jobs[i].aux_bytes = f.ipiv.data_ptr(), f.aux.data_ptr(), f.aux.numel()
    pb = lib.mba_lu_solve_prepare_batched_scratch_bytes(n_mat)
    pscratch = torch.empty(pb, dtype=torch.uint8, device=dev)
    st = lib.mba_lu_solve_prepare_batched(prec, jobs, n_mat, _lib.ptr(pscratch), pb, _lib.stream())
    _lib.check(st, "mba_lu_solve_prepare_batched")
    out[0]._keep = (keep, scratch, pscratch)
    return out


def check_batched(lus) -> None:
    """``check()`` of many factorisations with one device-to-host copy per
    batch of ``lu_factor_batched`` (after the stream has been synchronised)."""
    groups = {}
    for f in lus:
        b = getattr(f, "_batch", None)
        if b is None:
            f.check()
        else:
            groups.setdefault(id(b[0]), (b[0], []))[1].append(f)
    for status, fs in groups.values():
        host = status.cpu().tolist()
        for f in fs:
            _, i, n_mat = f._batch
            if host[i] != 0:
                raise _lib.MBAError("mba_lu_factor_batched: a wait timed out")
            f.info = int(host[n_mat + i])
            if f.info != 0:
                raise _lib.MBAError(f"mba_lu_factor: exactly zero pivot in column {f.info - 1}")


def convert_c128_c64_batched(srcs, dsts) -> None:
    """dst_i = complex64(src_i) for contiguous tensors of equal sizes, one launch."""
    lib = _lib.load()
    n = len(srcs)
    if n == 0:
        return
    if len(dsts) != n:
        raise ValueError("convert_c128_c64_batched: one destination per source")
    jobs = (_lib.ConvertJob * n)()
    for q, (a, b) in enumerate(zip(srcs, dsts)):
        if a.dtype != torch.complex128 or b.dtype != torch.complex64 or a.numel() != b.numel():
            raise ValueError("convert_c128_c64_batched: complex128 sources, complex64 destinations of equal size")
        if not (a.is_contiguous() and b.is_contiguous()):
            raise ValueError("convert_c128_c64_batched: contiguous tensors")
        jobs[q].src, jobs[q].dst, jobs[q].count = a.data_ptr(), b.data_ptr(), a.numel()
    sb = lib.mba_convert_c128_c64_batched_scratch_bytes(n)
    scratch = torch.empty(sb, dtype=torch.uint8, device=srcs[0].device)
    _lib.check(lib.mba_convert_c128_c64_batched(jobs, n, _lib.ptr(scratch), sb, _lib.stream()),
               "mba_convert_c128_c64_batched")


def refine_update_batched(Js, Ds) -> torch.Tensor:
    """J_i += D_i (complex128 J_i, complex64 D_i, [n_i, nrhs_i] contiguous), in
    one pass; returns the device tensor of max_c |D_i[:, c]| / |J_i[:, c]|
    (after the update) per system, float64."""
    lib = _lib.load()
    n = len(Js)
    if len(Ds) != n:
        raise ValueError("refine_update_batched: one correction per iterate")
    out = torch.empty(max(n, 1), dtype=torch.float64, device=Js[0].device if n else "cpu")
    if n == 0:
        return out[:0]
    jobs = (_lib.RefineJob * n)()
    for q, (J, D) in enumerate(zip(Js, Ds)):
        if J.dtype != torch.complex128 or D.dtype != torch.complex64 or J.dim() != 2 or J.shape != D.shape:
            raise ValueError("refine_update_batched: complex128 J and complex64 D of one [n, nrhs] shape")
        if J.stride(1) != 1 or D.stride(1) != 1:
            raise ValueError("refine_update_batched: rows must be contiguous")
        jobs[q].j, jobs[q].ldj, jobs[q].d, jobs[q].ldd = J.data_ptr(), J.stride(0), D.data_ptr(), D.stride(0)
        jobs[q].n, jobs[q].nrhs = J.shape[0], J.shape[1]
    sb = lib.mba_refine_update_batched_scratch_bytes(jobs, n)
    scratch = torch.empty(sb, dtype=torch.uint8, device=Js[0].device)
    _lib.check(lib.mba_refine_update_batched(jobs, n, _lib.ptr(out), _lib.ptr(scratch), sb, _lib.stream()),
               "mba_refine_update_batched")
    return out


def gemm_grouped(products, alpha: float = 1.0, beta: float = 0.0):
    """C_j = alpha A_j B_j + beta C_j for a list of (A_j, B_j, C_j) row-major
    complex matrices of one dtype (complex64: 3M product on MFMA; complex128:
    FP64 MFMA), all in one launch.  C_j is updated in place."""
    lib = _lib.load()
    if not products:
        return
    dt = products[0][0].dtype
    for A, B, C in products:
        if not (A.dtype == B.dtype == C.dtype == dt) or A.dim() != 2 or B.dim() != 2 or C.dim() != 2:
            raise ValueError("gemm_grouped: 2-D operands of one complex dtype")
        if A.shape[1] != B.shape[0] or C.shape != (A.shape[0], B.shape[1]):
            raise ValueError("gemm_grouped: shape mismatch")
        if any(T.numel() > 0 and T.stride(1) != 1 for T in (A, B, C)):
            raise ValueError("gemm_grouped: rows must be contiguous")
    products = [p for p in products if p[2].numel() > 0]  # nothing to write
    for A, B, C in products:
        if A.shape[1] == 0:  # empty sum: C = beta C
            C.mul_(beta)
    products = [p for p in products if p[0].shape[1] > 0]
    if not products:
        return
    n_job = len(products)
    jobs = (_lib.GemmJob * n_job)()
    for i, (A, B, C) in enumerate(products):
        jobs[i].a, jobs[i].b, jobs[i].c = A.data_ptr(), B.data_ptr(), C.data_ptr()
        jobs[i].m, jobs[i].n, jobs[i].k = A.shape[0], B.shape[1], A.shape[1]
        jobs[i].lda, jobs[i].ldb, jobs[i].ldc = A.stride(0), B.stride(0), C.stride(0)
        jobs[i].alpha, jobs[i].beta = alpha, beta
    sb = lib.mba_gemm_grouped_scratch_bytes(n_job)
    scratch = torch.empty(sb, dtype=torch.uint8, device=products[0][2].device)
    st = lib.mba_gemm_grouped(_lib.prec_of(dt), jobs, n_job, _lib.ptr(scratch), sb, _lib.stream())
    _lib.check(st, "mba_gemm_grouped")


def lu_solve_batched(lus, Bs, persistent: bool = False, overwrite: bool = False, err_out=None):
    """X_i = A_i^{-1} B_i for prepared factorisations lus[i] (one dtype) and
    right-hand sides Bs[i] [n_i, nrhs_i].  Default: two launches, one
    workgroup per job and 16 right-hand sides walking its 64-row blocks.
    ``persistent`` (complex64): all jobs in one persistent launch, a
    dataflow of 64-row block tasks inside each XCD
    (``mba_lu_solve_batched_persistent``); its error word is checked here
    (one synchronisation).  Returns new tensors (the inputs are not
    modified; ``overwrite``: a right-hand side that already has the factor's
    dtype and contiguous rows is solved in place and returned).  ``err_out``
    (persistent path, a one-element int32 device tensor): receives the
    launch's error word instead of the check here, so the call does not
    synchronise; the caller checks it later."""
    lib = _lib.load()
    n_job = len(lus)
    if n_job == 0:
        return []
    if len(Bs) != n_job or any(lu.lu.dtype != lus[0].lu.dtype for lu in lus):
        raise ValueError("lu_solve_batched: one right-hand side per factor, one dtype")
    jobs = (_lib.LuSolveJob * n_job)()
    Xs, keep = [], []
    for i, (lu, B) in enumerate(zip(lus, Bs)):
        X = B.reshape(lu.n, -1)
        if not (overwrite and X.dtype == lu.lu.dtype and X.is_contiguous()):
            X = X.to(lu.lu.dtype).clone().contiguous()
        W = torch.empty_like(X)
        jobs[i].lu, jobs[i].n, jobs[i].lda = lu.lu.data_ptr(), lu.n, lu.lu.stride(0)
        jobs[i].aux, jobs[i].b, jobs[i].ldb = lu.aux.data_ptr(), X.data_ptr(), X.stride(0)
        jobs[i].nrhs, jobs[i].work = X.shape[1], W.data_ptr()
        Xs.append(X)
        keep.append(W)
    if persistent:
        sb = lib.mba_lu_solve_batched_persistent_scratch_bytes(jobs, n_job)
        if sb == 0:
            raise ValueError("lu_solve_batched: too many blocks or right-hand sides for one launch")
        scratch = torch.empty(sb, dtype=torch.uint8, device=Xs[0].device)
        err = err_out if err_out is not None else torch.zeros(1, dtype=torch.int32, device=Xs[0].device)
        if err.dtype != torch.int32 or err.numel() < 1 or not err.is_cuda:
            raise ValueError("lu_solve_batched: err_out must be an int32 device tensor")
        st = lib.mba_lu_solve_batched_persistent(lus[0].prec, jobs, n_job, _lib.ptr(scratch), sb, _lib.ptr(err),
                                                 _lib.stream())
        _lib.check(st, "mba_lu_solve_batched_persistent")
        if err_out is None and int(err.item()) != 0:
            raise RuntimeError("mba_lu_solve_batched_persistent: a block wait timed out")
    else:
        sb = lib.mba_lu_solve_batched_scratch_bytes(n_job)
        scratch = torch.empty(sb, dtype=torch.uint8, device=Xs[0].device)
        st = lib.mba_lu_solve_batched(lus[0].prec, jobs, n_job, _lib.ptr(scratch), sb, _lib.stream())
        _lib.check(st, "mba_lu_solve_batched")
    return Xs


def residual_c128(Z: torch.Tensor, X: torch.Tensor, V: torch.Tensor) -> torch.Tensor:
    """r = V - Z X in FP64 (complex128 Z; X and V are taken in complex128)."""
    if Z.dim() != 2 or Z.shape[0] != Z.shape[1] or Z.dtype != torch.complex128 or Z.stride(1) != 1:
        raise ValueError("residual_c128: square complex128 Z with unit column stride")
    lib = _lib.load()
    n = Z.shape[0]
    X2 = X.reshape(n, -1).to(torch.complex128).contiguous()
    V2 = V.reshape(n, -1).to(torch.complex128).contiguous()
    if X2.shape != V2.shape:
        raise ValueError("residual_c128: X and V of one shape")
    R = torch.empty_like(V2)
    st = lib.mba_residual_c128(_lib.ptr(Z), n, Z.stride(0), _lib.ptr(X2), _lib.ptr(V2), _lib.ptr(R),
                               X2.shape[1], _lib.stream())
    _lib.check(st, "mba_residual_c128")
    return R.reshape(V.shape)


def to_c64(Z: torch.Tensor) -> torch.Tensor:
    """complex64 copy of a 2-D matrix (rounded from complex128 on the device;
    a complex64 input is copied as is).  The conversion kernel reads 16-byte
    elements, so anything but a complex128 matrix with unit column stride is
    rejected (or copied) before the launch."""
    if Z.dim() != 2:
        raise ValueError("to_c64: a 2-D matrix")
    if Z.dtype == torch.complex64:
        return Z.clone(memory_format=torch.contiguous_format)
    if Z.dtype != torch.complex128:
        raise ValueError(f"to_c64: complex128 or complex64 input, not {Z.dtype}")
    if Z.stride(1) != 1:
        Z = Z.contiguous()
    lib = _lib.load()
    m, n = Z.shape
    out = torch.empty((m, n), dtype=torch.complex64, device=Z.device)
    st = lib.mba_convert_c128_c64(_lib.ptr(Z), m, n, Z.stride(0), _lib.ptr(out), n, _lib.stream())
    _lib.check(st, "mba_convert_c128_c64")
    return out


@dataclass
class RefineInfo:
    iterations: int
    converged: bool
    corrections: list
    fallback: str = ""


def refine_solve(Z64: torch.Tensor, V: torch.Tensor, lu32: Optional[LUFactor] = None,
                 tol: float = 1e-8, maxit: int = 20, fallback_fp64: bool = True):
    """Mixed-precision solve of Z64 J = V to FP64 accuracy.

    J0 = LU32 \\ V; repeat r = V - Z64 J (FP64), d = LU32 \\ r, J += d until
    the error left is at most ``tol``: the relative correction c_k =
    ||d_k|| / ||J|| estimates the error before sweep k, and c_k / c_{k-1} the
    contraction of one sweep, so the error after it is about c_k^2 / c_{k-1}
    (J0 is the same iteration's first step from J = 0, a correction of
    relative size c_0 = 1).  If the correction stops shrinking
    (|d_k| > 0.5 |d_{k-1}| twice) or maxit is reached, falls back to an FP64
    LU of Z64 when ``fallback_fp64`` is set.  Returns (J complex128,
    RefineInfo)."""
    if lu32 is None:
        lu32 = LUFactor(to_c64(Z64), overwrite=True)
    V128 = V.to(torch.complex128)
    J = lu32.solve(V.to(torch.complex64)).to(torch.complex128)
    corr = []
    stalls = 0
    for it in range(1, maxit + 1):
        r = residual_c128(Z64, J, V128)
        d = lu32.solve(r.to(torch.complex64)).to(torch.complex128)
        J = J + d
        rel = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(J))
        prev = corr[-1] if corr else 1.0
        corr.append(rel)
        if min(rel, rel * rel / max(prev, 1e-300)) <= tol:
            return J, RefineInfo(it, True, corr)
        if len(corr) >= 2 and corr[-1] > 0.5 * corr[-2]:
            stalls += 1
            if stalls >= 2:
                break
    if not fallback_fp64:
        return J, RefineInfo(len(corr), False, corr)
    J = LUFactor(Z64).solve(V128)
    return J, RefineInfo(len(corr), False, corr, fallback="fp64-lu")
