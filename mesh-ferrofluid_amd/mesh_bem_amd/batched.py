"""Many meshes and many incident waves at once.

* ``BatchedMoM.assemble``: the EFIE face-edge blocks of every mesh in one
  assembly launch (``mba_mom_zmat_faceedge_ns_batched``: a 1-D grid over the
  tiles of all meshes) plus one launch for all near pairs, then a fixed-order
  gather per mesh into its RWG matrix.
* ``BatchedMoM.solve``: the FP32 LUs of all rounded matrices in lock step
  (GEMM tile variant picked by the autotuner for the size), the triangular
  solves of every mesh and right-hand side in two launches, and refinement
  whose residual R = V - Z J is an FP64 MFMA GEMM over all right-hand sides.
* ``GemmTuner``: times the FP32 GEMM tile variants on the LU's trailing-update
  shapes per size bucket and caches the fastest (optionally in a JSON file).
"""

import ctypes
import json
import os
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib
from .dense import (check_batched, convert_c128_c64_batched, gemm_grouped, lu_factor_batched, lu_solve_batched,
                    refine_update_batched, to_c64)
from .geometry import get_edge_unique
from .mom import MomMesh, rhs_ns, rhs_ns_many, rule_tensor

BUCKETS = (1024, 2048, 3072, 4096, 8192, 16384, 32768)


def _bucket(n: int) -> int:
    for b in BUCKETS:
        if n <= b:
            return b
    return BUCKETS[-1]


class GemmTuner:
    """Fastest FP32 GEMM tile variant per matrix-size bucket."""

    def __init__(self, path: Optional[str] = None, reps: int = 3):
        self.path = path
        self.reps = reps
        self.best: Dict[int, int] = {}
        self.timings: Dict[int, List[float]] = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.best = {int(k): int(v) for k, v in json.load(f).items()}

    def tune(self, n: int, device="cuda") -> int:
        b = _bucket(n)
        if b in self.best:
            return self.best[b]
        lib = _lib.load()
        m = max(256, b - 256)
        k = 256
        A = torch.randn((m, k), dtype=torch.complex64, device=device)
        B = torch.randn((k, m), dtype=torch.complex64, device=device)
        C = torch.randn((m, m), dtype=torch.complex64, device=device)
        times = []
        for v in range(6):
            def run():
                st = lib.mba_gemm_variant(_lib.MBA_F32, v, m, m, k, -1.0, _lib.ptr(A), k, _lib.ptr(B), m, 1.0,
                                          _lib.ptr(C), m, _lib.stream())
                _lib.check(st, "mba_gemm_variant")
            run()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(self.reps):
                run()
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) / self.reps)
        self.timings[b] = times
        self.best[b] = int(np.argmin(times))
        if self.path:
            os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump({str(k2): v2 for k2, v2 in self.best.items()}, f)
        return self.best[b]

    def apply(self, n: int):
        lib = _lib.load()
        _lib.check(lib.mba_gemm_set_variant(_lib.MBA_F32, self.tune(n)), "mba_gemm_set_variant")


class BatchLuTuner:
    """Fastest schedule of the persistent batched LU for a batch: super-panel
    width (4 or 8 panels of 32 columns, i.e. the trailing product's k = 128 or
    256) x trailing product (FP32 3M on MFMA, or the FP32-accurate split-FP16
    product).  Each candidate factors copies of the batch's own matrices
    (``reps`` timed runs after a warm-up, device events); the fastest is kept
    per batch shape (the sorted matrix orders), optionally in a JSON file."""

    CANDIDATES = ((8, 0), (8, 1), (4, 0), (4, 1))

    def __init__(self, path: Optional[str] = None, reps: int = 2):
        self.path = path
        self.reps = reps
        self.best: Dict[str, List[int]] = {}
        self.timings: Dict[str, List[float]] = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.best = {k: list(v) for k, v in json.load(f).items()}

    @staticmethod
    def key(mats) -> str:
        return ",".join(str(n) for n in sorted(int(A.shape[0]) for A in mats))

    def tune(self, mats) -> tuple:
        """mats: the batch's complex64 matrices (not modified)."""
        k = self.key(mats)
        if k in self.best:
            return tuple(self.best[k])
        times = []
        for cand in self.CANDIDATES:
            def run():
                lus = lu_factor_batched([A.clone() for A in mats], overwrite=True, persistent=True, options=cand)
                return lus
            run()
            torch.cuda.synchronize()
            ts = []
            for _ in range(self.reps):
                copies = [A.clone() for A in mats]
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                lu_factor_batched(copies, overwrite=True, persistent=True, options=cand)
                e1.record()
                torch.cuda.synchronize()
                ts.append(e0.elapsed_time(e1))
                del copies
            times.append(min(ts))
        self.timings[k] = times
        self.best[k] = list(self.CANDIDATES[int(np.argmin(times))])
        if self.path:
            os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump(self.best, f)
        return tuple(self.best[k])


@dataclass
class BatchResult:
    J: List[torch.Tensor]  # per mesh [E, D] complex128
    iterations: List[int]
    converged: List[bool]
    times: Dict[str, float] = field(default_factory=dict)


class BatchedMoM:
    def __init__(self, meshes: List[MomMesh], ks: List[float]):
        self.meshes = meshes
        self.ks = [float(k) for k in ks]
        dev = meshes[0].verts.device
        self.device = dev
        voff, foff = 0, 0
        vs, fs = [], []
        self.face_off = []
        for m in meshes:
            vs.append(m.verts.to(torch.float64))
            fs.append(m.faces.to(torch.int64) + voff)
            self.face_off.append(foff)
            voff += m.n_verts
            foff += m.n_faces
        self.verts = torch.cat(vs, 0).contiguous()
        self.faces = torch.cat(fs, 0).to(torch.int32).contiguous()
        self.total_faces = foff
        self.n_near = [int(m.near_ptr[-1].item()) for m in meshes]

    @staticmethod
    def build(geoms, ks, device="cuda"):
        """geoms: list of (verts [V,3], faces [F,3]) numpy arrays."""
        meshes = []
        for V, F in geoms:
            vt = torch.tensor(V, dtype=torch.float64)
            ft = torch.tensor(F, dtype=torch.int32)
            edges, fe, _ = get_edge_unique(vt, ft)
            meshes.append(MomMesh.build(vt, F, fe, edges.shape[0], device=device))
        return BatchedMoM(meshes, ks)

    def assemble(self, prec: int = _lib.MBA_F64) -> List[torch.Tensor]:
        """RWG matrices [E_m, E_m] of every mesh (complex128 for MBA_F64)."""
        lib = _lib.load()
        cdt = torch.complex64 if prec == _lib.MBA_F32 else torch.complex128
        rdt = torch.float32 if prec == _lib.MBA_F32 else torch.float64
        n = len(self.meshes)
        zfs = [torch.empty((3 * m.n_faces, 3 * m.n_faces), dtype=cdt, device=self.device) for m in self.meshes]
        desc = (_lib.NsMesh * n)()
        for i, (m, zf) in enumerate(zip(self.meshes, zfs)):
            desc[i].n_faces = m.n_faces
            desc[i].face_off = self.face_off[i]
            desc[i].wavenumber = self.ks[i]
            desc[i].near_ptr = m.near_ptr.data_ptr()
            desc[i].near_idx = m.near_idx.data_ptr()
            desc[i].zf = zf.data_ptr()
            desc[i].ldzf = zf.stride(0)
            desc[i].n_near = self.n_near[i]
        rule = rule_tensor(self.device)
        wsb = lib.mba_mom_batched_workspace_bytes(prec, self.total_faces, n)
        ws = torch.empty(wsb, dtype=torch.uint8, device=self.device)
        verts = self.verts.to(rdt).contiguous()
        st = lib.mba_mom_zmat_faceedge_ns_batched(prec, _lib.ptr(verts), _lib.ptr(self.faces), self.total_faces,
                                                  _lib.ptr(rule), rule.shape[0], desc, n, _lib.ptr(ws), wsb,
                                                  _lib.stream())
        _lib.check(st, "mba_mom_zmat_faceedge_ns_batched")
        # every mesh's RWG gather in one launch
        Zs = [torch.empty((m.n_edges, m.n_edges), dtype=cdt, device=self.device) for m in self.meshes]
        jobs = (_lib.GatherJob * n)()
        for i, (m, zf, Z) in enumerate(zip(self.meshes, zfs, Zs)):
            jobs[i].zf, jobs[i].ldzf = zf.data_ptr(), zf.stride(0)
            jobs[i].rwg_fe, jobs[i].n_edges = m.rwg_fe.data_ptr(), m.n_edges
            jobs[i].z, jobs[i].ldz = Z.data_ptr(), Z.stride(0)
        sb = lib.mba_mom_gather_rwg_batched_scratch_bytes(n)
        gs = torch.empty(sb, dtype=torch.uint8, device=self.device)
        _lib.check(lib.mba_mom_gather_rwg_batched(prec, jobs, n, _lib.ptr(gs), sb, _lib.stream()),
                   "mba_mom_gather_rwg_batched")
        del zfs
        return Zs

    def rhs(self, dirs: List[np.ndarray], pols: List[np.ndarray], prec=_lib.MBA_F64) -> List[torch.Tensor]:
        """Plane-wave right-hand sides V_m [E_m, ndir_m] of every mesh: the
        directions and polarisations go to the device in one copy each, the
        face-edge integrals and RWG sums of all meshes are one launch each."""
        lib = _lib.load()
        cdt = torch.complex64 if prec == _lib.MBA_F32 else torch.complex128
        rdt = torch.float32 if prec == _lib.MBA_F32 else torch.float64
        n = len(self.meshes)
        ds = [np.atleast_2d(np.asarray(d, dtype=np.float64)) for d in dirs]
        ps = [np.atleast_2d(np.asarray(p, dtype=np.float64)) for p in pols]
        if len(ds) != n or len(ps) != n or any(d.shape != p.shape or d.shape[1] != 3 for d, p in zip(ds, ps)):
            raise ValueError("rhs: dirs and pols [ndir, 3] per mesh")
        D = torch.from_numpy(np.ascontiguousarray(np.concatenate(ds))).to(self.device)
        P = torch.from_numpy(np.ascontiguousarray(np.concatenate(ps))).to(self.device)
        rule = rule_tensor(self.device)
        Vs = [torch.empty((m.n_edges, d.shape[0]), dtype=cdt, device=self.device) for m, d in zip(self.meshes, ds)]
        desc = (_lib.NsRhsMesh * n)()
        off = 0
        for i, (m, d, V) in enumerate(zip(self.meshes, ds, Vs)):
            desc[i].n_faces, desc[i].face_off, desc[i].wavenumber = m.n_faces, self.face_off[i], self.ks[i]
            desc[i].dirs, desc[i].pols = D[off].data_ptr(), P[off].data_ptr()
            desc[i].ndir = d.shape[0]
            desc[i].rwg_fe, desc[i].n_edges, desc[i].v = m.rwg_fe.data_ptr(), m.n_edges, V.data_ptr()
            off += d.shape[0]
        verts = self.verts.to(rdt).contiguous()
        wsb = lib.mba_mom_rhs_ns_batched_workspace_bytes(prec, self.total_faces, desc, n)
        ws = torch.empty(wsb, dtype=torch.uint8, device=self.device)
        _lib.check(lib.mba_mom_rhs_ns_batched(prec, _lib.ptr(verts), _lib.ptr(self.faces), self.total_faces,
                                              _lib.ptr(rule), rule.shape[0], desc, n, _lib.ptr(ws), wsb,
                                              _lib.stream()), "mba_mom_rhs_ns_batched")
        return Vs

    def solve(self, Zs: List[torch.Tensor], Vs: List[torch.Tensor], tuner: Optional[GemmTuner] = None,
              tol: float = 1e-8, maxit: int = 10, persistent: bool = True,
              lu_tuner: Optional[BatchLuTuner] = None, solve_persistent: bool = True) -> BatchResult:
        """Mixed precision for all meshes: the FP32 LUs of the whole batch
        (``lu_factor_batched``; ``persistent``: one persistent launch, its
        schedule from ``lu_tuner`` when given; otherwise in lock step, the
        trailing products of every matrix in one grouped launch per step and
        kind, GEMM tile variant of the largest size bucket from the tuner),
        then refinement sweeps J += LU32 \\ (V - Z J).  A sweep is one grouped
        FP64 MFMA launch for the residuals of every mesh still refining and the
        triangular solves of all their right-hand sides
        (``lu_solve_batched``: by default one launch, a
        dataflow of 64-row block tasks inside each XCD; ``solve_persistent``
        False: the two-launch block walks); one host synchronisation then
        reads the largest relative correction of every mesh.  The correction of
        sweep s estimates the error before it, and its ratio to the previous
        correction the contraction per sweep (the first solve is the same
        iteration's step from J = 0, a correction of relative size 1); a mesh
        stops once the error left after the sweep (correction x contraction)
        is at most ``tol``."""
        lib = _lib.load()
        if tuner is not None:
            tuner.apply(max(Z.shape[0] for Z in Zs))
        if all(Z.dtype == torch.complex128 and Z.is_contiguous() for Z in Zs):
            # one buffer and one conversion launch for the whole batch
            zsz = [Z.numel() for Z in Zs]
            zbuf = torch.empty(sum(zsz), dtype=torch.complex64, device=Zs[0].device)
            Z32 = [t.view(Z.shape) for t, Z in zip(torch.split(zbuf, zsz), Zs)]
            convert_c128_c64_batched(list(Zs), Z32)
        else:
            Z32 = [to_c64(Z) for Z in Zs]
        opts = lu_tuner.tune(Z32) if (persistent and lu_tuner is not None) else None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lus = lu_factor_batched(Z32, overwrite=True, persistent=persistent, options=opts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        check_batched(lus)
        n_mesh = len(Zs)
        dev = Zs[0].device
        V128s = [V.to(torch.complex128).contiguous() for V in Vs]
        shapes = [tuple(V.shape) for V in V128s]
        sizes = [V.numel() for V in V128s]

        def views(buf):
            return [t.view(sh) for t, sh in zip(torch.split(buf, sizes), shapes)]

        # one buffer per kind for the whole batch, a view per mesh
        Js = views(torch.zeros(sum(sizes), dtype=torch.complex128, device=dev))
        Rs = views(torch.empty(sum(sizes), dtype=torch.complex128, device=dev))
        B32s = views(torch.empty(sum(sizes), dtype=torch.complex64, device=dev))
        errs = torch.zeros(maxit + 1, dtype=torch.int32, device=dev)  # solve launches' error words
        convert_c128_c64_batched(V128s, B32s)
        Xs = lu_solve_batched(lus, B32s, persistent=solve_persistent, overwrite=True, err_out=errs[0:1])
        refine_update_batched(Js, Xs)  # J = LU32 \ V
        # the first solve is the iteration's first step from J = 0: a
        # correction of relative size 1
        its, conv, prev = [0] * n_mesh, [False] * n_mesh, [1.0] * n_mesh
        active = list(range(n_mesh))
        sweep = 0
        while active:
            sweep += 1
            torch._foreach_copy_([Rs[i] for i in active], [V128s[i] for i in active])
            gemm_grouped([(Zs[i], Js[i], Rs[i]) for i in active], alpha=-1.0, beta=1.0)
            convert_c128_c64_batched([Rs[i] for i in active], [B32s[i] for i in active])
            Ds = lu_solve_batched([lus[i] for i in active], [B32s[i] for i in active],
                                  persistent=solve_persistent, overwrite=True, err_out=errs[sweep:sweep + 1])
            rels = refine_update_batched([Js[i] for i in active], Ds).cpu().tolist()
            for i, r in zip(active, rels):
                its[i] += 1
                # error left after this sweep: its correction r times the
                # observed contraction r / (previous correction)
                est = min(r, r * r / max(prev[i], 1e-300))
                prev[i] = r
                conv[i] = est <= tol
            active = [i for i in active if not conv[i] and its[i] < maxit]
        if int(errs.max().item()) != 0:
            raise RuntimeError("mba_lu_solve_batched_persistent: a block wait timed out")
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if tuner is not None:
            _lib.check(lib.mba_gemm_set_variant(_lib.MBA_F32, 4), "mba_gemm_set_variant")
        return BatchResult(Js, its, conv, {"lu_s": t1 - t0, "solve_refine_s": t2 - t1})


def star_batch(n_meshes: int, seed0: int = 0, freqs=(5, 6, 7, 8, 9, 10, 11), n_dirs: int = 128,
               k_range=(1.0, 8.0)):
    """Seeded workload of many uneven meshes: star meshes on geodesic spheres of
    frequency freqs[i % len] (E = 30 n^2), ka ~ U(k_range), n_dirs waves each."""
    from .geometry import star_mesh
    from .mgn import random_waves

    geoms, ks, dirs, pols = [], [], [], []
    for i in range(n_meshes):
        s = seed0 + i
        V, F, _ = star_mesh(s, freqs[i % len(freqs)])
        rng = np.random.Generator(np.random.PCG64(7919 * s + 3))
        ks.append(float(rng.uniform(*k_range)))
        d, p = random_waves(rng, n_dirs)
        geoms.append((V, F))
        dirs.append(d)
        pols.append(p)
    return geoms, ks, dirs, pols
