#!/usr/bin/env python
"""Headline benchmark: mixed-precision RWG MoM solve at 30 720 unknowns on one MI355X.

Workload (one "step"): icosphere subdivision 5 (F = 20 480 triangles,
E = 30 720 RWG), ka = 10, x-polarised plane wave along +z, EFIE with the
7-point rule and analytic FP64 near-singular terms:
  FP64 assembly of Z (complex128, resident) and of the RHS,
  rounding to complex64, FP32 blocked LU (MFMA trailing update),
  iterative refinement with FP64 residuals to ||dJ|| <= 1e-8 ||J||.
The mesh tables are built before the timed region; every step re-assembles.

Printed value: whole-solve throughput in face pairs per second (Mpairs/s =
1e-6 F^2 / t_step).  Alongside: the plain FP64 solve on the same GPU
(FP64 assembly + torch.linalg.solve complex128), the relative error of the
mixed solution against it, FP32 assembly throughput, the roofline record of
the dominant kernel (the LU trailing GEMM, MFMA bound) from HIP events, and
the CPU baseline (the oracle's FP64 assembly on a sample of source faces and
LAPACK zgesv at a smaller order, both scaled).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mesh-ferrofluid_amd"))

METRIC = "MoM solve time @30k RWG, relerr<=1e-6 vs FP64, mixed vs FP64; Mpairs/s"
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name):
    if name == "c2":
        return dict(subdiv=5, k=10.0)
    if name == "c1":
        return dict(subdiv=3, k=1.0)
    if name == "small":
        return dict(subdiv=4, k=5.0)
    raise ValueError(name)


def cpu_baseline(V, F, k, E, target_s=12.0):
    """Oracle FP64 assembly on a window of source faces + LAPACK zgesv at order
    E_s, scaled to the full workload.  Returns the record."""
    import numpy as np
    import scipy.linalg

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from mesh_bem_amd.mom import near_pairs

    cores = len(os.sched_getaffinity(0))
    nF = len(F)
    ptr, idx = near_pairs(V, F)
    nb = 4
    t = 0.0
    while True:
        t0 = time.perf_counter()
        O.ns_zmat_faceedge_window(V, F, k, ptr, idx, 0, nb)
        t = time.perf_counter() - t0
        if t > 2.0 or nb >= nF:
            break
        nb = min(nF, nb * 4)
    t_asm = t * nF / nb
    Es = 3072
    rng = np.random.default_rng(0)
    A = rng.standard_normal((Es, Es)) + 1j * rng.standard_normal((Es, Es))
    b = rng.standard_normal(Es) + 1j * rng.standard_normal(Es)
    t0 = time.perf_counter()
    scipy.linalg.solve(A, b, check_finite=False)
    t_lu = (time.perf_counter() - t0) * (E / Es) ** 3
    t_cpu = t_asm + t_lu
    return {
        "value": 1e-6 * nF * nF / t_cpu,
        "unit": "Mpairs/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle FP64 EFIE assembly of {nb} of {nF} source faces ({t:.1f} s, scaled x{nF / nb:.0f})"
                   f" + scipy zgesv at E={Es} ({t_lu * (Es / E) ** 3:.1f} s, scaled by (E/{Es})^3);"
                   f" extrapolated full solve {t_cpu:.0f} s"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "small"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison")
    ap.add_argument("--no-extra", action="store_true", help="only the timed steps")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE={world}); "
              "the distributed layout is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from mesh_bem_amd import _lib
    from mesh_bem_amd.dense import LUFactor, refine_solve, to_c64
    from mesh_bem_amd.geometry import icosphere
    from mesh_bem_amd.mom import MomMesh, assemble_ns, rhs_ns

    dev = torch.device("cuda:0")
    wl = workload(args.workload)
    Vn, Fn = icosphere(wl["subdiv"])
    k = wl["k"]
    nF = len(Fn)
    log(f"workload {args.workload}: F={nF} k={k}; building mesh tables")
    mesh = MomMesh.build(torch.tensor(Vn, dtype=torch.float64), Fn, device=dev)
    E = mesh.n_edges
    dirs, pols = [[0.0, 0.0, 1.0]], [[1.0, 0.0, 0.0]]
    F64 = _lib.MBA_F64
    Z64 = torch.empty((E, E), dtype=torch.complex128, device=dev)
    stream = torch.cuda.current_stream()

    def step(profile=False):
        assemble_ns(mesh, k, F64, out=Z64)
        V = rhs_ns(mesh, k, F64, dirs, pols)
        lu = LUFactor(to_c64(Z64), overwrite=True, profile=profile)
        J, info = refine_solve(Z64, V, lu32=lu, tol=1e-8)
        return J, info, lu, V

    for i in range(args.warmup):
        log(f"warmup step {i + 1}/{args.warmup}")
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    for i in range(args.steps):
        J, info, lu, V = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    value = 1e-6 * nF * nF / (ms * 1e-3)
    log(f"mixed solve {ms:.1f} ms/step, refine iterations {info.iterations}, converged {info.converged}")
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 factor + c128 assembly/residual",
        "data": "synthetic",
        "config": {"workload": f"icosphere subdiv {wl['subdiv']} ({nF} triangles, {E} RWG), ka={k}, "
                               "EFIE 7-point rule, x-pol plane wave along +z, mixed-precision LU + FP64 refinement",
                   "faces": nF, "rwg": E, "ka": k, "parallelism": "single GPU"},
        "refine_iterations": info.iterations,
        "refine_converged": info.converged,
        "refine_fallback": info.fallback or None,
    }
    if not args.no_extra:
        # phase breakdown + roofline of the dominant kernel (trailing GEMM)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        assemble_ns(mesh, k, F64, out=Z64)
        V = rhs_ns(mesh, k, F64, dirs, pols)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        Z32 = to_c64(Z64)
        lu = LUFactor(Z32, overwrite=True, profile=True)
        torch.cuda.synchronize()
        tc = time.perf_counter()
        J, info = refine_solve(Z64, V, lu32=lu, tol=1e-8)
        torch.cuda.synchronize()
        td = time.perf_counter()
        prof = lu.profile
        gemm_ms = prof["phase_ms"]["gemm"]
        nl = max(1, prof["phase_launches"]["gemm"])
        achieved = prof["gemm_flops"] / (gemm_ms * 1e-3) / 1e12
        out["roofline"] = {
            "kernel": "k_cgemm_f32 (LU trailing update, complex-as-real v_mfma_f32_32x32x2f32)",
            "bound": "mfma",
            "achieved": round(achieved, 2),
            "peak": PEAK_FP32_MFMA_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_MFMA_TFLOPS, 4),
            "traffic": None,
            "launches": nl,
            "avg_launch_ms": round(gemm_ms / nl, 4),
            "flops_per_launch_avg": prof["gemm_flops"] / nl,
        }
        out["phases_s"] = {
            "assembly_fp64": round(tb - ta, 4),
            "lu_fp32": round(tc - tb, 4),
            "lu_phase_ms": {p: round(v, 2) for p, v in prof["phase_ms"].items()},
            "refine": round(td - tc, 4),
        }
        out["lu_fp32_tflops_effective"] = round((8.0 / 3.0) * E ** 3 / (tc - tb) / 1e12, 2)
        # FP32 assembly throughput (Mpairs/s = F^2 / t)
        Z32a = torch.empty((E, E), dtype=torch.complex64, device=dev)
        m32 = MomMesh(**{**mesh.__dict__, "verts": mesh.verts.to(torch.float32)})
        assemble_ns(m32, k, _lib.MBA_F32, out=Z32a)
        torch.cuda.synchronize()
        te = time.perf_counter()
        assemble_ns(m32, k, _lib.MBA_F32, out=Z32a)
        torch.cuda.synchronize()
        t32 = time.perf_counter() - te
        out["assembly_fp32_Mpairs_s"] = round(1e-6 * nF * nF / t32, 1)
        out["assembly_fp64_Mpairs_s"] = round(1e-6 * nF * nF / (tb - ta), 1)
        Zerr = (Z32a.to(torch.complex128) - Z64).abs().max().item() / Z64.abs().max().item()
        out["assembly_fp32_max_abs_err_rel_max"] = float(f"{Zerr:.3e}")
        del Z32a, Z32
        if not args.no_fp64:
            log("plain FP64 solve (complex128 assembly + torch.linalg.solve)")
            torch.cuda.synchronize()
            tf0 = time.perf_counter()
            assemble_ns(mesh, k, F64, out=Z64)
            V = rhs_ns(mesh, k, F64, dirs, pols)
            J64 = torch.linalg.solve(Z64, V)
            torch.cuda.synchronize()
            tf1 = time.perf_counter()
            rel = (torch.linalg.vector_norm(J - J64) / torch.linalg.vector_norm(J64)).item()
            out["fp64_solve_s"] = round(tf1 - tf0, 4)
            out["mixed_solve_s"] = round(ms * 1e-3, 4)
            out["speedup_vs_fp64"] = round((tf1 - tf0) / (ms * 1e-3), 3)
            out["relerr_vs_fp64"] = float(f"{rel:.3e}")
            log(f"fp64 solve {tf1 - tf0:.3f} s, relerr {rel:.2e}")
        if not args.no_cpu:
            log("CPU baseline (oracle sample)")
            out["cpu_baseline"] = cpu_baseline(Vn, Fn, k, E)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
