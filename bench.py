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
import math
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mesh-ferrofluid_amd"))

METRIC = "MoM solve time @30k RWG, relerr<=1e-6 vs FP64, mixed vs FP64; Mpairs/s"
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_FP16_MFMA_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pmc_traffic(name):
    """HBM bytes per launch of the dominant kernel, as measured by the
    FETCH_SIZE / WRITE_SIZE rocprofv3 runs summarised in profiles/<name>."""
    import re

    path = os.path.join(ROOT, "profiles", name)
    try:
        m = re.search(r"traffic_bytes_per_launch:\s*([0-9]+)", open(path).read())
        return int(m.group(1)) if m else None
    except OSError:
        return None


def workload(name):
    if name == "c5":
        return dict(freq=45, chi=3.0)
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
    Es = 8192 if E >= 8192 else E
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


def cpu_baseline_c5(F_full, chi, target_s=10.0):
    """Oracle P0 assembly (FP64) + LAPACK solve on a smaller droplet, scaled
    by (F/F_s)^2 (assembly) and (F/F_s)^3 (dense solve, the reference's
    torch.linalg.solve path)."""
    import numpy as np
    import scipy.linalg

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from mesh_bem_amd.geometry import noisy_ellipsoid
    from mesh_bem_amd.mom import near_pairs

    V, Fc, _ = noisy_ellipsoid(0, 24)
    Fs = len(Fc)
    ptr, idx = near_pairs(V, Fc)
    t0 = time.perf_counter()
    Vm, Km = O.laplace_p0(V, Fc, ptr, idx)
    t_asm = time.perf_counter() - t0
    A, n = O.face_areas_normals(V, Fc)
    t0 = time.perf_counter()
    scipy.linalg.solve((1 / chi + 0.5) * np.diag(A) + Km, A * n[:, 2], check_finite=False)
    t_sol = time.perf_counter() - t0
    t_cpu = t_asm * (F_full / Fs) ** 2 + t_sol * (F_full / Fs) ** 3
    return {
        "value": 1e-6 * F_full * F_full / t_cpu,
        "unit": "Mpairs/s",
        "cores": len(os.sched_getaffinity(0)),
        "kind": "port",
        "sample": (f"oracle FP64 P0 V+K' assembly at F={Fs} ({t_asm:.2f} s, scaled by (F/{Fs})^2) + scipy dgesv "
                   f"at F={Fs} ({t_sol:.2f} s, scaled by (F/{Fs})^3); extrapolated step {t_cpu:.0f} s"),
    }


def run_c5(args, dev):
    """Config 5: magnetic droplet loop.  Each step: GPU near-pair lists, FP32
    P0 V + K' assembly (FP64 analytic near terms), diagonal shift, GMRES
    (FP32 matrix, FP64 iterate) warm-started from the previous step, face
    potentials.  Value: face pairs per second of the whole step."""
    import numpy as np
    import torch

    from mesh_bem_amd.gmres import DenseOperator
    from mesh_bem_amd.magnetic import MagneticSolver, droplet_steps

    freq = 45 if args.workload == "c5" else 15
    chi = 3.0
    steps = args.steps
    ms = MagneticSolver(chi, tol=1e-6, restart=50)
    gen = droplet_steps(0, freq, steps + args.warmup)
    prev = None
    rec = []
    nF = None
    log(f"workload {args.workload}: geodesic frequency {freq}, {steps} steps after {args.warmup} warm-up")
    for s, V, Fc, c in gen:
        v = torch.tensor(V, device=dev)
        f = torch.tensor(Fc, dtype=torch.int32, device=dev)
        nF = len(Fc)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = ms.solve(v, f, x0=prev)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        prev = out.sigma
        if s >= args.warmup:
            rec.append((t1 - t0, out.assembly_s, out.solve_s, out.info.iterations, out.info.residual))
    t = np.array([r[0] for r in rec])
    ms_step = 1e3 * t.mean()
    # cold-start GMRES on the last system, for the warm-start comparison
    A, b, Vm, area = ms.system(v, f)
    from mesh_bem_amd.gmres import gmres

    torch.cuda.synchronize()
    tc = time.perf_counter()
    _, cold = gmres(A, b, restart=50, tol=1e-6)
    torch.cuda.synchronize()
    t_cold = time.perf_counter() - tc
    # matvec roofline (HBM): 4 F^2 bytes per product, HIP events on the launch stream
    op = DenseOperator(A)
    x = torch.randn(nF, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    for _ in range(3):
        op.apply(x, y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    nrep = 20
    e0.record()
    for _ in range(nrep):
        op.apply(x, y)
    e1.record()
    torch.cuda.synchronize()
    mv_ms = e0.elapsed_time(e1) / nrep
    gbs = 4.0 * nF * nF / (mv_ms * 1e-3) / 1e9
    out = {
        "metric": "C5 magnetic droplet step (P0 V+K' assembly + warm-started GMRES); Mpairs/s",
        "value": round(1e-6 * nF * nF / (ms_step * 1e-3), 3),
        "unit": "Mpairs/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 operators + f64 near terms / GMRES iterate",
        "data": "synthetic",
        "config": {"workload": f"noisy ellipsoid, geodesic frequency {freq} ({nF} triangles), axes 1:1:{c:.3f}, "
                               f"radial noise 0.02, step perturbation 0.002, chi={chi}, H0=z",
                   "faces": nF, "parallelism": "single GPU"},
        "assembly_ms_mean": round(1e3 * float(np.mean([r[1] for r in rec])), 3),
        "solve_ms_mean": round(1e3 * float(np.mean([r[2] for r in rec])), 3),
        "gmres_iterations_warm_mean": round(float(np.mean([r[3] for r in rec])), 2),
        "gmres_iterations_cold_last": cold.iterations,
        "gmres_cold_last_s": round(t_cold, 4),
        "gmres_final_residual_max": float(f"{max(r[4] for r in rec):.3e}"),
        "roofline": {"kernel": "k_mv32 (GMRES matvec, float32 [F,F])", "bound": "hbm",
                     "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": None,
                     "avg_launch_ms": round(mv_ms, 4), "bytes_per_launch": 4 * nF * nF},
    }
    if not args.no_cpu:
        log("CPU baseline (oracle sample)")
        out["cpu_baseline"] = cpu_baseline_c5(nF, chi)
    print(json.dumps(out), flush=True)


def run_field(args, dev):
    """Field radiation (post-processing of a MoM solution): scattered H of the
    vertex currents of the 30 720-RWG icosphere (10 242 vertices, ka = 10) on
    the demo's 256 x 256 sampling plane (65 536 points), complex64 currents
    (seeded random values: the cost does not depend on them).  One step = one
    radiate_H call over all 6.7e8 (point, vertex) pairs.  Roofline: FP32
    VALU, 61 flops per pair (offset 3, R^2 5, R 1, phase 1, kR 1, 1/R^3 2,
    complex factor 6, cross product 18, complex multiply-add 24) plus three
    transcendentals (rsqrt, sin, cos)."""
    import numpy as np
    import torch

    from mesh_bem_amd.geometry import icosphere
    from mesh_bem_amd.radiation import radiate_H, xz_grid

    subdiv, k, res = (5, 10.0, 256) if args.workload == "field" else (3, 4.0, 64)
    Vn, Fn = icosphere(subdiv)
    rng = np.random.Generator(np.random.PCG64(11))
    J = rng.normal(size=(len(Vn), 3)) + 1j * rng.normal(size=(len(Vn), 3))
    v = torch.tensor(Vn, dtype=torch.float64, device=dev)
    f = torch.tensor(Fn, dtype=torch.int64, device=dev)
    a = torch.zeros(len(Vn), dtype=torch.float64, device=dev)
    e1, e2 = v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 1]]
    a.index_add_(0, f.reshape(-1), (0.5 * torch.linalg.cross(e1, e2).norm(dim=1) / 3.0).repeat_interleave(3))
    pts = xz_grid(res, device=dev)
    P32, V32, A32 = pts.float(), v.float(), a.float()
    J32 = torch.tensor(J, dtype=torch.complex64, device=dev)
    n_pairs = pts.shape[0] * len(Vn)
    log(f"workload {args.workload}: {pts.shape[0]} points x {len(Vn)} vertices")
    for _ in range(args.warmup):
        radiate_H(k, P32, V32, J32, A32)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        H = radiate_H(k, P32, V32, J32, A32)
    ev1.record()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    ev_ms = ev0.elapsed_time(ev1) / args.steps
    J64 = J32.to(torch.complex128)
    radiate_H(k, pts, v, J64, a)
    torch.cuda.synchronize()
    t64 = time.perf_counter()
    H64 = radiate_H(k, pts, v, J64, a)
    torch.cuda.synchronize()
    ms64 = (time.perf_counter() - t64) * 1e3
    err32 = float((H.to(torch.complex128) - H64).norm() / H64.norm())
    flops = 61.0 * n_pairs
    tflops = flops / (ev_ms * 1e-3) / 1e12
    out = {
        "metric": "field radiation (scattered H from vertex currents on a sampling plane); Gpairs/s",
        "value": round(1e-9 * n_pairs / (ms * 1e-3), 3),
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"icosphere subdiv {subdiv} ({len(Vn)} vertices), ka={k}, {res}x{res} xz plane "
                               f"in [-2,2]^2, random complex64 vertex currents", "parallelism": "single GPU"},
        "fp64_ms": round(ms64, 3),
        "fp32_relerr_vs_fp64": float(f"{err32:.3e}"),
        "roofline": {"kernel": "k_radiate_h<float> + k_radiate_sum (61 flops + 3 transcendentals per pair)",
                     "bound": "valu", "achieved": round(tflops, 2), "peak": PEAK_FP32_MFMA_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(tflops / PEAK_FP32_MFMA_TFLOPS, 4), "traffic": None,
                     "avg_launch_ms": round(ev_ms, 4), "flops_per_launch": flops,
                     "transcendentals_per_s": float(f"{3 * n_pairs / (ev_ms * 1e-3):.3e}")},
    }
    if not args.no_cpu:
        log("CPU baseline (oracle sample)")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        sample = 2048 if args.workload == "field" else 256
        Pc = pts[:sample].cpu().numpy()
        tc = time.perf_counter()
        O.radiate_H(k, Pc, Vn, J, a.cpu().numpy())
        dt = time.perf_counter() - tc
        out["cpu_baseline"] = {"value": round(1e-9 * sample * len(Vn) / dt, 6), "unit": "Gpairs/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle.radiate_H (numpy, FP64) on the first {sample} grid points x all "
                                         f"{len(Vn)} vertices, {dt:.2f} s"}
    print(json.dumps(out), flush=True)


def run_remesh(args, dev):
    """Remeshing (remesh_and_transfer_velocity): the config-5 droplet geometry,
    geodesic frequency 45 (40 500 faces, 20 252 vertices), one Loop pass on
    the host (-> 162 000 faces, 81 002 vertices) and the nearest-vertex
    transfer of one seeded velocity field [1, 20 252, 3] on the GPU, FP32.
    One step = one transfer (1.64e9 (new, old) vertex pairs, the transfer
    runs on the subdivided mesh); the host subdivision and the quadric
    decimation to 90 % of the faces are timed beside it.  Roofline: FP32 VALU, 8 flops per pair
    (offset 3, squared length 5) plus the compare-select."""
    import numpy as np
    import torch

    from mesh_bem_amd.geometry import geodesic_sphere
    from mesh_bem_amd.remesh import decimate_quadric, loop_subdivide, transfer_velocity_nearest

    freq = 45 if args.workload == "remesh" else 9
    Vn, Fn = geodesic_sphere(freq)
    rng = np.random.Generator(np.random.PCG64(5))
    vel = torch.tensor(rng.normal(size=(1, len(Vn), 3)), dtype=torch.float32, device=dev)
    tc = time.perf_counter()
    nv, nf = loop_subdivide(Vn, Fn)
    t_sub = time.perf_counter() - tc
    tc = time.perf_counter()
    dv, df = decimate_quadric(nv, nf, int(len(nf) * 0.9))
    t_dec = time.perf_counter() - tc
    P = torch.tensor(nv, dtype=torch.float32, device=dev)
    Q = torch.tensor(np.asarray(Vn), dtype=torch.float32, device=dev)
    n_pairs = P.shape[0] * Q.shape[0]
    log(f"workload {args.workload}: {len(Fn)} -> {len(nf)} faces, {P.shape[0]} new x {Q.shape[0]} old vertices")
    for _ in range(args.warmup):
        transfer_velocity_nearest(P, Q, vel)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        out, nearest = transfer_velocity_nearest(P, Q, vel)
    ev1.record()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    ev_ms = ev0.elapsed_time(ev1) / args.steps
    flops = 8.0 * n_pairs
    tflops = flops / (ev_ms * 1e-3) / 1e12
    res = {
        "metric": "remesh velocity transfer (nearest old vertex of every Loop-subdivided vertex); Gpairs/s",
        "value": round(1e-9 * n_pairs / (ms * 1e-3), 3),
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"geodesic frequency {freq} sphere ({len(Fn)} faces) -> one Loop pass "
                               f"({len(nf)} faces), one seeded velocity field", "parallelism": "single GPU"},
        "loop_subdivide_host_ms": round(t_sub * 1e3, 2),
        "decimate_host_ms": round(t_dec * 1e3, 2),
        "decimated_faces": int(len(df)),
        "roofline": {"kernel": "k_nearest<float> + k_nearest_finish (8 flops + compare per pair)",
                     "bound": "valu", "achieved": round(tflops, 2), "peak": PEAK_FP32_MFMA_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(tflops / PEAK_FP32_MFMA_TFLOPS, 4), "traffic": None,
                     "avg_launch_ms": round(ev_ms, 4), "flops_per_launch": flops},
    }
    if not args.no_cpu:
        log("CPU baseline (numpy brute force sample)")
        sample = 2048 if args.workload == "remesh" else 256
        pc = nv[:sample]
        qc = np.asarray(Vn, dtype=np.float64)
        tcb = time.perf_counter()
        d = ((pc[:, None, :] - qc[None, :, :]) ** 2).sum(-1)
        d.argmin(axis=1)
        dt = time.perf_counter() - tcb
        res["cpu_baseline"] = {"value": round(1e-9 * sample * len(qc) / dt, 6), "unit": "Gpairs/s", "cores": 1,
                               "kind": "port",
                               "sample": f"numpy brute-force argmin (FP64) for the first {sample} new vertices x "
                                         f"all {len(qc)} old vertices, {dt:.2f} s"}
    print(json.dumps(res), flush=True)


def run_hd(args, dev):
    """Ferrofluid step, surface Helmholtz decomposition
    (MeshHelmholtzDecomposition.solve): noisy ellipsoid of geodesic frequency
    45 (40 500 faces, 20 252 vertices, the config-5 geometry), gaussQR = 4
    (16 points per face), one seeded random velocity field, FP32.  One step =
    one solve(): the dense sum over 20 252 targets x 648 000 source points
    (1.3e10 pairs) plus the O(V) normal projection.  Roofline: FP32 VALU, 29
    flops per pair (offset 3, R^2 5, 1/R^3 scaled 3, (n.u) d + (n x u) x d 12,
    accumulate 6) and one rsqrt."""
    import numpy as np
    import torch

    from mesh_bem_amd.ferrofluid import MeshHelmholtzDecomposition
    from mesh_bem_amd.geometry import noisy_ellipsoid

    freq, Q = (45, 4) if args.workload == "hd" else (15, 4)
    Vn, Fn, c = noisy_ellipsoid(0, freq)
    rng = np.random.Generator(np.random.PCG64(17))
    vel = rng.normal(size=(1, len(Vn), 3))
    v = torch.tensor(Vn, dtype=torch.float32, device=dev)
    f = torch.tensor(Fn, dtype=torch.int32, device=dev)
    u = torch.tensor(vel, dtype=torch.float32, device=dev)
    hd = MeshHelmholtzDecomposition(gaussQR=Q, device=dev, dtype=torch.float32)
    n_pairs = len(Vn) * len(Fn) * Q * Q
    log(f"workload {args.workload}: {len(Vn)} vertices x {len(Fn)} faces x {Q * Q} points")
    for _ in range(args.warmup):
        hd.solve(v, f, u)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        out = hd.solve(v, f, u)
    ev1.record()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # the pair sum alone, HIP events on the launch stream
    from mesh_bem_amd.ferrofluid import decomposition_velocity

    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        decomposition_velocity(v, f, u, Q)
    e1.record()
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1) / args.steps
    out64 = MeshHelmholtzDecomposition(gaussQR=Q, device=dev, dtype=torch.float64).solve(v.double(), f, u.double())
    err32 = float((out.double() - out64).norm() / out64.norm())
    flops = 29.0 * n_pairs
    tflops = flops / (k_ms * 1e-3) / 1e12
    res = {
        "metric": "ferrofluid step: surface Helmholtz decomposition (vertex x face-quadrature pair sum); Gpairs/s",
        "value": round(1e-9 * n_pairs / (ms * 1e-3), 3),
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"noisy ellipsoid geodesic frequency {freq} ({len(Fn)} faces, {len(Vn)} vertices), "
                               f"gaussQR={Q}, one random velocity field", "parallelism": "single GPU"},
        "fp32_relerr_vs_fp64": float(f"{err32:.3e}"),
        "roofline": {"kernel": "k_hd_sum<float> + k_hd_slices (29 flops + 1 rsqrt per pair)", "bound": "valu",
                     "achieved": round(tflops, 2), "peak": PEAK_FP32_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / PEAK_FP32_MFMA_TFLOPS, 4), "traffic": None,
                     "avg_launch_ms": round(k_ms, 4), "flops_per_launch": flops},
    }
    if not args.no_cpu:
        log("CPU baseline (oracle sample)")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        sample = 64
        tc = time.perf_counter()
        O.helmholtz_decomposition_u(Vn, Fn, vel, Q, targets=Vn[:sample], block=2)
        dt = time.perf_counter() - tc
        res["cpu_baseline"] = {"value": round(1e-9 * sample * len(Fn) * Q * Q / dt, 6), "unit": "Gpairs/s",
                               "cores": 1, "kind": "port",
                               "sample": f"oracle.helmholtz_decomposition_u (numpy, FP64) at the first {sample} "
                                         f"vertices x all {len(Fn) * Q * Q} source points, {dt:.2f} s"}
    print(json.dumps(res), flush=True)


def run_bem2d(args, dev):
    """2-D BEM Dirichlet solve (MeshBEMSolver2d.solve, the reference's
    src/mesh_BEM/mesh_BEM_solver_2d.py): Laplace, LINEAR shape functions,
    gaussQR = 8 (the 2-D demo's settings), FP32, on a closed ellipse of
    semi-axes 0.25 and 0.15 (logarithmic capacity 0.2: the 2-D log single layer
    is singular on curves of capacity 1) with 16 384 segments.  One step = single layer + double layer assembly
    ([F, F, 2, 2] each, F^2 Q^2 = 1.7e10 point pairs per layer), packing to
    [V, V], the right-hand side 1/2 M g + K^T g and the dense LU solve of the
    [V, V] system.  Value: segment pairs per second of the whole step
    (F^2 / t).  Roofline: the single-layer kernel, FP32 VALU, 29 flops + 1 log
    per point pair (offset 8, d^2 3, log scale 1, weight 3, LINEAR shape
    products 6, four accumulations 8)."""
    import numpy as np
    import torch

    from mesh_bem_amd.bem2d import MeshBEMSolver2d, ellipse_curve, layer2d
    from mesh_bem_amd.types import BEMType

    n = 16384 if args.workload == "bem2d" else 2048
    Q = 8
    Vn, Fn = ellipse_curve(n, 0.25, 0.15)
    g = Vn[:, 0] ** 2 - Vn[:, 1] ** 2
    v = torch.tensor(Vn, dtype=torch.float32, device=dev)
    f = torch.tensor(Fn, dtype=torch.int32, device=dev)
    gt = torch.tensor(g, dtype=torch.float32, device=dev).reshape(1, n, 1)
    flags = torch.zeros((1, n, 1), dtype=torch.uint8, device=dev)
    solver = MeshBEMSolver2d(Q, 1, int(BEMType.LAPLACE), 1.0, None, dtype=torch.float32, device=dev)
    log(f"workload {args.workload}: {n} segments, gaussQR={Q}")
    for _ in range(args.warmup):
        solver.solve(v, f, gt, torch.zeros_like(gt), flags)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        neu, _ = solver.solve(v, f, gt, torch.zeros_like(gt), flags)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        layer2d(v, f, Q, 1, 1)
    e1.record()
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1) / args.steps
    solver64 = MeshBEMSolver2d(Q, 1, int(BEMType.LAPLACE), 1.0, None, dtype=torch.float64, device=dev)
    neu64, _ = solver64.solve(v.double(), f, gt.double(), torch.zeros_like(gt).double(), flags)
    err32 = float((neu.double() - neu64).norm() / neu64.norm())
    pts = float(n) * n * Q * Q
    flops = 29.0 * pts
    tflops = flops / (k_ms * 1e-3) / 1e12
    res = {
        "metric": "2-D BEM Dirichlet solve (Laplace single + double layer, LINEAR, gaussQR 8); Mpairs/s",
        "value": round(1e-6 * n * n / (ms * 1e-3), 3),
        "unit": "Mpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"ellipse of semi-axes 0.25, 0.15 with {n} segments, Dirichlet data x^2 - y^2, "
                               f"gaussQR={Q}, LINEAR",
                   "segments": n, "parallelism": "single GPU"},
        "fp32_relerr_vs_fp64": float(f"{err32:.3e}"),
        "roofline": {"kernel": "k_bem2d_layer<float, 0> (single layer, 29 flops + 1 log per point pair)",
                     "bound": "valu", "achieved": round(tflops, 2), "peak": PEAK_FP32_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / PEAK_FP32_MFMA_TFLOPS, 4), "traffic": None,
                     "avg_launch_ms": round(k_ms, 4), "flops_per_launch": flops,
                     "point_pairs_per_s": round(pts / (k_ms * 1e-3), 1)},
    }
    if not args.no_cpu:
        import scipy.linalg

        log("CPU baseline (oracle sample)")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        ns = 512
        Vs, Fs = ellipse_curve(ns, 0.25, 0.15)
        gp, gw = O.gl_ref(Q)
        tc = time.perf_counter()
        O.bem2d_layer(Vs, Fs, gp, gw, 1, 1)
        O.bem2d_layer(Vs, Fs, gp, gw, 2, 1)
        t_asm = time.perf_counter() - tc
        nl = 2048
        rng = np.random.default_rng(0)
        A = rng.standard_normal((nl, nl)) + nl * np.eye(nl)
        tc = time.perf_counter()
        scipy.linalg.solve(A, rng.standard_normal(nl), check_finite=False)
        t_lu = time.perf_counter() - tc
        t_full = t_asm * (n / ns) ** 2 + t_lu * (n / nl) ** 3
        res["cpu_baseline"] = {"value": round(1e-6 * n * n / t_full, 4), "unit": "Mpairs/s",
                               "cores": len(os.sched_getaffinity(0)), "kind": "port",
                               "sample": f"oracle.bem2d_layer (numpy, FP64) single + double layer at {ns} segments "
                                         f"({t_asm:.2f} s, scaled by (F/{ns})^2) + scipy dgesv at {nl} ({t_lu:.2f} s, "
                                         f"scaled by (F/{nl})^3); extrapolated step {t_full:.1f} s"}
    print(json.dumps(res), flush=True)


def run_edges(args, dev):
    """Mesh topology: the unique-edge table of get_edge_unique
    (src/utils/mesh_utils.py:921-977) on the GPU (mba_edge_unique: keys,
    hipCUB radix sort of (key, slot), head flags, scan, fill) for an icosphere of
    subdivision 8 (1 310 720 faces).  Value: faces per second.  Roofline: HBM;
    algorithmic bytes 48 per face (faces read 12 B; edges, face-edge indices
    and adjacency written, 12 B each per face)."""
    import torch

    from mesh_bem_amd.geometry import get_edge_unique, icosphere

    sub = 8 if args.workload == "edges" else 5
    Vn, Fn = icosphere(sub)
    v = torch.tensor(Vn, device=dev)
    f = torch.tensor(Fn, dtype=torch.int32, device=dev)
    log(f"workload {args.workload}: icosphere subdiv {sub}, {len(Fn)} faces")
    for _ in range(args.warmup):
        get_edge_unique(v, f)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e, fe, adj = get_edge_unique(v, f)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    e_h, fe_h, adj_h = get_edge_unique(v.cpu(), f.cpu())
    exact = bool(torch.equal(e.cpu(), e_h) and torch.equal(fe.cpu(), fe_h) and torch.equal(adj.cpu(), adj_h))
    nbytes = 48.0 * len(Fn)
    gbs = nbytes / (ms * 1e-3) / 1e9
    res = {
        "metric": "unique-edge table (get_edge_unique) on the GPU; Mfaces/s",
        "value": round(1e-6 * len(Fn) / (ms * 1e-3), 2),
        "unit": "Mfaces/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32/int64", "data": "synthetic",
        "config": {"workload": f"icosphere subdivision {sub} ({len(Fn)} faces, {len(Vn)} vertices)",
                   "parallelism": "single GPU"},
        "bit_exact_vs_host": exact,
        "roofline": {"kernel": "mba_edge_unique pipeline (keys, radix sort, heads, scan, fill; host-timed step)",
                     "bound": "hbm", "achieved": round(gbs, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(gbs / PEAK_HBM_GBS, 5), "traffic": None, "bytes_per_step": nbytes},
    }
    if not args.no_cpu:
        tc = time.perf_counter()
        get_edge_unique(v.cpu(), f.cpu())
        dt = time.perf_counter() - tc
        res["cpu_baseline"] = {"value": round(1e-6 * len(Fn) / dt, 3), "unit": "Mfaces/s", "cores": 1, "kind": "port",
                               "sample": f"host numpy path of get_edge_unique (np.unique + stable argsort) on the "
                                         f"same mesh, {dt:.2f} s"}
    print(json.dumps(res), flush=True)


def run_pressure(args, dev):
    """Ferrofluid pressure step (MeshPressure.solve,
    src/mesh_ferrofluid/mesh_pressure.py:54-163) with the droplet demo's
    parameters (demo_3d_mesh_droplet.py:30-48: rho = 1300, g = 9.81, sigma =
    0.02, dt = 5e-4, gaussQR = 8, LINEAR, FP32, every vertex DIRICHLET) on a
    sphere of the demo's radius L / 2 (L = 5 * 2 pi / sqrt(rho g / sigma)),
    icosphere subdivision 4 (5 120 faces, 2 562 vertices), random velocity.
    One step = boundary data kernel, single + double layer of the reference's
    relation-typed rule, packing, RHS, dense solve, gradient / velocity update.
    Value: face pairs per second (F^2 / t).  Roofline: the single-layer
    kernel, FP32 VALU, 73 flops + sqrt + reciprocal per table row (LINEAR),
    Q^4 = 4096 rows for separate face pairs (counted for all F^2 pairs)."""
    import numpy as np
    import torch

    from mesh_bem_amd.ferrofluid import MeshPressure
    from mesh_bem_amd.geometry import icosphere
    from mesh_bem_amd.laplace import layer_ref

    sub = 4 if args.workload == "pressure" else 2
    Q = 8
    rho, g, sigma, dt = 1300.0, 9.81, 0.02, 5e-4
    L = 5 * (2 * math.pi / math.sqrt(rho * g / sigma))
    Vn, Fn = icosphere(sub)
    Vn = Vn * (L / 2)
    rng = np.random.default_rng(5)
    vel = torch.tensor(rng.standard_normal((1, len(Vn), 3)) * 1e-3, dtype=torch.float32, device=dev)
    v = torch.tensor(Vn, dtype=torch.float32, device=dev)
    f = torch.tensor(Fn, dtype=torch.int32, device=dev)
    flags = torch.zeros((1, len(Vn), 1), dtype=torch.uint8, device=dev)
    mp = MeshPressure(rho, g, sigma, 0.5 * math.pi, Q, 1, None, device=dev, dtype=torch.float32)
    log(f"workload {args.workload}: icosphere {sub} ({len(Fn)} faces), radius {L / 2:.4g}, gaussQR {Q}")
    for _ in range(args.warmup):
        mp.solve(v, f, dt, vel, flags)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = mp.solve(v, f, dt, vel, flags)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    layer_ref(v, f, Q, 1, 1)
    e1.record()
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1)
    nF = len(Fn)
    flops = 73.0 * Q ** 4 * nF * nF
    tflops = flops / (k_ms * 1e-3) / 1e12
    mp64 = MeshPressure(rho, g, sigma, 0.5 * math.pi, Q, 1, None, device=dev, dtype=torch.float64)
    out64 = mp64.solve(v.double(), f, dt, vel.double(), flags)
    err32 = float((out.double() - out64).norm() / out64.norm())
    res = {
        "metric": "ferrofluid pressure step (MeshPressure.solve: curvature + Laplace BEM + gradient); Mpairs/s",
        "value": round(1e-6 * nF * nF / (ms * 1e-3), 3),
        "unit": "Mpairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"droplet demo parameters, icosphere subdivision {sub} ({nF} faces) of radius "
                               f"{L / 2:.4g}, gaussQR={Q}, LINEAR, all vertices Dirichlet", "parallelism": "single GPU"},
        "fp32_relerr_vs_fp64": float(f"{err32:.3e}"),
        "roofline": {"kernel": "k_laplace_ref<float> (single layer, 73 flops + sqrt + rcp per table row)",
                     "bound": "valu", "achieved": round(tflops, 2), "peak": PEAK_FP32_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / PEAK_FP32_MFMA_TFLOPS, 4), "traffic": None,
                     "avg_launch_ms": round(k_ms, 4), "flops_per_launch": flops},
    }
    if not args.no_cpu:
        log("CPU baseline (oracle sample)")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        Vs, Fs = icosphere(2)
        Vs = Vs * (L / 2)
        tc = time.perf_counter()
        O.pressure_step(Vs, Fs, np.zeros_like(Vs), dt, rho, g, sigma, Q)
        t_s = time.perf_counter() - tc
        t_full = t_s * (nF / len(Fs)) ** 2
        res["cpu_baseline"] = {"value": round(1e-6 * nF * nF / t_full, 4), "unit": "Mpairs/s",
                               "cores": len(os.sched_getaffinity(0)), "kind": "port",
                               "sample": f"oracle.pressure_step (C/OpenMP layers + numpy + scipy, FP64) at "
                                         f"{len(Fs)} faces ({t_s:.2f} s), scaled by (F/{len(Fs)})^2 (the layers "
                                         f"dominate); extrapolated step {t_full:.1f} s"}
    print(json.dumps(res), flush=True)


def run_c3(args, dev):
    """Config 3: 64 seeded star meshes on geodesic spheres of frequency 5..11
    (750-3630 RWG), 128 plane waves each, ka ~ U(1, 8).  Step: FP64 assembly of
    all meshes (one launch for the far field of every mesh, one for all near
    pairs, a gather per mesh), the RHS, the FP32 LUs of all meshes in lock step
    (shared cooperative panel launches, autotuned GEMM tiles), multi-RHS
    solves and FP64-GEMM refinement to an estimated relative error of 1e-8."""
    import numpy as np
    import torch

    from mesh_bem_amd import _lib
    from mesh_bem_amd.batched import BatchedMoM, BatchLuTuner, GemmTuner, star_batch

    n_mesh, n_dir = (64, 128) if args.workload == "c3" else (8, 16)
    geoms, ks, dirs, pols = star_batch(n_mesh, seed0=0, n_dirs=n_dir)
    bm = BatchedMoM.build(geoms, ks, device=dev)
    Es = [m.n_edges for m in bm.meshes]
    tuner = GemmTuner()
    for E in sorted(set(Es)):
        tuner.tune(E)
    lu_tuner = BatchLuTuner() if args.batched_lu == "persistent" else None
    log(f"workload {args.workload}: {n_mesh} meshes, E in [{min(Es)}, {max(Es)}], {n_dir} waves; tuned "
        f"{tuner.best}")

    def step():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Zs = bm.assemble(_lib.MBA_F64)
        Vs = bm.rhs(dirs, pols)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = bm.solve(Zs, Vs, tuner=tuner, tol=1e-8, persistent=args.batched_lu == "persistent", lu_tuner=lu_tuner,
                       solve_persistent=args.batched_solve == "persistent")
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return res, t1 - t0, t2 - t1, Zs, Vs

    for i in range(args.warmup):
        step()
    ts = []
    for i in range(args.steps):
        res, ta, tsol, Zs, Vs = step()
        ts.append((ta, tsol))
    t_asm = float(np.mean([t[0] for t in ts]))
    t_sol = float(np.mean([t[1] for t in ts]))
    # accuracy vs the plain FP64 solve on the GPU (torch.linalg.solve) for 4 meshes
    errs = []
    for Z, V, J in list(zip(Zs, Vs, res.J))[:4]:
        ref = torch.linalg.solve(Z, V)
        errs.append(float(torch.linalg.vector_norm(J - ref) / torch.linalg.vector_norm(ref)))
    tot = t_asm + t_sol
    pairs = sum(len(g[1]) ** 2 for g in geoms)
    out = {
        "metric": "C3 batched multi-mesh multi-RHS mixed-precision solve; right-hand sides per second",
        "value": round(n_mesh * n_dir / tot, 2),
        "unit": "RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * tot, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c64 factors (" + ("one persistent batched-LU launch" if args.batched_lu == "persistent"
                                     else "lock-step batched LU") +
                  (", tuned trailing update" if lu_tuner is not None else ", FP32 MFMA 3M trailing updates") +
                  ") + c128 assembly/residual"),
        "batched_lu": args.batched_lu,
        "batched_solve": args.batched_solve,
        "value_definition": "right-hand sides per second: meshes x waves / (FP64 assembly + RHS + FP32 LUs "
                            "+ solves + FP64 refinement) per step",
        "data": "synthetic",
        "config": {"workload": f"{n_mesh} star meshes (geodesic frequency 5-11, RWG {min(Es)}-{max(Es)}), "
                               f"{n_dir} waves each, ka~U(1,8), EFIE 7-point rule", "parallelism": "single GPU"},
        "assembly_rhs_s": round(t_asm, 4),
        "assembly_Mpairs_s": round(1e-6 * pairs / t_asm, 1),
        "lu_s": round(res.times["lu_s"], 4),
        "solve_refine_s": round(res.times["solve_refine_s"], 4),
        "refine_iterations_max": int(max(res.iterations)),
        "all_converged": bool(all(res.converged)),
        "relerr_vs_fp64_max": float(f"{max(errs):.3e}"),
        "tuned_gemm_variant_per_bucket": {str(k): v for k, v in tuner.best.items()},
    }
    if lu_tuner is not None and lu_tuner.best:
        kk = next(iter(lu_tuner.best))
        out["tuned_batched_lu"] = {"super_panels": lu_tuner.best[kk][0],
                                   "trailing": ["fp32_3m", "split_fp16"][lu_tuner.best[kk][1]],
                                   "candidates_ms": {f"{c[0]}x32/{['3m', 'split'][c[1]]}": round(t, 2)
                                                     for c, t in zip(lu_tuner.CANDIDATES, lu_tuner.timings.get(kk, []))}}
    print(json.dumps(out), flush=True)


def run_c4(args, dev):
    """Config 4 (scaled down): learned warm start.  Training data from the
    solver: star meshes at geodesic frequency 8 (1920 RWG), ka ~ U(1, 4),
    128 plane waves each (160 meshes: 20 480 triples).  The network (6
    processor layers, hidden 128) trains on the GPU with the HIP kernels;
    held-out meshes are solved by GMRES (restart 50, 1e-6) from zero and from
    the scaled prediction, and by LU."""
    import numpy as np
    import torch

    from mesh_bem_amd import mgn as M
    from mesh_bem_amd.dense import LUFactor
    from mesh_bem_amd.gmres import gmres

    full = args.workload == "c4"
    n_mesh, n_dir, freq = (160, 128, 8) if full else (24, 32, 6)
    t0 = time.perf_counter()
    mds, smp = M.generate(n_mesh, freq, n_dir, seed0=1000, device=dev, keep_system=args.c4_res_weight > 0)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    log(f"generated {n_mesh * n_dir} triples in {t_gen:.1f} s")
    net = M.MeshGraphNet(device=dev, seed=0)
    t0 = time.perf_counter()
    hist = M.train(net, mds, smp, epochs=args.epochs, batch=16, lr=2e-3, log=log, res_weight=args.c4_res_weight)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    n_eval = 8 if full else 4
    # held-out meshes with all their waves (the span start uses every prediction
    # of the mesh); GMRES runs for the first n_gm waves
    n_gm = 4
    tm, ts = M.generate(n_eval, freq, n_dir, seed0=50000, device=dev, keep_system=True)
    rec = {"cold_it": [], "warm_it": [], "cold_s": [], "warm_s": [], "lu_s": [], "jerr": [], "r0": [],
           "span_it": [], "span_r0": []}
    # ablation: iterations from x0 = J + s (J0 - J), the network's own error
    # direction scaled by s (how accurate a prediction must be to pay)
    scales = [1.0, 0.3, 0.1, 0.03, 0.01, 1e-3, 1e-4]
    abl = {sc: [] for sc in scales}
    for md, s in zip(tm, ts):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        lu = LUFactor(s.Z32, overwrite=False)
        lu.solve(s.V32[:, :n_gm])
        torch.cuda.synchronize()
        rec["lu_s"].append((time.perf_counter() - ta) / n_gm)
        J0 = M.predict_currents(net, md, s.x)  # [E, n_dir]
        # span start: least-squares combination of all n_dir predictions,
        # min_c |b - Z P c| (normal equations in FP64)
        P = J0.to(torch.complex128)
        ZP = s.Z32.to(torch.complex128) @ P
        G = ZP.conj().T @ ZP
        for d in range(n_gm):
            b = s.V32[:, d]
            J = torch.complex(s.J[d][:, 0], s.J[d][:, 1])
            rec["jerr"].append(float((J0[:, d] - J).norm() / J.norm()))
            torch.cuda.synchronize()
            ta = time.perf_counter()
            _, ic = gmres(s.Z32, b, restart=50, tol=1e-6, maxiter=5000)
            torch.cuda.synchronize()
            tb = time.perf_counter()
            x0 = M.scaled_guess(s.Z32, b, J0[:, d])
            _, iw = gmres(s.Z32, b, x0=x0, restart=50, tol=1e-6, maxiter=5000)
            torch.cuda.synchronize()
            tc = time.perf_counter()
            r0 = b.to(torch.complex128) - (s.Z32.to(torch.complex128) @ x0)
            rec["r0"].append(float(r0.norm() / b.norm()))
            rec["cold_it"].append(ic.iterations)
            rec["warm_it"].append(iw.iterations)
            rec["cold_s"].append(tb - ta)
            rec["warm_s"].append(tc - tb)
            b64 = b.to(torch.complex128)
            c = torch.linalg.solve(G, ZP.conj().T @ b64)
            xs = P @ c
            rec["span_r0"].append(float((b64 - ZP @ c).norm() / b64.norm()))
            _, isp = gmres(s.Z32, b, x0=xs, restart=50, tol=1e-6, maxiter=5000)
            rec["span_it"].append(isp.iterations)
            Jd = J.to(torch.complex128)
            for sc in scales:
                _, ia = gmres(s.Z32, b, x0=Jd + sc * (J0[:, d].to(torch.complex128) - Jd), restart=50, tol=1e-6,
                              maxiter=5000)
                abl[sc].append(ia.iterations)
    # MFMA roofline of the network's largest GEMM (edge MLP, first layer), HIP events
    sel = [(i, d) for i in range(min(16, len(smp))) for d in range(1)]
    g = M.batch_graph([mds[i] for i, _ in sel], [smp[i].x[d] for i, d in sel], dev)
    H = net.H
    W1 = net.W("pe0_1")
    Ne = g.n_edges
    Y = torch.empty((Ne, H), dtype=torch.float32, device=dev)
    Z = torch.empty_like(Y)
    hbuf = torch.randn((g.n_nodes, H), device=dev)
    ebuf = torch.randn((Ne, H), device=dev)
    segs = [(hbuf, g.src, H), (hbuf, g.dst, H), (ebuf, None, H)]
    for _ in range(3):
        M.gemm(segs, True, 0, W1, 0, Ne, H, 3 * H + 1, Y, Z=Z, act=1)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        M.gemm(segs, True, 0, W1, 0, Ne, H, 3 * H + 1, Y, Z=Z, act=1)
    e1.record()
    torch.cuda.synchronize()
    g_ms = e0.elapsed_time(e1) / 20
    tf = 2.0 * Ne * H * (3 * H + 1) / (g_ms * 1e-3) / 1e12
    mean = lambda k: float(np.mean(rec[k]))
    samples_s = args.epochs * n_mesh * n_dir / t_train
    out = {
        "metric": "C4 learned warm start: GMRES iterations (warm / cold) and training throughput; samples/s",
        "value": round(samples_s, 2),
        "unit": "training samples/s",
        "n_gpus": 1,
        "steps": args.epochs,
        "warmup": 0,
        "ms_per_step": round(1e3 * t_train / max(1, args.epochs), 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 GEMM operands, fp32 accumulate/activations",
        "data": "synthetic (solver-generated)",
        "config": {"workload": f"{n_mesh} star meshes (geodesic frequency {freq}, {mds[0].E} RWG) x {n_dir} waves, "
                               f"ka~U(1,4); MGN 6 layers hidden 128 ({net.n_params} params); GMRES restart 50 tol 1e-6",
                   "parallelism": "single GPU"},
        "data_generation_s": round(t_gen, 2),
        "train_s": round(t_train, 2),
        "loss_history": [round(h, 4) for h in hist],
        "heldout_J_relerr_mean": round(mean("jerr"), 4),
        "heldout_r0_scaled_mean": round(mean("r0"), 4),
        "gmres_iterations_cold_mean": round(mean("cold_it"), 1),
        "gmres_iterations_warm_mean": round(mean("warm_it"), 1),
        "gmres_cold_s_mean": round(mean("cold_s"), 4),
        "gmres_warm_s_mean": round(mean("warm_s"), 4),
        "lu_s_per_rhs_mean": round(mean("lu_s"), 5),
        "span_start_r0_mean": round(mean("span_r0"), 4),
        "gmres_iterations_span_start_mean": round(mean("span_it"), 1),
        "ablation_iterations_vs_error_scale": {f"{sc:g}": round(float(np.mean(v)), 1) for sc, v in abl.items()},
        "train_res_weight": args.c4_res_weight,
        "roofline": {"kernel": "k_mgn_gemm (edge MLP layer 1, gathered [h_src, h_dst, e, 1] x W, BF16 MFMA)",
                     "bound": "mfma", "achieved": round(tf, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tf / 2500.0, 5), "traffic": None, "avg_launch_ms": round(g_ms, 4)},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "small", "c3", "c3small", "c4", "c4small", "c5",
                                                            "c5small", "field", "fieldsmall", "hd", "hdsmall", "bem2d", "bem2dsmall", "edges",
                                                            "edgessmall", "pressure", "pressuresmall", "remesh", "remeshsmall"])
    ap.add_argument("--epochs", type=int, default=4, help="training epochs (c4)")
    ap.add_argument("--c4-res-weight", type=float, default=0.0,
                    help="weight of the residual term |b - Z J_pred|^2 / |b|^2 in the training loss (c4)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison")
    ap.add_argument("--no-extra", action="store_true", help="only the timed steps")
    ap.add_argument("--lu-update", default="split_f16", choices=["fp32", "split_f16"],
                    help="trailing-update product of the complex64 LU (c2/c1/small)")
    ap.add_argument("--batched-lu", default="persistent", choices=["persistent", "lockstep"],
                    help="batched LU of c3: one persistent launch or the lock-step launches")
    ap.add_argument("--batched-solve", default="persistent", choices=["walk", "persistent"],
                    help="triangular solves of c3: per-job block walks (two launches) or one persistent "
                         "dataflow launch")
    ap.add_argument("--lu-panel", default="partial", choices=["partial", "tournament"],
                    help="panel pivoting of the complex64 LU (c2/c1/small)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE={world}); "
              "a multi-GPU run would shard the batch of meshes, one solve per GPU", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from mesh_bem_amd import _lib
    from mesh_bem_amd.dense import LUFactor, refine_solve, set_lu_panel, set_lu_update, to_c64
    from mesh_bem_amd.geometry import icosphere
    from mesh_bem_amd.mom import MomMesh, assemble_ns, rhs_ns

    dev = torch.device("cuda:0")
    if args.workload in ("c5", "c5small"):
        return run_c5(args, dev)
    if args.workload in ("c4", "c4small"):
        return run_c4(args, dev)
    if args.workload in ("c3", "c3small"):
        return run_c3(args, dev)
    if args.workload in ("field", "fieldsmall"):
        return run_field(args, dev)
    if args.workload in ("hd", "hdsmall"):
        return run_hd(args, dev)
    if args.workload in ("bem2d", "bem2dsmall"):
        return run_bem2d(args, dev)
    if args.workload in ("edges", "edgessmall"):
        return run_edges(args, dev)
    if args.workload in ("pressure", "pressuresmall"):
        return run_pressure(args, dev)
    if args.workload in ("remesh", "remeshsmall"):
        return run_remesh(args, dev)
    wl = workload(args.workload)
    Vn, Fn = icosphere(wl["subdiv"])
    k = wl["k"]
    nF = len(Fn)
    log(f"workload {args.workload}: F={nF} k={k}; building mesh tables")
    mesh = MomMesh.build(torch.tensor(Vn, dtype=torch.float64), Fn, device=dev)
    E = mesh.n_edges
    dirs, pols = [[0.0, 0.0, 1.0]], [[1.0, 0.0, 0.0]]
    F64 = _lib.MBA_F64
    set_lu_update(args.lu_update)
    set_lu_panel(args.lu_panel)
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
        "dtype": ("c64 factor (trailing update: FP32-accurate split-FP16 MFMA, 3 hi/lo piece products)"
                  if args.lu_update == "split_f16" else "c64 factor (FP32 MFMA trailing update)")
                 + " + c128 assembly/residual",
        "value_definition": "1e-6 F^2 / t_step: face pairs per second of the whole mixed solve "
                            "(FP64 assembly + FP32 LU + FP64 refinement); assembly-only rates are "
                            "assembly_fp64_Mpairs_s / assembly_fp32_Mpairs_s",
        "data": "synthetic",
        "config": {"workload": f"icosphere subdiv {wl['subdiv']} ({nF} triangles, {E} RWG), ka={k}, "
                               "EFIE 7-point rule, x-pol plane wave along +z, mixed-precision LU + FP64 refinement",
                   "faces": nF, "rwg": E, "ka": k, "parallelism": "single GPU"},
        "refine_iterations": info.iterations,
        "refine_converged": info.converged,
        "refine_fallback": info.fallback or None,
        "lu_update": args.lu_update,
        "lu_panel": args.lu_panel,
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
        if prof["phase_launches"]["split"] > 0:
            # split-FP16 trailing updates: achieved = algorithmic 8mnk flops;
            # peak = the scheme's FP32-equivalent MFMA peak (FP16 dense peak /
            # 3 piece products per real product), so frac = FP16 MFMA utilisation
            sp_ms = prof["phase_ms"]["split"]
            nl = max(1, prof["phase_launches"]["split"])
            achieved = prof["split_flops"] / (sp_ms * 1e-3) / 1e12
            peak = PEAK_FP16_MFMA_TFLOPS / 3.0
            out["roofline"] = {
                "kernel": "k_cgemm_h3 (LU trailing update: FP32-accurate complex product on "
                          "v_mfma_f32_32x32x16_f16, hi/lo split operands, 12 MFMAs per 32x32 block per 16 k; "
                          "achieved counts 8mnk flops)",
                "bound": "mfma",
                "achieved": round(achieved, 2),
                "peak": round(peak, 1),
                "peak_basis": "FP16 dense MFMA 2500 TFLOP/s / 3 piece products per real product",
                "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4),
                "traffic": pmc_traffic("traffic_c2_mixed_solve.md") if args.workload == "c2" else None,
                "traffic_source": "profiles/traffic_c2_mixed_solve.md (FETCH_SIZE x2.0 + WRITE_SIZE, copy-calibrated)",
                "launches": nl,
                "avg_launch_ms": round(sp_ms / nl, 4),
                "flops_per_launch_avg": prof["split_flops"] / nl,
                "fp32_products_and_splits_ms": round(prof["phase_ms"]["gemm"], 2),
            }
        else:
            gemm_ms = prof["phase_ms"]["gemm"]
            nl = max(1, prof["phase_launches"]["gemm"])
            achieved = prof["gemm_flops"] / (gemm_ms * 1e-3) / 1e12
            out["roofline"] = {
                "kernel": "k_cgemm3m_f32 (LU trailing update, 3M complex product on v_mfma_f32_32x32x2f32; "
                          "achieved counts 8mnk flops)",
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
        # FP32 assembly against the vector-ALU issue rates measured on this
        # device in the same run (FP32 FMA, v_sin / v_cos / v_rsq): the 7 x 7
        # point-pair body of k_ns_zf<float, 7> issues 1689 VALU instructions
        # per face pair on gfx950 (hipcc -O3 assembly of the unrolled loop), 147
        # of them v_rsq / v_sin / v_cos; the issue bound is the time the full-
        # rate ones take at the FMA rate plus the transcendentals at theirs.
        import ctypes

        from mesh_bem_amd.mom import zmat_faceedge_ns

        zf32 = torch.empty((3 * nF, 3 * nF), dtype=torch.complex64, device=dev)
        zmat_faceedge_ns(m32, k, _lib.MBA_F32, out=zf32)  # warm-up
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        zmat_faceedge_ns(m32, k, _lib.MBA_F32, out=zf32)  # records + far field + near terms, no gather
        ev1.record()
        torch.cuda.synchronize()
        tzf = ev0.elapsed_time(ev1) * 1e-3
        del zf32
        rates = (ctypes.c_double * 4)()
        if _lib.load().mba_measure_valu_rates(rates, 0, _lib.stream()) == 0:
            n_valu, n_trans = 1689, 147
            r_fma, r_trans = rates[0], (rates[1] + rates[2] + rates[3]) / 3.0
            bound_s = nF * nF * ((n_valu - n_trans) / r_fma + n_trans / r_trans)
            out["assembly_fp32_roofline"] = {
                "kernel": "k_ns_zf<float, 7> (timed with its records and near-term launches, without the gather)",
                "bound": "valu issue",
                "valu_instructions_per_face_pair": n_valu,
                "transcendentals_per_face_pair": n_trans,
                "measured_fma_per_s": float(f"{rates[0]:.4e}"),
                "measured_sin_cos_rsq_per_s": [float(f"{rates[q]:.4e}") for q in (1, 2, 3)],
                "issue_bound_ms": round(bound_s * 1e3, 2),
                "faceedge_assembly_ms": round(tzf * 1e3, 2),
                "frac": round(bound_s / tzf, 3),
                "algorithmic_tflops": round(nF * nF * 49 * 29 / tzf / 1e12, 2),
                "algorithmic_flops_definition": "49 point pairs x 29 flops per face pair",
                "rwg_assembly_ms_with_gather": round(t32 * 1e3, 2),
            }
        out["assembly_fp64_Mpairs_s"] = round(1e-6 * nF * nF / (tb - ta), 1)
        Zerr = (Z32a.to(torch.complex128) - Z64).abs().max().item() / Z64.abs().max().item()
        out["assembly_fp32_max_abs_err_rel_max"] = float(f"{Zerr:.3e}")
        del Z32a, Z32
        if args.lu_update != "fp32":
            # the same factorisation with the FP32 MFMA trailing update, for comparison
            set_lu_update("fp32")
            Z32 = to_c64(Z64)
            torch.cuda.synchronize()
            tg = time.perf_counter()
            lu2 = LUFactor(Z32, overwrite=True)
            torch.cuda.synchronize()
            th = time.perf_counter()
            _, info2 = refine_solve(Z64, V, lu32=lu2, tol=1e-8)
            out["lu_fp32_update_s"] = round(th - tg, 4)
            out["refine_iterations_fp32_update"] = info2.iterations
            set_lu_update(args.lu_update)
            del lu2, Z32
        if not args.no_fp64:
            # Plain FP64 comparators on the same GPU, each warmed up once
            # (library handles, kernel loading) and timed as the median of 3:
            # (a) FP64 assembly + torch.linalg.solve complex128 (rocSOLVER),
            # (b) FP64 assembly + this library's complex128 LU (FP64 MFMA
            #     trailing update) and its triangular solves.
            def fp64_torch():
                assemble_ns(mesh, k, F64, out=Z64)
                V = rhs_ns(mesh, k, F64, dirs, pols)
                return torch.linalg.solve(Z64, V)

            def fp64_own():
                assemble_ns(mesh, k, F64, out=Z64)
                V = rhs_ns(mesh, k, F64, dirs, pols)
                lu64 = LUFactor(Z64, overwrite=True)  # Z64 is re-assembled by every call
                return lu64.solve(V)

            def median_time(fn, reps=3):
                fn()
                torch.cuda.synchronize()
                ts, res = [], None
                for _ in range(reps):
                    t0 = time.perf_counter()
                    res = fn()
                    torch.cuda.synchronize()
                    ts.append(time.perf_counter() - t0)
                ts.sort()
                return ts[len(ts) // 2], ts, res

            log("plain FP64 solve (complex128 assembly + torch.linalg.solve): warm-up + 3 timed")
            t_torch, ts_torch, J64 = median_time(fp64_torch)
            rel = (torch.linalg.vector_norm(J - J64) / torch.linalg.vector_norm(J64)).item()
            del J64
            torch.cuda.empty_cache()
            log("plain FP64 solve (complex128 assembly + own complex128 LU): warm-up + 3 timed")
            t_own, ts_own, J64b = median_time(fp64_own)
            rel_own = (torch.linalg.vector_norm(J - J64b) / torch.linalg.vector_norm(J64b)).item()
            del J64b
            torch.cuda.empty_cache()
            t_best = min(t_torch, t_own)
            out["fp64_solve_s"] = round(t_best, 4)
            out["fp64_solve_torch_linalg_s"] = round(t_torch, 4)
            out["fp64_solve_torch_linalg_runs_s"] = [round(x, 4) for x in ts_torch]
            out["fp64_solve_own_c128_lu_s"] = round(t_own, 4)
            out["fp64_solve_own_c128_lu_runs_s"] = [round(x, 4) for x in ts_own]
            out["fp64_comparator"] = ("faster of: FP64 assembly + torch.linalg.solve complex128, "
                                      "FP64 assembly + own complex128 LU (FP64 MFMA); each warm-up + median of 3")
            out["mixed_solve_s"] = round(ms * 1e-3, 4)
            out["speedup_vs_fp64"] = round(t_best / (ms * 1e-3), 3)
            out["speedup_vs_torch_linalg_solve"] = round(t_torch / (ms * 1e-3), 3)
            out["relerr_vs_fp64"] = float(f"{rel:.3e}")
            out["relerr_vs_own_c128_lu"] = float(f"{rel_own:.3e}")
            log(f"fp64 solve: torch {t_torch:.3f} s, own c128 LU {t_own:.3f} s; relerr {rel:.2e} / {rel_own:.2e}")
        if not args.no_cpu:
            log("CPU baseline (oracle sample)")
            out["cpu_baseline"] = cpu_baseline(Vn, Fn, k, E)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
