"""Summaries of rocprofv3 results databases, and the copy-kernel calibration run.

  python -m mesh_bem_amd.profiling kernels <results.db> [--top N] [--cmd "..."]
      per-kernel dispatch statistics (calls, total / average / min / max
      duration, registers, LDS, scratch) as a markdown table.
  python -m mesh_bem_amd.profiling pmc <results.db> [--kernel SUBSTR] [--cmd "..."]
      per-kernel averages of the collected counters per dispatch.
  python -m mesh_bem_amd.profiling copy [--mib N] [--reps R]
      a plain device-to-device copy of N MiB (R times) on cuda:0: the
      reference workload whose FETCH_SIZE / WRITE_SIZE readings calibrate the
      counters (bytes moved are known exactly: N MiB read, N MiB written).
  python -m mesh_bem_amd.profiling mfma-rate [--n N] [--m M]
      what the matrix cores deliver on this device: torch.matmul (the vendor
      BLAS) FP16 / BF16 at N^3, and this library's split-FP16 complex product
      (mba_gemm_split_f16) at M x M x k for k = 512 and 1024, each timed with
      device events after a warm-up.
  python -m mesh_bem_amd.profiling valu-rate
      what the vector ALU issues on this device (mba_measure_valu_rates):
      FP32 FMA (not packed) and the hardware v_sin / v_cos / v_rsq per second.
"""

import argparse
import re
import sqlite3
import sys


def _demangle(name: str) -> str:
    """Itanium-demangled name (c++filt) for names the profiler left mangled."""
    import subprocess

    try:
        out = subprocess.run(["c++filt", name], capture_output=True, text=True, timeout=10).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        out = name
    if out and out != name:
        return out
    # c++filt does not know every builtin type (e.g. _Float16, "DF16_"): take
    # the nested name's components (length-prefixed identifiers after _ZN)
    if not name.startswith("_ZN"):
        return name
    comps, i = [], 3
    while i < len(name) and name[i].isdigit():
        j = i
        while j < len(name) and name[j].isdigit():
            j += 1
        n = int(name[i:j])
        comps.append(name[j:j + n])
        i = j + n
    comps = [c for c in comps if not c.startswith("_GLOBAL__N")]
    return "::".join(comps) if comps else name


def _short(name: str) -> str:
    name = name.strip()
    if name.startswith("_Z"):
        name = _demangle(name)
    name = name.replace("(anonymous namespace)::", "")
    name = re.sub(r"^void ", "", name)
    cut = name.find("(")
    if cut > 0:
        name = name[:cut]
    return name[:90]


def kernel_stats(db: str):
    c = sqlite3.connect(db)
    rows = c.execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration), "
        "max(vgpr_count), max(accum_vgpr_count), max(lds_size), max(scratch_size) "
        "from kernels group by name order by sum(duration) desc").fetchall()
    c.close()
    return rows


def print_kernels(db: str, top: int = 25, cmd: str = ""):
    rows = kernel_stats(db)
    tot = sum(r[2] for r in rows) or 1
    out = []
    if cmd:
        out.append(f"Command: `{cmd}`\n")
    out.append(f"Total GPU kernel time: {tot / 1e6:.2f} ms over {sum(r[1] for r in rows)} dispatches\n")
    out.append("| kernel | calls | total ms | % | avg us | min us | max us | VGPR | AGPR | LDS B | scratch B |")
    out.append("|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
    for r in rows[:top]:
        name, n, s, a, mn, mx, vg, ag, lds, scr = r
        out.append(f"| `{_short(name)}` | {n} | {s / 1e6:.3f} | {100 * s / tot:.1f} | {a / 1e3:.2f} | "
                   f"{mn / 1e3:.2f} | {mx / 1e3:.2f} | {vg} | {ag} | {lds} | {scr} |")
    print("\n".join(out))


def pmc_stats(db: str, kernel: str = ""):
    c = sqlite3.connect(db)
    cols = [d[1] for d in c.execute("pragma table_info(counters_collection)")]
    name_col = "kernel_name" if "kernel_name" in cols else "name"
    q = (f"select {name_col}, counter_name, count(*), avg(value), sum(value) from counters_collection "
         f"group by {name_col}, counter_name order by {name_col}")
    rows = c.execute(q).fetchall()
    c.close()
    if kernel:
        rows = [r for r in rows if kernel in r[0]]
    return rows


def print_pmc(db: str, kernel: str = "", cmd: str = ""):
    rows = pmc_stats(db, kernel)
    out = []
    if cmd:
        out.append(f"Command: `{cmd}`\n")
    out.append("| kernel | counter | dispatches | avg per dispatch | sum |")
    out.append("|---|---|---:|---:|---:|")
    for name, ctr, n, avg, s in rows:
        out.append(f"| `{_short(name)}` | {ctr} | {n} | {avg:.1f} | {s:.1f} |")
    print("\n".join(out))


def copy_run(mib: int = 1024, reps: int = 5):
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {reps}", flush=True)


def mfma_rate(n: int = 16384, m: int = 29696, reps: int = 5):
    import os
    import torch

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from mesh_bem_amd.dense import gemm_split_f16

    dev = torch.device("cuda:0")

    def timed(fn, r):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(r):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / r * 1e-3

    for dt in (torch.float16, torch.bfloat16):
        a = torch.randn(n, n, device=dev, dtype=dt)
        b = torch.randn(n, n, device=dev, dtype=dt)
        t = timed(lambda: torch.matmul(a, b), reps)
        print(f"torch.matmul {str(dt).split('.')[-1]} {n}^3: {t * 1e3:.2f} ms, {2 * n ** 3 / t / 1e12:.1f} TFLOP/s",
              flush=True)
        del a, b
    for k in (512, 1024):
        A = torch.randn(m, k, device=dev, dtype=torch.complex64)
        B = torch.randn(k, m, device=dev, dtype=torch.complex64)
        C = torch.randn(m, m, device=dev, dtype=torch.complex64)
        t = timed(lambda: gemm_split_f16(A, B, C, -1.0, 1.0), 3)
        tf = 8 * m * m * k / t / 1e12
        print(f"split-FP16 complex product {m}^2 x {k}: {t * 1e3:.2f} ms, {tf:.1f} TFLOP/s algorithmic "
              f"({3 * tf:.1f} TFLOP/s of FP16 MFMA work)", flush=True)
        del A, B, C


def valu_rate():
    import ctypes
    import os

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from mesh_bem_amd import _lib

    lib = _lib.load()
    out = (ctypes.c_double * 4)()
    _lib.check(lib.mba_measure_valu_rates(out, 0, _lib.stream()), "mba_measure_valu_rates")
    names = ("FP32 FMA (not packed)", "v_sin_f32", "v_cos_f32", "v_rsq_f32")
    for nm, r in zip(names, out):
        print(f"{nm}: {r:.4e} per second", flush=True)
    print(f"FP32 FMA = {2 * out[0] / 1e12:.1f} TFLOP/s", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    k.add_argument("--top", type=int, default=25)
    k.add_argument("--cmd", default="")
    p = sub.add_parser("pmc")
    p.add_argument("db")
    p.add_argument("--kernel", default="")
    p.add_argument("--cmd", default="")
    cp = sub.add_parser("copy")
    cp.add_argument("--mib", type=int, default=1024)
    cp.add_argument("--reps", type=int, default=5)
    mr = sub.add_parser("mfma-rate")
    mr.add_argument("--n", type=int, default=16384)
    mr.add_argument("--m", type=int, default=29696)
    sub.add_parser("valu-rate")
    a = ap.parse_args(argv)
    if a.what == "kernels":
        print_kernels(a.db, a.top, a.cmd)
    elif a.what == "pmc":
        print_pmc(a.db, a.kernel, a.cmd)
    elif a.what == "mfma-rate":
        mfma_rate(a.n, a.m)
    elif a.what == "valu-rate":
        valu_rate()
    else:
        copy_run(a.mib, a.reps)


if __name__ == "__main__":
    sys.exit(main())
