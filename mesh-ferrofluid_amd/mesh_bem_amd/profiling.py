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
    a = ap.parse_args(argv)
    if a.what == "kernels":
        print_kernels(a.db, a.top, a.cmd)
    elif a.what == "pmc":
        print_pmc(a.db, a.kernel, a.cmd)
    else:
        copy_run(a.mib, a.reps)


if __name__ == "__main__":
    sys.exit(main())
