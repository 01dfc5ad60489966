"""Build of the in-tree HIP library ``libmesh_bem_amd.so`` (gfx950).

``hipcc`` cross-compiles for gfx950 without a GPU.  Every HIP source is built
with packed FP32 instructions disabled (``-target-feature -packed-fp32-ops``);
the compiler otherwise forms v_pk_* FP32 operations from scalar float code.
"""

import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(ROOT, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libmesh_bem_amd.so")
SOURCES = ["mom_assemble.hip", "dense.hip", "laplace.hip", "gmres.hip", "mesh.hip", "mgn.hip", "radiate.hip", "helmholtz.hip", "gemm_h3.hip", "bem2d.hip", "topology.hip", "pressure.hip", "mesh_io.cpp", "remesh.hip", "loop_subdivide.cpp", "decimate.cpp", "rates.hip", "upload.hip"]
ARCH = os.environ.get("MBA_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False, jobs: int = 3) -> str:
    """Compile every source to an object (at most `jobs` compilers at once),
    link the shared library, return its path."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    build_dir = os.path.join(ROOT, "_build")
    os.makedirs(build_dir, exist_ok=True)
    cmds = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmds.append((src, obj, [hipcc, *HIPCC_FLAGS, "-c", path, "-o", obj]))

    def compile_one(item):
        src, _, cmd = item
        if verbose:
            print(" ".join(cmd), flush=True)
        # capture_output drains both pipes while the compiler runs, so a
        # compile that prints a lot cannot block on a full pipe
        return src, subprocess.run(cmd, capture_output=True)

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        results = list(pool.map(compile_one, cmds))
    failed = []
    for src, r in results:
        out = (r.stdout + r.stderr).decode(errors="replace")
        if r.returncode != 0:
            failed.append((src, out))
        elif verbose and out:
            print(out)
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    objs = [obj for _, obj, _ in cmds]
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
