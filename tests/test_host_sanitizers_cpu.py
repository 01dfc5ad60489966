"""The library's host C++ (mesh readers of LoadSingleMesh, src/utils/io.py:60-139;
Loop subdivision and quadric decimation behind remesh_and_transfer_velocity,
src/utils/mesh_utils.py:728-815) built with g++ -fsanitize=address,undefined
and run over the reader fixtures, malformed files and a hypothesis fuzz of
mutated files.  The sanitized build is a separate executable
(tests/host_asan_driver.cpp linked with the library's host sources); any
sanitizer report aborts it and fails the test.  CPU only."""

import os
import random
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mesh-ferrofluid_amd", "csrc")
HOST_SOURCES = ["mesh_io.cpp", "loop_subdivide.cpp", "decimate.cpp"]

hyp = pytest.importorskip("hypothesis")
from hypothesis import given, settings, strategies as st  # noqa: E402


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("asan") / "host_asan_driver")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", os.path.join(ROOT, "tests", "host_asan_driver.cpp"),
           *[os.path.join(CSRC, s) for s in HOST_SOURCES], "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("sanitized build failed:\n" + r.stderr[-4000:])
    return out


def run_driver(driver, paths):
    env = dict(os.environ)
    # leak checking needs ptrace, which containers may refuse; the readers'
    # allocations are freed by mba_mesh_free on every path
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    r = subprocess.run([driver, *paths], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, \
        r.stderr[-4000:]
    return [int(line.split(" ", 1)[0]) for line in r.stdout.splitlines()]


def icosphere_obj():
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
         (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    return v, f


def valid_files():
    v, f = icosphere_obj()
    obj = "".join(f"v {x} {y} {z}\n" for x, y, z in v) + "".join(f"f {a+1} {b+1} {c+1}\n" for a, b, c in f)
    ply = ("ply\nformat ascii 1.0\nelement vertex %d\nproperty float x\nproperty float y\nproperty float z\n"
           "element face %d\nproperty list uchar int vertex_indices\nend_header\n" % (len(v), len(f))
           + "".join(f"{x} {y} {z}\n" for x, y, z in v) + "".join(f"3 {a} {b} {c}\n" for a, b, c in f))
    plyb = (("ply\nformat binary_little_endian 1.0\nelement vertex %d\nproperty double x\nproperty double y\n"
             "property double z\nelement face %d\nproperty list uchar int vertex_indices\nend_header\n"
             % (len(v), len(f))).encode()
            + b"".join(struct.pack("<3d", *p) for p in v) + b"".join(struct.pack("<B3i", 3, *t) for t in f))
    js = '{"verts": %s, "faces": %s}' % ([list(p) for p in v], [list(t) for t in f])
    inn = "1\n%d %d\n" % (len(v), len(f)) + "".join(f"{x} {y} {z}\n" for x, y, z in v) + \
        "".join(f"{a+1} {b+1} {c+1}\n" for a, b, c in f)
    dae = ("<COLLADA><library_geometries><geometry id=\"g\"><mesh>"
           "<source id=\"p\"><float_array id=\"pa\" count=\"%d\">%s</float_array></source>"
           "<vertices id=\"vv\"><input semantic=\"POSITION\" source=\"#p\"/></vertices>"
           "<triangles count=\"%d\"><input semantic=\"VERTEX\" source=\"#vv\" offset=\"0\"/><p>%s</p></triangles>"
           "</mesh></geometry></library_geometries></COLLADA>"
           % (3 * len(v), " ".join(f"{c}" for p in v for c in p), len(f), " ".join(f"{i}" for t in f for i in t)))
    return {".obj": obj.encode(), ".ply": ply.encode(), "_b.ply": plyb, ".json": js.encode(), ".in": inn.encode(),
            ".dae": dae.encode()}


def test_readers_and_remeshing_valid_files(driver, tmp_path):
    paths = []
    for suffix, data in valid_files().items():
        p = tmp_path / ("m" + suffix)
        p.write_bytes(data)
        paths.append(str(p))
    assert run_driver(driver, paths) == [0] * len(paths)


def test_malformed_files(driver, tmp_path):
    cases = {
        "huge.in": b"1 1000000000000 1",
        "neg.in": b"1 -5 3",
        "wrap.obj": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4294967299\n",
        "nan.json": b'{"verts": [[0,0,0],[1,0,0],[0,1,0]], "faces": [[0, 1, 1e300]]}',
        "trunc.ply": b"ply\nformat binary_little_endian 1.0\nelement vertex 1000000\nproperty double x\nend_header\n\x00\x01",
        "badoff.dae": b'<COLLADA><geometry id="g"><mesh><source id="p"><float_array>0 0 0 1 0 0 0 1 0</float_array>'
                      b'</source><vertices id="v"><input semantic="POSITION" source="#p"/></vertices>'
                      b'<triangles><input semantic="VERTEX" source="#v" offset="2147483647"/><p>0 1 2</p>'
                      b'</triangles></mesh></geometry></COLLADA>',
        "empty.obj": b"",
    }
    paths = []
    for name, data in cases.items():
        p = tmp_path / name
        p.write_bytes(data)
        paths.append(str(p))
    codes = run_driver(driver, paths)
    assert all(c != 0 for c in codes[:-1]), codes


def mutate(data: bytes, rnd: random.Random) -> bytes:
    b = bytearray(data)
    for _ in range(rnd.randint(1, 6)):
        op = rnd.randrange(5)
        if not b:
            break
        i = rnd.randrange(len(b))
        if op == 0:
            b[i] = rnd.randrange(256)
        elif op == 1:
            del b[i:i + rnd.randint(1, 16)]
        elif op == 2:
            b[i:i] = rnd.choice([b"-1", b"99999999999", b"4294967296", b"1e400", b"nan", b" ", b"\n", b"[", b"]",
                                 b"<p>", b"</p>", b"3 ", b"\x00\xff\xff\xff"])
        elif op == 3:
            b = b[:i]
        else:
            j = rnd.randrange(len(b))
            b[i], b[j] = b[j], b[i]
    return bytes(b)


@settings(max_examples=25, deadline=None)
@given(seed=st.integers(0, 2 ** 31 - 1))
def test_fuzz_mutated_files(driver, tmp_path_factory, seed):
    rnd = random.Random(seed)
    d = tmp_path_factory.mktemp("fuzz")
    paths = []
    for k, (suffix, data) in enumerate(valid_files().items()):
        for r in range(4):
            p = d / f"f{k}_{r}{suffix}"
            p.write_bytes(mutate(data, rnd))
            paths.append(str(p))
    run_driver(driver, paths)


def test_loop_and_decimate_on_larger_mesh(driver, tmp_path):
    """Subdivision and decimation on a 1280-face sphere under the sanitizers."""
    import sys

    sys.path.insert(0, os.path.join(ROOT, "mesh-ferrofluid_amd"))
    from mesh_bem_amd.geometry import icosphere

    V, F = icosphere(3)
    p = tmp_path / "s.obj"
    p.write_text("".join(f"v {x:.17g} {y:.17g} {z:.17g}\n" for x, y, z in V)
                 + "".join(f"f {a+1} {b+1} {c+1}\n" for a, b, c in np.asarray(F)))
    assert run_driver(driver, [str(p)]) == [0]
