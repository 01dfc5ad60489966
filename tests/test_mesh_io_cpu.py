"""Mesh files through the library's host readers (mba_mesh_read /
mba_mesh_write_beast_in; no device work): LoadSingleMesh
(src/utils/io.py:60-139) and save_beast_in (demo/generate_in_file.py:12-40).

The reference reads .obj with pywavefront, .ply with open3d and .dae with
trimesh (all absent here, unpinned versions in readme.md); these tests pin the
readers with hand-written fixtures of known content (every index form of OBJ,
ascii / little- / big-endian PLY with extra properties, COLLADA triangles and
polylists over several geometries), polygons split into fans, the .in round
trip, and the error paths."""

import struct

import numpy as np
import pytest
import torch

from mesh_bem_amd._lib import MBAError
from mesh_bem_amd.io import LoadSingleMesh, read_mesh_parts, save_beast_in

SQUARE_V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1.0]], dtype=np.float64)


def test_obj_index_forms_and_fans(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text(
        "# comment\nmtllib x.mtl\no thing\n"
        + "".join(f"v {x} {y} {z}\n" for x, y, z in SQUARE_V)
        + "vt 0 0\nvt 1 0\nvn 0 0 1\n"
        "f 1/1/1 2/2/1 5/1/1\n"     # v/vt/vn
        "f 2//1 3//1 5//1\r\n"      # v//vn, CRLF
        "f -3 -2 -1\n"              # negative: vertices 3, 4, 5
        "f 1 2 3 4\n")              # quad -> (1 2 3), (1 3 4)
    v, f = LoadSingleMesh(str(p), dtype=torch.float64)
    assert v.dtype == torch.float64 and f.dtype == torch.int32
    assert np.array_equal(v.numpy(), SQUARE_V)
    assert f.tolist() == [[0, 1, 4], [1, 2, 4], [2, 3, 4], [0, 1, 2], [0, 2, 3]]


def test_ply_ascii_extra_properties_and_quad(tmp_path):
    p = tmp_path / "m.ply"
    hdr = ("ply\nformat ascii 1.0\ncomment x\nelement vertex 5\nproperty float x\nproperty float y\n"
           "property float z\nproperty uchar red\nproperty float nx\nelement face 2\n"
           "property list uchar int vertex_indices\nend_header\n")
    body = "".join(f"{x} {y} {z} 255 0.5\n" for x, y, z in SQUARE_V) + "4 0 1 2 3\n3 0 1 4\n"
    p.write_text(hdr + body)
    (v, f), = read_mesh_parts(str(p))
    assert np.array_equal(v, SQUARE_V)
    assert f.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 4]]


@pytest.mark.parametrize("endian", ["<", ">"])
def test_ply_binary(tmp_path, endian):
    rng = np.random.default_rng(1)
    V = rng.standard_normal((7, 3))
    F = np.array([[0, 1, 2], [2, 3, 4], [4, 5, 6], [6, 0, 3]], dtype=np.int32)
    fmt = "binary_little_endian" if endian == "<" else "binary_big_endian"
    hdr = (f"ply\nformat {fmt} 1.0\nelement vertex 7\nproperty double x\nproperty double y\nproperty double z\n"
           "property ushort flags\nelement face 4\nproperty list uchar uint vertex_index\n"
           "element edge 1\nproperty int vertex1\nproperty int vertex2\nend_header\n").encode()
    body = b"".join(struct.pack(endian + "dddH", *v, 3) for v in V)
    body += b"".join(struct.pack(endian + "BIII", 3, *f) for f in F)
    body += struct.pack(endian + "ii", 0, 1)
    p = tmp_path / "m.ply"
    p.write_bytes(hdr + body)
    v, f = LoadSingleMesh(str(p), dtype=torch.float64)
    assert np.array_equal(v.numpy(), V)
    assert np.array_equal(f.numpy(), F)
    # truncated body
    p.write_bytes(hdr + body[:40])
    with pytest.raises(MBAError):
        read_mesh_parts(str(p))


def test_json_nested_and_flat(tmp_path):
    p = tmp_path / "m.json"
    p.write_text('{"edges": [[0, 1]], "verts": [[0, 0, 0], [1, 0, 0], [0, 1e-3, 2.5]], "faces": [[0, 1, 2]]}')
    v, f = LoadSingleMesh(str(p), dtype=torch.float64)
    assert np.array_equal(v.numpy(), [[0, 0, 0], [1, 0, 0], [0, 1e-3, 2.5]]) and f.tolist() == [[0, 1, 2]]
    p.write_text('{"faces": [0, 1, 2], "verts": [0, 0, 0, 1, 0, 0, 0, 1, 0]}')
    v, f = LoadSingleMesh(str(p))
    assert v.shape == (3, 3) and f.tolist() == [[0, 1, 2]]


def test_beast_in_round_trip(tmp_path):
    from mesh_bem_amd.geometry import icosphere

    V, F = icosphere(2)
    V = V.astype(np.float32).astype(np.float64)
    p = tmp_path / "sphere.in"
    save_beast_in(str(p), V, F)
    lines = p.read_text().splitlines()
    assert lines[0] == "1" and lines[1] == f"{len(V)} {len(F)}"
    assert [int(x) for x in lines[2 + len(V)].split()] == (F[0] + 1).tolist()  # 1-based
    v, f = LoadSingleMesh(str(p), dtype=torch.float64)
    assert np.array_equal(v.numpy(), V) and np.array_equal(f.numpy(), F)
    # already 1-based input is written unchanged
    save_beast_in(str(p), V, F + 1)
    assert [int(x) for x in p.read_text().splitlines()[2 + len(V)].split()] == (F[0] + 1).tolist()


DAE = """<?xml version="1.0"?>
<COLLADA xmlns="http://www.collada.org/2005/11/COLLADASchema" version="1.4.1">
 <library_geometries>
  <geometry id="g0" name="tri">
   <mesh>
    <source id="g0-pos"><float_array id="g0-pos-a" count="9">0 0 0  1 0 0  0 1 0</float_array></source>
    <source id="g0-nrm"><float_array id="g0-nrm-a" count="3">0 0 1</float_array></source>
    <vertices id="g0-v"><input semantic="POSITION" source="#g0-pos"/></vertices>
    <triangles count="1"><input semantic="VERTEX" source="#g0-v" offset="0"/>
     <input semantic="NORMAL" source="#g0-nrm" offset="1"/><p>0 0 1 0 2 0</p></triangles>
   </mesh>
  </geometry>
  <geometry id="g1" name="poly">
   <mesh>
    <source id="g1-pos"><float_array id="g1-pos-a" count="15">0 0 0 1 0 0 1 1 0 0 1 0 0.5 0.5 1</float_array></source>
    <vertices id="g1-v"><input semantic="POSITION" source="#g1-pos"/></vertices>
    <polylist count="2"><input semantic="VERTEX" source="#g1-v" offset="0"/>
     <vcount>4 3</vcount><p>0 1 2 3 0 1 4</p></polylist>
   </mesh>
  </geometry>
 </library_geometries>
</COLLADA>
"""


def test_dae_scene_parts(tmp_path):
    p = tmp_path / "m.dae"
    p.write_text(DAE)
    vs, fs = LoadSingleMesh(str(p), dtype=torch.float64)
    assert len(vs) == 2 and len(fs) == 2
    assert np.array_equal(vs[0].numpy(), [[0, 0, 0], [1, 0, 0], [0, 1, 0]]) and fs[0].tolist() == [[0, 1, 2]]
    assert np.array_equal(vs[1].numpy(), SQUARE_V)
    assert fs[1].tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 4]]


def test_errors(tmp_path):
    with pytest.raises(RuntimeError, match="not found"):
        LoadSingleMesh(str(tmp_path / "missing.obj"))
    p = tmp_path / "m.stl"
    p.write_text("solid")
    with pytest.raises(MBAError):
        read_mesh_parts(str(p))
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")  # vertex 3 does not exist
    with pytest.raises(MBAError):
        read_mesh_parts(str(p))


@pytest.mark.parametrize("name,text", [
    # counts far beyond what the file holds: rejected before any allocation
    # (once aborted the process with std::bad_alloc)
    ("huge.in", "1 1000000000000 1"),
    ("huge2.in", "1 3 4611686018427387904\n0 0 0\n1 0 0\n0 1 0\n1 2 3\n"),
    # a 1-based face index past the vertex count / zero
    ("oob.in", "1\n3 1\n0 0 0\n1 0 0\n0 1 0\n1 2 4\n"),
    ("zero.in", "1\n3 1\n0 0 0\n1 0 0\n0 1 0\n0 1 2\n"),
    # indices that do not fit int32 must not wrap onto valid vertices
    ("wrap.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4294967299\n"),
    ("wrap.json", '{"verts": [[0,0,0],[1,0,0],[0,1,0]], "faces": [[0, 1, 4294967298]]}'),
    ("nan.json", '{"verts": [[0,0,0],[1,0,0],[0,1,0]], "faces": [[0, 1, 1e300]]}'),
])
def test_malformed_counts_and_indices_return_errors(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    with pytest.raises(MBAError):
        read_mesh_parts(str(p))


def test_ply_huge_list_count_is_an_error(tmp_path):
    p = tmp_path / "m.ply"
    hdr = ("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
           "element face 1\nproperty list uchar int vertex_indices\nend_header\n")
    p.write_text(hdr + "0 0 0\n1 0 0\n0 1 0\n3 0 1 4294967298\n")
    with pytest.raises(MBAError):
        read_mesh_parts(str(p))
