"""Mesh files: drop-in ``LoadSingleMesh`` (src/utils/io.py:60-139) and
``save_beast_in`` (demo/generate_in_file.py:12-40) on the library's native
readers (``mba_mesh_read`` / ``mba_mesh_write_beast_in``, host C++).

``LoadSingleMesh(obj_path, device, dtype)`` returns ``[verts [V, 3] dtype,
faces [F, 3] int32]`` on ``device`` like the reference, for .obj, .ply,
.json and .in files; a .dae file returns lists ``[[verts, ...], [faces, ...]]``
(one entry per geometry) as the reference does for a trimesh Scene.  A
missing path raises RuntimeError with the reference's message.
"""

import ctypes
import os
from typing import List

import numpy as np
import torch

from . import _lib


def read_mesh_parts(path: str):
    """[(verts [V, 3] float64 numpy, faces [F, 3] int32 numpy), ...] of a file."""
    lib = _lib.load()
    parts = ctypes.POINTER(_lib.MeshPart)()
    n = ctypes.c_int32(0)
    st = lib.mba_mesh_read(os.fsencode(path), ctypes.byref(parts), ctypes.byref(n))
    _lib.check(st, f"mba_mesh_read({path})")
    out = []
    try:
        for i in range(n.value):
            p = parts[i]
            v = np.ctypeslib.as_array(p.verts, shape=(p.n_verts * 3,)).copy() if p.n_verts else np.zeros(0)
            f = np.ctypeslib.as_array(p.faces, shape=(p.n_faces * 3,)).copy() if p.n_faces else np.zeros(0, np.int32)
            out.append((v.reshape(-1, 3), f.astype(np.int32).reshape(-1, 3)))
    finally:
        lib.mba_mesh_free(parts, n.value)
    return out


def LoadSingleMesh(obj_path: str, device: torch.device = torch.device("cpu"), dtype=torch.float32) -> List:
    """Load a mesh (src/utils/io.py:60-139): [verts, faces]."""
    if not os.path.exists(obj_path):
        raise RuntimeError(f"Mesh object path :{obj_path} not found, check your DataSet")
    parts = read_mesh_parts(str(obj_path))
    tens = [(torch.from_numpy(v).to(device).to(dtype), torch.from_numpy(f).to(device).to(torch.int32))
            for v, f in parts]
    if str(obj_path).lower().endswith(".dae"):
        if len(tens) == 1:
            return [tens[0][0], tens[0][1]]
        return [[t[0] for t in tens], [t[1] for t in tens]]
    return [tens[0][0], tens[0][1]]


def save_beast_in(filename: str, verts, faces) -> None:
    """Write the BEAST .in layout (demo/generate_in_file.py:12-40): "1",
    "NV NF", vertex lines, 1-based face lines."""
    lib = _lib.load()
    v = np.ascontiguousarray(np.asarray(verts.detach().cpu() if torch.is_tensor(verts) else verts, dtype=np.float64))
    f = np.ascontiguousarray(np.asarray(faces.detach().cpu() if torch.is_tensor(faces) else faces, dtype=np.int32))
    if v.ndim != 2 or v.shape[1] != 3 or f.ndim != 2 or f.shape[1] != 3:
        raise ValueError("save_beast_in: verts [N, 3], faces [M, 3]")
    st = lib.mba_mesh_write_beast_in(os.fsencode(filename), v.ctypes.data, v.shape[0], f.ctypes.data, f.shape[0])
    _lib.check(st, f"mba_mesh_write_beast_in({filename})")
