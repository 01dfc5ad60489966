"""Mesh geometry and topology on the host.

Drop-in counterparts of the reference helpers used by the BEM path:

* ``compute_face_normals`` / ``compute_face_areas`` / ``compute_vert_areas``
  (src/utils/mesh_utils.py:158-310): normal = normalise((v1 - v0) x (v2 - v1)),
  area = |...| / 2, lumped vertex area = sum of adjacent A_f / 3.
* ``get_edge_unique`` (src/utils/mesh_utils.py:921-977): local edge e of face f
  joins faces[f][(e+1)%3] and faces[f][(e+2)%3]; edge key ``max*V + min``.
  The reference de-duplicates with ``torch.unique(sorted=False)`` whose order is
  backend dependent; here the keys are sorted (the order torch returns on the
  GPU), so the numbering is deterministic.
* ``compute_panel_relation`` (src/utils/mesh_utils.py:700-725): count of shared
  vertex indices, computed from the vertex->face incidence in O(F) memory
  instead of an [F, F, 9] temporary.

Synthetic meshes for the workloads: icosphere, geodesic sphere of frequency n,
spherical-harmonic star meshes and noisy ellipsoids.
"""

from typing import List, Tuple

import numpy as np
import torch


# ---------------------------------------------------------------- per-face data
def _face_vectors(vertices: torch.Tensor, faces: torch.Tensor):
    f = faces.to(torch.int64)
    v0 = vertices[..., f[:, 0], :]
    v1 = vertices[..., f[:, 1], :]
    v2 = vertices[..., f[:, 2], :]
    return v0, v1, v2


def _segment_vector(vertices: torch.Tensor, faces: torch.Tensor) -> torch.Tensor:
    f = faces.to(torch.int64)
    return vertices[..., f[:, 1], :] - vertices[..., f[:, 0], :]


def compute_face_normals(vertices: torch.Tensor, faces: torch.Tensor) -> torch.Tensor:
    """[F,3] unit normals normalise((v1-v0) x (v2-v1)) of triangles; for 2-D
    segments ([F,2] faces, [V,2] vertices) normalise(e_y, -e_x), e = v1 - v0
    (src/utils/mesh_utils.py:158-197)."""
    if faces.shape[-1] == 2:
        e = _segment_vector(vertices, faces)
        return torch.nn.functional.normalize(torch.stack((e[..., 1], -e[..., 0]), dim=-1), dim=-1)
    v0, v1, v2 = _face_vectors(vertices, faces)
    n = torch.cross(v1 - v0, v2 - v1, dim=-1)
    return torch.nn.functional.normalize(n, dim=-1)


def compute_face_areas(
    vertices: torch.Tensor, faces: torch.Tensor, keepdim: bool = True
) -> torch.Tensor:
    """[F,1] (keepdim) or [F] triangle areas; segment lengths for 2-D
    segments (src/utils/mesh_utils.py:200-239)."""
    if faces.shape[-1] == 2:
        return _segment_vector(vertices, faces).norm(dim=-1, keepdim=keepdim)
    v0, v1, v2 = _face_vectors(vertices, faces)
    a = 0.5 * torch.cross(v1 - v0, v2 - v1, dim=-1).norm(dim=-1, keepdim=keepdim)
    return a


def compute_vert_areas(
    vertices: torch.Tensor, faces: torch.Tensor, keepdim: bool = True
) -> torch.Tensor:
    """[V,1] lumped vertex areas: sum over adjacent faces of A_f / nv (nv = 3
    vertices per triangle, 2 per segment; src/utils/mesh_utils.py:277-310)."""
    nv = faces.shape[-1]
    fa = compute_face_areas(vertices, faces, keepdim=False)
    va = torch.zeros(vertices.shape[-2], dtype=vertices.dtype, device=vertices.device)
    idx = faces.to(torch.int64).reshape(-1)
    va.index_add_(0, idx, (fa[:, None] / float(nv)).expand(-1, nv).reshape(-1))
    return va[:, None] if keepdim else va


def compute_vert_normals(vertices: torch.Tensor, faces: torch.Tensor) -> torch.Tensor:
    """[V,3] area-weighted vertex normals (src/utils/mesh_utils.py:242-274)."""
    fn = compute_face_normals(vertices, faces)
    fa = compute_face_areas(vertices, faces, keepdim=True)
    vn = torch.zeros_like(vertices)
    idx = faces.to(torch.int64).reshape(-1)
    nv, dim = faces.shape[-1], vertices.shape[-1]
    vn.index_add_(0, idx, (fa * fn)[:, None, :].expand(-1, nv, -1).reshape(-1, dim))
    return torch.nn.functional.normalize(vn, dim=-1)


# ---------------------------------------------------------------- topology
def get_edge_unique(verts: torch.Tensor, faces: torch.Tensor) -> List[torch.Tensor]:
    """Unique edges of a triangle mesh.

    Returns ``edges [E,2]`` as (max, min) vertex pairs, ``face_edge_indices
    [3F]`` (edge index of local edge e of face f at 3f+e, dtype of ``faces``)
    and ``edge_adjacent_faces [E,2]`` (lower face index first).
    """
    N_verts = verts.shape[0]
    if faces.is_cuda:
        return _edge_unique_device(faces, N_verts)
    f = faces.detach().to("cpu").to(torch.int64).numpy()
    F = f.shape[0]
    a = np.stack([f[:, 1], f[:, 2], f[:, 0]], axis=1).reshape(-1)  # (e+1)%3
    b = np.stack([f[:, 2], f[:, 0], f[:, 1]], axis=1).reshape(-1)  # (e+2)%3
    key = np.maximum(a, b) * N_verts + np.minimum(a, b)
    uniq, inverse = np.unique(key, return_inverse=True)
    edges = np.stack([uniq // N_verts, uniq % N_verts], axis=1)
    E = uniq.size
    fe_face = np.repeat(np.arange(F), 3)
    adj = np.full((E, 2), -1, dtype=np.int64)
    order = np.argsort(inverse, kind="stable")
    inv_sorted = inverse[order]
    first = np.ones(inv_sorted.size, dtype=bool)
    first[1:] = inv_sorted[1:] != inv_sorted[:-1]
    adj[inv_sorted[first], 0] = fe_face[order[first]]
    second = ~first
    adj[inv_sorted[second], 1] = fe_face[order[second]]
    dev = faces.device
    return (
        torch.from_numpy(edges).to(faces.dtype).to(dev),
        torch.from_numpy(inverse.astype(np.int64)).to(faces.dtype).to(dev),
        torch.from_numpy(adj).to(faces.dtype).to(dev),
    )


def _edge_unique_device(faces: torch.Tensor, n_verts: int) -> List[torch.Tensor]:
    """get_edge_unique of a device mesh on the GPU (mba_edge_unique): the same
    edges, order and adjacency as the host path above, bit for bit."""
    from . import _lib

    lib = _lib.load()
    dev = faces.device
    F = faces.shape[0]
    ft = faces.to(torch.int32).contiguous()
    wsb = lib.mba_edge_unique_workspace_bytes(F, n_verts)
    if wsb == 0:
        raise ValueError("get_edge_unique: mesh too large for the device path")
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    edges = torch.empty((3 * F, 2), dtype=torch.int32, device=dev)
    fe = torch.empty(3 * F, dtype=torch.int32, device=dev)
    adj = torch.empty((3 * F, 2), dtype=torch.int32, device=dev)
    ne = torch.zeros(1, dtype=torch.int64, device=dev)
    st = lib.mba_edge_unique(_lib.ptr(ft), F, n_verts, _lib.ptr(edges), _lib.ptr(fe), _lib.ptr(adj), _lib.ptr(ne),
                             _lib.ptr(ws), wsb, _lib.stream())
    _lib.check(st, "mba_edge_unique")
    E = int(ne.item())
    return edges[:E].to(faces.dtype), fe.to(faces.dtype), adj[:E].to(faces.dtype)


def rwg_face_edges(face_edge_indices, n_edges: int) -> np.ndarray:
    """[E,2] int32: the two face-edge slots (3f+e) of every RWG, ascending.

    Raises ValueError unless every edge has exactly two incident faces (a closed
    2-manifold), which the RWG basis requires.
    """
    fe = np.asarray(
        face_edge_indices.detach().cpu().numpy()
        if isinstance(face_edge_indices, torch.Tensor)
        else face_edge_indices
    ).astype(np.int64).reshape(-1)
    counts = np.bincount(fe, minlength=n_edges)
    if counts.size != n_edges or np.any(counts != 2):
        bad = int(np.count_nonzero(counts != 2))
        raise ValueError(f"mesh is not a closed 2-manifold: {bad} edges without exactly 2 faces")
    order = np.argsort(fe, kind="stable")
    return order.reshape(n_edges, 2).astype(np.int32)


def check_orientation(faces: np.ndarray, rwg_fe: np.ndarray) -> None:
    """Raise ValueError unless the two faces of every RWG edge traverse it in
    opposite directions (a consistently oriented surface).  The RWG sign of
    local edge e of face f is + iff faces[f][(e+1)%3] < faces[f][(e+2)%3]
    (src/mesh_BEM/cuda/mesh_MOM_3d.cu:154-165); on a consistently oriented
    surface exactly one face of every edge is T+, else the basis is invalid."""
    f = np.asarray(faces, dtype=np.int64)
    slots = np.asarray(rwg_fe, dtype=np.int64)
    fi, ei = slots // 3, slots % 3
    a = f[fi, (ei + 1) % 3]
    b = f[fi, (ei + 2) % 3]
    plus = a < b  # [E, 2]
    bad = int(np.count_nonzero(plus[:, 0] == plus[:, 1]))
    if bad:
        raise ValueError(f"inconsistently oriented mesh: {bad} edges traversed in the same direction by both faces")


def compute_panel_relation(faces: torch.Tensor) -> torch.Tensor:
    """[F,F] uint8 number of shared vertex indices (0..3 for triangles, 0..2
    for segments), in O(F) extra memory."""
    f = faces.detach().cpu().to(torch.int64).numpy()
    F = f.shape[0]
    nv = int(f.max()) + 1 if F else 0
    rel = np.zeros((F, F), dtype=np.uint8)
    # vertex -> faces incidence
    vf_face = np.repeat(np.arange(F), f.shape[1])
    vf_vert = f.reshape(-1)
    order = np.argsort(vf_vert, kind="stable")
    starts = np.searchsorted(vf_vert[order], np.arange(nv + 1))
    for v in range(nv):
        fs = vf_face[order[starts[v] : starts[v + 1]]]
        if fs.size:
            rel[np.ix_(fs, fs)] += 1
    return torch.from_numpy(rel).to(faces.device)


def check_mesh(verts: np.ndarray, faces: np.ndarray) -> None:
    """Raise ValueError for index-out-of-range or zero-area faces (area at most
    1e-12 of the squared longest edge: collinear corners up to rounding)."""
    if faces.size and (faces.min() < 0 or faces.max() >= verts.shape[0]):
        raise ValueError("face index out of range")
    v0, v1, v2 = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    a = 0.5 * np.linalg.norm(np.cross(v1 - v0, v2 - v1), axis=1)
    l2 = np.maximum.reduce([((v1 - v0) ** 2).sum(1), ((v2 - v1) ** 2).sum(1), ((v0 - v2) ** 2).sum(1)])
    if np.any(~(a > 1e-12 * l2)):
        raise ValueError("zero-area face")


# ---------------------------------------------------------------- generators
_ICO_FACES = [
    (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
    (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
    (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
    (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1),
]


def _icosahedron() -> Tuple[np.ndarray, np.ndarray]:
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array(
        [
            (-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
            (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
            (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1),
        ],
        dtype=np.float64,
    )
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v, np.array(_ICO_FACES, dtype=np.int64)


def icosphere(subdiv: int, radius: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    """Icosphere: V = 10*4^k + 2, E = 30*4^k, F = 20*4^k, outward orientation.

    Each face is split into 4 children stored consecutively (depth-first), so
    face indices are spatially local.
    """
    v, f = _icosahedron()
    verts = [tuple(p) for p in v]
    faces = [tuple(x) for x in f]
    for _ in range(subdiv):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                p = (np.array(verts[i]) + np.array(verts[j])) * 0.5
                p /= np.linalg.norm(p)
                cache[key] = len(verts)
                verts.append(tuple(p))
            return cache[key]

        nf = []
        for (a, b, c) in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius
    Fc = np.array(faces, dtype=np.int64)
    return V, _orient_outward(V, Fc)


def geodesic_sphere(freq: int, radius: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    """Class-I geodesic sphere of frequency n: F = 20 n^2, E = 30 n^2, V = 10 n^2 + 2."""
    v, f = _icosahedron()
    n = int(freq)
    vid = {}
    verts = []

    def vertex(p):
        p = p / np.linalg.norm(p)
        key = tuple(np.round(p, 12))
        if key not in vid:
            vid[key] = len(verts)
            verts.append(p)
        return vid[key]

    faces = []
    for (a, b, c) in f:
        A, B, C = v[a], v[b], v[c]
        idx = {}
        for i in range(n + 1):
            for j in range(n + 1 - i):
                k = n - i - j
                idx[(i, j)] = vertex((i * A + j * B + k * C) / n)
        for i in range(n):
            for j in range(n - i):
                faces.append((idx[(i, j)], idx[(i + 1, j)], idx[(i, j + 1)]))
                if i + j < n - 1:
                    faces.append((idx[(i + 1, j)], idx[(i + 1, j + 1)], idx[(i, j + 1)]))
    V = np.array(verts) * radius
    Fc = np.array(faces, dtype=np.int64)
    return V, _orient_outward(V, Fc)


def _orient_outward(V: np.ndarray, F: np.ndarray) -> np.ndarray:
    """Flip faces whose normal points toward the origin (star-shaped meshes)."""
    v0, v1, v2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    n = np.cross(v1 - v0, v2 - v1)
    c = (v0 + v1 + v2) / 3.0
    flip = np.einsum("ij,ij->i", n, c) < 0
    F = F.copy()
    F[flip] = F[flip][:, [0, 2, 1]]
    return F


def _real_sph_harm_basis(theta: np.ndarray, phi: np.ndarray, lmax: int) -> np.ndarray:
    """Real spherical harmonics Y_lm for l <= lmax, [npts, (lmax+1)^2]."""
    from scipy.special import sph_harm_y

    cols = []
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), theta, phi)
            if m < 0:
                cols.append(np.sqrt(2.0) * (-1) ** m * y.imag)
            elif m == 0:
                cols.append(y.real)
            else:
                cols.append(np.sqrt(2.0) * (-1) ** m * y.real)
    return np.stack(cols, axis=1)


def star_mesh(seed: int, freq: int, lmax: int = 6, amp: float = None):
    """Star-shaped closed mesh r = 1 + A s(theta,phi)/max|s| on a geodesic sphere.

    s = sum_{l<=lmax,m} c_lm Y_lm with c_lm ~ N(0,1), A ~ U(0, 0.3) (numpy PCG64).
    Returns (verts, faces, amplitude).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    V, F = geodesic_sphere(freq)
    theta = np.arccos(np.clip(V[:, 2], -1.0, 1.0))
    phi = np.arctan2(V[:, 1], V[:, 0])
    Y = _real_sph_harm_basis(theta, phi, lmax)
    c = rng.standard_normal(Y.shape[1])
    s = Y @ c
    A = rng.uniform(0.0, 0.3) if amp is None else amp
    r = 1.0 + A * s / np.max(np.abs(s))
    V = V * r[:, None]
    return V, _orient_outward(V, F), A


def noisy_ellipsoid(seed: int, freq: int, noise: float = 0.02):
    """Ellipsoid with axes 1:1:c, c ~ U(1,2), plus radial N(0, noise^2) perturbation."""
    rng = np.random.Generator(np.random.PCG64(seed))
    V, F = geodesic_sphere(freq)
    c = rng.uniform(1.0, 2.0)
    V = V * np.array([1.0, 1.0, c])
    nrm = V / np.linalg.norm(V, axis=1, keepdims=True)
    V = V + nrm * rng.normal(0.0, noise, size=(V.shape[0], 1))
    return V, _orient_outward(V, F), c


def morton_order(points: np.ndarray) -> np.ndarray:
    """Permutation sorting points along a 3-D Morton (Z-order) curve."""
    p = np.asarray(points, dtype=np.float64)
    lo = p.min(axis=0)
    span = np.maximum(p.max(axis=0) - lo, 1e-30)
    q = np.clip(((p - lo) / span * 1023.0).astype(np.int64), 0, 1023)

    def spread(x):
        x = x & 0x3FF
        x = (x | (x << 16)) & 0x30000FF
        x = (x | (x << 8)) & 0x300F00F
        x = (x | (x << 4)) & 0x30C30C3
        x = (x | (x << 2)) & 0x9249249
        return x

    code = spread(q[:, 0]) | (spread(q[:, 1]) << 1) | (spread(q[:, 2]) << 2)
    return np.argsort(code, kind="stable")
