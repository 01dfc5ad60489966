"""Loop subdivision (host C++, the refinement step of
remesh_and_transfer_velocity, src/utils/mesh_utils.py:728-815).

open3d (the reference's implementation) is absent, so parity with it is not
established; Loop's masks are pinned by what the mathematics fixes: counts
and topology (V + E vertices, 4F faces, two faces per edge, Euler
characteristic, orientation), a regular planar grid reproduced (interior
vertices fixed, interior edge vertices at midpoints, a straight boundary kept),
the icosahedron's symmetric closed forms (|v'| of an old vertex, edge
vertices on the edge's symmetry axis) and the boundary masks on one triangle.
"""

import math

import numpy as np
import pytest

from mesh_bem_amd import _lib
from mesh_bem_amd.geometry import icosphere
from mesh_bem_amd.remesh import average_edge_length, loop_subdivide


def _edges(f):
    e = np.sort(np.vstack([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    u, c = np.unique(e, axis=0, return_counts=True)
    return u, c


def test_icosahedron_counts_topology_and_closed_forms():
    v, f = icosphere(0)
    v = np.asarray(v, dtype=np.float64)
    f = np.asarray(f, dtype=np.int32)
    v = v / np.linalg.norm(v, axis=1, keepdims=True)
    E = len(_edges(f)[0])
    nv, nf = loop_subdivide(v, f)
    assert nv.shape == (len(v) + E, 3) and nf.shape == (4 * len(f), 3)
    u, c = _edges(nf)
    assert np.all(c == 2)  # closed 2-manifold
    assert len(nv) - len(u) + len(nf) == 2  # Euler characteristic of the sphere
    # orientation: every directed edge once
    d = np.vstack([nf[:, [0, 1]], nf[:, [1, 2]], nf[:, [2, 0]]])
    assert len(np.unique(d, axis=0)) == len(d)
    # old vertex, valence 5: sum of neighbours = sqrt(5) v for the unit icosahedron
    beta = (0.625 - (0.375 + 0.25 * math.cos(2 * math.pi / 5)) ** 2) / 5
    r_old = 1 - 5 * beta + math.sqrt(5) * beta
    assert np.allclose(np.linalg.norm(nv[: len(v)], axis=1), r_old, atol=1e-14)
    assert np.allclose(nv[: len(v)] / r_old, v, atol=1e-14)
    # edge vertices (one per old unique edge, ascending) lie on the edge's
    # two-fold axis: parallel to v0 + v1
    old_edges = _edges(f)[0]
    for i, (a, b) in enumerate(old_edges):
        w = nv[len(v) + i]
        s = v[a] + v[b]
        assert np.linalg.norm(np.cross(w, s)) < 1e-14 * np.linalg.norm(s)


def test_planar_grid_reproduced():
    n = 6
    xs, ys = np.meshgrid(np.arange(n + 1, dtype=np.float64), np.arange(n + 1, dtype=np.float64), indexing="ij")
    v = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)], axis=1)
    idx = lambda i, j: i * (n + 1) + j  # noqa: E731
    f = []
    for i in range(n):
        for j in range(n):
            f.append([idx(i, j), idx(i + 1, j), idx(i + 1, j + 1)])
            f.append([idx(i, j), idx(i + 1, j + 1), idx(i, j + 1)])
    f = np.array(f, dtype=np.int32)
    nv, nf = loop_subdivide(v, f)
    assert np.all(nv[:, 2] == 0.0)  # stays planar
    # every interior vertex has valence 6 on this grid: b = 1/16, the old
    # vertex is the mean of its ring -> unchanged
    inner = [idx(i, j) for i in range(1, n) for j in range(1, n)]
    assert np.allclose(nv[inner], v[inner], atol=1e-13)
    # straight boundaries: non-corner boundary vertices stay (3/4 v + 1/8 (b0 + b1))
    side = [idx(0, j) for j in range(1, n)] + [idx(n, j) for j in range(1, n)]
    assert np.allclose(nv[side], v[side], atol=1e-13)
    # every edge vertex of this regular grid is the edge midpoint (interior: the
    # two opposite vertices are symmetric about it; boundary: midpoint rule)
    old_edges = _edges(f)[0]
    mid = 0.5 * (v[old_edges[:, 0]] + v[old_edges[:, 1]])
    assert np.allclose(nv[len(v):], mid, atol=1e-13)
    assert average_edge_length(nv, nf) < average_edge_length(v, f)


def test_single_triangle_boundary_masks():
    v = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.3, 0.8, 0.1]])
    f = np.array([[0, 1, 2]], dtype=np.int32)
    nv, nf = loop_subdivide(v, f)
    for a in range(3):
        others = [b for b in range(3) if b != a]
        assert np.allclose(nv[a], 0.75 * v[a] + 0.125 * (v[others[0]] + v[others[1]]), atol=1e-15)
    old_edges = _edges(f)[0]
    assert np.allclose(nv[3:], 0.5 * (v[old_edges[:, 0]] + v[old_edges[:, 1]]), atol=1e-15)
    assert nf.shape == (4, 3)


def test_invalid_faces_raise():
    v = np.zeros((3, 3))
    with pytest.raises(_lib.MBAError):
        loop_subdivide(v, np.array([[0, 1, 3]], dtype=np.int32))
    with pytest.raises(_lib.MBAError):
        loop_subdivide(v, np.array([[0, 1, 1]], dtype=np.int32))


def _check_closed_manifold(v, f, chi):
    u, c = _edges(f)
    assert np.all(c == 2)
    assert len(v) - len(u) + len(f) == chi
    d = np.vstack([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    assert len(np.unique(d, axis=0)) == len(d)  # consistent orientation
    a = 0.5 * np.linalg.norm(np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]), axis=1)
    assert np.all(a > 0)


def test_decimation_of_a_sphere_keeps_a_closed_manifold_near_the_sphere():
    """Quadric decimation (Garland-Heckbert; open3d, the reference's
    implementation, is absent): the target is reached, the closed 2-manifold,
    its Euler characteristic and orientation survive, and the collapsed
    vertices stay close to the sphere the mesh samples."""
    from mesh_bem_amd.remesh import decimate_quadric

    v, f = icosphere(3)
    v = np.asarray(v, dtype=np.float64)
    f = np.asarray(f, dtype=np.int32)
    target = int(len(f) * 0.9)
    dv, df = decimate_quadric(v, f, target)
    assert target - 2 <= len(df) <= target
    _check_closed_manifold(dv, df, 2)
    r = np.linalg.norm(dv, axis=1)
    assert np.all(np.abs(r - 1.0) < 0.02)
    # the surviving vertices that never moved are exactly input vertices
    moved = ~np.any(np.all(np.isclose(dv[:, None, :], v[None, :, :], atol=0, rtol=0), axis=2), axis=1)
    assert moved.sum() <= (len(f) - len(df)) // 2  # at most one moved vertex per collapse
    # nothing to do when the target is already met
    sv, sf = decimate_quadric(v, f, len(f))
    assert np.array_equal(sv, v) and np.array_equal(sf, f)


def test_decimation_of_a_plane_stays_planar_and_keeps_the_boundary():
    from mesh_bem_amd.remesh import decimate_quadric

    n = 8
    xs, ys = np.meshgrid(np.arange(n + 1, dtype=np.float64), np.arange(n + 1, dtype=np.float64), indexing="ij")
    v = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)], axis=1)
    idx = lambda i, j: i * (n + 1) + j  # noqa: E731
    f = []
    for i in range(n):
        for j in range(n):
            f.append([idx(i, j), idx(i + 1, j), idx(i + 1, j + 1)])
            f.append([idx(i, j), idx(i + 1, j + 1), idx(i, j + 1)])
    f = np.array(f, dtype=np.int32)
    dv, df = decimate_quadric(v, f, len(f) // 2)
    assert len(df) <= len(f) // 2
    assert np.all(dv[:, 2] == 0.0)  # zero-error collapses stay in the plane
    # boundary vertices are kept in place: all 4n of them survive
    on_boundary = (dv[:, 0] == 0) | (dv[:, 0] == n) | (dv[:, 1] == 0) | (dv[:, 1] == n)
    assert on_boundary.sum() == 4 * n
    # the decimated mesh still covers the square: areas add up to n^2
    a = 0.5 * np.abs(np.cross(dv[df[:, 1]] - dv[df[:, 0]], dv[df[:, 2]] - dv[df[:, 0]])[:, 2])
    assert abs(a.sum() - n * n) < 1e-9


def test_shapes_are_checked_before_any_native_call():
    """Wrong operand shapes raise ValueError in the host layer, before the
    native code would read past an array."""
    import torch

    from mesh_bem_amd.remesh import decimate_quadric, loop_subdivide, transfer_velocity_nearest

    v = np.zeros((4, 3))
    f = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    with pytest.raises(ValueError):
        decimate_quadric(v.reshape(-1), f, 1)
    with pytest.raises(ValueError):
        decimate_quadric(v, f[:, :2], 1)
    with pytest.raises(ValueError):
        loop_subdivide(v[:, :2], f)
    with pytest.raises(ValueError):
        transfer_velocity_nearest(torch.zeros(5, 3), torch.zeros(4, 3), torch.zeros(1, 3, 3))
    with pytest.raises(ValueError):
        transfer_velocity_nearest(torch.zeros(5, 2), torch.zeros(4, 3), torch.zeros(1, 4, 3))
