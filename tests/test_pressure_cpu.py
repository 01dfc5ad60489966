"""CPU checks of the pressure-step oracle (MeshPressure.solve,
src/mesh_ferrofluid/mesh_pressure.py:54-163): pytorch3d's cot Laplacian and
the reference's signed mean curvature, vertex normals and the tangential
gradient, each against what the formula does not restate:

* the mean curvature of a sphere of radius R: the reference divides the cot
  Laplacian by the full one-ring area (pytorch3d's inv_areas), three times the
  barycentric area, so kappa -> -1/(3R) (sign: the Laplacian points inward);
* linear precision of the cot Laplacian: sum_i L_ij (x_i - x_j) = 0 at interior
  vertices of a planar mesh; cot of an equilateral triangle's angles = 1/sqrt(3);
* the tangential gradient of a linear function a . x is the projection of a on
  each face plane (exact on flat triangles);
* vertex normals of a sphere converge to the radial direction (first order on
  the icosphere's uneven triangles)."""

import math

import numpy as np

import oracle as O
from mesh_bem_amd.geometry import icosphere


def test_sphere_mean_curvature():
    R = 0.37
    spread = []
    for sub in (1, 2, 3):
        V, F = icosphere(sub)
        k = O.signed_mean_curvature_p3d(R * V, F)
        assert abs(k.mean() + 1.0 / (3.0 * R)) < 0.02 / R
        spread.append(k.std())
    assert spread[2] < 0.6 * spread[1] < 0.4 * spread[0]


def _planar_grid(n=6):
    xs = np.linspace(0.0, 1.0, n)
    X, Y = np.meshgrid(xs, xs)
    rng = np.random.default_rng(2)
    V = np.stack([X.ravel(), Y.ravel(), np.zeros(n * n)], 1)
    inner = (V[:, 0] > 0) & (V[:, 0] < 1) & (V[:, 1] > 0) & (V[:, 1] < 1)
    V[inner, :2] += 0.03 * rng.standard_normal((inner.sum(), 2))
    F = []
    for i in range(n - 1):
        for j in range(n - 1):
            a, b, c, d = i * n + j, i * n + j + 1, (i + 1) * n + j, (i + 1) * n + j + 1
            F += [[a, b, d], [a, d, c]]
    return V, np.array(F, dtype=np.int32), inner


def test_cot_laplacian_linear_precision_and_equilateral():
    V, F, inner = _planar_grid()
    L, inv = O.cot_laplacian_p3d(V, F)
    lapx = L.T @ V - L.sum(0)[:, None] * V
    assert np.abs(lapx[inner]).max() < 1e-12
    # equilateral triangle: each cot = 1/sqrt(3), L symmetric, inv = 1 / area
    T = np.array([[0, 0, 0], [1, 0, 0], [0.5, math.sqrt(3) / 2, 0]])
    L3, inv3 = O.cot_laplacian_p3d(T, np.array([[0, 1, 2]]))
    off = L3[~np.eye(3, dtype=bool)]
    assert np.allclose(off, 1 / math.sqrt(3)) and np.allclose(L3, L3.T)
    assert np.allclose(inv3, 1.0 / (math.sqrt(3) / 4))


def test_tangent_gradient_of_linear_function():
    V, F = icosphere(2)
    a = np.array([0.3, -1.2, 0.7])
    g = O.vert_tangent_gradient(V @ a, V, F)
    A, n = O.face_areas_normals(V, F)
    proj = a[None, :] - (n @ a)[:, None] * n  # exact per face
    ref = np.zeros_like(V)
    for t in range(3):
        np.add.at(ref, F[:, t], proj / 3.0)
    assert np.allclose(g, ref, atol=1e-12)


def test_vertex_normals_radial():
    errs = []
    for sub in (2, 3, 4):
        V, F = icosphere(sub)
        vn = O.vertex_normals_area(V, F)
        errs.append(np.abs(vn - V / np.linalg.norm(V, axis=1, keepdims=True)).max())
    # first order on the icosphere's uneven triangles: halves per subdivision
    assert errs[0] < 0.03 and errs[1] < 0.6 * errs[0] and errs[2] < 0.6 * errs[1]
