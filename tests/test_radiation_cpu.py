"""Oracle of the field radiation (demo/demo_Helmholtz_MOM_3d.py:246-257) pinned
by what the mathematics fixes, independently of the expression it restates:

* one vertex current is a Hertzian dipole: H = (A J x r^) (jk + 1/r) e^{-jkr} / (4 pi r);
* gradG_y (src/utils/functions.py:844-868) is the gradient of G with respect to
  the source point (central finite differences of G);
* superposition over vertices, and a point on a vertex drops that vertex only
  (the R < 1e-6 mask).
"""

import math

import numpy as np

import oracle as O


def test_single_vertex_is_hertzian_dipole():
    rng = np.random.default_rng(1)
    for k in (0.5, 3.0, 10.0):
        v = rng.normal(size=(1, 3))
        J = rng.normal(size=(1, 3)) + 1j * rng.normal(size=(1, 3))
        A = np.array([0.37])
        P = rng.normal(size=(20, 3)) * 2
        H = O.radiate_H(k, P, v, J, A)
        d = P - v
        r = np.linalg.norm(d, axis=1, keepdims=True)
        Hd = np.cross(A[0] * J, d / r) * (1j * k + 1 / r) * np.exp(-1j * k * r) / (4 * math.pi * r)
        assert np.abs(H - Hd).max() <= 1e-13 * np.abs(Hd).max()


def test_gradG_y_is_source_gradient():
    rng = np.random.default_rng(2)
    h = 1e-6
    for k in (1.0, 7.0):
        p1 = rng.normal(size=(1, 3))
        p2 = rng.normal(size=(1, 3))
        g = O.green_gradG_y(k, p1, p2)[0]
        fd = np.array([(O.green_G(k, p1, p2 + h * e) - O.green_G(k, p1, p2 - h * e)) / (2 * h) for e in np.eye(3)])
        assert np.abs(g - fd[:, 0]).max() <= 1e-8 * np.abs(g).max()


def test_superposition_and_mask():
    rng = np.random.default_rng(3)
    V = rng.normal(size=(7, 3))
    J = rng.normal(size=(7, 3)) + 1j * rng.normal(size=(7, 3))
    A = rng.uniform(0.1, 1.0, 7)
    P = np.concatenate([rng.normal(size=(5, 3)) * 3, V[2:3]])  # last point sits on vertex 2
    H = O.radiate_H(2.0, P, V, J, A)
    parts = sum(O.radiate_H(2.0, P, V[i:i + 1], J[i:i + 1], A[i:i + 1]) for i in range(7))
    assert np.allclose(H, parts, rtol=1e-13, atol=0)
    keep = [i for i in range(7) if i != 2]
    assert np.allclose(H[-1], O.radiate_H(2.0, P[-1:], V[keep], J[keep], A[keep])[0], rtol=1e-13, atol=0)
