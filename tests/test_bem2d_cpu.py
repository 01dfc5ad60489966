"""CPU checks of the 2-D BEM oracle (oracle.bem2d_*: the reference's
src/mesh_BEM/utils/integrations_2d.py, BEM_utils.py:9-138 and 314-357,
rhs_layer_2d.py, mesh_BEM_solver_2d.py), each against something the formula
does not restate:

* Gauss's identity: for a closed polygon, sum over test segments of the
  double layer of a source segment is -L/2 (outward normals);
* the SAME_EDGE map's closed form: the reference integrates the half square
  y < x of a segment with itself, (L^2 / 2 pi)(3/4 - log(L) / 2) (PLANAR);
* LINEAR shape functions sum to one: summing the LINEAR block over (iy, ix)
  gives the PLANAR entry;
* the Helmholtz kernel against scipy's Hankel function: the reference uses the
  large-argument form of H^(2), so G = -(j/4) H_0^(2)(k d) only to O(1 / k d);
  grad_y G against central differences of G in y to the same order;
* the reference demo's known answer (demo_Laplace_2d.py:39-78): Dirichlet data
  x^2 - y^2 on a circle of radius R gives the Neumann trace 2 (x^2 - y^2) / R,
  approached at first order in h by the LINEAR solve.
"""

import math

import numpy as np
import pytest

import oracle as O
from mesh_bem_amd.bem2d import ellipse_curve


def test_double_layer_gauss_identity():
    V, F = ellipse_curve(64, 1.0, 0.6)
    gp, gw = O.gl_ref(8)
    K = O.bem2d_layer(V, F, gp, gw, 2, 0)
    L, _ = O.seg_areas_normals(V, F)
    col = K[:, :, 0, 0].sum(axis=0)
    assert np.allclose(col, -0.5 * L, rtol=2e-3)
    # a flipped orientation flips the sign
    K2 = O.bem2d_layer(V, F[:, ::-1], gp, gw, 2, 0)
    assert np.allclose(K2[:, :, 0, 0].sum(axis=0), 0.5 * L, rtol=2e-3)


@pytest.mark.parametrize("Lscale", [0.01, 0.3, 2.0])
def test_self_term_closed_form(Lscale):
    V, F = ellipse_curve(32, Lscale, Lscale)
    gp, gw = O.gl_ref(12)
    Vs = O.bem2d_layer(V, F, gp, gw, 1, 0)
    L, _ = O.seg_areas_normals(V, F)
    l = L[0]
    exact = (l * l / (2 * math.pi)) * (0.75 - 0.5 * math.log(l))
    assert abs(Vs[0, 0, 0, 0] - exact) <= 2e-3 * abs(exact) + 1e-3 * l * l


@pytest.mark.parametrize("layer", [1, 2])
@pytest.mark.parametrize("bem_type", [0, 1])
def test_linear_sums_to_planar(layer, bem_type):
    V, F = ellipse_curve(20, 1.0, 0.5)
    gp, gw = O.gl_ref(6)
    P = O.bem2d_layer(V, F, gp, gw, layer, 0, bem_type, 3.0)
    Lm = O.bem2d_layer(V, F, gp, gw, layer, 1, bem_type, 3.0)
    assert np.allclose(Lm.sum(axis=(2, 3)), P[:, :, 0, 0], rtol=1e-12, atol=1e-14)


def test_helmholtz_kernel_vs_scipy():
    scipy_special = pytest.importorskip("scipy.special")
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (50, 2))
    y = rng.uniform(-1, 1, (50, 2))
    k = 200.0
    d = np.linalg.norm(x - y, axis=1)
    G = O.green2d(1, k, x, y)
    exact = -0.25j * scipy_special.hankel2(0, k * d)
    assert np.all(np.abs(G - exact) <= 0.3 / (k * d) * np.abs(exact))
    # grad_y G against central differences in y
    h = 1e-6
    fd = np.zeros((50, 2), dtype=complex)
    for c in range(2):
        e = np.zeros(2)
        e[c] = h
        fd[:, c] = (O.green2d(1, k, x, y + e) - O.green2d(1, k, x, y - e)) / (2 * h)
    gr = O.grad_green2d_y(1, k, x, y)
    assert np.all(np.linalg.norm(gr - fd, axis=1) <= 1.0 / (k * d) * np.linalg.norm(fd, axis=1))
    # Laplace: the gradient is exact
    fdL = np.stack([(O.green2d(0, 1, x, y + np.array([h, 0])) - O.green2d(0, 1, x, y - np.array([h, 0]))) / (2 * h),
                    (O.green2d(0, 1, x, y + np.array([0, h])) - O.green2d(0, 1, x, y - np.array([0, h]))) / (2 * h)], 1)
    assert np.allclose(O.grad_green2d_y(0, 1, x, y), fdL, rtol=1e-6, atol=1e-8)


def test_demo_known_answer_converges():
    R = 0.005
    errs = []
    for n in (64, 128, 256):
        V, F = ellipse_curve(n, R, R)
        g = V[:, 0] ** 2 - V[:, 1] ** 2
        t, _, _ = O.bem2d_dirichlet(V, F, g, 8, 1)
        ex = 2.0 * g / R
        errs.append(np.linalg.norm(t - ex) / np.linalg.norm(ex))
    assert errs[0] < 0.45 and errs[2] < 0.1
    assert errs[1] < 0.6 * errs[0] and errs[2] < 0.6 * errs[1]


def test_pack_planar_spreads_quarters():
    V, F = ellipse_curve(6, 1.0, 1.0)
    U = np.arange(36, dtype=float).reshape(6, 6, 1, 1)
    A = O.bem2d_pack(0, U, F, 6)
    assert np.isclose(A.sum(), U.sum())
    # vertex 0 belongs to segments 5 (slot 1) and 0 (slot 0)
    assert np.isclose(A[0, 0], (U[0, 0] + U[0, 5] + U[5, 0] + U[5, 5])[0, 0] / 4)
