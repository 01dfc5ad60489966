"""Oracle of the surface Helmholtz decomposition
(src/mesh_ferrofluid/mesh_Hemholtz_decomposition.py:52-143), pinned by the
Helmholtz representation: for a field that is divergence- and curl-free
inside a closed surface, sum W (n . u) grad G + W (n x u) x grad G over the
surface reproduces the field at interior points (u = grad phi - curl A with
phi, A the single layers of n . u and n x u).  A constant field and the
linear harmonic gradient (2x, -2y, 0) are interpolated exactly by the P1
vertex data, so the only error is the Gauss rule's (the points stay 0.5 from
the surface).  Also: the product's tensor rule equals the oracle's.
"""

import numpy as np
import torch

import oracle as O
from mesh_bem_amd.ferrofluid import tensor_rule
from mesh_bem_amd.geometry import icosphere


def _interior(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 3))
    return X / np.linalg.norm(X, axis=1, keepdims=True) * rng.uniform(0, 0.5, (n, 1))


def test_reproduces_constant_and_harmonic_fields_inside():
    V, F = icosphere(2)
    X = _interior(16, 0)
    U = np.array([0.3, -1.0, 0.5])
    u = O.helmholtz_decomposition_u(V, F, np.tile(U, (len(V), 1)), 4, targets=X)[0]
    assert np.abs(u - U).max() < 1e-6
    vel = np.stack([2 * V[:, 0], -2 * V[:, 1], 0 * V[:, 0]], 1)
    ex = np.stack([2 * X[:, 0], -2 * X[:, 1], 0 * X[:, 0]], 1)
    u2 = O.helmholtz_decomposition_u(V, F, vel, 4, targets=X)[0]
    assert np.abs(u2 - ex).max() < 1e-6 * np.abs(ex).max()


def test_tangential_part_kept():
    """new_vel keeps the tangential velocity and replaces the normal part."""
    V, F = icosphere(1)
    rng = np.random.default_rng(1)
    vel = rng.normal(size=(1, len(V), 3))
    out = O.helmholtz_decomposition_solve(V, F, vel, 3)[0]
    vn = O.vert_normals(V, F)
    tan = lambda a: a - (a * vn).sum(-1, keepdims=True) * vn
    assert np.allclose(tan(out), tan(vel[0]), rtol=0, atol=1e-12)


def test_tensor_rule_matches_oracle():
    import torch

    for Q in (2, 3, 4):
        R = tensor_rule(Q, dtype=torch.float64, device="cpu").numpy()
        p, w = O.gl_ref(Q)
        assert np.array_equal(R[:, 0], np.tile(p[None, :], (Q, 1)).reshape(-1))
        assert np.array_equal(R[:, 1], (p[:, None] * p[None, :]).reshape(-1))
        assert np.allclose(R[:, 2], (w[:, None] * w[None, :]).reshape(-1), rtol=1e-15, atol=0)


def test_magnetic_pressure_ellipsoid_closed_form():
    """P_mag of the surface charge of a magnetised ellipsoid (oracle P0 solve):
    inside, H is uniform, H0 / (1 + N_z chi), so P_mag = chi |H_in|^2 / 2
    (chi n_z^2 + 1) at every surface point (mu0 = 1).  Converges at O(h);
    dropping the tangential field or the chi |H|^2 term fails."""
    import oracle as O
    from mesh_bem_amd.geometry import geodesic_sphere
    from mesh_bem_amd.magnetic import magnetic_pressure

    c, chi = 1.5, 3.0
    V, F = geodesic_sphere(10)
    V = V * np.array([1.0, 1.0, c])
    ptr, idx = O.near_pairs_bruteforce(V, F)
    Vo, Ko = O.laplace_p0(V, F, ptr, idx)
    A, n = O.face_areas_normals(V, F)
    sig = np.linalg.solve((1 / chi + 0.5) * np.diag(A) + Ko, A * n[:, 2])
    phi = (Vo @ sig) / A
    pm = magnetic_pressure(torch.tensor(V), torch.tensor(F), torch.tensor(sig), torch.tensor(phi), chi).numpy()
    _, _, Nc = O.demag_factors(1.0, 1.0, c)
    hin = 1.0 / (1.0 + Nc * chi)
    ne = V / np.array([1.0, 1.0, c * c])
    ne /= np.linalg.norm(ne, axis=1, keepdims=True)
    ref = 0.5 * chi * hin ** 2 * (chi * ne[:, 2] ** 2 + 1.0)
    assert np.linalg.norm(pm - ref) / np.linalg.norm(ref) < 0.025
    # planted: no tangential field (the potential taken as the applied field's own)
    pm_bad = magnetic_pressure(torch.tensor(V), torch.tensor(F), torch.tensor(sig),
                               torch.tensor((V[F].mean(1)[:, 2])), chi).numpy()
    assert np.linalg.norm(pm_bad - ref) / np.linalg.norm(ref) > 0.2
