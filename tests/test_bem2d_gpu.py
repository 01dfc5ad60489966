"""GPU parity of the 2-D BEM path (bem2d.hip through the C ABI) with the CPU
oracle's restatement of the reference's 2-D solver
(src/mesh_BEM/utils/integrations_2d.py, BEM_utils.py:314-357,
rhs_layer_2d.py, mesh_BEM_solver_2d.py).

Tolerances: FP64 layers 1e-11 relative to the largest entry; FP32 layers
2e-5 (Laplace log kernel and Helmholtz phase in FP32); FP64 solves 1e-8
relative (the single-layer matrices of these curves have condition numbers
up to ~1e4)."""

import numpy as np
import pytest
import torch

import oracle as O
from mesh_bem_amd.bem2d import MeshBEMSolver2d, ellipse_curve, layer2d, pack2d
from mesh_bem_amd.types import BEMType, BoundaryType

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 2e-5)])
@pytest.mark.parametrize("bem_type,k", [(0, 1.0), (1, 7.0), (2, 3.0)])
@pytest.mark.parametrize("layer", [1, 2])
@pytest.mark.parametrize("order_type", [0, 1])
def test_layer_matches_oracle(dev, dtype, tol, bem_type, k, layer, order_type):
    print(f"bem2d layer dtype={dtype} bem={bem_type} layer={layer} order={order_type}", flush=True)
    V, F = ellipse_curve(40, 1.0, 0.55)
    V = V + 0.01 * np.random.default_rng(3).standard_normal(V.shape)
    gaussQR = 6
    gp, gw = O.gl_ref(gaussQR)
    ref = O.bem2d_layer(V, F, gp, gw, layer, order_type, bem_type, k)
    out = layer2d(torch.tensor(V, dtype=dtype, device=dev), torch.tensor(F, device=dev), gaussQR, layer,
                  order_type, bem_type, k).cpu().numpy()
    assert out.shape == ref.shape
    assert _rel(out, ref) < tol


@pytest.mark.parametrize("order_type", [0, 1])
@pytest.mark.parametrize("cplx", [False, True])
def test_pack_matches_oracle(dev, order_type, cplx):
    V, F = ellipse_curve(30, 1.0, 0.7)
    rng = np.random.default_rng(1)
    nphi = 2 if order_type else 1
    U = rng.standard_normal((30, 30, nphi, nphi))
    if cplx:
        U = U + 1j * rng.standard_normal(U.shape)
    out = pack2d(torch.tensor(U, device=dev), torch.tensor(F, device=dev), 30).cpu().numpy()
    assert np.allclose(out, O.bem2d_pack(order_type, U, F, 30), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("bem_type,k", [(0, 1.0), (1, 5.0), (2, 5.0)])
def test_dirichlet_solve_matches_oracle(dev, bem_type, k):
    n = 96
    V, F = ellipse_curve(n, 0.8, 0.5)
    g = V[:, 0] ** 2 - V[:, 1] ** 2 + 0.3 * V[:, 0]
    t_ref, _, _ = O.bem2d_dirichlet(V, F, g, 8, 1, bem_type, k)
    solver = MeshBEMSolver2d(8, 1, bem_type, k, None, dtype=torch.float64, device=dev)
    Vt = torch.tensor(V, dtype=torch.float64, device=dev)
    gt = torch.tensor(g, dtype=torch.float64, device=dev).reshape(1, n, 1)
    flags = torch.full((1, n, 1), int(BoundaryType.DIRICHLET), dtype=torch.uint8, device=dev)
    neu, dirich = solver.solve(Vt, torch.tensor(F, device=dev), gt, torch.zeros_like(gt), flags)
    assert neu.shape == (1, n, 1) and dirich.shape == (1, n, 1)
    t = neu.reshape(-1).cpu().numpy()
    assert np.linalg.norm(t - t_ref) / np.linalg.norm(t_ref) < 1e-8
    assert torch.equal(dirich.reshape(-1).cpu(), gt.reshape(-1).cpu())


def test_demo_circle_known_answer(dev):
    """The reference demo (demo_Laplace_2d.py:100-160) in float32: 256 segments
    of a circle of radius 0.005, gaussQR 8, LINEAR; the Neumann trace of
    x^2 - y^2 is 2 (x^2 - y^2) / R (first-order accurate)."""
    R = 0.005
    V, F = ellipse_curve(256, R, R)
    g = V[:, 0] ** 2 - V[:, 1] ** 2
    solver = MeshBEMSolver2d(8, 1, int(BEMType.LAPLACE), 1.0, None, dtype=torch.float32, device=dev)
    gt = torch.tensor(g, dtype=torch.float32, device=dev).reshape(1, -1, 1)
    flags = torch.zeros_like(gt, dtype=torch.uint8)
    neu, _ = solver.solve(torch.tensor(V, dtype=torch.float32, device=dev), torch.tensor(F, device=dev), gt,
                          torch.zeros_like(gt), flags)
    ex = 2 * g / R
    t = neu.reshape(-1).double().cpu().numpy()
    assert np.linalg.norm(t - ex) / np.linalg.norm(ex) < 0.1


def test_neumann_raises(dev):
    V, F = ellipse_curve(16)
    solver = MeshBEMSolver2d(4, 1, 0, 1.0, None, dtype=torch.float64, device=dev)
    Vt = torch.tensor(V, device=dev)
    flags = torch.full((1, 16, 1), int(BoundaryType.NEUMANN), dtype=torch.uint8, device=dev)
    z = torch.zeros((1, 16, 1), dtype=torch.float64, device=dev)
    with pytest.raises(NotImplementedError):
        solver.solve(Vt, torch.tensor(F, device=dev), z, z, flags)
