"""GPU parity of the pressure step (pressure.hip through the C ABI, with the
Laplace solve of MeshBEMSolver3d.solve) against the oracle's restatement of
MeshPressure.solve (src/mesh_ferrofluid/mesh_pressure.py:54-163).  The
demo's setting: every vertex DIRICHLET (demo_3d_mesh_droplet.py:140-150),
rho = 1300, sigma = 0.02, dt = 5e-4.  Tolerances: FP64 1e-10 relative for the
boundary data, 1e-8 for the velocity update (the Laplace system has condition
~1e3); FP32 boundary data 1e-5."""

import numpy as np
import pytest
import torch

import oracle as O
from mesh_bem_amd.ferrofluid import MeshPressure
from mesh_bem_amd.geometry import icosphere, star_mesh

pytestmark = pytest.mark.gpu


def _mesh(kind):
    if kind == "sphere":
        V, F = icosphere(1)
        return 0.01 * V, F
    V, F = star_mesh(3, 3)[:2]
    return 0.01 * V, F


@pytest.mark.parametrize("kind", ["sphere", "star"])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_boundary_data_matches_oracle(dev, kind, dtype, tol):
    V, F = _mesh(kind)
    rng = np.random.default_rng(0)
    vel = rng.standard_normal((1, len(V), 3)) * 0.01
    pm = rng.standard_normal((1, len(V))) * 0.1
    mp = MeshPressure(1300.0, 9.81, 0.02, 0.5 * np.pi, 4, 1, None, device=dev, dtype=dtype)
    flags = torch.zeros((1, len(V), 1), dtype=torch.uint8, device=dev)
    flags[0, ::5] = 1  # some NEUMANN vertices for the data kernel
    vn, kap, pd, dn = mp.boundary_data(torch.tensor(V, dtype=dtype, device=dev), torch.tensor(F, device=dev), 5e-4,
                                       torch.tensor(vel, dtype=dtype, device=dev), flags,
                                       torch.tensor(pm, dtype=dtype, device=dev))
    vn_ref = O.vertex_normals_area(V, F)
    k_ref = O.signed_mean_curvature_p3d(V, F)
    p_ref = 1300.0 * 9.81 * V[:, 1] - pm[0] + 0.02 * k_ref
    fl = flags.reshape(-1).cpu().numpy()
    p_ref[fl == 1] = 0.0
    dn_ref = (1300.0 / 5e-4) * (vel[0] * vn_ref).sum(1)
    dn_ref[fl == 0] = 0.0
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()
    assert rel(vn.cpu().numpy(), vn_ref) < tol
    assert rel(kap.cpu().numpy(), k_ref) < tol
    assert rel(pd[0].cpu().numpy(), p_ref) < tol
    assert rel(dn[0].cpu().numpy(), dn_ref) < tol


@pytest.mark.parametrize("kind", ["sphere", "star"])
def test_pressure_step_matches_oracle(dev, kind):
    V, F = _mesh(kind)
    rng = np.random.default_rng(1)
    vel = rng.standard_normal((len(V), 3)) * 0.01
    mp = MeshPressure(1300.0, 9.81, 0.02, 0.5 * np.pi, 4, 1, None, device=dev, dtype=torch.float64)
    flags = torch.zeros((1, len(V), 1), dtype=torch.uint8, device=dev)
    out = mp.solve(torch.tensor(V, device=dev), torch.tensor(F, device=dev), 5e-4,
                   torch.tensor(vel[None], device=dev), flags)
    ref, p_ref, dpdn_ref = O.pressure_step(V, F, vel, 5e-4, 1300.0, 9.81, 0.02, 4)
    assert out.shape == (1, len(V), 3)
    assert np.abs(mp.last["p"][0].cpu().numpy() - p_ref).max() / np.abs(p_ref).max() < 1e-10
    assert np.linalg.norm(mp.last["dpdn"][0].cpu().numpy() - dpdn_ref) / np.linalg.norm(dpdn_ref) < 1e-8
    assert np.linalg.norm(out[0].cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-8


def test_triple_junction_and_neumann_raise(dev):
    V, F = _mesh("sphere")
    mp = MeshPressure(1300.0, 9.81, 0.02, 0.5 * np.pi, 2, 1, None, device=dev, dtype=torch.float64)
    vt, ft = torch.tensor(V, device=dev), torch.tensor(F, device=dev)
    vel = torch.zeros((1, len(V), 3), dtype=torch.float64, device=dev)
    flags = torch.zeros((1, len(V), 1), dtype=torch.uint8, device=dev)
    flags[0, 0] = 6
    with pytest.raises(RuntimeError):
        mp.solve(vt, ft, 5e-4, vel, flags)
    flags[0, 0] = 1
    with pytest.raises(NotImplementedError):
        mp.solve(vt, ft, 5e-4, vel, flags)
