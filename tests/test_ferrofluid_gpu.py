"""Surface Helmholtz-decomposition kernel (mba_helmholtz_decomp) and the
drop-in MeshHelmholtzDecomposition.solve against the oracle.

* FP64 at the vertices for two velocity fields (icosphere, 642 vertices):
  1e-12 relative; FP32: 1e-5.
* A size that splits the sources into slices (2 562 vertices, 81 920 source
  points): FP64 vs oracle on a subset, bit-identical repeat calls.
* Interior reproduction through the kernel (constant field): 1e-6.
* new_vel of solve() vs the oracle's (FP64): 1e-12.
"""

import numpy as np
import pytest
import torch

import oracle as O
from mesh_bem_amd.ferrofluid import MeshHelmholtzDecomposition, decomposition_velocity
from mesh_bem_amd.geometry import icosphere

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_vertices_fp64_fp32(dev):
    V, F = icosphere(3)
    rng = np.random.default_rng(3)
    vel = rng.normal(size=(2, len(V), 3))
    ref = O.helmholtz_decomposition_u(V, F, vel, 4)
    T = lambda a, dt: torch.tensor(a, dtype=dt, device=dev)
    Fi = torch.tensor(F, dtype=torch.int32, device=dev)
    u64 = decomposition_velocity(T(V, torch.float64), Fi, T(vel, torch.float64), 4)
    u32 = decomposition_velocity(T(V, torch.float32), Fi, T(vel, torch.float32), 4)
    e64, e32 = rel(u64.cpu().numpy(), ref), rel(u32.cpu().numpy(), ref)
    print(f"  FP64 {e64:.2e}  FP32 {e32:.2e}")
    assert e64 < 1e-12 and e32 < 1e-5


def test_sliced_and_deterministic(dev):
    V, F = icosphere(4)
    rng = np.random.default_rng(4)
    vel = rng.normal(size=(1, len(V), 3))
    sel = np.arange(0, len(V), 20)
    ref = O.helmholtz_decomposition_u(V, F, vel, 4, targets=V[sel], block=8)
    args = (torch.tensor(V, device=dev), torch.tensor(F, dtype=torch.int32, device=dev), torch.tensor(vel, device=dev), 4)
    u1 = decomposition_velocity(*args)
    u2 = decomposition_velocity(*args)
    assert torch.equal(u1, u2)
    e = rel(u1.cpu().numpy()[:, sel], ref)
    print(f"  sliced FP64 {e:.2e}")
    assert e < 1e-12


def test_interior_reproduction(dev):
    V, F = icosphere(2)
    rng = np.random.default_rng(5)
    X = rng.normal(size=(32, 3))
    X = X / np.linalg.norm(X, axis=1, keepdims=True) * rng.uniform(0, 0.5, (32, 1))
    U = np.array([0.3, -1.0, 0.5])
    u = decomposition_velocity(torch.tensor(V, device=dev), torch.tensor(F, dtype=torch.int32, device=dev),
                               torch.tensor(np.tile(U, (len(V), 1)), device=dev), 4, targets=torch.tensor(X, device=dev))
    assert np.abs(u[0].cpu().numpy() - U).max() < 1e-6


def test_solve_dropin(dev):
    V, F = icosphere(2)
    rng = np.random.default_rng(6)
    vel = rng.normal(size=(1, len(V), 3))
    hd = MeshHelmholtzDecomposition(gaussQR=4, device=dev, dtype=torch.float64)
    out = hd.solve(torch.tensor(V, device=dev), torch.tensor(F, dtype=torch.int64, device=dev), torch.tensor(vel, device=dev))
    assert out.shape == (1, len(V), 3)
    assert rel(out.cpu().numpy(), O.helmholtz_decomposition_solve(V, F, vel, 4)) < 1e-12
