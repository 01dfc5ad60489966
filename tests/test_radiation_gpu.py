"""Field radiation kernel (mba_radiate_h) against the oracle.

* FP64 on an icosphere (162 vertices) at random points, grid points and a
  point exactly on a vertex (masked pair): 1e-12 relative (norm over all
  points); FP32 at points at least 0.05 from the surface: 2e-5.
* A vertex count that splits into several vertex slices (10 242 vertices,
  512 points): FP64 vs oracle, and two calls give identical bits (the slice
  sums are added in a fixed order).
* The demo's argument shapes (pm [Nm,1,3], Js [V,3], vert_areas [V,1]).
"""

import numpy as np
import pytest
import torch

import oracle as O
from mesh_bem_amd.geometry import icosphere
from mesh_bem_amd.radiation import radiate_H, xz_grid

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _case(subdiv, n_pts, seed):
    V, F = icosphere(subdiv)
    rng = np.random.default_rng(seed)
    J = rng.normal(size=(len(V), 3)) + 1j * rng.normal(size=(len(V), 3))
    A = O.vert_areas(V, F)
    P = rng.normal(size=(n_pts, 3)) * 1.5
    return V, J, A, P


def test_fp64_and_fp32_match_oracle(dev):
    V, J, A, P = _case(2, 300, 5)
    grid = xz_grid(16, device="cpu").numpy()
    P = np.concatenate([P, grid, V[7:8]])
    k = 4.0
    ref = O.radiate_H(k, P, V, J, A)
    T = lambda a, dt: torch.tensor(a, dtype=dt, device=dev)
    H64 = radiate_H(k, T(P, torch.float64), T(V, torch.float64), T(J, torch.complex128), T(A, torch.float64))
    e64 = rel(H64.cpu().numpy(), ref)
    far = np.min(np.linalg.norm(P[:, None, :] - V[None], axis=-1), axis=1) > 0.05
    H32 = radiate_H(k, T(P[far], torch.float32), T(V, torch.float32), T(J, torch.complex64), T(A, torch.float32))
    e32 = rel(H32.cpu().numpy(), ref[far])
    print(f"  FP64 {e64:.2e}  FP32 {e32:.2e}")
    assert e64 < 1e-12 and e32 < 2e-5


def test_sliced_sum_matches_oracle_and_is_deterministic(dev):
    V, J, A, P = _case(5, 512, 6)
    k = 10.0
    sel = np.arange(0, 512, 8)
    ref = O.radiate_H(k, P[sel], V, J, A)
    T = lambda a, dt: torch.tensor(a, dtype=dt, device=dev)
    args = (T(P, torch.float64), T(V, torch.float64), T(J, torch.complex128), T(A, torch.float64))
    H1 = radiate_H(k, *args)
    H2 = radiate_H(k, *args)
    assert torch.equal(H1, H2)
    e = rel(H1.cpu().numpy()[sel], ref)
    print(f"  sliced FP64 {e:.2e}")
    assert e < 1e-12


def test_demo_shapes(dev):
    V, J, A, P = _case(1, 40, 7)
    pm = torch.tensor(P, device=dev).reshape(-1, 1, 3)
    H = radiate_H(2.0, pm, torch.tensor(V, device=dev), torch.tensor(J, device=dev),
                  torch.tensor(A, device=dev).reshape(-1, 1))
    assert H.shape == (40, 3) and H.dtype == torch.complex128
    assert rel(H.cpu().numpy(), O.radiate_H(2.0, P, V, J, A)) < 1e-12
