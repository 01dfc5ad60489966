"""GPU nearest-vertex velocity transfer (mba_nearest_transfer, the step
mesh_utils.py:800-815 of remesh_and_transfer_velocity) against a numpy
brute-force argmin over coordinate-difference distances (first index on
ties), and the remesh drop-in end to end."""

import numpy as np
import pytest
import torch

from mesh_bem_amd.geometry import icosphere
from mesh_bem_amd.remesh import remesh_and_transfer_velocity, transfer_velocity_nearest

pytestmark = pytest.mark.gpu


def _brute(p, q):
    d = ((p[:, None, :] - q[None, :, :]) ** 2).sum(-1)
    return d.argmin(axis=1)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("np_,nq,batch", [(1, 1, 1), (300, 77, 2), (5000, 3001, 3), (700, 9000, 0)])
def test_nearest_transfer_matches_brute_force(dev, dtype, np_, nq, batch):
    print(f"nearest transfer {dtype} np={np_} nq={nq} B={batch}", flush=True)
    rng = np.random.default_rng(np_ + nq)
    q = rng.standard_normal((nq, 3))
    p = rng.standard_normal((np_, 3))
    if nq > 10:
        q[5] = q[3]  # an exact duplicate: the smaller index must win
        p[0] = q[3]
    q = q.astype(np.float32 if dtype == torch.float32 else np.float64)
    p = p.astype(q.dtype)
    vel = rng.standard_normal((batch, nq, 3)).astype(q.dtype)
    out, nearest = transfer_velocity_nearest(torch.tensor(p, device=dev), torch.tensor(q, device=dev),
                                             torch.tensor(vel, device=dev))
    ref = _brute(p.astype(np.float64), q.astype(np.float64)) if dtype == torch.float64 else _brute(p, q)
    got = nearest.cpu().numpy()
    if dtype == torch.float64:
        assert np.array_equal(got, ref)
    else:
        # float32 distances: an index may differ only where the two distances tie in float32
        d = lambda i, j: ((p[i].astype(np.float64) - q[j]) ** 2).sum()  # noqa: E731
        for i in np.flatnonzero(got != ref):
            assert abs(d(i, got[i]) - d(i, ref[i])) <= 1e-5 * max(d(i, ref[i]), 1e-30)
    if nq > 10:
        assert got[0] == 3
    assert out.shape == (batch, np_, 3)
    if batch:
        assert np.array_equal(out.cpu().numpy(), vel[:, got, :])


def test_remesh_drop_in(dev):
    v, f = icosphere(1)
    verts = torch.tensor(np.asarray(v), dtype=torch.float32, device=dev)
    faces = torch.tensor(np.asarray(f, dtype=np.int32), device=dev)
    vel = torch.randn((2, verts.shape[0], 3), dtype=torch.float32, device=dev)
    nv, nf, nvel = remesh_and_transfer_velocity(verts, faces, vel, target_edge_length=0.1)
    # icosphere(1) mean edge ~0.55: three passes (the reference's cap), 80 * 4^3
    # faces, then the decimation to 90 % of them
    target = int(80 * 64 * 0.9)
    assert target - 2 <= nf.shape[0] <= target and nf.shape[1] == 3
    assert nv.shape[1] == 3 and nvel.shape == (2, nv.shape[0], 3)
    # Loop's edge vertices sit on the bisector of their end points, so exact
    # ties are common: the chosen vertex must be at the minimal distance (to
    # float32 rounding) and the velocity must be its velocity
    p = nv.cpu().numpy().astype(np.float64)
    q = verts.cpu().numpy().astype(np.float64)
    d = ((p[:, None, :] - q[None, :, :]) ** 2).sum(-1)
    _, nearest = transfer_velocity_nearest(nv, verts, vel[:, :, :])
    got = nearest.cpu().numpy()
    dmin = d.min(axis=1)
    assert np.all(d[np.arange(len(p)), got] <= dmin * (1 + 1e-5) + 1e-12)
    assert np.array_equal(nvel.cpu().numpy(), vel.cpu().numpy()[:, got, :])
