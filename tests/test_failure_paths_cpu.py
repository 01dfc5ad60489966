"""Failure paths forced from the host, without a GPU: malformed meshes are
rejected before any device work (index out of range, zero-area face,
non-manifold edge, inconsistent orientation -- the RWG sign convention of
src/mesh_BEM/cuda/mesh_MOM_3d.cu:154-165 needs one T+ and one T- per edge);
network checkpoints and generated triples round-trip through files that load
without executing anything (torch.load(weights_only=True), np.load without
pickles)."""

import numpy as np
import pytest
import torch

from mesh_bem_amd.geometry import icosphere
from mesh_bem_amd.mgn import MeshGraphNet, Sample, load_samples, save_samples
from mesh_bem_amd.mom import MomMesh


def build(V, F):
    return MomMesh.build(torch.tensor(V, dtype=torch.float64), np.asarray(F), device="cpu")


def test_good_mesh_builds():
    V, F = icosphere(1)
    m = build(V, F)
    assert m.n_edges == 3 * len(F) // 2


def test_inconsistent_orientation_rejected():
    V, F = icosphere(1)
    F = np.array(F)
    F[5] = F[5][::-1]
    with pytest.raises(ValueError, match="oriented"):
        build(V, F)


def test_zero_area_face_rejected():
    V, F = icosphere(1)
    V = np.array(V)
    F = np.array(F)
    a, b, c = F[0]
    V[c] = 0.5 * (V[a] + V[b])  # collinear corners
    with pytest.raises(ValueError, match="zero-area"):
        build(V, F)


def test_index_out_of_range_and_non_manifold_rejected():
    V, F = icosphere(1)
    F = np.array(F)
    bad = F.copy()
    bad[3, 1] = len(V)
    with pytest.raises(ValueError):
        build(V, bad)
    with pytest.raises(ValueError, match="manifold"):
        build(V, F[:-1])  # an open surface: edges with one face


def test_network_checkpoint_round_trip(tmp_path):
    net = MeshGraphNet(hidden=16, layers=2, device="cpu", seed=3)
    net.m.normal_()
    net.v.uniform_()
    net.step_count = 17
    p = str(tmp_path / "net.pt")
    net.save(p, {"epochs_done": 4, "history": torch.tensor([1.0, 0.5])})
    other = MeshGraphNet(hidden=16, layers=2, device="cpu", seed=9)
    extra = other.load(p)
    assert torch.equal(other.params, net.params) and torch.equal(other.m, net.m) and torch.equal(other.v, net.v)
    assert other.step_count == 17 and int(extra["epochs_done"]) == 4
    with pytest.raises(ValueError):
        MeshGraphNet(hidden=32, layers=2, device="cpu").load(p)


def test_generated_triples_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    smp = [Sample(i, 1.5 + i, rng.normal(size=(3, 3)), rng.normal(size=(3, 3)),
                  torch.randn(3, 10, 7), torch.randn(3, 10, 2)) for i in range(3)]
    p = str(tmp_path / "shard.npz")
    save_samples(p, smp)
    back = load_samples(p)
    for a, b in zip(smp, back):
        assert a.mesh == b.mesh and a.k == b.k
        assert np.array_equal(a.dirs, b.dirs) and np.array_equal(a.pols, b.pols)
        assert torch.equal(a.x, b.x) and torch.equal(a.J, b.J)


def test_to_c64_checks_dtype_before_launch():
    """The rounding kernel reads 16-byte elements: a complex64 matrix is
    copied as is and anything but complex128 / complex64 is rejected, both
    before any library call (so this runs without a GPU)."""
    import torch

    from mesh_bem_amd.dense import to_c64

    z = torch.arange(6, dtype=torch.float32).reshape(2, 3).to(torch.complex64)
    c = to_c64(z)
    assert c.dtype == torch.complex64 and torch.equal(c, z) and c.data_ptr() != z.data_ptr()
    with pytest.raises(ValueError):
        to_c64(torch.zeros((2, 2), dtype=torch.float64))
    with pytest.raises(ValueError):
        to_c64(torch.zeros(4, dtype=torch.complex128))
