"""GPU edge table (mba_edge_unique, replacing get_edge_unique,
src/utils/mesh_utils.py:921-977) against the host path of the same function
and the oracle's brute-force edge list: bit-exact (integer work)."""

import numpy as np
import pytest
import torch

import oracle as O
from mesh_bem_amd.geometry import get_edge_unique, icosphere, star_mesh

pytestmark = pytest.mark.gpu


def _meshes():
    V, F = icosphere(3)
    yield "icosphere3", V, F
    rng = np.random.default_rng(4)
    perm = rng.permutation(len(V))
    inv = np.argsort(perm)
    yield "relabelled", V[perm], inv[F].astype(np.int32)
    Vs, Fs = star_mesh(5, 9)[:2]
    yield "star", Vs, Fs
    yield "open", V, F[: len(F) // 2]  # boundary edges have one face


@pytest.mark.parametrize("case", range(4))
def test_edge_unique_matches_host(dev, case):
    name, V, F = list(_meshes())[case]
    print(f"edges {name}", flush=True)
    vt = torch.tensor(V, dtype=torch.float64)
    ft = torch.tensor(np.asarray(F, dtype=np.int32))
    e_h, fe_h, adj_h = get_edge_unique(vt, ft)
    e_d, fe_d, adj_d = get_edge_unique(vt.to(dev), ft.to(dev))
    assert torch.equal(e_d.cpu(), e_h) and torch.equal(fe_d.cpu(), fe_h) and torch.equal(adj_d.cpu(), adj_h)
    if name != "open":
        ref_e, ref_fe = O.edges_bruteforce(np.asarray(F), len(V))
        assert np.array_equal(e_d.cpu().numpy(), ref_e) and np.array_equal(fe_d.cpu().numpy(), ref_fe)
    else:
        assert (adj_d[:, 1] == -1).any()


def test_edge_unique_large(dev):
    V, F = icosphere(6)  # 81 920 faces
    ft = torch.tensor(F, dtype=torch.int32)
    e_h, fe_h, adj_h = get_edge_unique(torch.tensor(V), ft)
    e_d, fe_d, adj_d = get_edge_unique(torch.tensor(V, device=dev), ft.to(dev))
    assert e_d.shape[0] == 3 * len(F) // 2
    assert torch.equal(e_d.cpu(), e_h) and torch.equal(fe_d.cpu(), fe_h) and torch.equal(adj_d.cpu(), adj_h)
