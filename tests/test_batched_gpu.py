"""Batched multi-mesh, multi-wave path on the GPU.

The one-launch assembly over several uneven meshes must give exactly the
matrices of the single-mesh assembly (same kernel body, same records, same
gather order), and the mixed-precision multi-RHS solve must agree with a
complex128 LAPACK-equivalent solve (torch.linalg.solve on the FP64 matrix) to
1e-9."""

import numpy as np
import pytest
import torch

import oracle as O
from mesh_bem_amd import _lib
from mesh_bem_amd.batched import BatchedMoM, GemmTuner, star_batch
from mesh_bem_amd.mom import assemble_ns

pytestmark = pytest.mark.gpu


def test_batched_assembly_equals_single(dev):
    geoms, ks, dirs, pols = star_batch(5, seed0=3, freqs=(3, 5, 4), n_dirs=6)
    bm = BatchedMoM.build(geoms, ks, device=dev)
    for prec in (_lib.MBA_F64, _lib.MBA_F32):
        Zs = bm.assemble(prec)
        for m, k, Z in zip(bm.meshes, ks, Zs):
            mm = m if prec == _lib.MBA_F64 else type(m)(**{**m.__dict__, "verts": m.verts.to(torch.float32)})
            Z1 = assemble_ns(mm, k, prec)
            assert torch.equal(Z, Z1)


def test_batched_matches_oracle_and_solves(dev):
    geoms, ks, dirs, pols = star_batch(3, seed0=11, freqs=(3, 4), n_dirs=8)
    bm = BatchedMoM.build(geoms, ks, device=dev)
    Zs = bm.assemble(_lib.MBA_F64)
    Vs = bm.rhs(dirs, pols)
    V0, F0 = geoms[0]
    Zo, Vo, _, _ = O.ns_system(bm.meshes[0].verts.cpu().numpy(), F0, ks[0], dirs[0], pols[0])
    assert np.linalg.norm(Zs[0].cpu().numpy() - Zo) / np.linalg.norm(Zo) < 1e-11
    tuner = GemmTuner()
    res = bm.solve(Zs, Vs, tuner=tuner, tol=1e-10)
    assert all(res.converged)
    for Z, V, J in zip(Zs, Vs, res.J):
        ref = torch.linalg.solve(Z.cpu(), V.cpu())
        e = (torch.linalg.vector_norm(J.cpu() - ref) / torch.linalg.vector_norm(ref)).item()
        print(f"  E={Z.shape[0]} relerr {e:.2e}")
        assert e < 1e-9
    assert set(tuner.best.values()) <= set(range(6))
