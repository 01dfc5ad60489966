"""GPU tests of the FP32-accurate split-FP16 complex GEMM (power-of-two scaled
hi/lo FP16 pieces, three piece products with FP32 accumulation) and of the
complex64 LU whose trailing update uses it.

References: an FP64 CPU product of the same complex64 inputs; the error bound
of the scheme is about 3 * 2^-22 per term relative to the row maximum of A
times the column maximum of B (2^-22: the hi + lo representation; the
dropped lo.lo term), i.e. within a small factor of an FP32 product.
"""

import numpy as np
import pytest
import scipy.linalg
import torch

from mesh_bem_amd.dense import LUFactor, gemm_split_f16, refine_solve, set_lu_update, set_split_tile

pytestmark = pytest.mark.gpu


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def ref_product(A, B):
    return A.cpu().numpy().astype(np.complex128) @ B.cpu().numpy().astype(np.complex128)


@pytest.mark.parametrize("M,N,K", [(256, 128, 16), (300, 200, 100), (1, 5, 3), (513, 130, 256),
                                   (1000, 1000, 256), (77, 333, 1), (257, 129, 17), (700, 300, 512), (260, 140, 333)])
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_split_gemm_matches_fp64(dev, M, N, K, beta):
    print(f"split gemm M={M} N={N} K={K} beta={beta}", flush=True)
    rng = np.random.default_rng(M * 7 + N + K)
    A = torch.tensor(crand(rng, M, K), dtype=torch.complex64, device=dev)
    B = torch.tensor(crand(rng, K, N), dtype=torch.complex64, device=dev)
    C = torch.tensor(crand(rng, M, N), dtype=torch.complex64, device=dev)
    c_in = C.cpu().numpy().astype(np.complex128)
    out = gemm_split_f16(A, B, C, alpha=-1.0, beta=beta).cpu().numpy()
    ref = -ref_product(A, B) + beta * c_in
    scale = np.abs(A.cpu().numpy()).max() * np.abs(B.cpu().numpy()).max() * K + beta * np.abs(c_in).max()
    assert np.abs(out - ref).max() / scale < 2e-6
    # normwise: within a small factor of FP32 rounding
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert rel < 1e-6, rel


def test_split_gemm_wide_dynamic_range(dev):
    """Rows of A and columns of B spanning 10^-15 .. 10^15, entries within a
    row spanning 10^-6 .. 1: the per-row / per-column power-of-two scaling
    keeps the error at the 2^-22 level of each row / column maximum."""
    rng = np.random.default_rng(11)
    M, N, K = 300, 260, 256
    a = crand(rng, M, K) * 10.0 ** rng.uniform(-6, 0, (M, K))
    b = crand(rng, K, N) * 10.0 ** rng.uniform(-6, 0, (K, N))
    a *= 10.0 ** rng.uniform(-15, 15, (M, 1))
    b *= 10.0 ** rng.uniform(-15, 15, (1, N))
    A = torch.tensor(a, dtype=torch.complex64, device=dev)
    B = torch.tensor(b, dtype=torch.complex64, device=dev)
    out = gemm_split_f16(A, B).cpu().numpy()
    ref = ref_product(A, B)
    a32, b32 = np.abs(A.cpu().numpy()), np.abs(B.cpu().numpy())
    bound = K * a32.max(axis=1)[:, None] * b32.max(axis=0)[None, :]
    assert np.all(np.abs(out - ref) <= 1e-6 * bound)


def test_split_gemm_strided_views(dev):
    rng = np.random.default_rng(5)
    big = torch.tensor(crand(rng, 600, 600), dtype=torch.complex64, device=dev)
    A = big[10:340, 3:259]
    B = big[100:356, 150:551]
    C = torch.zeros((330, 401), dtype=torch.complex64, device=dev)
    gemm_split_f16(A, B, C, 1.0, 0.0)
    ref = ref_product(A, B)
    assert np.linalg.norm(C.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-6


def test_split_gemm_rejects_large_k(dev):
    A = torch.zeros((4, 1100), dtype=torch.complex64, device=dev)
    B = torch.zeros((1100, 4), dtype=torch.complex64, device=dev)
    from mesh_bem_amd._lib import MBAError

    with pytest.raises(MBAError):
        gemm_split_f16(A, B)


@pytest.mark.parametrize("n", [2600, 3000])
def test_lu_with_split_update(dev, n):
    """complex64 LU whose trailing updates (blocks >= 2048) use the split
    product: backward error at the FP32 level, and FP64 refinement reaches the
    LAPACK complex128 solution."""
    print(f"lu split update n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = crand(rng, n, n) + 4 * np.sqrt(n) * np.eye(n) * (rng.uniform(0.5, 1.0, n))
    b = crand(rng, n, 1)
    Z64 = torch.tensor(a, dtype=torch.complex128, device=dev)
    set_lu_update("split_f16")
    try:
        f = LUFactor(Z64.to(torch.complex64))
    finally:
        set_lu_update("fp32")
    x = f.solve(torch.tensor(b, dtype=torch.complex64, device=dev)).cpu().numpy()
    a32 = Z64.to(torch.complex64).cpu().numpy().astype(np.complex128)
    bwd = np.linalg.norm(a32 @ x - b) / (np.linalg.norm(a32) * np.linalg.norm(x))
    assert bwd < 50 * 2.4e-7 * np.sqrt(n), bwd
    f32 = LUFactor(Z64.to(torch.complex64))
    x32 = f32.solve(torch.tensor(b, dtype=torch.complex64, device=dev)).cpu().numpy()
    assert np.linalg.norm(x - x32) / np.linalg.norm(x32) < 1e-4
    V = torch.tensor(b[:, 0], dtype=torch.complex128, device=dev)
    J, info = refine_solve(Z64, V, lu32=f, tol=1e-12)
    ref = scipy.linalg.solve(a, b[:, 0])
    assert info.converged
    assert np.linalg.norm(J.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-11


@pytest.mark.parametrize("tile", [0, 1])
@pytest.mark.parametrize("M,N,K", [(700, 300, 512), (129, 257, 48), (1000, 1000, 256)])
def test_split_gemm_tiles(dev, tile, M, N, K):
    """Both workgroup tiles (256 x 128 with 8 waves, 128 x 128 with 4 waves)."""
    rng = np.random.default_rng(M + N + K + tile)
    A = torch.tensor(crand(rng, M, K), dtype=torch.complex64, device=dev)
    B = torch.tensor(crand(rng, K, N), dtype=torch.complex64, device=dev)
    C = torch.tensor(crand(rng, M, N), dtype=torch.complex64, device=dev)
    c_in = C.cpu().numpy().astype(np.complex128)
    set_split_tile(tile)
    try:
        out = gemm_split_f16(A, B, C, alpha=-1.0, beta=1.0).cpu().numpy()
    finally:
        set_split_tile(0)
    ref = -ref_product(A, B) + c_in
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-6


@pytest.mark.parametrize("M,N,K", [(700, 300, 512), (1000, 1000, 256)])
def test_split_gemm_tiles_bit_identical(dev, M, N, K):
    """Each accumulator receives its products in the same order whatever the
    tile, so both tiles give bit-identical results."""
    rng = np.random.default_rng(M * N + K)
    A = torch.tensor(crand(rng, M, K), dtype=torch.complex64, device=dev)
    B = torch.tensor(crand(rng, K, N), dtype=torch.complex64, device=dev)
    C0 = torch.tensor(crand(rng, M, N), dtype=torch.complex64, device=dev)
    outs = []
    try:
        for tile in (0, 1):
            set_split_tile(tile)
            outs.append(gemm_split_f16(A, B, C0.clone(), alpha=-1.0, beta=1.0))
    finally:
        set_split_tile(0)
    assert torch.equal(outs[0], outs[1])
