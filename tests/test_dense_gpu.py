"""GPU tests of the dense kernels: MFMA complex GEMM, LU + solve, FP64 residual,
mixed-precision refinement.

References: GEMM against an FP64 CPU product of the same inputs; LU against
LAPACK (scipy zgesv) and the backward error ||A x - b|| / (||A|| ||x||).
"""

import numpy as np
import pytest
import scipy.linalg
import torch

from mesh_bem_amd import _lib
from mesh_bem_amd.dense import LUFactor, gemm, refine_solve, residual_c128, to_c64

pytestmark = pytest.mark.gpu


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("dtype,tol", [(torch.complex64, 2e-6), (torch.complex128, 1e-14)])
@pytest.mark.parametrize("M,N,K", [(128, 64, 16), (200, 77, 33), (1, 5, 3), (513, 130, 128)])
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_gemm(dev, dtype, tol, M, N, K, beta):
    rng = np.random.default_rng(M * 7 + N + K)
    a, b, c = crand(rng, M, K), crand(rng, K, N), crand(rng, M, N)
    A = torch.tensor(a, dtype=dtype, device=dev)
    B = torch.tensor(b, dtype=dtype, device=dev)
    C = torch.tensor(c, dtype=dtype, device=dev)
    c_in = C.cpu().numpy().astype(np.complex128)  # gemm updates C in place
    out = gemm(A, B, C, alpha=-1.0, beta=beta).cpu().numpy()
    ref = -(A.cpu().numpy().astype(np.complex128) @ B.cpu().numpy().astype(np.complex128)) + beta * c_in
    scale = np.abs(a).max() * np.abs(b).max() * K + beta * np.abs(c).max()
    assert np.abs(out - ref).max() / scale < tol


def test_gemm_strided_views(dev):
    """Sub-matrix views with a leading dimension larger than the width."""
    rng = np.random.default_rng(5)
    big = torch.tensor(crand(rng, 300, 300), dtype=torch.complex64, device=dev)
    A = big[10:140, 3:70]
    B = big[100:167, 150:251]
    C = torch.zeros((130, 101), dtype=torch.complex64, device=dev)
    gemm(A, B, C, 1.0, 0.0)
    ref = A.cpu().numpy().astype(np.complex128) @ B.cpu().numpy().astype(np.complex128)
    assert np.abs(C.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-5


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (7, 2), (129, 1), (300, 3), (1000, 40)])
def test_lu_solve_matches_lapack(dev, dtype, n, nrhs):
    print(f"lu n={n} nrhs={nrhs} dtype={dtype}", flush=True)
    rng = np.random.default_rng(n)
    a = crand(rng, n, n)
    b = crand(rng, n, nrhs)
    A = torch.tensor(a, dtype=dtype, device=dev)
    f = LUFactor(A)
    x = f.solve(torch.tensor(b, dtype=dtype, device=dev)).cpu().numpy()
    a64 = A.cpu().numpy().astype(np.complex128)
    ref = scipy.linalg.solve(a64, b)
    eps = 6e-8 if dtype == torch.complex64 else 1.2e-16
    bwd = np.linalg.norm(a64 @ x - b) / (np.linalg.norm(a64) * np.linalg.norm(x))
    assert bwd < 50 * eps * np.sqrt(n)
    cond = np.linalg.cond(a64)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 50 * eps * cond * np.sqrt(n)


def test_lu_pivoting_needed(dev):
    """A matrix with a zero leading entry only factors with row interchanges."""
    a = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 9]], dtype=np.complex128)
    f = LUFactor(torch.tensor(a, device=dev))
    x = f.solve(torch.ones(3, dtype=torch.complex128, device=dev)).cpu().numpy()
    assert np.allclose(a @ x, np.ones(3), atol=1e-13)
    assert f.ipiv[0].item() == 2


def test_lu_multi_panel_large(dev):
    """n spanning many panels (and a ragged last panel)."""
    n = 2500
    rng = np.random.default_rng(1)
    a = crand(rng, n, n)
    A = torch.tensor(a, dtype=torch.complex64, device=dev)
    f = LUFactor(A)
    b = crand(rng, n, 1)
    x = f.solve(torch.tensor(b, dtype=torch.complex64, device=dev)).cpu().numpy()
    a64 = A.cpu().numpy().astype(np.complex128)
    bwd = np.linalg.norm(a64 @ x - b) / (np.linalg.norm(a64) * np.linalg.norm(x))
    assert bwd < 1e-5


def test_residual_c128(dev):
    rng = np.random.default_rng(2)
    z, x, v = crand(rng, 333, 333), crand(rng, 333, 2), crand(rng, 333, 2)
    Z = torch.tensor(z, device=dev)
    r = residual_c128(Z, torch.tensor(x, device=dev), torch.tensor(v, device=dev)).cpu().numpy()
    assert np.abs(r - (v - z @ x)).max() < 1e-12 * np.abs(z @ x).max()


def test_refinement_reaches_fp64(dev):
    """FP32 factor + FP64 residual converges to the FP64 LAPACK solution."""
    n = 800
    rng = np.random.default_rng(3)
    a = crand(rng, n, n) + 3 * np.sqrt(n) * np.eye(n)
    b = crand(rng, n)
    Z = torch.tensor(a, device=dev)
    J, info = refine_solve(Z, torch.tensor(b, device=dev), tol=1e-12)
    ref = scipy.linalg.solve(a, b)
    assert info.converged and not info.fallback
    assert np.linalg.norm(J.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-11


def test_refinement_fallback_on_ill_conditioning(dev):
    """An FP32-singular system cannot refine; the solve falls back to FP64 LU."""
    n = 200
    rng = np.random.default_rng(4)
    U, _ = np.linalg.qr(crand(rng, n, n))
    s = np.logspace(0, -10, n)
    a = (U * s) @ U.conj().T
    b = crand(rng, n)
    J, info = refine_solve(torch.tensor(a, device=dev), torch.tensor(b, device=dev), tol=1e-12)
    assert info.fallback == "fp64-lu"
    ref = scipy.linalg.solve(a, b)
    assert np.linalg.norm(a @ J.cpu().numpy() - b) / np.linalg.norm(b) < 1e-5
    assert np.isfinite(ref).all()


def test_to_c64(dev):
    z = torch.randn(5, 7, dtype=torch.complex128, device=dev)
    assert torch.equal(to_c64(z), z.to(torch.complex64))


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("shape", [(300, 200, 256), (129, 65, 17), (1000, 1000, 128)])
def test_gemm_variants(dev, variant, shape):
    """Every FP32 GEMM tile variant (incl. the 3M ones) against a complex128 product."""
    from mesh_bem_amd import _lib

    M_, N_, K_ = shape
    g = torch.Generator().manual_seed(variant)
    A = torch.randn((M_, K_), dtype=torch.complex64, generator=g)
    B = torch.randn((K_, N_), dtype=torch.complex64, generator=g)
    C = torch.randn((M_, N_), dtype=torch.complex64, generator=g)
    Ad, Bd, Cd = A.to(dev), B.to(dev), C.to(dev)
    lib = _lib.load()
    st = lib.mba_gemm_variant(_lib.MBA_F32, variant, M_, N_, K_, -1.0, _lib.ptr(Ad), K_, _lib.ptr(Bd), N_, 0.5,
                              _lib.ptr(Cd), N_, _lib.stream())
    _lib.check(st)
    ref = -(A.to(torch.complex128) @ B.to(torch.complex128)) + 0.5 * C.to(torch.complex128)
    scale = (A.abs().to(torch.float64) @ B.abs().to(torch.float64)).max()
    err = (Cd.cpu().to(torch.complex128) - ref).abs().max() / scale
    assert err < 3e-6, (variant, float(err))


def test_gemm_3m_row_tiles_bit_identical(dev):
    """The 3M FP32 product takes 128-row tiles on grids of at least two
    workgroups per CU and 64-row tiles below that; each element's MFMA
    sequence is the same, so a product computed on 128-row tiles and a
    sub-block of it computed alone (64-row tiles) agree bit for bit."""
    from mesh_bem_amd import _lib

    cus = torch.cuda.get_device_properties(0).multi_processor_count
    M = 128 * 16
    N = 64 * ((2 * cus + 15) // 16 + 1)  # (M / 128) x (N / 64) >= 2 x CUs tiles
    K = 96
    assert (M // 128) * (N // 64) >= 2 * cus
    g = torch.Generator().manual_seed(21)
    A = torch.randn((M, K), dtype=torch.complex64, generator=g).to(dev)
    B = torch.randn((K, N), dtype=torch.complex64, generator=g).to(dev)
    lib = _lib.load()
    C = torch.zeros((M, N), dtype=torch.complex64, device=dev)
    _lib.check(lib.mba_gemm_variant(_lib.MBA_F32, 4, M, N, K, 1.0, _lib.ptr(A), K, _lib.ptr(B), N, 0.0, _lib.ptr(C),
                                    N, _lib.stream()))
    m, n = 200, 300  # (2 x 5) 128-row tiles: the 64-row form
    Cs = torch.zeros((m, n), dtype=torch.complex64, device=dev)
    _lib.check(lib.mba_gemm_variant(_lib.MBA_F32, 4, m, n, K, 1.0, _lib.ptr(A), K, _lib.ptr(B), N, 0.0,
                                    _lib.ptr(Cs), n, _lib.stream()))
    torch.cuda.synchronize()
    assert torch.equal(C[:m, :n], Cs)
    ref = A.cpu().to(torch.complex128) @ B.cpu().to(torch.complex128)
    assert float((C.cpu().to(torch.complex128) - ref).abs().max() / ref.abs().max()) < 1e-5


@pytest.mark.parametrize("n", [1, 63, 64, 65, 130, 700])
@pytest.mark.parametrize("nrhs", [1, 2, 3, 4])
def test_lu_solve_persistent_matches_per_block(dev, n, nrhs):
    """complex64 solves with <= 4 right-hand sides: the persistent triangular
    solves (one launch per triangle) agree with the per-block launches and
    with LAPACK, for ragged last blocks, and through a strided b (ldb > nrhs)
    on repeated calls with the same factor."""
    from mesh_bem_amd.dense import set_lu_solve

    print(f"persistent solve n={n} nrhs={nrhs}", flush=True)
    rng = np.random.default_rng(100 * n + nrhs)
    a = crand(rng, n, n)
    A = torch.tensor(a, dtype=torch.complex64, device=dev)
    f = LUFactor(A)
    lib = _lib.load()
    a64 = A.cpu().numpy().astype(np.complex128)
    rhs = [crand(rng, n, nrhs) for _ in range(2)]
    try:
        outs = {}
        for kind in ("per_block", "persistent"):
            set_lu_solve(kind)
            for rep in range(2):
                b = rhs[rep]
                wide = torch.zeros((n, nrhs + 3), dtype=torch.complex64, device=dev)
                wide[:, :nrhs] = torch.tensor(b, dtype=torch.complex64, device=dev)
                wsb = lib.mba_lu_solve_workspace_bytes(f.prec, n, nrhs)
                ws = torch.full((wsb,), 0x7f, dtype=torch.uint8, device=dev)  # stale workspace contents
                st = lib.mba_lu_solve(f.prec, _lib.ptr(f.lu), n, f.lu.stride(0), _lib.ptr(f.aux),
                                      _lib.ptr(wide), nrhs, wide.stride(0), _lib.ptr(ws), wsb, _lib.stream())
                _lib.check(st, "mba_lu_solve")
                w = wide.cpu().numpy()
                assert np.all(w[:, nrhs:] == 0), "columns beyond nrhs must stay untouched"
                x = w[:, :nrhs].astype(np.complex128)
                assert np.isfinite(x).all()
                ref = scipy.linalg.solve(a64, b)
                cond = np.linalg.cond(a64)
                assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 50 * 6e-8 * cond * np.sqrt(n)
                outs[(kind, rep)] = x
        # same factor, same order of operations up to FP32 summation order
        for rep in range(2):
            d = np.linalg.norm(outs[("persistent", rep)] - outs[("per_block", rep)])
            assert d / np.linalg.norm(outs[("per_block", rep)]) < 1e-4
    finally:
        set_lu_solve("persistent")


def _interchange(a, ipiv):
    p = a.copy()
    for j, q in enumerate(ipiv):
        if q != j:
            p[[j, q]] = p[[q, j]]
    return p


@pytest.mark.parametrize("n,update", [(1000, "fp32"), (2100, "fp32"), (4500, "split_f16")])
def test_lu_factors_reproduce_pa(dev, n, update):
    """complex64 partial-pivoting LU: P A = L U to FP32 accuracy, a check of
    every factor entry including the U rows the panel triangular solve
    forms (panels of 128 rows, inverted 64-row diagonal blocks on MFMA; n =
    1000 leaves a ragged last group of columns, 4500 runs the split-FP16
    trailing update over several super panels)."""
    from mesh_bem_amd.dense import set_lu_update

    print(f"P A = L U n={n} update={update}", flush=True)
    rng = np.random.default_rng(n + 5)
    a = crand(rng, n, n)
    A = torch.tensor(a, dtype=torch.complex64, device=dev)
    a64 = A.cpu().numpy().astype(np.complex128)
    set_lu_update(update)
    try:
        f = LUFactor(A)
    finally:
        set_lu_update("fp32")
    assert f.info == 0
    lu = f.lu.cpu().numpy().astype(np.complex128)
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    pa = _interchange(a64, f.ipiv.cpu().numpy().astype(np.int64))
    err = np.linalg.norm(pa - L @ U) / np.linalg.norm(a64)
    print(f"||PA - LU|| / ||A|| = {err:.3e}", flush=True)
    # measured 1.5e-5 / 3.1e-5 / 6.8e-5 (the per-column substitution it
    # replaced: 1.4e-5 / 3.1e-5 / 6.9e-5; LAPACK cgetrf at n = 2100: 1.5e-5)
    assert err < 40 * 6e-8 * np.sqrt(n)


@pytest.mark.parametrize("n,update", [(2200, "fp32"), (3000, "fp32"), (4500, "split_f16")])
def test_lu_tournament_panels(dev, n, update):
    """complex64 LU with tournament-pivoting panels (64 columns, >= 2048 rows;
    the later, shorter panels use partial pivoting): ipiv is a valid
    interchange sequence, the first panel's pivots equal the oracle's
    tournament on the same values, P A = L U to the accuracy of the
    partial-pivoting factorisation of the same matrix, and solves meet the
    FP32 backward-error bound."""
    import oracle as O
    from mesh_bem_amd.dense import set_lu_panel, set_lu_update

    print(f"tournament LU n={n} update={update}", flush=True)
    rng = np.random.default_rng(n + 11)
    a = crand(rng, n, n)
    A = torch.tensor(a, dtype=torch.complex64, device=dev)
    a64 = A.cpu().numpy().astype(np.complex128)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    set_lu_update(update)
    try:
        set_lu_panel("tournament")
        ft = LUFactor(A)
        set_lu_panel("partial")
        fp = LUFactor(A)
    finally:
        set_lu_panel("partial")
        set_lu_update("fp32")
    assert ft.info == 0
    ipiv = ft.ipiv.cpu().numpy().astype(np.int64)
    assert np.all(ipiv >= np.arange(n)) and np.all(ipiv < n)
    _, ip_o = O.tslu_pivots(a64[:, :64], n_cu=cus)
    assert np.array_equal(ipiv[:64], np.array(ip_o))
    errs = {}
    for name, f in (("tournament", ft), ("partial", fp)):
        lu = f.lu.cpu().numpy().astype(np.complex128)
        L = np.tril(lu, -1) + np.eye(n)
        U = np.triu(lu)
        pa = _interchange(a64, f.ipiv.cpu().numpy().astype(np.int64))
        errs[name] = np.linalg.norm(pa - L @ U) / np.linalg.norm(a64)
    print(errs, flush=True)
    assert errs["tournament"] < 4 * errs["partial"] + 1e-7
    b = crand(rng, n, 1)
    x = ft.solve(torch.tensor(b, dtype=torch.complex64, device=dev)).cpu().numpy().astype(np.complex128)
    bwd = np.linalg.norm(a64 @ x - b) / (np.linalg.norm(a64) * np.linalg.norm(x))
    assert bwd < 50 * 6e-8 * np.sqrt(n)


def test_refinement_with_tournament_panels(dev):
    """Mixed-precision solve with tournament panels refines to the FP64
    LAPACK solution."""
    from mesh_bem_amd.dense import set_lu_panel

    n = 3000
    rng = np.random.default_rng(7)
    a = crand(rng, n, n) + 3 * np.sqrt(n) * np.eye(n)
    b = crand(rng, n)
    try:
        set_lu_panel("tournament")
        J, info = refine_solve(torch.tensor(a, device=dev), torch.tensor(b, device=dev), tol=1e-12)
    finally:
        set_lu_panel("partial")
    ref = scipy.linalg.solve(a, b)
    assert info.converged and not info.fallback
    assert np.linalg.norm(J.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-11


@pytest.mark.parametrize("n", [64, 200, 700, 2500])
def test_lu_pivots_match_lapack(dev, n):
    """complex128 partial pivoting picks LAPACK's pivot rows (zgetrf's ipiv,
    the factorisation inside torch.linalg.solve): the panel's arg-max over
    |re| + |im| with ties to the lower row reproduces the sequence exactly on
    random matrices (multi-panel at n = 700, 2500)."""
    print(f"ipiv vs LAPACK n={n}", flush=True)
    rng = np.random.default_rng(900 + n)
    a = crand(rng, n, n)
    f = LUFactor(torch.tensor(a, dtype=torch.complex128, device=dev))
    ref = scipy.linalg.lu_factor(a)[1]
    got = f.ipiv.cpu().numpy().astype(np.int64)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("n", [1, 63, 3000, 40000])
def test_pivots_to_perm_sequential(dev, n):
    """mba_lu_pivots_to_perm equals the interchanges applied one by one
    (position i ends holding original row perm[i]) for LAPACK-style pivots
    p_j >= j with about half no-ops, near pivots and repeated far rows; n =
    40000 exceeds the in-LDS size and takes the global-memory path."""
    print(f"pivots_to_perm n={n}", flush=True)
    rng = np.random.default_rng(77 + n)
    j = np.arange(n)
    kind = rng.integers(0, 4, n)
    far = j + rng.integers(0, n - j)
    near = np.minimum(j + rng.integers(0, 3, n), n - 1)
    rep = np.minimum(j + (n - j) // 2, n - 1)  # many columns share this far row
    piv = np.where(kind < 2, j, np.where(kind == 2, near, np.where(rng.random(n) < 0.5, far, rep)))
    piv = np.maximum(piv, j).astype(np.int32)
    want = np.arange(n, dtype=np.int64)
    for i in range(n):
        p = piv[i]
        want[i], want[p] = want[p], want[i]
    ip = torch.tensor(piv, device=dev)
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    _lib.check(_lib.load().mba_lu_pivots_to_perm(_lib.ptr(ip), n, _lib.ptr(perm), _lib.stream()),
               "mba_lu_pivots_to_perm")
    assert np.array_equal(perm.cpu().numpy().astype(np.int64), want)


def test_lu_async_workgroup_cap(dev):
    """The asynchronous LU with its panels capped at a few workgroups (the
    panel width then shrinks so the capped workgroups' rows fit the LDS):
    n = 4096 with 16 workgroups per panel (64-column panels) factors to FP32
    backward error and solves like the uncapped factorisation (the pivots
    part ways at the first near-tie that the other blocking rounds
    differently, so they are not compared); a cap that leaves even 16-column
    panels too tall for the LDS is refused before any launch."""
    from mesh_bem_amd.dense import LUFactor

    print("LU with max_wg = 16 at n = 4096", flush=True)
    g = torch.Generator(device="cpu").manual_seed(8)
    n = 4096
    A = torch.complex(torch.randn((n, n), generator=g), torch.randn((n, n), generator=g)).to(torch.complex64).to(dev)
    lu = LUFactor(A, asynchronous=True, max_wg=16)
    torch.cuda.synchronize()
    lu.check()
    ref = LUFactor(A)
    same = float((lu.ipiv == ref.ipiv).float().mean())
    b = torch.complex(torch.randn(n, generator=g), torch.randn(n, generator=g)).to(torch.complex64).to(dev)
    x, x1 = lu.solve(b), ref.solve(b)
    A64, b64 = A.to(torch.complex128), b.to(torch.complex128)
    r = float((A64 @ x.to(torch.complex128) - b64).norm() / (A64.norm() * x.to(torch.complex128).norm()))
    x64 = torch.linalg.solve(A64, b64)
    e = float((x.to(torch.complex128) - x64).norm() / x64.norm())
    e1 = float((x1.to(torch.complex128) - x64).norm() / x64.norm())
    print(f"  backward error {r:.2e}, pivots equal {same:.4f}, forward error {e:.2e} (uncapped {e1:.2e})")
    assert r < 1e-6 and e < 10 * e1 + 1e-6
    # 2 workgroups of 8192 rows: 8192 x 16 columns x 8 bytes exceed the LDS
    big = torch.empty((16384, 16384), dtype=torch.complex64, device=dev)
    with pytest.raises(_lib.MBAError):
        LUFactor(big, overwrite=True, asynchronous=True, max_wg=2)
