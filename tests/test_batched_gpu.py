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
from mesh_bem_amd.mom import assemble_ns, rhs_ns

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
        # right-hand sides of all meshes in one pass: the per-mesh values exactly
        Vs = bm.rhs(dirs, pols, prec)
        for m, k, d, p, V in zip(bm.meshes, ks, dirs, pols, Vs):
            mm = m if prec == _lib.MBA_F64 else type(m)(**{**m.__dict__, "verts": m.verts.to(torch.float32)})
            assert torch.equal(V, rhs_ns(mm, k, prec, d, p))


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


@pytest.mark.parametrize("persistent", [True, False])
def test_lu_factor_batched_matches_single(dev, persistent):
    """Batched LU of uneven matrices (one persistent launch, or lock step):
    every factor solves its system like the single-matrix LU (same pivots in
    exact ties aside) to FP32 accuracy."""
    from mesh_bem_amd.dense import LUFactor, lu_factor_batched

    print(f"batched LU persistent={persistent}", flush=True)
    rng = np.random.default_rng(4)
    sizes = [300, 1000, 77, 640, 1301, 256]
    mats = [torch.tensor(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * n ** 0.5 * np.eye(n),
                         dtype=torch.complex64, device=dev) for n in sizes]
    lus = lu_factor_batched([A.clone() for A in mats], persistent=persistent)
    torch.cuda.synchronize()
    for A, lu in zip(mats, lus):
        lu.check()
        b = torch.tensor(rng.standard_normal(A.shape[0]) + 0j, dtype=torch.complex64, device=dev)
        x = lu.solve(b)
        x1 = LUFactor(A.clone()).solve(b)
        r = (A.to(torch.complex128) @ x.to(torch.complex128) - b.to(torch.complex128)).norm() / b.norm()
        assert float(r) < 1e-5
        assert torch.equal(lu.ipiv, LUFactor(A.clone()).ipiv)
        assert float((x - x1).norm() / x1.norm()) < 1e-5


@pytest.mark.parametrize("prec,persistent", [("f32", True), ("f32", False), ("f64", True)])
def test_lu_solve_batched_matches_single(dev, prec, persistent):
    """Batched triangular solves (one persistent launch, a dataflow of 64-row
    block tasks per XCD, or two launches walking the blocks) of several
    factorisations with ragged
    sizes (not multiples of the 64-row block) and right-hand-side counts (not
    multiples of the 16- or 64-column groups) agree with the per-matrix solve."""
    from mesh_bem_amd.dense import LUFactor, lu_solve_batched

    print(f"batched solves {prec} persistent={persistent}", flush=True)

    cdt = torch.complex64 if prec == "f32" else torch.complex128
    bar = 1e-5 if prec == "f32" else 1e-12
    rng = np.random.default_rng(6)
    cases = [(300, 1), (1000, 17), (77, 128), (64, 16), (1, 3), (513, 40)]
    mats, lus, Bs = [], [], []
    for n, r in cases:
        A = torch.tensor(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * n ** 0.5 * np.eye(n),
                         dtype=cdt, device=dev)
        mats.append(A)
        lus.append(LUFactor(A.clone()))
        Bs.append(torch.tensor(rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r)), dtype=cdt,
                               device=dev))
    B0 = [B.clone() for B in Bs]
    Xs = lu_solve_batched(lus, Bs, persistent=persistent)
    torch.cuda.synchronize()
    for A, lu, B, Bk, X in zip(mats, lus, Bs, B0, Xs):
        assert torch.equal(B, Bk)
        X1 = lu.solve(B)
        assert X.shape == X1.shape
        e = float((X - X1).norm() / X1.norm())
        r = float((A.to(torch.complex128) @ X.to(torch.complex128) - B.to(torch.complex128)).norm() / B.norm())
        print(f"  n={A.shape[0]} nrhs={B.shape[1]}: vs single {e:.2e}, residual {r:.2e}")
        assert e < bar and r < bar


def test_lu_solve_batched_persistent_many_jobs(dev):
    """The persistent dataflow solve on a batch larger than its eight XCD
    queues (48 jobs, 40-2100 rows, 1-200 right-hand sides, so up to four
    groups of 64 per job and chains of 33 blocks) agrees with the two-launch
    block walks and leaves a small residual; a second call on fresh scratch
    gives identical results (the task order is fixed, each block's sum is
    taken in the same order)."""
    from mesh_bem_amd.dense import LUFactor, lu_solve_batched

    print("persistent dataflow solves, 48 jobs", flush=True)
    rng = np.random.default_rng(11)
    mats, lus, Bs = [], [], []
    for q in range(48):
        n = int(rng.integers(40, 2101))
        r = int(rng.integers(1, 201))
        A = torch.tensor(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * n ** 0.5 * np.eye(n),
                         dtype=torch.complex64, device=dev)
        mats.append(A)
        lus.append(LUFactor(A.clone()))
        Bs.append(torch.tensor(rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r)),
                               dtype=torch.complex64, device=dev))
    Xw = lu_solve_batched(lus, Bs, persistent=False)
    Xp = lu_solve_batched(lus, Bs, persistent=True)
    Xp2 = lu_solve_batched(lus, Bs, persistent=True)
    torch.cuda.synchronize()
    worst_e, worst_r = 0.0, 0.0
    for A, B, X0, X1, X2 in zip(mats, Bs, Xw, Xp, Xp2):
        assert torch.equal(X1, X2)
        e = float((X1 - X0).norm() / X0.norm())
        r = float((A.to(torch.complex128) @ X1.to(torch.complex128) - B.to(torch.complex128)).norm() / B.norm())
        worst_e, worst_r = max(worst_e, e), max(worst_r, r)
    print(f"  worst vs walks {worst_e:.2e}, worst residual {worst_r:.2e}")
    assert worst_e < 1e-5 and worst_r < 1e-5


def test_staged_tables_many_calls_in_flight(dev):
    """Host job tables go through the library's pinned staging ring (64 slots
    per device, a slot reused once its copy has completed): 200 grouped
    products queued without a synchronisation, each with its own table
    (two uploads per call, so the ring wraps six times), all give the
    product of their own operands."""
    from mesh_bem_amd.dense import gemm_grouped

    print("200 grouped products in flight", flush=True)
    g = torch.Generator(device="cpu").manual_seed(5)
    ops = []
    for q in range(200):
        m, n, k = 17 + q % 13, 9 + q % 7, 5 + q % 11
        A = torch.randn((m, k), generator=g, dtype=torch.complex64).to(dev)
        B = torch.randn((k, n), generator=g, dtype=torch.complex64).to(dev)
        ops.append((A, B, torch.zeros((m, n), dtype=torch.complex64, device=dev)))
    refs = [A.cpu().to(torch.complex128) @ B.cpu().to(torch.complex128) for A, B, _ in ops]
    torch.cuda.synchronize()
    for A, B, C in ops:  # no host copy or synchronisation between the calls
        gemm_grouped([(A, B, C)], alpha=1.0, beta=0.0)
    outs = [C for _, _, C in ops]
    torch.cuda.synchronize()
    worst = max(float((C.cpu().to(torch.complex128) - R).norm() / R.norm()) for C, R in zip(outs, refs))
    print(f"  worst relative error {worst:.2e}")
    assert worst < 1e-5


def test_batched_refinement_helpers(dev):
    """One-launch helpers of the batched refinement: complex128 -> complex64
    conversions equal torch's elementwise conversion bit for bit; J += D
    equals torch's J + D.to(complex128) bit for bit, and the per-system max
    column ratio |D_c| / |J_c| matches an FP64 torch computation (1e-12),
    including an empty system, a zero column (ratio 0) and a zero J column
    with a nonzero correction (ratio inf)."""
    from mesh_bem_amd.dense import convert_c128_c64_batched, refine_update_batched

    print("batched conversion and refinement update", flush=True)
    g = torch.Generator(device="cpu").manual_seed(3)
    shapes = [(300, 7), (1, 1), (2049, 128), (64, 65), (0, 5), (513, 200)]

    def rnd(sh, dt):
        return torch.complex(torch.randn(sh, generator=g, dtype=torch.float64),
                             torch.randn(sh, generator=g, dtype=torch.float64)).to(dt).to(dev)

    srcs = [rnd(sh, torch.complex128) for sh in shapes]
    dsts = [torch.empty(sh, dtype=torch.complex64, device=dev) for sh in shapes]
    convert_c128_c64_batched(srcs, dsts)
    for a, b in zip(srcs, dsts):
        assert torch.equal(b, a.to(torch.complex64))

    Js = [rnd(sh, torch.complex128) for sh in shapes]
    Ds = [rnd(sh, torch.complex64) * 1e-3 for sh in shapes]
    Js[0][:, 3] = 0
    Ds[0][:, 3] = 0  # zero column: ratio 0
    Js[3][:, 5] = -Ds[3][:, 5].to(torch.complex128)  # J + D = 0 there, D != 0: ratio inf
    ref = [J + D.to(torch.complex128) for J, D in zip(Js, Ds)]
    want = []
    for R, D in zip(ref, Ds):
        if R.shape[0] == 0:
            want.append(0.0)
            continue
        dn = torch.linalg.vector_norm(D.to(torch.complex128), dim=0)
        jn = torch.linalg.vector_norm(R, dim=0)
        r = torch.where(jn > 0, dn / jn, torch.where(dn > 0, torch.full_like(dn, float("inf")), torch.zeros_like(dn)))
        want.append(float(r.max()))
    got = refine_update_batched(Js, Ds).cpu().tolist()
    for J, R in zip(Js, ref):
        assert torch.equal(J, R)
    print(f"  ratios {got}")
    for a, b in zip(got, want):
        if b == float("inf"):
            assert a == float("inf")
        else:
            assert abs(a - b) <= 1e-12 * max(b, 1e-300)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("variant", [4, 5])
def test_gemm_grouped_matches_torch(dev, prec, variant):
    """One grouped launch of ragged products (empty, single-row, tile-edge and
    strided operands) equals each product computed alone (FP64 on the CPU);
    FP32 runs both tile orders (variant 5 deals tiles to XCDs in runs)."""
    from mesh_bem_amd.dense import gemm_grouped

    cdt = torch.complex64 if prec == "f32" else torch.complex128
    bar = 1e-5 if prec == "f32" else 1e-13
    rng = np.random.default_rng(9)
    shapes = [(300, 77, 129), (1, 5, 3), (128, 64, 16), (0, 4, 4), (513, 200, 1), (64, 1, 257), (17, 0, 9)]
    cplx = lambda *s: torch.tensor(rng.standard_normal(s) + 1j * rng.standard_normal(s), dtype=cdt, device=dev)
    prods, refs = [], []
    for m, n, k in shapes:
        A = cplx(m, k + 3)[:, :k]  # strided rows
        B, C = cplx(k, n), cplx(m, n)
        refs.append((-1.0 * A.cpu().to(torch.complex128) @ B.cpu().to(torch.complex128)
                     + 0.5 * C.cpu().to(torch.complex128)))
        prods.append((A, B, C))
    lib = _lib.load()
    _lib.check(lib.mba_gemm_set_variant(_lib.MBA_F32, variant), "mba_gemm_set_variant")
    try:
        gemm_grouped(prods, alpha=-1.0, beta=0.5)
        torch.cuda.synchronize()
    finally:
        _lib.check(lib.mba_gemm_set_variant(_lib.MBA_F32, 4), "mba_gemm_set_variant")
    for (A, B, C), ref in zip(prods, refs):
        if ref.numel() == 0:
            continue
        e = float((C.cpu().to(torch.complex128) - ref).norm() / ref.norm())
        assert e < bar, (A.shape, B.shape, e)


def test_lu_factor_batched_tile_variants(dev):
    """The batched LU with grouped products (variants 4, 5) and with the
    per-matrix fallback (variant 0) gives the same pivots and accurate factors."""
    from mesh_bem_amd.dense import lu_factor_batched

    rng = np.random.default_rng(12)
    sizes = [700, 300, 1030]
    mats = [torch.tensor(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * n ** 0.5 * np.eye(n),
                         dtype=torch.complex64, device=dev) for n in sizes]
    lib = _lib.load()
    piv = []
    try:
        for v in (4, 5, 0):
            _lib.check(lib.mba_gemm_set_variant(_lib.MBA_F32, v), "mba_gemm_set_variant")
            lus = lu_factor_batched([A.clone() for A in mats])
            torch.cuda.synchronize()
            for A, lu in zip(mats, lus):
                lu.check()
                b = torch.tensor(rng.standard_normal(A.shape[0]) + 0j, dtype=torch.complex64, device=dev)
                x = lu.solve(b)
                r = (A.to(torch.complex128) @ x.to(torch.complex128) - b.to(torch.complex128)).norm() / b.norm()
                assert float(r) < 1e-5, (v, A.shape[0], float(r))
            piv.append([lu.ipiv.cpu() for lu in lus])
    finally:
        _lib.check(lib.mba_gemm_set_variant(_lib.MBA_F32, 4), "mba_gemm_set_variant")
    for p in piv[1:]:
        assert all(torch.equal(a, b) for a, b in zip(piv[0], p))


def _perm_of(ipiv):
    """Row permutation of LAPACK-style sequential interchanges (0-based)."""
    p = np.arange(len(ipiv))
    for j, q in enumerate(ipiv):
        p[j], p[q] = p[q], p[j]
    return p


def _check_factor(A, lu, tol):
    """P A = L U to `tol` (relative Frobenius) and partial pivoting's bound on
    the multipliers (|l_ij| <= |u_jj| in the |re| + |im| sense, so |l| <= sqrt 2)."""
    n = A.shape[0]
    LU = lu.lu.to(torch.complex128)
    L = torch.tril(LU, -1) + torch.eye(n, dtype=LU.dtype, device=LU.device)
    U = torch.triu(LU)
    perm = torch.tensor(_perm_of(lu.ipiv.cpu().numpy().astype(np.int64)), device=A.device)
    A64 = A.to(torch.complex128)
    err = (torch.linalg.matrix_norm(A64[perm] - L @ U) / torch.linalg.matrix_norm(A64)).item()
    lmax = float(torch.tril(LU, -1).abs().max()) if n > 1 else 0.0
    assert err < tol, (n, err)
    assert lmax <= 2 ** 0.5 * (1 + 1e-5), (n, lmax)
    return err


def test_lu_batched_persistent_random_pivoting(dev):
    """One persistent launch for random complex64 matrices that need row
    interchanges, orders around the 32-column panel and the 256-column super
    panel (1 ... 2049): P A = L U to FP32 accuracy with partial pivoting's
    multiplier bound, and solutions as accurate as the lock-step factors'."""
    from mesh_bem_amd.dense import lu_factor_batched

    rng = np.random.default_rng(17)
    sizes = [1, 31, 32, 33, 255, 256, 257, 700, 1301, 2049]
    print(f"persistent batched LU, random matrices {sizes}", flush=True)
    mats = [torch.tensor(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)), dtype=torch.complex64,
                         device=dev) for n in sizes]
    lp = lu_factor_batched([A.clone() for A in mats], persistent=True)
    ll = lu_factor_batched([A.clone() for A in mats], persistent=False)
    torch.cuda.synchronize()
    for A, fp, fl in zip(mats, lp, ll):
        fp.check()
        fl.check()
        n = A.shape[0]
        e = _check_factor(A, fp, 1e-4)
        b = torch.tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n), dtype=torch.complex64, device=dev)
        A64, b64 = A.to(torch.complex128), b.to(torch.complex128)
        rp = ((A64 @ fp.solve(b).to(torch.complex128) - b64).norm() / b64.norm()).item()
        rl = ((A64 @ fl.solve(b).to(torch.complex128) - b64).norm() / b64.norm()).item()
        print(f"  n={n} backward {e:.2e} residual persistent {rp:.2e} lock-step {rl:.2e}")
        assert rp < max(10 * rl, 1e-5)


def test_lu_batched_persistent_split_panel_phases(dev):
    """Forty 4000 x 4000 matrices: the first panels' rows exceed the LDS of
    the resident workgroups together, so each of those panel steps runs as two
    panel phases; every factor still solves its system to FP32 accuracy."""
    from mesh_bem_amd.dense import lu_factor_batched

    rng = np.random.default_rng(23)
    n, nm = 4000, 40
    print(f"persistent batched LU, {nm} x {n}", flush=True)
    base = torch.tensor(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)), dtype=torch.complex64,
                        device=dev)
    # distinct matrices: column scalings and a per-matrix diagonal shift of the base
    mats = []
    for i in range(nm):
        s = torch.tensor(rng.uniform(0.5, 2.0, n), dtype=torch.float32, device=dev)
        mats.append((base * s[None, :] + (0.1 * i) * torch.eye(n, dtype=torch.complex64, device=dev)).contiguous())
    lus = lu_factor_batched([A.clone() for A in mats], persistent=True)
    torch.cuda.synchronize()
    for i, (A, f) in enumerate(zip(mats, lus)):
        f.check()
        b = torch.ones(n, dtype=torch.complex64, device=dev)
        A64, x = A.to(torch.complex128), f.solve(b).to(torch.complex128)
        # normwise backward error of the solution
        r = ((A64 @ x - b.to(torch.complex128)).norm() /
             (torch.linalg.matrix_norm(A64) * x.norm() + b.norm())).item()
        assert r < 1e-5, (i, r)
        if i in (0, nm - 1):
            _check_factor(A, f, 1e-4)


def test_lu_batched_persistent_complex128_takes_lock_step(dev):
    """complex128 through the persistent entry point factors in lock step:
    pivots equal the single-matrix complex128 LU's."""
    from mesh_bem_amd.dense import LUFactor, lu_factor_batched

    rng = np.random.default_rng(5)
    mats = [torch.tensor(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)), dtype=torch.complex128,
                         device=dev) for n in (90, 300)]
    lus = lu_factor_batched([A.clone() for A in mats], persistent=True)
    torch.cuda.synchronize()
    for A, f in zip(mats, lus):
        f.check()
        assert torch.equal(f.ipiv, LUFactor(A.clone()).ipiv)


def test_lu_batched_persistent_schedules(dev):
    """Every schedule the batch tuner can pick -- super panels of 4 or 8
    32-column panels, FP32 3M or split-FP16 trailing product -- gives factors
    with P A = L U to FP32 accuracy (ragged orders around 32, 128, 256)."""
    from mesh_bem_amd.batched import BatchLuTuner
    from mesh_bem_amd.dense import lu_factor_batched

    rng = np.random.default_rng(29)
    sizes = [31, 129, 257, 700, 1301]
    mats = [torch.tensor(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)), dtype=torch.complex64,
                         device=dev) for n in sizes]
    for cand in BatchLuTuner.CANDIDATES:
        print(f"persistent LU schedule {cand}", flush=True)
        lus = lu_factor_batched([A.clone() for A in mats], persistent=True, options=cand)
        torch.cuda.synchronize()
        for A, f in zip(mats, lus):
            f.check()
            _check_factor(A, f, 1e-4)


def test_batch_lu_tuner_picks_and_caches(dev, tmp_path):
    """The tuner times its candidates on the batch, keeps the fastest per batch
    shape (also in its JSON file) and answers from the cache afterwards."""
    from mesh_bem_amd.batched import BatchLuTuner

    rng = np.random.default_rng(31)
    mats = [torch.tensor(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)), dtype=torch.complex64,
                         device=dev) for n in (300, 900, 1500)]
    path = str(tmp_path / "lu_tuner.json")
    t = BatchLuTuner(path=path, reps=1)
    c = t.tune(mats)
    assert c in BatchLuTuner.CANDIDATES and len(t.timings[t.key(mats)]) == len(BatchLuTuner.CANDIDATES)
    t2 = BatchLuTuner(path=path)
    assert t2.tune(mats) == c and not t2.timings
