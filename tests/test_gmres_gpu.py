"""FP32 GMRES building blocks and solver on the GPU.

Building blocks against numpy in FP64 (bars from FP32 rounding: matvec 1e-6
relative, inner products 1e-6, updates 1e-6).  The solver must reach the
requested relative residual, checked with an FP64 residual on the host, and
agree with LAPACK's solution to the accuracy the residual implies.
"""

import numpy as np
import pytest
import torch

from mesh_bem_amd import _lib
from mesh_bem_amd.gmres import _Blocks, gmres, matvec
from mesh_bem_amd.geometry import icosphere
from mesh_bem_amd.mom import MomMesh, assemble_ns, rhs_ns

pytestmark = pytest.mark.gpu


def crand(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("n,ld", [(1, 1), (7, 9), (128, 128), (1000, 1003), (2051, 2052)])
@pytest.mark.parametrize("cplx", [True, False])
def test_mv32(dev, n, ld, cplx):
    rng = np.random.default_rng(n)
    A = crand(rng, n, ld) if cplx else rng.standard_normal((n, ld)).astype(np.float32)
    x = crand(rng, n) if cplx else rng.standard_normal(n).astype(np.float32)
    b = crand(rng, n) if cplx else rng.standard_normal(n).astype(np.float32)
    At = torch.from_numpy(A).to(dev)[:, :n]
    xt = torch.from_numpy(x).to(dev)
    bt = torch.from_numpy(b).to(dev)
    ref = A[:, :n].astype(np.complex128) @ x.astype(np.complex128)
    y = torch.empty_like(bt)
    lib = _lib.load()
    fld = _lib.MBA_COMPLEX if cplx else _lib.MBA_REAL
    st = lib.mba_mv32(fld, _lib.ptr(At), n, n, ld, 0, _lib.ptr(xt), None, _lib.ptr(y), _lib.stream())
    _lib.check(st)
    scale = np.abs(A[:, :n]).astype(np.float64) @ np.abs(x).astype(np.float64)
    assert np.max(np.abs(y.cpu().numpy() - ref) / scale) < 1e-6
    st = lib.mba_mv32(fld, _lib.ptr(At), n, n, ld, 0, _lib.ptr(xt), _lib.ptr(bt), _lib.ptr(y), _lib.stream())
    _lib.check(st)
    assert np.max(np.abs(y.cpu().numpy() - (b - ref)) / (scale + np.abs(b))) < 1e-6
    # FP64 vectors over the FP32 matrix: exact products, FP64 sums
    x64, b64 = xt.to(torch.complex128 if cplx else torch.float64), bt.to(torch.complex128 if cplx else torch.float64)
    y64 = torch.empty_like(b64)
    st = lib.mba_mv32(fld, _lib.ptr(At), n, n, ld, 1, _lib.ptr(x64), _lib.ptr(b64), _lib.ptr(y64),
                      _lib.stream())
    _lib.check(st)
    assert np.max(np.abs(y64.cpu().numpy() - (b - ref)) / (scale + np.abs(b))) < 1e-13


@pytest.mark.parametrize("cplx", [True, False])
@pytest.mark.parametrize("n,nvec", [(5, 3), (3000, 17), (30720, 51)])
def test_dots_update_scale(dev, cplx, n, nvec):
    rng = np.random.default_rng(nvec)
    dt = torch.complex64 if cplx else torch.float32
    mk = (lambda *s: crand(rng, *s)) if cplx else (lambda *s: rng.standard_normal(s).astype(np.float32))
    Vb = mk(nvec + 1, n)
    w = mk(n)
    Vt = torch.from_numpy(Vb).to(dev)
    wt = torch.from_numpy(w).to(dev)
    blk = _Blocks(n, nvec + 1, _lib.MBA_COMPLEX if cplx else _lib.MBA_REAL, dev)
    out = torch.empty((nvec + 1, 2), dtype=torch.float64, device=dev)
    blk.dots(Vt, nvec, wt, out)
    o = out.cpu().numpy()
    ref = np.conj(Vb[:nvec].astype(np.complex128)) @ w.astype(np.complex128)
    got = o[:nvec, 0] + 1j * o[:nvec, 1]
    scale = np.abs(Vb[:nvec]).astype(np.float64) @ np.abs(w).astype(np.float64)
    assert np.max(np.abs(got - ref) / scale) < 1e-6
    nrm = np.sum(np.abs(w.astype(np.complex128)) ** 2)
    assert abs(o[nvec, 0] - nrm) / nrm < 1e-6 and o[nvec, 1] == 0.0
    # deterministic: a second call gives identical bits
    out2 = torch.empty_like(out)
    blk.dots(Vt, nvec, wt, out2)
    assert torch.equal(out, out2)
    coef = torch.from_numpy(np.stack([ref.real, ref.imag], 1).copy()).to(dev)
    w2 = wt.clone()
    blk.update(Vt, nvec, coef, -1.0, w2)
    c = ref if cplx else ref.real
    wref = w.astype(np.complex128) - Vb[:nvec].astype(np.complex128).T @ c
    sc = np.abs(w) + np.abs(Vb[:nvec]).astype(np.float64).T @ np.abs(c)
    assert np.max(np.abs(w2.cpu().numpy() - wref) / sc) < 1e-6
    dst = torch.empty_like(wt)
    nr = out[nvec:nvec + 1].contiguous()
    blk.scale(wt, nr, dst)
    np.testing.assert_allclose(dst.cpu().numpy(), w / np.sqrt(nrm), rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("cplx", [True, False])
def test_gmres_well_conditioned(dev, cplx):
    rng = np.random.default_rng(1)
    n = 700
    if cplx:
        A = np.eye(n) + 0.3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2 * n)
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        dt = torch.complex64
    else:
        A = np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
        b = rng.standard_normal(n)
        dt = torch.float32
    At = torch.tensor(A, dtype=dt, device=dev)
    bt = torch.tensor(b, dtype=dt, device=dev)
    x, info = gmres(At, bt, restart=20, tol=1e-6)
    xs = x.cpu().numpy().astype(np.complex128)
    res = np.linalg.norm(b - A @ xs) / np.linalg.norm(b)
    print(f"  it={info.iterations} restarts={info.restarts} res={res:.2e} info.res={info.residual:.2e}")
    assert info.converged and res < 2e-6
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(xs - xr) / np.linalg.norm(xr) < 1e-5


def test_gmres_efie_sphere(dev):
    """EFIE system of a 1920-RWG sphere (ka = 1): residual 1e-6 within restart cycles."""
    V, F = icosphere(3)
    mesh = MomMesh.build(torch.tensor(V, dtype=torch.float32), F, device=dev)
    Z = assemble_ns(mesh, 1.0, _lib.MBA_F32)
    b = rhs_ns(mesh, 1.0, _lib.MBA_F32, [[0, 0, 1.0]], [[1.0, 0, 0]])[:, 0].contiguous()
    x, info = gmres(Z, b, restart=50, tol=1e-6, maxiter=3000)
    Zc = Z.cpu().numpy().astype(np.complex128)
    bc = b.cpu().numpy().astype(np.complex128)
    res = np.linalg.norm(bc - Zc @ x.cpu().numpy()) / np.linalg.norm(bc)
    print(f"  EFIE it={info.iterations} restarts={info.restarts} res={res:.2e}")
    assert info.converged and res < 2e-6
    y = matvec(Z, x)
    assert np.linalg.norm(y.cpu().numpy() - Zc @ x.cpu().numpy()) / np.linalg.norm(bc) < 1e-5


def test_gmres_warm_start_exact(dev):
    rng = np.random.default_rng(2)
    n = 300
    A = np.eye(n) + 0.2 * rng.standard_normal((n, n)) / np.sqrt(n)
    xr = rng.standard_normal(n)
    b = A @ xr
    At = torch.tensor(A, dtype=torch.float32, device=dev)
    bt = torch.tensor(b, dtype=torch.float32, device=dev)
    x, info = gmres(At, bt, x0=torch.tensor(xr, dtype=torch.float64, device=dev), tol=1e-5)
    assert info.iterations == 0 and info.converged
    # the iterate is FP64: a converged solve reaches residuals below FP32 resolution
    x, info = gmres(At, bt, tol=1e-10)
    res = np.linalg.norm(A.astype(np.float32).astype(np.float64) @ x.cpu().numpy() - bt.double().cpu().numpy())
    assert info.converged and res / np.linalg.norm(b) < 2e-10


def test_nonfinite_warm_start_restarts_from_zero(dev):
    rng = np.random.default_rng(5)
    n = 400
    A = np.eye(n) + 0.3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2 * n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    At = torch.tensor(A, dtype=torch.complex64, device=dev)
    bt = torch.tensor(b, dtype=torch.complex64, device=dev)
    x0 = torch.zeros(n, dtype=torch.complex128, device=dev)
    x0[7] = complex(float("nan"), 0.0)
    x, info = gmres(At, bt, x0=x0, restart=30, tol=1e-8)
    xz, infoz = gmres(At, bt, restart=30, tol=1e-8)
    assert info.fallback == "x0-nonfinite" and info.converged and not infoz.fallback
    assert torch.equal(x, xz)
    x0 = torch.full((n,), float("inf"), dtype=torch.complex128, device=dev)
    x, info = gmres(At, bt, x0=x0, restart=30, tol=1e-8)
    assert info.fallback == "x0-nonfinite" and torch.equal(x, xz)


@pytest.mark.parametrize("cplx", [True, False])
def test_unconverged_gmres_falls_back_to_lu(dev, cplx):
    """An iteration cap far below what the system needs leaves GMRES
    unconverged; fallback="lu" finishes with a complex64 LU and FP64-residual
    refinement to the same tolerance."""
    rng = np.random.default_rng(6)
    n = 500
    # spread spectrum: slow for GMRES with a small restart
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = U @ np.diag(np.linspace(0.01, 1.0, n)) @ U.T + 0.05 * rng.standard_normal((n, n)) / np.sqrt(n)
    if cplx:
        A = A + 0.05j * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    dt = torch.complex64 if cplx else torch.float32
    At = torch.tensor(A, dtype=dt, device=dev)
    bt = torch.tensor(b, dtype=dt, device=dev)
    x, info = gmres(At, bt, restart=5, tol=1e-9, maxiter=20)
    assert not info.converged and not info.fallback
    x, info = gmres(At, bt, restart=5, tol=1e-9, maxiter=20, fallback="lu")
    Af = At.cpu().numpy().astype(np.complex128)
    res = np.linalg.norm(bt.cpu().numpy() - Af @ x.cpu().numpy()) / np.linalg.norm(b)
    print(f"  LU fallback: info {info.fallback} converged {info.converged} residual {res:.2e}")
    assert info.fallback == "lu" and info.converged and res < 1e-9
