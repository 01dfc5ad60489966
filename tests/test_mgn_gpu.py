"""Mesh-graph network kernels on the GPU against the CPU oracle.

* GEMM with gathered / concatenated operands, transposes, split-K and every
  epilogue option vs numpy with the same BF16 operand rounding (FP32
  accumulation: 1e-5 relative).
* Forward output, loss and all parameter gradients of the full network vs the
  oracle with BF16 emulation.  Bars 2e-3 (output, gradients) and 1e-2 per
  weight block: the GPU keeps intermediates in FP32 where the oracle keeps
  FP64, and those last-bit differences flip individual BF16 roundings of later
  operands (rounding the oracle's intermediates to FP32 alone moves its output
  by ~5e-4 relative on this graph).
* Training on solver-generated data lowers the loss, and the learned warm start
  lowers GMRES iterations on held-out meshes.
"""

import numpy as np
import pytest
import torch

from mgn_oracle import Net, bf16
from mesh_bem_amd import mgn as M

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


@pytest.mark.parametrize("case", ["rows", "gather_ones", "trans_split", "btrans_epilogue"])
def test_gemm_modes(dev, case):
    rng = np.random.default_rng(abs(hash(case)) % 1000)
    f32 = lambda *s: rng.normal(size=s).astype(np.float32)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    if case == "rows":
        M_, K, N = 300, 70, 130
        A, B = f32(M_, K), f32(K, N)
        Y = torch.empty((M_, N), device=dev)
        M.gemm([(T(A), None, K)], False, 0, T(B), 0, M_, N, K, Y)
        ref = bf16(A) @ bf16(B)
    elif case == "gather_ones":
        M_, n1, n2, w = 257, 40, 50, 33
        P1, P2, P3 = f32(n1, w), f32(n2, w + 5), f32(M_, 17)
        i1 = rng.integers(0, n1, M_).astype(np.int32)
        i2 = rng.integers(0, n2, M_).astype(np.int32)
        K = w + w + 17 + 1
        N = 128
        B = f32(K, N)
        Y = torch.empty((M_, N), device=dev)
        P2t = T(P2)
        M.gemm([(T(P1), T(i1), w), (P2t, T(i2), w), (T(P3), None, 17)], True, 0, T(B), 0, M_, N, K, Y)
        X = np.concatenate([P1[i1], P2[i2][:, :w], P3, np.ones((M_, 1), np.float32)], 1)
        ref = bf16(X) @ bf16(B)
    elif case == "trans_split":
        R_, K1, N = 9000, 45, 96
        X = f32(R_, K1)
        D = f32(R_, N)
        Y = torch.empty((K1 + 1, N), device=dev)
        M.gemm([(T(X), None, K1)], True, 1, T(D), 0, K1 + 1, N, R_, Y, split=7)
        Xa = np.concatenate([X, np.ones((R_, 1), np.float32)], 1)
        ref = bf16(Xa).T @ bf16(D)
    else:
        M_, K, N = 200, 128, 60
        A = f32(M_, K)
        W = f32(N + 1, K)  # Bop[k, n] = W[n, k]
        Zp = f32(M_, N)
        R = f32(M_, N)
        Y0 = f32(M_, N)
        Y = T(Y0.copy())
        Z = torch.empty((M_, N), device=dev)
        M.gemm([(T(A), None, K)], False, 0, T(W), 1, M_, N, K, Y, Z=Z, R=T(R), Zp=T(Zp), act=1, accumulate=1)
        v = bf16(A) @ bf16(W[:N]).T
        s = _sig(Zp.astype(np.float64))
        v = v * s * (1 + Zp * (1 - s))
        assert rel(Z.cpu().numpy(), v) < 1e-5
        v = v * _sig(v) + R
        ref = Y0 + v
    e = rel(Y.cpu().numpy(), ref)
    print(f"  {case}: rel {e:.2e}")
    assert e < 1e-5


def test_segsum_and_adam(dev):
    rng = np.random.default_rng(5)
    rows, nin, w = 50, 80, 7
    ptr = np.concatenate([[0], np.cumsum(rng.integers(0, 4, rows))]).astype(np.int64)
    idx = rng.integers(0, nin, ptr[-1]).astype(np.int32)
    inp = rng.normal(size=(nin, 12)).astype(np.float32)
    out0 = rng.normal(size=(rows, w)).astype(np.float32)
    out = torch.tensor(out0, device=dev)
    M.segsum(torch.tensor(ptr, device=dev), torch.tensor(idx, device=dev), rows, torch.tensor(inp, device=dev), 3, w,
             out, 1)
    ref = out0.astype(np.float64).copy()
    for i in range(rows):
        for q in range(ptr[i], ptr[i + 1]):
            ref[i] += inp[idx[q], 3:3 + w]
    assert rel(out.cpu().numpy(), ref) < 1e-6
    net = M.MeshGraphNet(hidden=8, layers=1, nf=3, ef=2, device=dev)
    p0 = net.params.cpu().numpy().astype(np.float64)
    g = rng.normal(size=p0.shape)
    net.grads.copy_(torch.tensor(g, dtype=torch.float32))
    net.adam(lr=1e-2)
    m = 0.1 * g
    v = 0.001 * g * g
    ref = p0 - 1e-2 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    assert rel(net.params.cpu().numpy(), ref) < 1e-6


def _graphs(rng, sizes, nf, ef):
    xs, srcs, dsts, goff = [], [], [], [0]
    for nn in sizes:
        ne = 3 * nn
        srcs.append(rng.integers(0, nn, ne) + goff[-1])
        dsts.append(rng.integers(0, nn, ne) + goff[-1])
        goff.append(goff[-1] + nn)
    src, dst = np.concatenate(srcs), np.concatenate(dsts)
    x_n = rng.normal(size=(goff[-1], nf)).astype(np.float32)
    x_e = rng.normal(size=(len(src), ef)).astype(np.float32)
    return x_n, x_e, src, dst, np.array(goff)


def test_network_forward_backward_match_oracle(dev):
    rng = np.random.default_rng(11)
    nf, ef, H, L = M.NODE_FEATURES, M.EDGE_FEATURES, 128, 2
    x_n, x_e, src, dst, goff = _graphs(rng, [70, 90, 40], nf, ef)
    net = M.MeshGraphNet(hidden=H, layers=L, device=dev, seed=3)
    g = M.make_graph(x_n, x_e, src, dst, goff, dev)
    t = rng.normal(size=(goff[-1], 2)).astype(np.float32)
    y, c = net.forward(g)
    lo, dy = net.loss(y, torch.tensor(t, device=dev), g)
    net.backward(g, c, dy)
    flat = net.params.cpu().numpy().astype(np.float64)
    on = Net(nf, ef, H, L, bf=True)
    gd = dict(x_n=x_n.astype(np.float64), x_e=x_e.astype(np.float64), src=src, dst=dst)
    yo, co = on.forward(flat, gd)
    Lo, dyo = on.loss(yo, t.astype(np.float64), goff)
    go = on.backward(flat, gd, co, dyo)
    ey = rel(y.cpu().numpy(), yo)
    el = abs(float(lo.mean()) - Lo) / Lo
    eg = rel(net.grads.cpu().numpy(), go)
    print(f"  forward {ey:.2e}  loss {el:.2e}  grads {eg:.2e}")
    assert ey < 2e-3 and el < 1e-4 and eg < 2e-3
    # per-layer gradient agreement, so no block hides behind the global norm
    from mgn_oracle import unpack
    G1 = unpack(net.grads.cpu().numpy().astype(np.float64), net.shapes)
    G2 = unpack(go, net.shapes)
    for name in G2:
        assert rel(G1[name], G2[name]) < 1e-2, name


def test_training_and_warm_start(dev):
    """Small-scale learned warm start: 20 training meshes x 32 waves (ka in [1, 2.5]),
    3 held-out meshes.  The trained network must beat the zero guess on
    held-out currents (relative error < 1) and GMRES must converge from its
    optimally scaled prediction; the iteration counts of both starts are
    reported (on EFIE systems the prediction error is spread over the whole
    spectrum and saves few or no iterations)."""
    from mesh_bem_amd.gmres import gmres

    mds, smp = M.generate(20, 5, 32, seed0=100, k_range=(1.0, 2.5), device=dev)
    net = M.MeshGraphNet(device=dev, seed=1)
    hist = M.train(net, mds, smp, epochs=6, batch=16, lr=2e-3)
    print(f"  loss history {['%.3f' % h for h in hist]}")
    assert hist[-1] < 0.7 * hist[1] and hist[-1] < 0.01 * hist[0]
    tm, ts = M.generate(3, 5, 2, seed0=900, k_range=(1.0, 2.5), device=dev, keep_system=True)
    cold, warm, jerr = [], [], []
    for md, s in zip(tm, ts):
        J0 = M.predict_currents(net, md, s.x)
        for d in range(s.x.shape[0]):
            b = s.V32[:, d]
            J = torch.complex(s.J[d][:, 0], s.J[d][:, 1])
            jerr.append(float((J0[:, d] - J).norm() / J.norm()))
            _, ic = gmres(s.Z32, b, restart=50, tol=1e-6)
            _, iw = gmres(s.Z32, b, x0=M.scaled_guess(s.Z32, b, J0[:, d]), restart=50, tol=1e-6)
            assert ic.converged and iw.converged
            cold.append(ic.iterations)
            warm.append(iw.iterations)
    print(f"  held-out J relative error {np.mean(jerr):.3f}; GMRES iterations cold {np.mean(cold):.1f} "
          f"warm {np.mean(warm):.1f}")
    assert max(jerr) < 1.0


def test_training_resumes_exactly_from_checkpoint(dev, tmp_path):
    """A run interrupted after its first epoch and resumed from the checkpoint
    ends with the parameters and Adam state of the uninterrupted run."""
    mds, smp = M.generate(3, 3, 4, seed0=40, k_range=(1.0, 2.0), device=dev)
    ref = M.MeshGraphNet(hidden=32, layers=2, device=dev, seed=1)
    h_ref = M.train(ref, mds, smp, epochs=2, batch=4, lr=2e-3, seed=5)

    class Interrupt(Exception):
        pass

    def stop_in_epoch_2(msg):
        if msg.startswith("epoch 2/"):
            raise Interrupt

    ck = str(tmp_path / "net.pt")
    net = M.MeshGraphNet(hidden=32, layers=2, device=dev, seed=1)
    with pytest.raises(Interrupt):
        M.train(net, mds, smp, epochs=2, batch=4, lr=2e-3, seed=5, checkpoint=ck, log=stop_in_epoch_2)
    fresh = M.MeshGraphNet(hidden=32, layers=2, device=dev, seed=77)
    h = M.train(fresh, mds, smp, epochs=2, batch=4, lr=2e-3, seed=5, checkpoint=ck)
    assert h == h_ref
    assert torch.equal(fresh.params, ref.params) and torch.equal(fresh.m, ref.m) and torch.equal(fresh.v, ref.v)
    assert fresh.step_count == ref.step_count
