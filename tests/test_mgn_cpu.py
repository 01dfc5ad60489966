"""The mesh-graph network oracle pinned on the CPU: its hand-written backward
pass against central finite differences of its forward pass (FP64, no BF16
rounding), and its BF16 rounding against known bit patterns."""

import numpy as np

from mgn_oracle import Net, bf16, layer_shapes


def tiny_graph(rng, nn=7, ne=16, nf=5, ef=3):
    src = rng.integers(0, nn, ne)
    dst = rng.integers(0, nn, ne)
    return dict(x_n=rng.normal(size=(nn, nf)), x_e=rng.normal(size=(ne, ef)), src=src, dst=dst)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    nf, ef, H, L = 5, 3, 6, 2
    net = Net(nf, ef, H, L, bf=False)
    n = sum((k + 1) * m for _, k, m in layer_shapes(nf, ef, H, L))
    flat = rng.normal(scale=0.4, size=n)
    g = tiny_graph(rng, nf=nf, ef=ef)
    goff = np.array([0, 3, 7])
    t = rng.normal(size=(7, 2))
    y, c = net.forward(flat, g)
    L0, dy = net.loss(y, t, goff)
    grad = net.backward(flat, g, c, dy)
    eps = 1e-6
    idx = rng.choice(n, 60, replace=False)
    for i in idx:
        fp = flat.copy()
        fp[i] += eps
        fm = flat.copy()
        fm[i] -= eps
        lp = net.loss(net.forward(fp, g)[0], t, goff)[0]
        lm = net.loss(net.forward(fm, g)[0], t, goff)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_bf16_rounding_bits():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.140625], dtype=np.float32)
    r = bf16(x)
    # ties to even at 1 + 2^-8 (halfway between 1 and 1 + 2^-7): rounds to 1
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == 1.0 and r[3] == 1.0 + 2 ** -7
    assert r[4] == -2.5 and r[5] == 3.140625
