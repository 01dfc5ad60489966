"""CPU FP64 oracle of the mesh-graph network -- TEST INFRASTRUCTURE ONLY.

Forward pass, loss and hand-written backward pass of the learned warm-start
network, in numpy.  With ``bf16=True`` every GEMM operand is rounded to BF16
(round to nearest even) exactly where the HIP kernels round it, so the GPU
forward/backward can be checked to FP32-accumulation accuracy; with
``bf16=False`` the backward is pinned by central finite differences of the
forward (tests/test_mgn_cpu.py).  The network has no reference counterpart.

Architecture (shared layout with mesh_bem_amd/mgn.py): node and edge encoders
MLP(in -> H -> H); L processor layers
    e' = e + MLP_e([h[src], h[dst], e]),  agg_i = sum_{dst(e) = i} e',
    h' = h + MLP_n([h, agg]);
decoder MLP(H -> H -> 2).  MLP(x) = silu(x W1 + b1) W2 + b2; a weight block is
[(K + 1), N] with the bias as its last row.
"""

import numpy as np


def layer_shapes(nf, ef, hidden, layers):
    H = hidden
    sh = [("enc_n1", nf, H), ("enc_n2", H, H), ("enc_e1", ef, H), ("enc_e2", H, H)]
    for l in range(layers):
        sh += [(f"pe{l}_1", 3 * H, H), (f"pe{l}_2", H, H), (f"pn{l}_1", 2 * H, H), (f"pn{l}_2", H, H)]
    sh += [("dec1", H, H), ("dec2", H, 2)]
    return sh


def unpack(flat, shapes):
    out, o = {}, 0
    for name, k, n in shapes:
        sz = (k + 1) * n
        out[name] = flat[o:o + sz].reshape(k + 1, n)
        o += sz
    return out


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * _sig(z)


def dsilu(z):
    s = _sig(z)
    return s * (1.0 + z * (1.0 - s))


class Net:
    def __init__(self, nf, ef, hidden, layers, bf=False):
        self.shapes = layer_shapes(nf, ef, hidden, layers)
        self.L = layers
        self.H = hidden
        self.bf = bf

    def _r(self, x):
        return bf16(x) if self.bf else x

    def _lin(self, X, W):
        Xa = np.concatenate([X, np.ones((X.shape[0], 1))], 1)
        return self._r(Xa) @ self._r(W)

    def _mlp(self, X, W1, W2):
        z1 = self._lin(X, W1)
        a1 = silu(z1)
        return self._lin(a1, W2), (X, z1, a1)

    def _mlp_back(self, dout, cache, W1, W2, g1, g2):
        X, z1, a1 = cache
        r = self._r
        g2 += r(np.concatenate([a1, np.ones((a1.shape[0], 1))], 1)).T @ r(dout)
        dz1 = (r(dout) @ r(W2[:-1]).T) * dsilu(z1)
        g1 += r(np.concatenate([X, np.ones((X.shape[0], 1))], 1)).T @ r(dz1)
        return r(dz1) @ r(W1[:-1]).T

    def forward(self, flat, g):
        """g: dict with x_n [Nn, nf], x_e [Ne, ef], src, dst [Ne]. Returns (y [Nn, 2], caches)."""
        P = unpack(flat, self.shapes)
        h, cn = self._mlp(g["x_n"], P["enc_n1"], P["enc_n2"])
        e, ce = self._mlp(g["x_e"], P["enc_e1"], P["enc_e2"])
        caches = {"enc_n": cn, "enc_e": ce}
        src, dst = g["src"], g["dst"]
        Nn = h.shape[0]
        for l in range(self.L):
            Xe = np.concatenate([h[src], h[dst], e], 1)
            de, c1 = self._mlp(Xe, P[f"pe{l}_1"], P[f"pe{l}_2"])
            e = e + de
            agg = np.zeros((Nn, self.H))
            np.add.at(agg, dst, e)
            Xn = np.concatenate([h, agg], 1)
            dh, c2 = self._mlp(Xn, P[f"pn{l}_1"], P[f"pn{l}_2"])
            h = h + dh
            caches[f"pe{l}"] = c1
            caches[f"pn{l}"] = c2
        y, cd = self._mlp(h, P["dec1"], P["dec2"])
        caches["dec"] = cd
        return y, caches

    @staticmethod
    def loss(y, t, goff):
        """Mean over graphs of |y - t|^2 / |t|^2 and its gradient."""
        G = len(goff) - 1
        L = 0.0
        dy = np.zeros_like(y)
        for k in range(G):
            a, b = goff[k], goff[k + 1]
            d = y[a:b] - t[a:b]
            den = np.sum(t[a:b] ** 2)
            L += np.sum(d ** 2) / den / G
            dy[a:b] = 2.0 * d / (den * G)
        return L, dy

    def backward(self, flat, g, caches, dy):
        P = unpack(flat, self.shapes)
        grads = np.zeros_like(flat)
        Gd = unpack(grads, self.shapes)
        src, dst = g["src"], g["dst"]
        H = self.H
        dh = self._mlp_back(dy, caches["dec"], P["dec1"], P["dec2"], Gd["dec1"], Gd["dec2"])
        de = np.zeros((len(src), H))
        for l in reversed(range(self.L)):
            dXn = self._mlp_back(dh, caches[f"pn{l}"], P[f"pn{l}_1"], P[f"pn{l}_2"], Gd[f"pn{l}_1"],
                                 Gd[f"pn{l}_2"])
            dh = dh + dXn[:, :H]
            dagg = dXn[:, H:]
            de = de + dagg[dst]
            dXe = self._mlp_back(de, caches[f"pe{l}"], P[f"pe{l}_1"], P[f"pe{l}_2"], Gd[f"pe{l}_1"],
                                 Gd[f"pe{l}_2"])
            de = de + dXe[:, 2 * H:]
            np.add.at(dh, src, dXe[:, :H])
            np.add.at(dh, dst, dXe[:, H:2 * H])
        self._mlp_back(de, caches["enc_e"], P["enc_e1"], P["enc_e2"], Gd["enc_e1"], Gd["enc_e2"])
        self._mlp_back(dh, caches["enc_n"], P["enc_n1"], P["enc_n2"], Gd["enc_n1"], Gd["enc_n2"])
        return grads
