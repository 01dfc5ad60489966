"""Learned warm start: a mesh-graph network predicting RWG currents.

Graph: one node per RWG function, directed edges between the RWG functions of
each triangle (6 per face).  Node inputs: edge geometry in a local frame, the
wavenumber, the tested incident field V_e, the incident field at the edge
midpoint and the physical-optics estimate of the current crossing the edge.
Edge inputs: midpoint offsets and the T+/T- roles of the shared face.
Output: the RWG coefficient J_e (re, im) of the EFIE solution for a unit
plane wave.

Network (no reference counterpart): node / edge encoders MLP(in -> H -> H),
``layers`` processor layers
    e' = e + MLP_e([h[src], h[dst], e]),  agg_i = sum_{dst(e) = i} e',
    h' = h + MLP_n([h, agg]),
decoder MLP(H -> H -> 2); MLP(x) = silu(x W1 + b1) W2 + b2.  Every GEMM
(forward and backward) runs on the HIP BF16-MFMA kernel with gathered /
concatenated operands (``mba_mgn_gemm``); aggregations are deterministic
segment sums; the loss is the mean relative L2 error per sample; training
uses Adam, all on the GPU.  The prediction seeds GMRES (FP32 matrix).
"""

import ctypes
import os
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from . import _lib

NODE_FEATURES = 28
EDGE_FEATURES = 6


def layer_shapes(nf=NODE_FEATURES, ef=EDGE_FEATURES, hidden=128, layers=6):
    H = hidden
    sh = [("enc_n1", nf, H), ("enc_n2", H, H), ("enc_e1", ef, H), ("enc_e2", H, H)]
    for l in range(layers):
        sh += [(f"pe{l}_1", 3 * H, H), (f"pe{l}_2", H, H), (f"pn{l}_1", 2 * H, H), (f"pn{l}_2", H, H)]
    sh += [("dec1", H, H), ("dec2", H, 2)]
    return sh


# ----------------------------------------------------------------- graphs
@dataclass
class Graph:
    """A batch of graphs on the device (node rows grouped per graph)."""

    x_n: torch.Tensor  # [Nn, nf] float32
    x_e: torch.Tensor  # [Ne, ef] float32
    src: torch.Tensor  # [Ne] int32
    dst: torch.Tensor  # [Ne] int32
    dst_ptr: torch.Tensor  # [Nn+1] int64: incoming edges of node i (CSR, ascending edge id)
    dst_idx: torch.Tensor  # [Ne] int32
    src_ptr: torch.Tensor  # [Nn+1] int64: outgoing edges
    src_idx: torch.Tensor  # [Ne] int32
    goff: torch.Tensor  # [G+1] int64 node offsets per graph
    eye_ptr: torch.Tensor  # [Ne+1] int64 (0..Ne), for per-edge gathers

    @property
    def n_nodes(self):
        return self.x_n.shape[0]

    @property
    def n_edges(self):
        return self.x_e.shape[0]


def _csr(keys: np.ndarray, n: int):
    order = np.argsort(keys, kind="stable")
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ptr, keys + 1, 1)
    return np.cumsum(ptr), order.astype(np.int32)


def make_graph(x_n, x_e, src, dst, goff, device) -> Graph:
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    Nn = int(goff[-1])
    dp, di = _csr(dst, Nn)
    sp, si = _csr(src, Nn)
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a)).to(device=device, dtype=dt)
    return Graph(
        x_n=x_n.to(device=device, dtype=torch.float32).contiguous() if torch.is_tensor(x_n) else t(x_n, torch.float32),
        x_e=x_e.to(device=device, dtype=torch.float32).contiguous() if torch.is_tensor(x_e) else t(x_e, torch.float32),
        src=t(src, torch.int32), dst=t(dst, torch.int32), dst_ptr=t(dp, torch.int64), dst_idx=t(di, torch.int32),
        src_ptr=t(sp, torch.int64), src_idx=t(si, torch.int32), goff=t(np.asarray(goff), torch.int64),
        eye_ptr=torch.arange(len(src) + 1, dtype=torch.int64, device=device))


# ----------------------------------------------------------------- kernels
def _vmat(segs, ones):
    v = _lib.MgnVMat()
    for q, (t, idx, w) in enumerate(segs):
        v.p[q] = t.data_ptr()
        v.idx[q] = idx.data_ptr() if idx is not None else None
        v.ld[q] = t.stride(0)
        v.w[q] = w
    v.nseg = len(segs)
    v.ones = 1 if ones else 0
    return v


def gemm(segs, ones, a_trans, B, b_trans, M, N, K, Y, Z=None, R=None, Zp=None, act=0, accumulate=0,
         split=1):
    lib = _lib.load()
    A = _vmat(segs, ones)
    part = None
    if split > 1:
        part = torch.empty(split * M * N, dtype=torch.float32, device=Y.device)
    st = lib.mba_mgn_gemm(ctypes.byref(A), int(a_trans), _lib.ptr(B), B.stride(0) if B.dim() == 2 else N,
                          int(b_trans), M, N, K, _lib.ptr(Y), Y.stride(0),
                          _lib.ptr(Z) if Z is not None else None, _lib.ptr(R) if R is not None else None,
                          _lib.ptr(Zp) if Zp is not None else None, int(act), int(accumulate), int(split),
                          _lib.ptr(part) if part is not None else None, _lib.stream())
    _lib.check(st, "mba_mgn_gemm")


def segsum(ptr, idx, rows, inp, coff, width, out, accumulate):
    lib = _lib.load()
    st = lib.mba_mgn_segsum(_lib.ptr(ptr), _lib.ptr(idx), rows, _lib.ptr(inp), inp.stride(0), coff, width,
                            _lib.ptr(out), out.stride(0), int(accumulate), _lib.stream())
    _lib.check(st, "mba_mgn_segsum")


def _split_for(rows):
    return int(min(256, max(1, rows // 2048)))


# ----------------------------------------------------------------- network
class MeshGraphNet:
    def __init__(self, hidden=128, layers=6, nf=NODE_FEATURES, ef=EDGE_FEATURES, device="cuda", seed=0):
        self.H, self.L, self.nf, self.ef = hidden, layers, nf, ef
        self.shapes = layer_shapes(nf, ef, hidden, layers)
        self.offsets = {}
        o = 0
        for name, k, n in self.shapes:
            self.offsets[name] = (o, k, n)
            o += (k + 1) * n
        self.n_params = o
        rng = np.random.Generator(np.random.PCG64(seed))
        flat = np.zeros(o, dtype=np.float32)
        for name, k, n in self.shapes:
            off = self.offsets[name][0]
            gain = 0.1 if name == "dec2" else (0.5 if name.endswith("_2") and name[0] == "p" else 1.0)
            w = rng.normal(0.0, gain * np.sqrt(2.0 / k), size=(k, n))
            flat[off:off + k * n] = w.reshape(-1)
        self.device = torch.device(device)
        self.params = torch.from_numpy(flat).to(self.device)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step_count = 0

    def W(self, name, flat=None):
        o, k, n = self.offsets[name]
        t = self.params if flat is None else flat
        return t[o:o + (k + 1) * n].view(k + 1, n)

    def G(self, name):
        return self.W(name, self.grads)

    # ------------------------------------------------------------ building blocks
    def _mlp(self, segs, rows, name1, name2, out, R=None, store=True):
        W1, W2 = self.W(name1), self.W(name2)
        k_in = W1.shape[0] - 1
        H = W1.shape[1]
        z1 = torch.empty((rows, H), dtype=torch.float32, device=self.device)
        a1 = torch.empty_like(z1)
        gemm(segs, True, 0, W1, 0, rows, H, k_in + 1, a1, Z=z1, act=1)
        gemm([(a1, None, H)], True, 0, W2, 0, rows, W2.shape[1], H + 1, out, R=R)
        return (segs, rows, z1, a1, name1, name2) if store else None

    def _mlp_back(self, dout, cache, dX=None, dX_accumulate=False):
        segs, rows, z1, a1, name1, name2 = cache
        W1, W2 = self.W(name1), self.W(name2)
        g1, g2 = self.G(name1), self.G(name2)
        k_in, H, N2 = W1.shape[0] - 1, W1.shape[1], W2.shape[1]
        sp = _split_for(rows)
        gemm([(a1, None, H)], True, 1, dout, 0, H + 1, N2, rows, g2, split=sp)
        dz1 = torch.empty((rows, H), dtype=torch.float32, device=self.device)
        gemm([(dout, None, N2)], False, 0, W2, 1, rows, H, N2, dz1, Zp=z1)
        gemm(segs, True, 1, dz1, 0, k_in + 1, H, rows, g1, split=sp)
        if dX is None:
            return None
        gemm([(dz1, None, H)], False, 0, W1, 1, rows, k_in, H, dX, accumulate=int(dX_accumulate))
        return dX

    # ------------------------------------------------------------ forward / backward
    def forward(self, g: Graph, store=True):
        H, dev = self.H, self.device
        Nn, Ne = g.n_nodes, g.n_edges
        c = {}
        h = torch.empty((Nn, H), dtype=torch.float32, device=dev)
        e = torch.empty((Ne, H), dtype=torch.float32, device=dev)
        c["enc_n"] = self._mlp([(g.x_n, None, self.nf)], Nn, "enc_n1", "enc_n2", h, store=store)
        c["enc_e"] = self._mlp([(g.x_e, None, self.ef)], Ne, "enc_e1", "enc_e2", e, store=store)
        for l in range(self.L):
            e2 = torch.empty_like(e)
            c[f"pe{l}"] = self._mlp([(h, g.src, H), (h, g.dst, H), (e, None, H)], Ne, f"pe{l}_1", f"pe{l}_2",
                                    e2, R=e, store=store)
            agg = torch.empty_like(h)
            segsum(g.dst_ptr, g.dst_idx, Nn, e2, 0, H, agg, 0)
            h2 = torch.empty_like(h)
            c[f"pn{l}"] = self._mlp([(h, None, H), (agg, None, H)], Nn, f"pn{l}_1", f"pn{l}_2", h2, R=h,
                                    store=store)
            h, e = h2, e2
        y = torch.empty((Nn, 2), dtype=torch.float32, device=dev)
        c["dec"] = self._mlp([(h, None, H)], Nn, "dec1", "dec2", y, store=store)
        return y, c

    def backward(self, g: Graph, c, dy: torch.Tensor):
        H, dev = self.H, self.device
        Nn, Ne = g.n_nodes, g.n_edges
        dh = torch.empty((Nn, H), dtype=torch.float32, device=dev)
        self._mlp_back(dy, c["dec"], dh)
        de = torch.zeros((Ne, H), dtype=torch.float32, device=dev)
        for l in reversed(range(self.L)):
            dXn = torch.empty((Nn, 2 * H), dtype=torch.float32, device=dev)
            self._mlp_back(dh, c[f"pn{l}"], dXn)
            # dh += dXn[:, :H] ; de += dagg[dst]
            dh_new = dXn[:, :H].contiguous()
            dh_new += dh
            segsum(g.eye_ptr, g.dst, Ne, dXn, H, H, de, 1)
            dXe = torch.empty((Ne, 3 * H), dtype=torch.float32, device=dev)
            self._mlp_back(de, c[f"pe{l}"], dXe)
            de_new = dXe[:, 2 * H:].contiguous()
            de_new += de
            segsum(g.src_ptr, g.src_idx, Nn, dXe, 0, H, dh_new, 1)
            segsum(g.dst_ptr, g.dst_idx, Nn, dXe, H, H, dh_new, 1)
            dh, de = dh_new, de_new
        self._mlp_back(de, c["enc_e"], None)
        self._mlp_back(dh, c["enc_n"], None)

    def loss(self, y, target, g: Graph, want_grad=True):
        lib = _lib.load()
        G = g.goff.shape[0] - 1
        lo = torch.empty(G, dtype=torch.float32, device=self.device)
        dy = torch.empty_like(y) if want_grad else None
        st = lib.mba_mgn_loss(_lib.ptr(y), _lib.ptr(target), _lib.ptr(g.goff), G, _lib.ptr(lo),
                              _lib.ptr(dy) if dy is not None else None, _lib.stream())
        _lib.check(st, "mba_mgn_loss")
        return lo, dy

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        lib = _lib.load()
        self.step_count += 1
        st = lib.mba_mgn_adam(_lib.ptr(self.params), _lib.ptr(self.grads), _lib.ptr(self.m), _lib.ptr(self.v),
                              self.n_params, lr, b1, b2, eps, self.step_count, _lib.stream())
        _lib.check(st, "mba_mgn_adam")

    def train_step(self, g: Graph, target: torch.Tensor, lr=1e-3):
        y, c = self.forward(g)
        lo, dy = self.loss(y, target, g)
        self.backward(g, c, dy)
        self.adam(lr)
        return lo

    def predict(self, g: Graph) -> torch.Tensor:
        y, _ = self.forward(g, store=False)
        return y

    # ------------------------------------------------------------- checkpoints
    def state_dict(self) -> dict:
        """Parameters, Adam moments and step count (tensors only, so the file
        loads with torch.load(weights_only=True))."""
        return {
            "config": torch.tensor([self.H, self.L, self.nf, self.ef], dtype=torch.int64),
            "params": self.params.detach().cpu(),
            "adam_m": self.m.detach().cpu(),
            "adam_v": self.v.detach().cpu(),
            "step_count": torch.tensor(self.step_count, dtype=torch.int64),
        }

    def load_state_dict(self, sd: dict) -> None:
        cfg = [int(x) for x in sd["config"].tolist()]
        if cfg != [self.H, self.L, self.nf, self.ef]:
            raise ValueError(f"checkpoint of a network with (hidden, layers, nf, ef) = {cfg}")
        for name in ("params", "adam_m", "adam_v"):
            if sd[name].numel() != self.n_params:
                raise ValueError(f"checkpoint {name}: {sd[name].numel()} values, network has {self.n_params}")
        self.params.copy_(sd["params"].to(self.device))
        self.m.copy_(sd["adam_m"].to(self.device))
        self.v.copy_(sd["adam_v"].to(self.device))
        self.step_count = int(sd["step_count"])

    def save(self, path: str, extra: Optional[dict] = None) -> None:
        """Checkpoint to ``path`` (written to a temporary file, then renamed)."""
        sd = self.state_dict()
        for k, v in (extra or {}).items():
            sd[f"extra_{k}"] = torch.as_tensor(v)
        tmp = path + ".tmp"
        torch.save(sd, tmp)
        os.replace(tmp, path)

    def load(self, path: str) -> dict:
        """Restore a checkpoint; returns its extra entries."""
        sd = torch.load(path, map_location="cpu", weights_only=True)
        self.load_state_dict(sd)
        return {k[len("extra_"):]: v for k, v in sd.items() if k.startswith("extra_")}


# ----------------------------------------------------------------- features
@dataclass
class MeshData:
    """Per-mesh graph topology and geometric features (device)."""

    verts: np.ndarray
    faces: np.ndarray
    E: int
    src: torch.Tensor  # [4E] int32 (device)
    dst: torch.Tensor
    dptr: torch.Tensor  # CSR of incoming edges [E+1] int64, idx int32
    didx: torch.Tensor
    sptr: torch.Tensor  # CSR of outgoing edges
    sidx: torch.Tensor
    x_e: torch.Tensor  # [4E, EF]
    geo: dict  # per-RWG frames (torch, device)


def mesh_data(verts, faces, face_edge_indices, n_edges, device) -> MeshData:
    from .geometry import rwg_face_edges

    V = np.asarray(verts, dtype=np.float64)
    F = np.asarray(faces, dtype=np.int64)
    fe = np.asarray(face_edge_indices, dtype=np.int64).reshape(-1)
    E = int(n_edges)
    rfe = rwg_face_edges(fe, E).astype(np.int64)
    f_of = rfe // 3
    le = rfe % 3
    plus = F[f_of, (le + 1) % 3] < F[f_of, (le + 2) % 3]  # [E, 2] sign of each slot
    tp = np.where(plus[:, 0], 0, 1)
    fp = f_of[np.arange(E), tp]
    fm = f_of[np.arange(E), 1 - tp]
    lp = le[np.arange(E), tp]
    a = V[F[fp, (lp + 1) % 3]]
    b = V[F[fp, (lp + 2) % 3]]
    mid = 0.5 * (a + b)
    t = b - a
    l = np.linalg.norm(t, axis=1, keepdims=True)
    t = t / l
    P = V[F]
    nf = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 1])
    nf /= np.linalg.norm(nf, axis=1, keepdims=True)
    nb = nf[fp] + nf[fm]
    nb /= np.linalg.norm(nb, axis=1, keepdims=True)
    cdir = np.cross(t, nb)
    cm = P[fm].mean(1) - P[fp].mean(1)
    cdir *= np.sign(np.sum(cm * cdir, axis=1, keepdims=True))
    hmean = float(l.mean())
    # graph edges: ordered pairs of the RWG functions of each face
    eid = fe.reshape(-1, 3)
    src, dst, sgn_s, sgn_d = [], [], [], []
    rwg_plus = np.zeros((len(F), 3))
    for e in range(E):
        for s in range(2):
            rwg_plus[f_of[e, s], le[e, s]] = 1.0 if s == tp[e] else -1.0
    for i in range(3):
        for j in range(3):
            if i != j:
                src.append(eid[:, i])
                dst.append(eid[:, j])
                sgn_s.append(rwg_plus[:, i])
                sgn_d.append(rwg_plus[:, j])
    src = np.concatenate(src)
    dst = np.concatenate(dst)
    d = (mid[dst] - mid[src]) / hmean
    x_e = np.concatenate([d, np.linalg.norm(d, axis=1, keepdims=True), np.concatenate(sgn_s)[:, None],
                          np.concatenate(sgn_d)[:, None]], 1)
    T = lambda x: torch.as_tensor(x, dtype=torch.float64, device=device)
    I = lambda x, dt: torch.as_tensor(np.ascontiguousarray(x)).to(device=device, dtype=dt)
    geo = dict(mid=T(mid), t=T(t), c=T(cdir), n=T(nb), l=T(l[:, 0] / hmean), hmean=hmean)
    dp, di = _csr(dst, E)
    sp, si = _csr(src, E)
    return MeshData(V, F, E, I(src, torch.int32), I(dst, torch.int32), I(dp, torch.int64), I(di, torch.int32),
                    I(sp, torch.int64), I(si, torch.int32), torch.as_tensor(x_e, dtype=torch.float32, device=device),
                    geo)


def node_features(md: MeshData, k: float, dirs, pols, V_rhs: torch.Tensor) -> torch.Tensor:
    """[D, E, NODE_FEATURES] float32 for D plane waves (unit amplitude)."""
    g = md.geo
    dev = g["mid"].device
    kd = torch.as_tensor(np.asarray(dirs, dtype=np.float64), device=dev)  # [D,3]
    pd = torch.as_tensor(np.asarray(pols, dtype=np.float64), device=dev)
    D, E = kd.shape[0], md.E
    ph = torch.exp(-1j * k * (g["mid"] @ kd.T).T)  # [D, E]
    Ein = pd[:, None, :] * ph[..., None]  # [D, E, 3]
    Hin = torch.cross(kd[:, None, :].expand(D, E, 3).to(Ein.dtype), Ein, dim=-1)
    n = g["n"].to(Ein.dtype)[None].expand(D, E, 3)
    Jpo = 2.0 * torch.cross(n, Hin, dim=-1)
    lit = ((g["n"] @ kd.T).T < 0).to(Jpo.dtype)  # [D, E]
    Ipo = (Jpo * g["c"].to(Jpo.dtype)[None]).sum(-1) * lit
    loc = lambda vec, fr: (vec * g[fr].to(vec.dtype)[None]).sum(-1)
    Vr = V_rhs.T.to(torch.complex128)  # [D, E]
    vs = torch.sqrt((Vr.abs() ** 2).mean(dim=1, keepdim=True))
    Vn = Vr / vs
    feats = [
        g["mid"][None].expand(D, E, 3),
        g["t"][None].expand(D, E, 3), g["c"][None].expand(D, E, 3), g["n"][None].expand(D, E, 3),
        g["l"][None, :, None].expand(D, E, 1),
        torch.full((D, E, 1), float(k), dtype=torch.float64, device=dev),
        (k * g["l"] * g["hmean"])[None, :, None].expand(D, E, 1),
        Vn.real[..., None], Vn.imag[..., None],
        Ipo.real[..., None], Ipo.imag[..., None],
    ]
    for fr in ("t", "c", "n"):
        q = loc(Ein, fr)
        feats += [q.real[..., None], q.imag[..., None]]
    for fr in ("t", "c", "n"):
        feats.append(loc(kd[:, None, :].expand(D, E, 3), fr)[..., None])
    x = torch.cat(feats, -1).to(torch.float32)
    assert x.shape[-1] == NODE_FEATURES, x.shape
    return x.contiguous()


def batch_graph(mds: List[MeshData], xs: List[torch.Tensor], device) -> Graph:
    """One batch from per-sample node features xs[i] [E_i, nf] on meshes mds[i]
    (per-mesh CSR tables shifted and concatenated on the device)."""
    no, eo = 0, 0
    goff = [0]
    src, dst, dptr, didx, sptr, sidx, xes = [], [], [], [], [], [], []
    for md in mds:
        ne = md.src.shape[0]
        src.append(md.src + no)
        dst.append(md.dst + no)
        dptr.append(md.dptr[:-1] + eo)
        sptr.append(md.sptr[:-1] + eo)
        didx.append(md.didx + eo)
        sidx.append(md.sidx + eo)
        xes.append(md.x_e)
        no += md.E
        eo += ne
        goff.append(no)
    tail = torch.tensor([eo], dtype=torch.int64, device=device)
    return Graph(x_n=torch.cat(xs, 0).contiguous(), x_e=torch.cat(xes, 0).contiguous(), src=torch.cat(src),
                 dst=torch.cat(dst), dst_ptr=torch.cat(dptr + [tail]), dst_idx=torch.cat(didx),
                 src_ptr=torch.cat(sptr + [tail]), src_idx=torch.cat(sidx),
                 goff=torch.tensor(goff, dtype=torch.int64, device=device),
                 eye_ptr=torch.arange(eo + 1, dtype=torch.int64, device=device))


# ----------------------------------------------------------------- data from the solver
@dataclass
class Sample:
    mesh: int
    k: float
    dirs: np.ndarray
    pols: np.ndarray
    x: torch.Tensor  # [D, E, nf]
    J: torch.Tensor  # [D, E, 2] float32 target (FP64 solve)
    Z32: Optional[torch.Tensor] = None  # complex64 system (GMRES evaluation, residual loss)
    Zh32: Optional[torch.Tensor] = None  # its conjugate transpose
    V32: Optional[torch.Tensor] = None


def random_waves(rng, D):
    d = rng.normal(size=(D, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.normal(size=(D, 3))
    p = r - np.sum(r * d, axis=1, keepdims=True) * d
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    return d, p


def generate(n_meshes: int, freq: int, n_dirs: int, seed0: int, k_range=(1.0, 4.0), device="cuda",
             keep_system=False):
    """Solver-generated training data: star meshes (seeds seed0..), ka ~ U(k_range),
    n_dirs plane waves each; J from the FP64 EFIE solve (assembly + LU on the GPU)."""
    from .dense import LUFactor
    from .geometry import get_edge_unique, star_mesh
    from .mom import MomMesh, assemble_ns, rhs_ns

    mds, samples = [], []
    for i in range(n_meshes):
        s = seed0 + i
        V, F, _ = star_mesh(s, freq)
        rng = np.random.Generator(np.random.PCG64(10007 * s + 1))
        k = float(rng.uniform(*k_range))
        dirs, pols = random_waves(rng, n_dirs)
        vt = torch.tensor(V, dtype=torch.float64)
        ft = torch.tensor(F, dtype=torch.int32)
        edges, fe, _ = get_edge_unique(vt, ft)
        mesh = MomMesh.build(vt, F, fe, edges.shape[0], device=device)
        Z = assemble_ns(mesh, k, _lib.MBA_F64)
        Vr = rhs_ns(mesh, k, _lib.MBA_F64, dirs, pols)
        J = LUFactor(Z.clone(), overwrite=True).solve(Vr)  # [E, D]
        md = mesh_data(V, F, fe.numpy(), edges.shape[0], device)
        x = node_features(md, k, dirs, pols, Vr)
        Jt = torch.stack([J.real.T, J.imag.T], -1).to(torch.float32).contiguous()
        smp = Sample(i, k, dirs, pols, x, Jt)
        if keep_system:
            smp.Z32 = Z.to(torch.complex64)
            smp.Zh32 = smp.Z32.conj().T.contiguous()
            smp.V32 = Vr.to(torch.complex64).contiguous()
        mds.append(md)
        samples.append(smp)
    return mds, samples


def save_samples(path: str, samples: List["Sample"]) -> None:
    """Generated triples (mesh index, ka, waves, features, FP64-solve currents)
    as one .npz shard; the systems (Z32 ...) are not stored."""
    arrays = {}
    for i, smp in enumerate(samples):
        arrays[f"s{i}_meta"] = np.array([smp.mesh, smp.k], dtype=np.float64)
        arrays[f"s{i}_dirs"] = np.asarray(smp.dirs)
        arrays[f"s{i}_pols"] = np.asarray(smp.pols)
        arrays[f"s{i}_x"] = smp.x.detach().cpu().numpy()
        arrays[f"s{i}_J"] = smp.J.detach().cpu().numpy()
    arrays["count"] = np.array([len(samples)], dtype=np.int64)
    tmp = path + ".tmp.npz"
    np.savez(tmp, **arrays)
    os.replace(tmp, path)


def load_samples(path: str, device="cpu") -> List["Sample"]:
    with np.load(path, allow_pickle=False) as z:
        n = int(z["count"][0])
        out = []
        for i in range(n):
            mesh, k = z[f"s{i}_meta"]
            out.append(Sample(int(mesh), float(k), z[f"s{i}_dirs"], z[f"s{i}_pols"],
                              torch.from_numpy(z[f"s{i}_x"]).to(device), torch.from_numpy(z[f"s{i}_J"]).to(device)))
    return out


def _residual_grad(y, sel, samples, goff, res_weight):
    """Adds res_weight * d/dy of mean_g |b_g - Z_g J_g|^2 / |b_g|^2 to nothing and
    returns (grad [Nn, 2], per-sample residual^2).  J_g = y_re + i y_im of sample g."""
    from .gmres import DenseOperator

    G = len(sel)
    grad = torch.empty_like(y)
    res2 = torch.empty(G, dtype=torch.float32, device=y.device)
    for q, (i, d) in enumerate(sel):
        s = samples[i]
        a, b = int(goff[q]), int(goff[q + 1])
        x = torch.complex(y[a:b, 0], y[a:b, 1]).contiguous()
        bvec = s.V32[:, d].contiguous()
        w = torch.empty_like(x)
        DenseOperator(s.Z32).apply(x, w, bvec)  # w = b - Z x = -r
        gz = torch.empty_like(x)
        DenseOperator(s.Zh32).apply(w, gz)  # Z^H w = -Z^H r
        nb = float((bvec.abs() ** 2).sum())
        sc = -2.0 * res_weight / (nb * G)
        grad[a:b, 0] = sc * gz.real
        grad[a:b, 1] = sc * gz.imag
        res2[q] = (w.abs() ** 2).sum() / nb
    return grad, res2


def train(net: MeshGraphNet, mds, samples, epochs=5, batch=16, lr=1e-3, seed=0, log=None, lr_final=1e-4,
          res_weight=0.0, checkpoint: Optional[str] = None):
    """Mini-batch training over all (mesh, direction) pairs.  Loss per sample:
    |J_pred - J|^2 / |J|^2 + res_weight |b - Z J_pred|^2 / |b|^2 (the second
    term needs samples generated with keep_system=True).  Returns the history
    of the mean loss per epoch.

    checkpoint: path of a checkpoint written after every epoch (network,
    Adam state, epochs done, loss history); when it exists, training resumes
    after its last epoch.  The batch order of epoch e depends only on (seed,
    e), so a resumed run repeats the uninterrupted one exactly."""
    pairs = [(i, d) for i, s in enumerate(samples) for d in range(s.x.shape[0])]
    hist = []
    n_steps = epochs * ((len(pairs) + batch - 1) // batch)
    step = 0
    ep0 = 0
    if checkpoint and os.path.exists(checkpoint):
        extra = net.load(checkpoint)
        ep0 = int(extra["epochs_done"])
        hist = [float(h) for h in extra["history"].tolist()][:ep0]
        step = ep0 * ((len(pairs) + batch - 1) // batch)
        if log:
            log(f"resumed from {checkpoint} after epoch {ep0}")
    for ep in range(ep0, epochs):
        order = np.random.Generator(np.random.PCG64(seed * 1000003 + ep)).permutation(len(pairs))
        tot, nb = 0.0, 0
        for b0 in range(0, len(order), batch):
            sel = [pairs[j] for j in order[b0:b0 + batch]]
            g = batch_graph([mds[i] for i, _ in sel], [samples[i].x[d] for i, d in sel], net.device)
            tgt = torch.cat([samples[i].J[d] for i, d in sel], 0)
            cur = lr * (lr_final / lr) ** (step / max(1, n_steps - 1))
            if res_weight > 0:
                y, c = net.forward(g)
                lo, dy = net.loss(y, tgt, g)
                goff = g.goff.cpu().numpy()
                gr, r2 = _residual_grad(y, sel, samples, goff, res_weight)
                dy += gr
                lo = lo + res_weight * r2
                net.backward(g, c, dy)
                net.adam(cur)
            else:
                lo = net.train_step(g, tgt, cur)
            step += 1
            tot += float(lo.mean())
            nb += 1
        hist.append(tot / nb)
        if log:
            log(f"epoch {ep + 1}/{epochs}: mean relative L2^2 {hist[-1]:.4f}")
        if checkpoint:
            net.save(checkpoint, {"epochs_done": ep + 1, "history": torch.tensor(hist, dtype=torch.float64)})
    return hist


def predict_currents(net: MeshGraphNet, md: MeshData, x: torch.Tensor) -> torch.Tensor:
    """Complex64 J [E, D] for node features x [D, E, nf]."""
    D = x.shape[0]
    g = batch_graph([md] * D, [x[d] for d in range(D)], net.device)
    y = net.predict(g).view(D, md.E, 2)
    return torch.complex(y[..., 0], y[..., 1]).T.contiguous()


def scaled_guess(Z32: torch.Tensor, b: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """alpha x with alpha = argmin |b - alpha Z x| (one matvec): the prediction
    never starts GMRES from a residual above |b|."""
    from .gmres import DenseOperator

    x = x.to(torch.complex64).contiguous()
    y = torch.empty_like(x)
    DenseOperator(Z32).apply(x, y)
    den = (y.abs() ** 2).sum()
    if float(den) == 0.0:
        return torch.zeros_like(x)
    alpha = torch.vdot(y.to(torch.complex128), b.to(torch.complex128)) / den
    return (alpha * x.to(torch.complex128)).to(torch.complex128)
