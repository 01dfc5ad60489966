"""CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module, and only as the checker.  The product path
(mesh-ferrofluid_amd/) never imports it.

The numeric kernels live in ``mom_oracle.c`` (built into ``liborc.so`` by the
Makefile next to this file); this module wraps them with numpy and restates
the host-side pieces of the reference in numpy:

* face areas / normals / vertex areas   -- src/utils/mesh_utils.py:158-310
* Laplace relation point tables        -- src/mesh_BEM/utils/integrations_3d.py:18-249,
  fed by the rank-3 rule of src/mesh_BEM/single_layer/abstract_single_layer.py:49-63
  and the tensorisation of integrations_3d.py:354-364
* panel relations (brute force)        -- src/utils/mesh_utils.py:700-725
* pack_matrix_tight                    -- src/utils/BEM_utils.py:314-357
* Dirichlet RHS / solve                -- src/mesh_BEM/rhs_layer/rhs_layer_3d.py:74-203,
  src/mesh_BEM/mesh_BEM_solver_3d.py:197-373
* solve_MOM pipeline                   -- src/mesh_BEM/mesh_BEM_solver_3d.py:96-195
* dense solves: LAPACK zgesv/dgesv through scipy (what torch.linalg.solve
  calls on the CPU; mesh_BEM_solver_3d.py:174, 346).

Parity status per function is listed in oracle/README.md.
"""

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        D = ctypes.POINTER(ctypes.c_double)
        I32 = ctypes.POINTER(ctypes.c_int32)
        I64 = ctypes.POINTER(ctypes.c_int64)
        U8 = ctypes.POINTER(ctypes.c_uint8)
        i64, i32, dbl = ctypes.c_int64, ctypes.c_int, ctypes.c_double
        L.orc_ref_zmat_faceedge.argtypes = [D, I32, D, D, D, D, i32, D, D, i32, dbl, i64, D]
        L.orc_ref_rhs_faceedge.argtypes = [D, I32, D, D, D, D, i32, dbl, i64, D]
        L.orc_ref_zmat_faceedge_terms.argtypes = [D, I32, D, D, D, D, i32, D, D, i32, dbl, i64, i32, D]
        L.orc_ref_rhs_faceedge_terms.argtypes = [D, I32, D, D, D, D, i32, dbl, i64, i32, D]
        L.orc_interp_faceedge.argtypes = [D, I32, I32, D, D, i64, i64, i64, D]
        L.orc_tri_potentials.argtypes = [D, D, D, D, D, D, D]
        L.orc_ns_zmat_faceedge.argtypes = [D, I32, D, D, D, i32, dbl, i64, I64, I32, D]
        L.orc_ns_zmat_faceedge_window.argtypes = [D, I32, D, D, D, i32, dbl, i64, I64, I32, i64, i64, D]
        L.orc_ns_zmat_faceedge_window.restype = None
        L.orc_ns_rhs_faceedge.argtypes = [D, I32, D, D, D, i32, dbl, i64, D, D, i64, D]
        L.orc_gather_rwg.argtypes = [D, i64, I32, i64, D]
        L.orc_gather_rhs.argtypes = [D, I32, i64, i64, D]
        L.orc_laplace_layer.argtypes = [D, I32, D, D, i64, U8, ctypes.POINTER(D), I64, i32, i32, D]
        L.orc_laplace_p1_ns.argtypes = [D, I32, i64, D, D, i32, I64, I32, D, D]
        L.orc_laplace_layer_ex.argtypes = [D, I32, D, D, i64, U8, ctypes.POINTER(D), I64, i32, i32, i32, D]
        L.orc_tri_S0_grad.argtypes = [D, D, D, D, D, D]
        L.orc_laplace_p0.argtypes = [D, I32, i64, D, D, i32, I64, I32, D, D]
        for f in (L.orc_ref_zmat_faceedge, L.orc_ref_rhs_faceedge, L.orc_ref_zmat_faceedge_terms,
                  L.orc_ref_rhs_faceedge_terms, L.orc_interp_faceedge,
                  L.orc_tri_potentials, L.orc_ns_zmat_faceedge, L.orc_ns_rhs_faceedge,
                  L.orc_gather_rwg, L.orc_gather_rhs, L.orc_laplace_layer, L.orc_laplace_layer_ex,
                  L.orc_laplace_p1_ns,
                  L.orc_tri_S0_grad,
                  L.orc_laplace_p0):
            f.restype = None
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _i64(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _fi32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ------------------------------------------------------------------ geometry
def face_areas_normals(verts, faces):
    """Areas [F] and unit normals [F,3]: normalise((v1-v0) x (v2-v1))."""
    v = _f64(verts)
    f = np.asarray(faces, dtype=np.int64)
    v0, v1, v2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    c = np.cross(v1 - v0, v2 - v1)
    nrm = np.linalg.norm(c, axis=1)
    return 0.5 * nrm, c / nrm[:, None]


def vert_areas(verts, faces):
    a, _ = face_areas_normals(verts, faces)
    va = np.zeros(len(verts))
    np.add.at(va, np.asarray(faces, dtype=np.int64).reshape(-1), np.repeat(a / 3.0, 3))
    return va


def edges_bruteforce(faces, n_verts):
    """Sorted-key edge table and face-edge map by a dictionary walk."""
    keys = {}
    f = np.asarray(faces, dtype=np.int64)
    fe = np.zeros(3 * len(f), dtype=np.int64)
    allk = sorted({max(f[i, (e + 1) % 3], f[i, (e + 2) % 3]) * n_verts
                   + min(f[i, (e + 1) % 3], f[i, (e + 2) % 3])
                   for i in range(len(f)) for e in range(3)})
    for idx, k in enumerate(allk):
        keys[k] = idx
    for i in range(len(f)):
        for e in range(3):
            a, b = f[i, (e + 1) % 3], f[i, (e + 2) % 3]
            fe[3 * i + e] = keys[max(a, b) * n_verts + min(a, b)]
    edges = np.array([(k // n_verts, k % n_verts) for k in allk], dtype=np.int64)
    return edges, fe


def rwg_fe(fe_idx, E):
    order = np.argsort(np.asarray(fe_idx, dtype=np.int64), kind="stable")
    return order.reshape(E, 2).astype(np.int32)


def gl_ref(N):
    """Reference GL table on [0,1] in FP64 from float32-rounded nodes."""
    x, w = np.polynomial.legendre.leggauss(N)
    x = x.astype(np.float32).astype(np.float64)
    w = w.astype(np.float32).astype(np.float64)
    return (x + 1.0) / 2.0, w / 2.0


def radon7():
    s15 = math.sqrt(15.0)
    pts = [(1 / 3, 1 / 3, 1 / 3)]
    ws = [9 / 40]
    for a, wa in (((6 - s15) / 21, (155 - s15) / 1200), ((6 + s15) / 21, (155 + s15) / 1200)):
        b = 1 - 2 * a
        pts += [(a, a, b), (a, b, a), (b, a, a)]
        ws += [wa] * 3
    return np.array(pts), 0.5 * np.array(ws)


def near_pairs_bruteforce(verts, faces, eta=2.0):
    """CSR over source faces b of test faces a with |o_a - o_b| < eta (rho_a + rho_b)."""
    v = _f64(verts)
    f = np.asarray(faces, dtype=np.int64)
    P = v[f]  # [F,3,3]
    o = P.mean(axis=1)
    rho = np.linalg.norm(P - o[:, None, :], axis=2).max(axis=1)
    D = np.linalg.norm(o[:, None, :] - o[None, :, :], axis=2)
    near = D < eta * (rho[:, None] + rho[None, :])
    ptr = np.zeros(len(f) + 1, dtype=np.int64)
    idx = []
    for b in range(len(f)):
        a = np.nonzero(near[:, b])[0]
        idx.append(a)
        ptr[b + 1] = ptr[b] + a.size
    return ptr, np.concatenate(idx).astype(np.int32)


# ------------------------------------------------------------------ MoM, ref mode
# Terms of the reference's CFIE (mesh_MOM_3d.cu:426-492): identity 1/2 f.f on
# the self pair, MFIE n x (f x grad G), EFIE vector potential jk f.f G, EFIE
# scalar potential -(j/k) div div G; RHS E_inc and n x H_inc (.cu:316-326).
ZT_ID, ZT_K, ZT_VEC, ZT_SCA = 1, 2, 4, 8
ZT_ALL, ZT_EFIE, ZT_MFIE = 15, ZT_VEC | ZT_SCA, ZT_ID | ZT_K
RT_E, RT_H = 1, 2
RT_ALL = 3


def ref_zmat_faceedge(verts, faces, gpx, gwx, gpy, gwy, k, terms=ZT_ALL):
    v, f = _f64(verts), _fi32(faces)
    A, n = face_areas_normals(v, f)
    gpx, gwx, gpy, gwy = map(_f64, (gpx, gwx, gpy, gwy))
    F = len(f)
    out = np.zeros((3 * F, 3 * F), dtype=np.complex128)
    lib().orc_ref_zmat_faceedge_terms(_d(v), _i32(f), _d(A), _d(_f64(n)), _d(gpx), _d(gwx), gpx.size,
                                      _d(gpy), _d(gwy), gpy.size, float(k), F, int(terms),
                                      _d(out.view(np.float64)))
    return out


def ref_rhs_faceedge(verts, faces, gp, gw, k, terms=RT_ALL):
    v, f = _f64(verts), _fi32(faces)
    A, n = face_areas_normals(v, f)
    gp, gw = _f64(gp), _f64(gw)
    out = np.zeros((len(f), 3), dtype=np.complex128)
    lib().orc_ref_rhs_faceedge_terms(_d(v), _i32(f), _d(A), _d(_f64(n)), _d(gp), _d(gw), gp.size,
                                     float(k), len(f), int(terms), _d(out.view(np.float64)))
    return out


def interp_faceedge(verts, faces, fe_idx, I):
    """[B, 3F, 3] complex face-vertex currents from RWG coefficients I [B, E]."""
    v, f, fe = _f64(verts), _fi32(faces), _fi32(fe_idx)
    A, _ = face_areas_normals(v, f)
    I = np.ascontiguousarray(np.atleast_2d(I), dtype=np.complex128)
    B, E = I.shape
    out = np.zeros((B, 3 * len(f), 3), dtype=np.complex128)
    lib().orc_interp_faceedge(_d(v), _i32(f), _i32(fe), _d(A), _d(I.view(np.float64)), B, E,
                              len(f), _d(out.view(np.float64)))
    return out


def vertex_sum(face_vert_J, faces, n_verts):
    """vert_Js[b, v] = sum over (f, t) with faces[f][t] == v (mesh_BEM_solver_3d.py:186-192)."""
    B = face_vert_J.shape[0]
    out = np.zeros((B, n_verts, 3), dtype=face_vert_J.dtype)
    idx = np.asarray(faces, dtype=np.int64).reshape(-1)
    for b in range(B):
        np.add.at(out[b], idx, face_vert_J[b])
    return out


def gather_rwg(ZF, rfe):
    F3 = ZF.shape[0]
    E = rfe.shape[0]
    ZF = np.ascontiguousarray(ZF, dtype=np.complex128)
    Z = np.zeros((E, E), dtype=np.complex128)
    lib().orc_gather_rwg(_d(ZF.view(np.float64)), F3 // 3, _i32(_fi32(rfe)), E, _d(Z.view(np.float64)))
    return Z


def gather_rhs(VF, rfe):
    VF = np.ascontiguousarray(VF.reshape(VF.shape[0], -1), dtype=np.complex128)
    E = rfe.shape[0]
    nd = VF.shape[1]
    V = np.zeros((E, nd), dtype=np.complex128)
    lib().orc_gather_rhs(_d(VF.view(np.float64)), _i32(_fi32(rfe)), E, nd, _d(V.view(np.float64)))
    return V


def ref_solve_MOM(verts, faces, gaussQR, k, zterms=ZT_ALL, rterms=RT_ALL, gaussQR_y=None):
    """FP64 restatement of MeshBEMSolver3d.solve_MOM (mesh_BEM_solver_3d.py:96-195).

    zterms / rterms select parts of the CFIE (all by default, as the
    reference); gaussQR_y overrides the source rule order (the reference uses
    gaussQR - 1).  Returns (vert_Js [1, V, 3], I [E], Z [E, E], V [E], edges, fe_idx)."""
    v, f = _f64(verts), np.asarray(faces, dtype=np.int64)
    edges, fe = edges_bruteforce(f, len(v))
    E = len(edges)
    gpx, gwx = gl_ref(gaussQR)
    gpy, gwy = gl_ref(gaussQR - 1 if gaussQR_y is None else gaussQR_y)
    ZF = ref_zmat_faceedge(v, f, gpx, gwx, gpy, gwy, k, zterms)
    rfe = rwg_fe(fe, E)
    Z = gather_rwg(ZF, rfe)
    VF = ref_rhs_faceedge(v, f, gpx, gwx, k, rterms)
    V = gather_rhs(VF.reshape(-1, 1), rfe)[:, 0]
    import scipy.linalg

    Icoef = scipy.linalg.solve(Z, V)
    fvj = interp_faceedge(v, f, fe, Icoef[None, :])
    vj = vertex_sum(fvj, f, len(v))
    return vj, Icoef, Z, V, edges, fe


# ------------------------------------------------------------------ MoM, ns mode
def tri_potentials(P0, P1, P2, r):
    S0 = np.zeros(1)
    S1 = np.zeros(3)
    rho = np.zeros(3)
    lib().orc_tri_potentials(_d(_f64(P0)), _d(_f64(P1)), _d(_f64(P2)), _d(_f64(r)), _d(S0), _d(S1), _d(rho))
    return S0[0], S1, rho


def ns_zmat_faceedge(verts, faces, k, near_ptr, near_idx, rule=None):
    v, f = _f64(verts), _fi32(faces)
    A, _ = face_areas_normals(v, f)
    lam, w = radon7() if rule is None else rule
    lam, w = _f64(lam), _f64(w)
    F = len(f)
    out = np.zeros((3 * F, 3 * F), dtype=np.complex128)
    lib().orc_ns_zmat_faceedge(_d(v), _i32(f), _d(A), _d(lam), _d(w), w.size, float(k), F,
                               _i64(np.ascontiguousarray(near_ptr, dtype=np.int64)),
                               _i32(_fi32(near_idx)), _d(out.view(np.float64)))
    return out


def ns_zmat_faceedge_window(verts, faces, k, near_ptr, near_idx, b_lo, b_hi, rule=None):
    """Columns of source faces [b_lo, b_hi): [3F, 3 (b_hi - b_lo)] complex."""
    v, f = _f64(verts), _fi32(faces)
    A, _ = face_areas_normals(v, f)
    lam, w = radon7() if rule is None else rule
    F = len(f)
    out = np.zeros((3 * F, 3 * (b_hi - b_lo)), dtype=np.complex128)
    lib().orc_ns_zmat_faceedge_window(_d(v), _i32(f), _d(A), _d(_f64(lam)), _d(_f64(w)), len(w),
                                      float(k), F, _i64(np.ascontiguousarray(near_ptr, dtype=np.int64)),
                                      _i32(_fi32(near_idx)), int(b_lo), int(b_hi),
                                      _d(out.view(np.float64)))
    return out


def ns_rhs_faceedge(verts, faces, k, dirs, pols, rule=None):
    v, f = _f64(verts), _fi32(faces)
    A, _ = face_areas_normals(v, f)
    lam, w = radon7() if rule is None else rule
    dirs = _f64(np.atleast_2d(dirs))
    pols = _f64(np.atleast_2d(pols))
    out = np.zeros((3 * len(f), dirs.shape[0]), dtype=np.complex128)
    lib().orc_ns_rhs_faceedge(_d(v), _i32(f), _d(A), _d(_f64(lam)), _d(_f64(w)), w.size, float(k),
                              len(f), _d(dirs), _d(pols), dirs.shape[0], _d(out.view(np.float64)))
    return out


def ns_system(verts, faces, k, dirs, pols, eta=2.0):
    """FP64 EFIE system (Z [E,E], V [E, ndir]) in the sorted-key edge numbering."""
    v, f = _f64(verts), np.asarray(faces, dtype=np.int64)
    edges, fe = edges_bruteforce(f, len(v))
    E = len(edges)
    ptr, idx = near_pairs_bruteforce(v, f, eta)
    ZF = ns_zmat_faceedge(v, f, k, ptr, idx)
    rfe = rwg_fe(fe, E)
    Z = gather_rwg(ZF, rfe)
    V = gather_rhs(ns_rhs_faceedge(v, f, k, dirs, pols), rfe)
    return Z, V, edges, fe


# ------------------------------------------------------------------ Laplace, ref mode
def rank3_rule(gp, gw):
    """(r1, r2, w) of abstract_single_layer.py:49-63: index a*Q + b -> (gp[b], gp[a] gp[b]), gw[a] gw[b]."""
    Q = gp.size
    r1 = np.tile(gp[None, :], (Q, 1)).reshape(-1)
    r2 = (gp[:, None] * gp[None, :]).reshape(-1)
    w = (gw[None, :] * gw[:, None]).reshape(-1)
    return r1, r2, w


def laplace_tables(gp, gw):
    """Point tables per relation type (0..3): rows (r1x, r2x, r1y, r2y, weight*jacobian),
    fed with the rank-3 rule of abstract_single_layer.py:49-63 (the reference's input)."""
    r1, r2, w = rank3_rule(_f64(gp), _f64(gw))
    return laplace_tables_from(r1, r2, w)


def laplace_tables_from(r1, r2, w):
    """Tables from a 2-D rule (r1, r2, w): integrations_3d.py:354-364 tensorises
    it into (xsi, eta1) x (eta2, eta3); get_rands_and_jacobian_3d (:18-249)
    maps the 4-D points per relation type."""
    r1, r2, w = _f64(r1), _f64(r2), _f64(w)
    Q2 = r1.size
    xsi = np.tile(r1[None, :], (Q2, 1)).reshape(-1)
    eta1 = np.tile(r2[None, :], (Q2, 1)).reshape(-1)
    eta2 = np.tile(r1[:, None], (1, Q2)).reshape(-1)
    eta3 = np.tile(r2[:, None], (1, Q2)).reshape(-1)
    wyx = (np.tile(w[:, None], (1, Q2)) * np.tile(w[None, :], (Q2, 1))).reshape(-1)

    maps = {}
    maps[0] = ([(xsi, eta1 * xsi, eta2, eta2 * eta3)], [xsi * eta2])
    j1 = xsi ** 3 * eta2
    maps[1] = ([(xsi, xsi * eta1, xsi * eta2, xsi * eta2 * eta3),
                (xsi * eta2, xsi * eta2 * eta3, xsi, xsi * eta1)], [j1, j1])
    ws = [
        (xsi, -xsi * eta1 * eta2, xsi * eta1 * (1 - eta2), xsi * eta1 * eta3),
        (xsi, -xsi * eta1 * eta2 * eta3, xsi * eta1 * eta2 * (1 - eta3), xsi * eta1),
        (xsi * (1 - eta1 * eta2), xsi * eta1 * eta2, xsi * eta1 * eta2 * eta3, xsi * eta1 * (1 - eta2)),
        (xsi * (1 - eta1 * eta2 * eta3), xsi * eta1 * eta2 * eta3, xsi * eta1, xsi * eta1 * eta2 * (1 - eta3)),
        (xsi * (1 - eta1 * eta2 * eta3), xsi * eta1 * eta2 * eta3, xsi * eta1 * eta2, xsi * eta1 * (1 - eta2 * eta3)),
    ]
    je = [xsi ** 3 * eta1 ** 2] + [xsi ** 3 * eta1 ** 2 * eta2] * 4
    maps[2] = ([(a, d, a + b, c) for (a, b, c, d) in ws], je)
    M = np.array([
        [[1, 0, 0, 0], [1, -1, 1, 0], [0, 0, 0, 1], [0, 0, 1, 0]],
        [[1, 0, 0, 0], [0, 1, -1, 1], [0, 0, 1, 0], [0, 0, 0, 1]],
        [[1, 0, 0, -1], [0, 1, 0, -1], [0, 0, 0, -1], [0, 0, 1, -1]],
    ], dtype=np.float64)
    wv = np.stack([xsi, xsi * eta1, xsi * eta1 * eta2, xsi * eta1 * eta2 * eta3], axis=0)
    sf = []
    for d in range(3):
        xz = M[d] @ wv
        sf.append((xz[0], xz[1], xz[0] - xz[2], xz[1] - xz[3]))
    js = xsi ** 3 * eta1 ** 2 * eta2
    maps[3] = (sf, [js] * 3)
    tabs = []
    for r in range(4):
        pts, jacs = maps[r]
        rows = []
        for (a, b, c, d), jj in zip(pts, jacs):
            rows.append(np.stack([a, b, c, d, wyx * jj], axis=1))
        tabs.append(np.ascontiguousarray(np.concatenate(rows, axis=0)))
    return tabs


def proper_tables(N):
    """Sauter-Schwab tables from a genuine 4-D tensor Gauss-Legendre rule of
    order N (FP64 nodes): (xsi, eta1) and (eta2, eta3) independent, no
    collapse.  The reference instead feeds its collapsed rank-3 rule
    (``laplace_tables``)."""
    gp, gw = np.polynomial.legendre.leggauss(N)
    gp, gw = (gp + 1.0) / 2.0, gw / 2.0
    return laplace_tables_from(np.repeat(gp, N), np.tile(gp, N), np.repeat(gw, N) * np.tile(gw, N))


# Fixes of the reference's Laplace rule (orc_laplace_layer_ex): shared vertex /
# edge first for the Sauter-Schwab maps; the left-out same-face sub-domains.
LAP_REORDER, LAP_SYMMETRY = 1, 2
LAP_EXACT = LAP_REORDER | LAP_SYMMETRY


def panel_relation_bruteforce(faces):
    f = np.asarray(faces, dtype=np.int64)
    return (f[:, None, :, None] == f[None, :, None, :]).sum(axis=(2, 3)).astype(np.uint8)


def laplace_layer(verts, faces, gp, gw, layer, order_type, rel=None, tabs=None, fix=0):
    """Unpacked [F, F, nphi, nphi] single (layer=1) or double (layer=2) layer.

    fix = 0 and tabs from (gp, gw) reproduce the reference; fix = LAP_EXACT with
    tabs = proper_tables(N) gives the exact Galerkin layers."""
    v, f = _f64(verts), _fi32(faces)
    A, n = face_areas_normals(v, f)
    F = len(f)
    if rel is None:
        rel = panel_relation_bruteforce(f)
    rel = np.ascontiguousarray(rel, dtype=np.uint8)
    if tabs is None:
        tabs = laplace_tables(gp, gw)
    nphi = 3 if order_type == 1 else 1
    P = ctypes.POINTER(ctypes.c_double)
    tp = (P * 4)(*[_d(t) for t in tabs])
    nt = np.array([t.shape[0] for t in tabs], dtype=np.int64)
    out = np.zeros((F, F, nphi, nphi))
    lib().orc_laplace_layer_ex(_d(v), _i32(f), _d(A), _d(_f64(n)), F,
                               rel.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), tp, _i64(nt),
                               int(layer), nphi, int(fix), _d(out))
    return out


def pack_matrix_tight(order_type, U, faces, n_verts):
    """[Fy, Fx, nphi, nphi] -> [V, V] with A[vert(fy,iy), vert(fx,ix)] += U (BEM_utils.py:314-357)."""
    f = np.asarray(faces, dtype=np.int64)
    F = len(f)
    if order_type == 0:
        U = np.repeat(np.repeat(U.reshape(F, F, 1, 1), 3, axis=2), 3, axis=3) / 9.0
    src = U.transpose(0, 2, 1, 3).reshape(-1)
    rows = np.repeat(f.reshape(-1), 3 * F)
    cols = np.tile(f.reshape(-1), 3 * F)
    out = np.zeros(n_verts * n_verts, dtype=U.dtype)
    np.add.at(out, rows * n_verts + cols, src)
    return out.reshape(n_verts, n_verts)


def consistent_mass(verts, faces):
    """P1 mass matrix M_ij = int phi_i phi_j (A/12 (1 + delta_ij) per face pair of corners)."""
    v, f = _f64(verts), np.asarray(faces, dtype=np.int64)
    A, _ = face_areas_normals(v, f)
    M = np.zeros((len(v), len(v)))
    loc = (np.ones((3, 3)) + np.eye(3)) / 12.0
    for a in range(3):
        for b in range(3):
            np.add.at(M, (f[:, a], f[:, b]), A * loc[a, b])
    return M


def ref_laplace_dirichlet(verts, faces, g, gaussQR, order_type=1, tabs=None, fix=0, mass="lumped"):
    """FP64 restatement of MeshBEMSolver3d.solve for an all-Dirichlet batch of one.

    g: [V] Dirichlet data.  Returns Neumann data t [V] solving V t = 1/2 M g + K^T g.
    tabs / fix replace the reference's layer rule (``laplace_layer``); mass
    "lumped" is the reference's 1/2 g vert_area (rhs_layer_3d.py:127),
    "consistent" the Galerkin P1 mass matrix."""
    v, f = _f64(verts), np.asarray(faces, dtype=np.int64)
    nV = len(v)
    gp, gw = gl_ref(gaussQR)
    rel = panel_relation_bruteforce(f)
    if tabs is None:
        tabs = laplace_tables(gp, gw)
    Vu = laplace_layer(v, f, gp, gw, 1, order_type, rel, tabs, fix)
    Ku = laplace_layer(v, f, gp, gw, 2, order_type, rel, tabs, fix)
    Vt = pack_matrix_tight(order_type, Vu, f, nV)
    if order_type == 0:
        Ku = np.repeat(np.repeat(Ku.reshape(len(f), len(f), 1, 1), 3, axis=2), 3, axis=3) / 9.0
    gf = np.asarray(g)[f]  # [F, 3] = g at (fy, iy)
    src = np.einsum("yxij,yi->xj", Ku, gf)  # sum over fy, iy
    gk2 = np.zeros(nV)
    np.add.at(gk2, f.reshape(-1), src.reshape(-1))
    if mass == "lumped":
        rhs = 0.5 * np.asarray(g) * vert_areas(v, f) + gk2
    else:
        rhs = 0.5 * consistent_mass(v, f) @ np.asarray(g) + gk2
    import scipy.linalg

    return scipy.linalg.solve(Vt, rhs), Vt, rhs


# ------------------------------------------------------------------ Laplace, ns mode (P0)
def tri_S0_grad(P0, P1, P2, r):
    S0 = np.zeros(1)
    G = np.zeros(3)
    lib().orc_tri_S0_grad(_d(_f64(P0)), _d(_f64(P1)), _d(_f64(P2)), _d(_f64(r)), _d(S0), _d(G))
    return S0[0], G


def laplace_p0(verts, faces, near_ptr, near_idx, rule=None):
    """(V, Kt) [F, F] P0 Galerkin single layer and adjoint double layer."""
    v, f = _f64(verts), _fi32(faces)
    lam, w = radon7() if rule is None else rule
    F = len(f)
    Vm = np.zeros((F, F))
    Km = np.zeros((F, F))
    lib().orc_laplace_p0(_d(v), _i32(f), F, _d(_f64(lam)), _d(_f64(w)), len(w),
                         _i64(np.ascontiguousarray(near_ptr, dtype=np.int64)), _i32(_fi32(near_idx)),
                         _d(Vm), _d(Km))
    return Vm, Km


def laplace_p1_ns(verts, faces, near_ptr, near_idx, rule=None):
    """NS P1 layers (Vu, Ku) [F, F, 3, 3] in the layout of laplace_layer:
    7-point rule on far pairs, outer 7-point rule x analytic inner integral on
    near pairs."""
    v, f = _f64(verts), _fi32(faces)
    lam, w = radon7() if rule is None else rule
    F = len(f)
    Vu = np.zeros((F, F, 3, 3))
    Ku = np.zeros((F, F, 3, 3))
    lib().orc_laplace_p1_ns(_d(v), _i32(f), F, _d(_f64(lam)), _d(_f64(w)), len(w),
                            _i64(np.ascontiguousarray(near_ptr, dtype=np.int64)), _i32(_fi32(near_idx)),
                            _d(Vu), _d(Ku))
    return Vu, Ku


def dirichlet_solve_from_layers(verts, faces, g, Vu, Ku, mass="consistent"):
    """V t = 1/2 M g + K g from unpacked P1 layers (the Dirichlet branch of
    mesh_BEM_solver_3d.py:197-373 with the P1 mass matrix, or the reference's
    lumped 1/2 g vert_area).  Returns (t, V packed, rhs)."""
    v, f = _f64(verts), np.asarray(faces, dtype=np.int64)
    nV = len(v)
    Vt = pack_matrix_tight(1, Vu, f, nV)
    src = np.einsum("yxij,yi->xj", Ku, np.asarray(g)[f])
    gk = np.zeros(nV)
    np.add.at(gk, f.reshape(-1), src.reshape(-1))
    if mass == "lumped":
        rhs = 0.5 * np.asarray(g) * vert_areas(v, f) + gk
    else:
        rhs = 0.5 * consistent_mass(v, f) @ np.asarray(g) + gk
    import scipy.linalg

    return scipy.linalg.solve(Vt, rhs), Vt, rhs


def demag_factors(a, b, c):
    """Demagnetising factors (N_a, N_b, N_c) of an ellipsoid with semi-axes a, b, c
    (Osborn 1945): N_i = (abc/2) int_0^inf ds / ((s + i^2) sqrt((s+a^2)(s+b^2)(s+c^2)))."""
    from scipy.integrate import quad

    def N(i2):
        g = lambda s: 1.0 / ((s + i2) * math.sqrt((s + a * a) * (s + b * b) * (s + c * c)))
        return 0.5 * a * b * c * quad(g, 0, np.inf, limit=200)[0]

    return N(a * a), N(b * b), N(c * c)


# ------------------------------------------------------------------ field radiation
def green_G(k, p1, p2):
    """G = e^{-jkR} / (4 pi R), 0 where R < 1e-6 (src/utils/functions.py:806-841,
    called with pos_k False: the wavenumber enters as -k)."""
    tk = -k
    d = np.linalg.norm(np.asarray(p1, dtype=np.float64) - np.asarray(p2, dtype=np.float64), axis=-1)
    inv = np.where(d < 1e-6, 0.0, 1.0 / np.where(d < 1e-6, 1.0, d))
    kx = tk * d
    return (np.cos(kx) + 1j * np.sin(kx)) / (4 * math.pi) * inv


def green_gradG_y(k, p1, p2):
    """d G / d p2 as written in src/utils/functions.py:844-868:
    G * inv_dist * (1 - j kx) * ray_d with kx = -k R and ray_d = (p1 - p2) / R."""
    p1 = np.asarray(p1, dtype=np.float64)
    p2 = np.asarray(p2, dtype=np.float64)
    tk = -k
    diff = p1 - p2
    d = np.linalg.norm(diff, axis=-1, keepdims=True)
    ray = diff / np.maximum(d, 1e-12)
    inv = np.where(d < 1e-6, 0.0, 1.0 / np.where(d < 1e-6, 1.0, d))
    kx = tk * d
    Gv = green_G(k, p1, p2)[..., None]
    return Gv * inv * (1 - 1j * kx) * ray


def radiate_H(k, points, verts, vert_Js, vert_areas, block=256):
    """Scattered H of vertex currents at points, the demo's expression
    (demo/demo_Helmholtz_MOM_3d.py:246-257):
        -(vert_areas * cross(Js, -gradG_y(k, p1=pm, p2=verts))).sum(-2)
    evaluated over blocks of points.  points [Nm,3], verts [V,3], vert_Js [V,3]
    complex, vert_areas [V].  Returns [Nm,3] complex128."""
    P = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    V = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    J = np.asarray(vert_Js, dtype=np.complex128).reshape(-1, 3)
    A = np.asarray(vert_areas, dtype=np.float64).reshape(-1, 1)
    out = np.empty((len(P), 3), dtype=np.complex128)
    for b0 in range(0, len(P), block):
        pm = P[b0:b0 + block, None, :]
        g = green_gradG_y(k, pm, V[None, :, :])
        out[b0:b0 + block] = -(A[None] * np.cross(np.broadcast_to(J[None], g.shape), -g)).sum(axis=-2)
    return out


# ------------------------------------------------------------------ ferrofluid: Helmholtz decomposition
def vert_normals(verts, faces):
    """Area-weighted vertex normals, normalised (src/utils/mesh_utils.py:242-274)."""
    a, n = face_areas_normals(verts, faces)
    vn = np.zeros((len(verts), 3))
    f = np.asarray(faces, dtype=np.int64)
    for c in range(3):
        np.add.at(vn, f[:, c], a[:, None] * n)
    return vn / np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-12)


def helmholtz_decomposition_u(verts, faces, vel, gaussQR, targets=None, block=64):
    """u1 + u2 of MeshHelmholtzDecomposition.solve
    (src/mesh_ferrofluid/mesh_Hemholtz_decomposition.py:52-126) at targets
    (default: the vertices), literally: tensor rule r1 = p[b], r2 = p[a] p[b],
    w = w[a] w[b] (:33-48); weights 2 A_f w r1 (BEM_utils.py:254-310, PLANAR);
    points and velocities (1-r1) x1 + (r1-r2) x2 + r2 x3 (BEM_utils.py:210-251);
    grad_G_y_3d(x=y, y=target) = (y - target) / (4 pi R^3) (BEM_utils.py:141-172);
    u1 = sum W (n . u) g, u2 = sum W (n x u) x g.  vel [B, V, 3] -> [B, Nt, 3]."""
    v = _f64(verts)
    f = np.asarray(faces, dtype=np.int64)
    vel = np.asarray(vel, dtype=np.float64)
    if vel.ndim == 2:
        vel = vel[None]
    p, w = gl_ref(gaussQR)
    r1 = np.tile(p[None, :], (gaussQR, 1)).reshape(-1)
    r2 = (p[:, None] * p[None, :]).reshape(-1)
    ww = (w[:, None] * w[None, :]).reshape(-1)
    area, n = face_areas_normals(v, f)
    W = 2.0 * area[:, None] * ww[None, :] * r1[None, :]  # [F, Q2]
    c0, c1, c2 = (1 - r1)[None, :, None], (r1 - r2)[None, :, None], r2[None, :, None]
    y = c0 * v[f[:, 0]][:, None] + c1 * v[f[:, 1]][:, None] + c2 * v[f[:, 2]][:, None]  # [F, Q2, 3]
    tg = v if targets is None else np.asarray(targets, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((vel.shape[0], len(tg), 3))
    for b in range(vel.shape[0]):
        ub = vel[b]
        uy = c0 * ub[f[:, 0]][:, None] + c1 * ub[f[:, 1]][:, None] + c2 * ub[f[:, 2]][:, None]  # [F, Q2, 3]
        ndu = (n[:, None, :] * uy).sum(-1)  # [F, Q2]
        ncu = np.cross(np.broadcast_to(n[:, None, :], uy.shape), uy)  # [F, Q2, 3]
        for t0 in range(0, len(tg), block):
            x = tg[t0:t0 + block]
            d = y[None] - x[:, None, None, :]  # [T, F, Q2, 3]
            R = np.linalg.norm(d, axis=-1, keepdims=True)
            g = (1.0 / 4.0 / math.pi) * (d / R) / (R * R)
            u1 = (W[None, :, :, None] * ndu[None, :, :, None] * g).sum(axis=(1, 2))
            u2 = (np.cross(np.broadcast_to(ncu[None], g.shape), g) * W[None, :, :, None]).sum(axis=(1, 2))
            out[b, t0:t0 + block] = u1 + u2
    return out


def helmholtz_decomposition_solve(verts, faces, vel, gaussQR):
    """new_vel of MeshHelmholtzDecomposition.solve (mesh_Hemholtz_decomposition.py:128-143):
    (I - n n^T) vel + (n . (u1 + u2)) n with area-weighted vertex normals."""
    u = helmholtz_decomposition_u(verts, faces, vel, gaussQR)
    vn = vert_normals(verts, faces)
    vel = np.asarray(vel, dtype=np.float64)
    if vel.ndim == 2:
        vel = vel[None]
    un = (vn[None] * u).sum(-1, keepdims=True)
    return vel - (vel * vn[None]).sum(-1, keepdims=True) * vn[None] + un * vn[None]


# ---------------------------------------------------------------- 2-D BEM (line segments)
# "Faces" are segments [F, 2] of a closed curve, vertices [V, 2].
def seg_areas_normals(verts, faces):
    """Segment lengths |v1 - v0| and normals (e_y, -e_x) / |e|, e = v1 - v0
    (src/utils/mesh_utils.py:174-183, 216-223)."""
    V = np.asarray(verts, dtype=np.float64)
    F = np.asarray(faces, dtype=np.int64)
    e = V[F[:, 1]] - V[F[:, 0]]
    L = np.linalg.norm(e, axis=1)
    n = np.stack([e[:, 1], -e[:, 0]], axis=1) / L[:, None]
    return L, n


def seg_vert_areas(verts, faces):
    """Vertex 'areas': sum of half the lengths of the two incident segments
    (mesh_utils.py:277-310 with dim = 2)."""
    L, _ = seg_areas_normals(verts, faces)
    va = np.zeros(len(verts))
    F = np.asarray(faces, dtype=np.int64)
    for t in range(2):
        np.add.at(va, F[:, t], L / 2.0)
    return va


def hankel2_ref(order, z):
    """The reference's Hankel function of the second kind (BEM_utils.py:9-31):
    the large-argument form sqrt(2 / (pi z)) exp(-j (z - pi/4 - order pi/2))."""
    th = math.pi / 4.0 + order * (math.pi / 2.0)
    return np.sqrt(2.0 / math.pi / z) * (np.cos(z - th) - 1j * np.sin(z - th))


def green2d(bem_type, k, x, y):
    """G_2d (BEM_utils.py:34-69): LAPLACE (bem_type 0) (1 / 2 pi) log(1 / |x - y|);
    Helmholtz 0.25 (Im H - j Re H) with H = hankel2_ref(0, k |x - y|)."""
    d = np.linalg.norm(x - y, axis=-1)
    if bem_type == 0:
        return (1.0 / (2.0 * math.pi)) * np.log(1.0 / d)
    h = hankel2_ref(0, k * d)
    return 0.25 * (h.imag - 1j * h.real)


def grad_green2d_y(bem_type, k, x, y):
    """grad_G_y_2d (BEM_utils.py:107-138): LAPLACE (1 / 2 pi) (x - y) / |x - y|^2;
    Helmholtz 0.25 (Im H1 - j Re H1) k (x - y) / |x - y|."""
    d = np.linalg.norm(x - y, axis=-1, keepdims=True)
    direc = (x - y) / d
    if bem_type == 0:
        return (1.0 / (2.0 * math.pi)) * direc / d
    h = hankel2_ref(1, k * d)
    return 0.25 * (h.imag - 1j * h.real) * (k * direc)


def panel_relation_2d(faces):
    """Shared vertex indices of two segments: 0 SEPARATE, 1 SAME_VERTEX,
    2 SAME_EDGE (mesh_utils.py:700-725 with dim = 2)."""
    F = np.asarray(faces, dtype=np.int64)
    return (F[:, None, :, None] == F[None, :, None, :]).sum(axis=(2, 3)).astype(np.uint8)


def bem2d_layer(verts, faces, gp, gw, layer, order_type, bem_type=0, k=1.0):
    """Unpacked [F, F, nphi, nphi] 2-D layer, row = test segment fy, column =
    source segment fx (integrations_2d.py:18-232):
      out[fy, fx, iy, ix] = A_y A_x sum_{a,b} w_a w_b J phi_iy(r1y) phi_ix(r1x) f(x, y)
    x = (1 - r1x) X0 + r1x X1 on fx, y = (1 - r1y) Y0 + r1y Y1 on fy, r1x = gp[a],
    r1y = gp[b] and J = 1 except SAME_EDGE: r1y = gp[a] gp[b], J = gp[a] (the
    half square y < x, as the reference maps it).  f = G (layer 1) or
    grad_y G . n_fy (layer 2)."""
    V = np.asarray(verts, dtype=np.float64)
    Fc = np.asarray(faces, dtype=np.int64)
    nF = len(Fc)
    L, nrm = seg_areas_normals(V, Fc)
    rel = panel_relation_2d(Fc)
    gp = np.asarray(gp, dtype=np.float64)
    gw = np.asarray(gw, dtype=np.float64)
    Q = gp.size
    ra = np.tile(gp, Q)          # x point index a (fast)
    rb = np.repeat(gp, Q)        # y point index b (slow)
    w = np.tile(gw, Q) * np.repeat(gw, Q)
    nphi = 2 if order_type == 1 else 1
    cplx = bem_type != 0
    out = np.zeros((nF, nF, nphi, nphi), dtype=np.complex128 if cplx else np.float64)
    for r in (0, 1, 2):
        fy, fx = np.nonzero(rel == r)
        if fy.size == 0:
            continue
        r1x = ra
        r1y = ra * rb if r == 2 else rb
        jac = ra if r == 2 else np.ones_like(ra)
        X0, X1 = V[Fc[fx, 0]], V[Fc[fx, 1]]
        Y0, Y1 = V[Fc[fy, 0]], V[Fc[fy, 1]]
        x = (1 - r1x)[None, :, None] * X0[:, None, :] + r1x[None, :, None] * X1[:, None, :]
        y = (1 - r1y)[None, :, None] * Y0[:, None, :] + r1y[None, :, None] * Y1[:, None, :]
        if layer == 1:
            f = green2d(bem_type, k, x, y)
        else:
            f = (grad_green2d_y(bem_type, k, x, y) * nrm[fy][:, None, :]).sum(-1)
        phix = [1 - r1x, r1x] if nphi == 2 else [np.ones_like(r1x)]
        phiy = [1 - r1y, r1y] if nphi == 2 else [np.ones_like(r1y)]
        s = (L[fy] * L[fx])[:, None]
        for iy in range(nphi):
            for ix in range(nphi):
                out[fy, fx, iy, ix] = (s * f * (w * jac * phiy[iy] * phix[ix])[None, :]).sum(1)
    return out


def bem2d_pack(order_type, U, faces, n_verts):
    """pack_matrix_tight with dim = 2 (BEM_utils.py:314-357): PLANAR entries
    spread as U / 4 over the 2 x 2 vertex pairs."""
    Fc = np.asarray(faces, dtype=np.int64)
    nF = len(Fc)
    if order_type == 0:
        U = np.repeat(np.repeat(U.reshape(nF, nF, 1, 1), 2, 2), 2, 3) / 4.0
    A = np.zeros((n_verts, n_verts), dtype=U.dtype)
    for iy in range(2):
        for ix in range(2):
            np.add.at(A, (Fc[:, iy][:, None], Fc[:, ix][None, :]), U[:, :, iy, ix])
    return A


def bem2d_dirichlet(verts, faces, g, gaussQR, order_type=1, bem_type=0, k=1.0):
    """MeshBEMSolver2d.solve with Dirichlet data only (mesh_BEM_solver_2d.py:77-250,
    rhs_layer_2d.py:66-130): V t = 1/2 M g + K^T g (LAPLACE, HELMHOLTZ_BEM) or
    V t = M g (HELMHOLTZ_MOM: gk = 2 gk1).  Returns (t, V, rhs)."""
    import scipy.linalg

    gp, gw = gl_ref(gaussQR)
    Vu = bem2d_layer(verts, faces, gp, gw, 1, order_type, bem_type, k)
    Ku = bem2d_layer(verts, faces, gp, gw, 2, order_type, bem_type, k)
    nV = len(verts)
    Vt = bem2d_pack(order_type, Vu, faces, nV)
    Kt = bem2d_pack(order_type, Ku, faces, nV)
    va = seg_vert_areas(verts, faces)
    g = np.asarray(g).reshape(-1)
    if bem_type == 2:
        rhs = g * va
    else:
        rhs = 0.5 * g * va + Kt.T @ g
    t = scipy.linalg.solve(Vt, rhs)
    return t, Vt, rhs


# ---------------------------------------------------------------- ferrofluid pressure step
def cot_laplacian_p3d(verts, faces, eps=1e-12):
    """pytorch3d.ops.cot_laplacian (the dependency of mesh_pressure.py:111-112;
    pytorch3d from git, unpinned in readme.md:14), restated from its published
    algorithm: Heron areas with the squared area clamped at eps, cot of the
    angle at vertex k = (sum of the other two squared sides - own squared
    side) / (4 area), L[i, j] = sum over faces of the cot opposite edge (i, j)
    (symmetric), inv_areas = 1 / (sum of the areas of the faces at a vertex).
    Returns (L dense [V, V], inv_areas [V])."""
    V = np.asarray(verts, dtype=np.float64)
    F = np.asarray(faces, dtype=np.int64)
    v0, v1, v2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    A = np.linalg.norm(v1 - v2, axis=1)
    B = np.linalg.norm(v0 - v2, axis=1)
    C = np.linalg.norm(v0 - v1, axis=1)
    s = 0.5 * (A + B + C)
    area = np.sqrt(np.maximum(s * (s - A) * (s - B) * (s - C), eps))
    A2, B2, C2 = A * A, B * B, C * C
    cot = np.stack([(B2 + C2 - A2) / area, (A2 + C2 - B2) / area, (A2 + B2 - C2) / area], 1) / 4.0
    nV = len(V)
    L = np.zeros((nV, nV))
    np.add.at(L, (F[:, 1], F[:, 2]), cot[:, 0])
    np.add.at(L, (F[:, 2], F[:, 0]), cot[:, 1])
    np.add.at(L, (F[:, 0], F[:, 1]), cot[:, 2])
    L = L + L.T
    asum = np.zeros(nV)
    for t in range(3):
        np.add.at(asum, F[:, t], area)
    inv = np.where(asum > 0, 1.0 / np.where(asum > 0, asum, 1.0), 0.0)
    return L, inv


def vertex_normals_area(verts, faces):
    """compute_vert_normals (mesh_utils.py:242-274): normalise(sum_f A_f n_f)."""
    V = np.asarray(verts, dtype=np.float64)
    F = np.asarray(faces, dtype=np.int64)
    A, n = face_areas_normals(V, F)
    vn = np.zeros_like(V)
    for t in range(3):
        np.add.at(vn, F[:, t], A[:, None] * n)
    return vn / np.linalg.norm(vn, axis=1, keepdims=True)


def signed_mean_curvature_p3d(verts, faces):
    """kappa of mesh_pressure.py:113-124: lv_j = 0.5 inv_area_j sum_i L_ij (v_i - v_j),
    kappa = 0.5 |lv| sign(lv . n_j)."""
    V = np.asarray(verts, dtype=np.float64)
    L, inv = cot_laplacian_p3d(V, faces)
    lv = 0.5 * inv[:, None] * (L.T @ V - L.sum(axis=0)[:, None] * V)
    vn = vertex_normals_area(V, faces)
    return 0.5 * np.linalg.norm(lv, axis=1) * np.sign((lv * vn).sum(1))


def vert_tangent_gradient(p, verts, faces):
    """compute_mesh_vert_tangent_gradient (mesh_utils.py:501-615) for p [V]:
    per face sum_t p_t (n x e_t) / (2 A), e_t = v_{t+2} - v_{t+1}, shared by
    its three vertices with weight 1/3."""
    V = np.asarray(verts, dtype=np.float64)
    F = np.asarray(faces, dtype=np.int64)
    A, n = face_areas_normals(V, F)
    g = np.zeros((len(F), 3))
    for t in range(3):
        e = V[F[:, (t + 2) % 3]] - V[F[:, (t + 1) % 3]]
        g += p[F[:, t]][:, None] * np.cross(n, e)
    g = g / 2.0 / A[:, None]
    out = np.zeros_like(V)
    for t in range(3):
        np.add.at(out, F[:, t], g / 3.0)
    return out


def pressure_step(verts, faces, vel, dt, rho, g, sigma, gaussQR, order_type=1, p_mag=None):
    """MeshPressure.solve with every vertex DIRICHLET (the droplet demo,
    demo_3d_mesh_droplet.py:140-150): p = rho g v_y - P_mag + sigma kappa,
    Laplace Dirichlet solve for dp/dn (ref_laplace_dirichlet), new_vel =
    vel - (dt / rho)(dpdn n + tangential grad p).  vel [V, 3]."""
    V = np.asarray(verts, dtype=np.float64)
    kap = signed_mean_curvature_p3d(V, faces)
    p = rho * g * V[:, 1] + sigma * kap
    if p_mag is not None:
        p = p - np.asarray(p_mag).reshape(-1)
    dpdn = np.asarray(ref_laplace_dirichlet(V, faces, p, gaussQR, order_type)[0]).reshape(-1)
    vn = vertex_normals_area(V, faces)
    grad = dpdn[:, None] * vn + vert_tangent_gradient(p, V, faces)
    return vel - (dt / rho) * grad, p, dpdn


# ---------------------------------------------------------------- tournament pivoting
# No reference counterpart: the reference's dense solve (torch.linalg.solve,
# src/mesh_BEM/mesh_BEM_solver_3d.py:174) factors with partial pivoting.  The
# product's optional tournament-pivoting panel follows CALU's tall-skinny
# selection (Grigori, Demmel, Xiang, SIAM J. Matrix Anal. Appl. 32 (2011)
# 1317-1350) on a fixed binary tree; this restates it for the checker.

def _gepp_select(rows, ids, npiv):
    """Partial pivoting on a copy of ``rows`` [r, c] (complex); ties go to the
    smaller id.  Returns the positions (into rows) of the npiv pivots in order
    and the eliminated copy."""
    F = np.array(rows, dtype=np.complex128, copy=True)
    ids = np.asarray(ids)
    free = np.ones(F.shape[0], bool)
    order = []
    for j in range(npiv):
        mag = np.where(free, np.abs(F[:, j].real) + np.abs(F[:, j].imag), -1.0)
        best = mag.max()
        cand = np.flatnonzero(mag == best)
        s = int(cand[np.argmin(ids[cand])])
        order.append(s)
        free[s] = False
        piv = F[s, j]
        rest = np.flatnonzero(free)
        if rest.size:
            l = F[rest, j] / piv if piv != 0 else np.zeros(rest.size, complex)
            F[rest, j + 1:] -= l[:, None] * F[s, j + 1:][None, :]
            F[rest, j] = l
    return order, F


def tslu_pivots(panel, n_cu=256, nb=64):
    """Tournament pivot rows of an [m, nb] panel (m >= 2 nb): G = min(n_cu,
    m // nb) leaves of R = ceil(m / G) consecutive rows (G recomputed so none
    is empty); a leaf keeps all its rows when it has <= nb of them and G > 1,
    else the nb rows partial pivoting picks; level l merges leaf g with
    g + 2^l by partial pivoting on the stacked original rows (own candidates
    first).  Returns (pivot ids in order, LAPACK-style interchanges ipiv
    [nb] with row j swapped with ipiv[j] >= j in sequence)."""
    panel = np.asarray(panel)
    m = panel.shape[0]
    G = min(n_cu, m // nb)
    R = -(-m // G)
    G = -(-m // R)
    cands = []
    for g in range(G):
        ids = list(range(g * R, min(m, (g + 1) * R)))
        if G == 1 or len(ids) > nb:
            order, _ = _gepp_select(panel[ids], ids, min(nb, len(ids)))
            ids = [ids[i] for i in order]
        cands.append(ids)
    step = 1
    while step < G:
        for g in range(0, G, 2 * step):
            if g + step < G:
                ids = cands[g] + cands[g + step]
                order, _ = _gepp_select(panel[ids], ids, min(nb, len(ids)))
                cands[g] = [ids[i] for i in order]
        step *= 2
    piv = cands[0]
    rows = list(range(m))  # rows[p] = original row now at place p
    where = {r: r for r in range(m)}
    ipiv = []
    for j, r in enumerate(piv):
        src = where[r]
        ipiv.append(src)
        a = rows[j]
        rows[j], rows[src] = r, a
        where[r], where[a] = j, src
    return piv, ipiv
