"""Pins of the oracle's reference-mode functions, term by term.

The reference's CFIE (src/mesh_BEM/cuda/mesh_MOM_3d.cu:332-496, RHS :244-330)
and its Laplace layers (src/mesh_BEM/utils/integrations_3d.py:252-452) are
restated literally in oracle/mom_oracle.c; the GPU parity tests compare the
HIP kernels with that restatement.  These tests tie the restatement to the
mathematics, each with a check that a dropped or sign-flipped term fails:

* EFIE part: on well-separated face pairs the reference's collapsed rule and
  the NS 7-point rule both converge to the same smooth integral, and
  the NS EFIE is itself pinned (Mie series, analytic potentials), so
  the reference's vector- and scalar-potential terms must agree with it to
  quadrature accuracy.  The RHS E term likewise against the NS RHS.
* identity term: the closed-form Gram integral of two RWG functions.
* MFIE part: its current must equal the (independent) NS EFIE current
  to discretisation accuracy, and its bistatic RCS the Mie series.
* Laplace layers: fed a proper (non-collapsed) 4-D rule and the two fixes of
  the reference's maps (shared vertex first, left-out same-face sub-domains),
  the layer code reproduces exact Galerkin integrals (tests/golden/
  laplace_pairs.npz: closed-form inner integral, adaptive outer integral) to
  1e-9 for separate, common-vertex, common-edge and identical pairs; the
  reference's rule enters only through the tables.  The composed Dirichlet
  solve with exact layers reaches the reference demo's known answer
  (demo/demo_Laplace_3d.py:39-78).
"""

import math
import os

import numpy as np
import pytest
import scipy.linalg

import oracle as O
from mesh_bem_amd.geometry import icosphere

HERE = os.path.dirname(os.path.abspath(__file__))
KH = np.array([math.cos(math.pi / 6), math.sin(math.pi / 6), 0.0])  # reference incidence (.cu:24-70)
EZ = np.array([0.0, 0.0, 1.0])


def mesh_tables(V, F):
    f = np.asarray(F, dtype=np.int64)
    edges, fe = O.edges_bruteforce(f, len(V))
    return f, fe, O.rwg_fe(fe, len(edges))


def far_mask(V, f, eta=4.0):
    """Face-edge mask of face pairs farther apart than eta (rho_a + rho_b)."""
    ptr, idx = O.near_pairs_bruteforce(V, f, eta)
    nF = len(f)
    near = np.zeros((nF, nF), dtype=bool)
    for b in range(nF):
        near[idx[ptr[b]:ptr[b + 1]], b] = True
    return np.repeat(np.repeat(~near, 3, 0), 3, 1)


@pytest.fixture(scope="module")
def sphere2():
    V, F = icosphere(2)
    f, fe, rfe = mesh_tables(V, F)
    return V, f, fe, rfe


# ------------------------------------------------------------------ EFIE part
def test_ref_efie_terms_match_ns_rule_on_far_pairs(sphere2):
    V, f, fe, rfe = sphere2
    k = 1.0
    gpx, gwx = O.gl_ref(4)
    gpy, gwy = O.gl_ref(3)
    ptr, idx = O.near_pairs_bruteforce(V, f, 2.0)
    ZN = O.ns_zmat_faceedge(V, f, k, ptr, idx)
    m = far_mask(V, f)
    assert m.sum() > 0.4 * m.size  # about half of all pairs at 320 faces
    scale = np.abs(ZN[m]).max()

    def rel(Z):
        return np.abs(Z - ZN)[m].max() / scale

    Zvec = O.ref_zmat_faceedge(V, f, gpx, gwx, gpy, gwy, k, O.ZT_VEC)
    Zsca = O.ref_zmat_faceedge(V, f, gpx, gwx, gpy, gwy, k, O.ZT_SCA)
    Zefie = O.ref_zmat_faceedge(V, f, gpx, gwx, gpy, gwy, k, O.ZT_EFIE)
    assert rel(Zefie) < 1e-6
    assert np.abs(Zefie - (Zvec + Zsca)).max() < 1e-14 * np.abs(Zefie).max()
    # planted mistakes: a dropped or sign-flipped potential term
    assert rel(Zsca) > 1e-3 and rel(Zvec) > 0.5
    assert rel(Zvec - Zsca) > 0.5 and rel(Zsca - Zvec) > 1e-3


def test_ref_rhs_e_term_matches_ns_rule_rhs(sphere2):
    V, f, fe, rfe = sphere2
    k = 1.5
    gp, gw = O.gl_ref(4)
    RE = O.ref_rhs_faceedge(V, f, gp, gw, k, O.RT_E).reshape(-1)
    RN = O.ns_rhs_faceedge(V, f, k, [KH], [EZ])[:, 0]
    assert np.abs(RE - RN).max() < 1e-6 * np.abs(RN).max()
    # the H term of the same call is n x H_inc with H_inc = khat x E_inc
    RH = O.ref_rhs_faceedge(V, f, gp, gw, k, O.RT_H).reshape(-1)
    RA = O.ref_rhs_faceedge(V, f, gp, gw, k).reshape(-1)
    assert np.abs(RA - RE - RH).max() < 1e-14 * np.abs(RA).max()
    assert np.abs(RH - RN).max() > 0.1 * np.abs(RN).max()


# ------------------------------------------------------------------ identity term
def test_ref_identity_term_is_the_rwg_gram_integral():
    """1/2 int_T f_i . f_j with f = c (x - P) has the closed form
    1/2 c_i c_j (A/12 (sum_k |P_k|^2 + |sum_k P_k|^2) - A c.(P_i + P_j) + A P_i.P_j)
    (second moment int x.x = A/12 (sum |P_k|^2 + |sum P_k|^2)); zero off the
    self pair."""
    V, F = icosphere(1)
    f = np.asarray(F, dtype=np.int64)
    gpx, gwx = O.gl_ref(4)
    gpy, gwy = O.gl_ref(3)
    Z = O.ref_zmat_faceedge(V, f, gpx, gwx, gpy, gwy, 1.0, O.ZT_ID)
    A, _ = O.face_areas_normals(V, f)
    nF = len(f)
    ref = np.zeros((3 * nF, 3 * nF))
    for t in range(nF):
        P = V[f[t]]
        cen = P.mean(0)
        mxx = A[t] / 12 * ((P ** 2).sum() + (P.sum(0) ** 2).sum())
        c = []
        for e in range(3):
            i1, i2 = f[t][(e + 1) % 3], f[t][(e + 2) % 3]
            c.append(np.linalg.norm(V[i2] - V[i1]) / (2 * A[t]) * (1 if i1 < i2 else -1))
        for a in range(3):
            for b in range(3):
                g = mxx - A[t] * cen @ (P[a] + P[b]) + A[t] * P[a] @ P[b]
                ref[3 * t + a, 3 * t + b] = 0.5 * c[a] * c[b] * g
    # float32-rounded Gauss nodes: exact only to ~1e-7
    assert np.abs(Z.imag).max() == 0.0
    assert np.abs(Z.real - ref).max() < 1e-6 * np.abs(ref).max()
    assert np.abs(2 * Z.real - ref).max() > 0.5 * np.abs(ref).max()


# ------------------------------------------------------------------ MFIE part
@pytest.fixture(scope="module")
def mfie_parts(sphere2):
    V, f, fe, rfe = sphere2
    k = 1.0
    gpx, gwx = O.gl_ref(4)
    gpy, gwy = O.gl_ref(3)
    Zid = O.gather_rwg(O.ref_zmat_faceedge(V, f, gpx, gwx, gpy, gwy, k, O.ZT_ID), rfe)
    Zk = O.gather_rwg(O.ref_zmat_faceedge(V, f, gpx, gwx, gpy, gwy, k, O.ZT_K), rfe)
    Vh = O.gather_rhs(O.ref_rhs_faceedge(V, f, gpx, gwx, k, O.RT_H).reshape(-1, 1), rfe)[:, 0]
    ptr, idx = O.near_pairs_bruteforce(V, f, 2.0)
    Zn = O.gather_rwg(O.ns_zmat_faceedge(V, f, k, ptr, idx), rfe)
    Vn = O.gather_rhs(O.ns_rhs_faceedge(V, f, k, [KH], [EZ]), rfe)[:, 0]
    return k, Zid, Zk, Vh, np.linalg.solve(Zn, Vn)


def test_ref_mfie_current_matches_ns_rule_efie_current(mfie_parts):
    """Two independent formulations (MFIE with the reference's rule; EFIE with
    the 7-point rule and analytic singular terms) give one current."""
    k, Zid, Zk, Vh, Ins = mfie_parts

    def rel(Z, v):
        return np.linalg.norm(np.linalg.solve(Z, v) - Ins) / np.linalg.norm(Ins)

    assert rel(Zid + Zk, Vh) < 0.03
    # planted mistakes
    assert rel(Zid - Zk, Vh) > 0.5  # K sign
    assert rel(Zid, Vh) > 0.3  # K dropped
    assert rel(Zid + Zk, -Vh) > 1.5  # H sign
    assert rel(2 * Zid + Zk, Vh) > 0.3  # identity factor


def mie_amplitudes(ka, theta):
    """(S1, S2) of a PEC sphere (Bohren-Huffman amplitude functions;
    a_n = psi_n'/xi_n', b_n = psi_n/xi_n)."""
    from scipy.special import spherical_jn, spherical_yn

    nmax = int(ka + 4 * ka ** (1 / 3) + 10)
    mu = math.cos(theta)
    S1 = S2 = 0j
    p0, p1 = 0.0, 1.0
    for n in range(1, nmax + 1):
        jn, jp = spherical_jn(n, ka), spherical_jn(n, ka, derivative=True)
        hn = jn + 1j * spherical_yn(n, ka)
        hp = jp + 1j * spherical_yn(n, ka, derivative=True)
        an, bn = (jn + ka * jp) / (hn + ka * hp), jn / hn
        tau = n * mu * p1 - (n + 1) * p0
        c = (2 * n + 1) / (n * (n + 1))
        S1 += c * (an * p1 + bn * tau)
        S2 += c * (an * tau + bn * p1)
        p0, p1 = p1, ((2 * n + 1) * mu * p1 - (n + 1) * p0) / n
    return S1, S2


def bistatic_errors(V, F, fe, I, k):
    """Relative error of sigma(theta) / (pi a^2) against Mie in the E plane
    (|S2|^2) and H plane (|S1|^2) at 0, 45, 90, 135, 180 degrees."""
    lam, w = O.radon7()
    A, _ = O.face_areas_normals(V, F)
    hp = np.cross(KH, EZ)
    errs = []
    for plane, pol in (("E", EZ), ("H", hp)):
        for th in np.deg2rad([0, 45, 90, 135, 180]):
            rh = math.cos(th) * KH + math.sin(th) * pol
            N = np.zeros(3, dtype=complex)
            for t in range(len(F)):
                P = V[F[t]]
                X = lam @ P
                ph = w * np.exp(1j * k * (X @ rh))
                for e in range(3):
                    i1, i2 = F[t][(e + 1) % 3], F[t][(e + 2) % 3]
                    c = np.linalg.norm(V[i2] - V[i1]) / (2 * A[t]) * (1 if i1 < i2 else -1)
                    N += I[fe[3 * t + e]] * 2 * A[t] * c * (ph[:, None] * (X - P[e])).sum(0)
            Nt = N - rh * (rh @ N)
            s_mom = k ** 2 * np.linalg.norm(Nt) ** 2 / (4 * np.pi ** 2)
            S1, S2 = mie_amplitudes(k, th)
            errs.append(s_mom / (4 * abs(S2 if plane == "E" else S1) ** 2 / k ** 2) - 1)
    return np.array(errs)


def test_mie_amplitudes_backscatter_consistent():
    from test_oracle_cpu import mie_backscatter

    for ka in (0.5, 1.0, 3.0):
        S1, S2 = mie_amplitudes(ka, math.pi)
        assert abs(S1 + S2) < 1e-9 * abs(S1)
        assert abs(4 * abs(S1) ** 2 / ka ** 2 / mie_backscatter(ka) - 1) < 1e-12


def test_ref_mfie_bistatic_rcs_matches_mie(sphere2, mfie_parts):
    V, f, fe, rfe = sphere2
    k, Zid, Zk, Vh, _ = mfie_parts
    e = bistatic_errors(V, f, fe, np.linalg.solve(Zid + Zk, Vh), k)
    assert np.abs(e).max() < 0.08, e
    e_flip = bistatic_errors(V, f, fe, np.linalg.solve(Zid - Zk, Vh), k)
    assert np.abs(e_flip).max() > 0.5


def test_ref_cfie_is_the_sum_of_its_parts(sphere2):
    V, f, fe, rfe = sphere2
    gpx, gwx = O.gl_ref(3)
    gpy, gwy = O.gl_ref(2)
    parts = [O.ref_zmat_faceedge(V, f, gpx, gwx, gpy, gwy, 2.0, t) for t in (O.ZT_ID, O.ZT_K, O.ZT_VEC, O.ZT_SCA)]
    full = O.ref_zmat_faceedge(V, f, gpx, gwx, gpy, gwy, 2.0)
    assert np.abs(full - sum(parts)).max() < 1e-13 * np.abs(full).max()


# ------------------------------------------------------------------ Laplace layers
GOLD = np.load(os.path.join(HERE, "golden", "laplace_pairs.npz"))
PAIRS = ["separate", "vertex", "edge", "face"]


def layer_block(name, layer, tabs, fix):
    Y, X, rel = GOLD[f"{name}_Y"], GOLD[f"{name}_X"], int(GOLD[f"{name}_rel"])
    verts = np.concatenate([Y, X])
    faces = np.array([[0, 1, 2], [3, 4, 5]])
    R = np.array([[3, rel], [rel, 3]], dtype=np.uint8)
    U = O.laplace_layer(verts, faces, None, None, layer, 1, rel=R, tabs=tabs, fix=fix)
    return U[0, 0] if name == "face" else U[0, 1]


@pytest.mark.parametrize("name", PAIRS)
@pytest.mark.parametrize("layer", [1, 2])
def test_laplace_layer_with_exact_rule_matches_golden(name, layer):
    """Sauter-Schwab maps of the restatement, a proper 10^4 Gauss rule and the
    two fixes: the exact Galerkin block (golden) to 1e-9."""
    ref = GOLD[f"{name}_{'V' if layer == 1 else 'K'}"]
    U = layer_block(name, layer, O.proper_tables(10), O.LAP_EXACT)
    if name == "face" and layer == 2:
        assert np.abs(U).max() < 1e-12  # (x - y).n_y = 0 on a flat face
        return
    assert np.abs(U - ref).max() < 1e-9 * np.abs(ref).max(), np.abs(U - ref).max() / np.abs(ref).max()


def test_laplace_planted_kernel_mistakes_fail_golden():
    ref = GOLD["edge_K"]
    U = layer_block("edge", 2, O.proper_tables(10), O.LAP_EXACT)
    assert np.abs(-U - ref).max() > 0.5 * np.abs(ref).max()  # normal sign
    assert np.abs(U.T - ref).max() > 0.05 * np.abs(ref).max()  # test / trial swapped
    # the reference's same-face maps cover half of T x T
    Uh = layer_block("face", 1, O.proper_tables(10), O.LAP_REORDER)
    assert np.abs(Uh + Uh.T - GOLD["face_V"]).max() < 1e-9 * np.abs(GOLD["face_V"]).max()
    assert np.abs(2 * Uh - GOLD["face_V"]).max() > 1e-3 * np.abs(GOLD["face_V"]).max()


def test_golden_generator_reproduces_one_block():
    import importlib.util

    spec = importlib.util.spec_from_file_location("mk", os.path.join(HERE, "golden", "make_laplace_pairs.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    X = [GOLD["separate_X"][i] for i in range(3)]
    Y = [GOLD["separate_Y"][i] for i in range(3)]
    assert np.abs(mk.block(X, Y, 1, eps=1e-9) - GOLD["separate_V"]).max() < 1e-8 * np.abs(GOLD["separate_V"]).max()


def test_reference_rule_enters_only_through_the_tables():
    """The reference's double collapse is the rank-3 rule fed to the same map
    builder; its missing vertex reordering changes only common-vertex and
    common-edge blocks; the layer code is the one pinned above."""
    V, F = icosphere(1)
    f = np.asarray(F, dtype=np.int64)
    gp, gw = O.gl_ref(3)
    t_ref = O.laplace_tables(gp, gw)
    t_from = O.laplace_tables_from(*O.rank3_rule(gp, gw))
    for a, b in zip(t_ref, t_from):
        assert np.array_equal(a, b)
    rel = O.panel_relation_bruteforce(f)
    for layer in (1, 2):
        U0 = O.laplace_layer(V, f, gp, gw, layer, 1, rel, t_ref, 0)
        U1 = O.laplace_layer(V, f, gp, gw, layer, 1, rel, t_ref, O.LAP_REORDER)
        same = (rel == 0) | (rel == 3)
        assert np.array_equal(U0[same], U1[same])
        if layer == 1:
            assert np.abs(U0[~same] - U1[~same]).max() > 1e-6 * np.abs(U0).max()


def test_pack_matrix_tight_index_map():
    """Separable blocks U[fy, fx, iy, ix] = a(fy, iy) b(fx, ix) pack into the
    outer product of the vertex sums: rows = row face's vertices."""
    V, F = icosphere(1)
    f = np.asarray(F, dtype=np.int64)
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal((len(f), 3)), rng.standard_normal((len(f), 3))
    U = a[:, None, :, None] * b[None, :, None, :]
    A = O.pack_matrix_tight(1, U, f, len(V))
    sa, sb = np.zeros(len(V)), np.zeros(len(V))
    np.add.at(sa, f.reshape(-1), a.reshape(-1))
    np.add.at(sb, f.reshape(-1), b.reshape(-1))
    assert np.abs(A - np.outer(sa, sb)).max() < 1e-12 * np.abs(A).max()


# ------------------------------------------------------------------ Dirichlet solve
def test_dirichlet_solve_with_exact_layers_reaches_known_answer():
    """f = x^2 - y^2 on the unit sphere has dn f = 2 (x^2 - y^2) (the reference
    demo's known answer).  The composed solve V t = 1/2 M g + K g with exact
    layers and the consistent P1 mass is within 2 % at 320 faces; dropping K or
    flipping its sign fails.  The reference's lumped mass (1/2 g vert_area,
    rhs_layer_3d.py:127) and its rule do not reach it (recorded, not fixed)."""
    V, F = icosphere(2)
    g = V[:, 0] ** 2 - V[:, 1] ** 2
    exact = 2 * g

    def err(t):
        return np.linalg.norm(t - exact) / np.linalg.norm(exact)

    tabs = O.proper_tables(6)
    t, Vt, rhs = O.ref_laplace_dirichlet(V, F, g, 3, 1, tabs, O.LAP_EXACT, mass="consistent")
    assert err(t) < 0.02
    half_mg = 0.5 * O.consistent_mass(V, F) @ g
    kg = rhs - half_mg
    assert err(scipy.linalg.solve(Vt, half_mg)) > 0.2  # K dropped
    assert err(scipy.linalg.solve(Vt, half_mg - kg)) > 0.3  # K sign
    t_lumped, _, _ = O.ref_laplace_dirichlet(V, F, g, 3, 1, tabs, O.LAP_EXACT)
    assert err(t_lumped) > 0.2
    t_ref, _, _ = O.ref_laplace_dirichlet(V, F, g, 3, 1)
    assert err(t_ref) > 1.0


# ------------------------------------------------------------------ NS P1 layers
def test_ns_p1_layers_against_exact_layers_and_known_answer():
    """The NS P1 rule (7-point rule on far pairs; rule on T_x times
    closed-form inner integral on near pairs) against the exact Galerkin
    layers: far blocks to the rule's accuracy, near blocks within the outer
    rule's error; its Dirichlet solve reaches the demo's known answer as the
    exact layers do."""
    V, F = icosphere(2)
    f = np.asarray(F, dtype=np.int64)
    ptr, idx = O.near_pairs_bruteforce(V, f)
    Vu, Ku = O.laplace_p1_ns(V, f, ptr, idx)
    rel = O.panel_relation_bruteforce(f)
    tabs = O.proper_tables(8)
    Ve = O.laplace_layer(V, f, None, None, 1, 1, rel, tabs, O.LAP_EXACT)
    Ke = O.laplace_layer(V, f, None, None, 2, 1, rel, tabs, O.LAP_EXACT)
    near = np.zeros((len(f), len(f)), dtype=bool)
    for b in range(len(f)):
        near[b, idx[ptr[b]:ptr[b + 1]]] = True
    assert np.array_equal(near, near.T)
    assert np.abs(Vu - Ve)[~near].max() < 1e-4 * np.abs(Ve[~near]).max()
    assert np.abs(Ku - Ke)[~near].max() < 5e-4 * np.abs(Ke[~near]).max()
    assert np.abs(Vu - Ve)[near].max() < 0.03 * np.abs(Ve).max()
    assert np.abs(Ku - Ke)[near].max() < 0.05 * np.abs(Ke).max()
    g = V[:, 0] ** 2 - V[:, 1] ** 2
    t, _, _ = O.dirichlet_solve_from_layers(V, f, g, Vu, Ku)
    assert np.linalg.norm(t - 2 * g) / np.linalg.norm(2 * g) < 0.02
    # planted: the double layer with the opposite normal
    t_bad, _, _ = O.dirichlet_solve_from_layers(V, f, g, Vu, -Ku)
    assert np.linalg.norm(t_bad - 2 * g) / np.linalg.norm(2 * g) > 0.3
