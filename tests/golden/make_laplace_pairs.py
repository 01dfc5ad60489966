"""Writes tests/golden/laplace_pairs.npz: exact P1 Galerkin single- and
double-layer blocks of four triangle pairs (separate, common vertex, common
edge, identical), the integrals behind src/mesh_BEM/utils/integrations_3d.py:252-452
for G = 1/(4 pi R) and dG/dn_y = (x - y).n_y / (4 pi R^3).

Method (independent of the Sauter-Schwab maps it checks): the inner
integral over the y-triangle in closed form -- int phi/R from the analytic
potentials S0 = int 1/R, S1 = int (y - rho)/R (Graglia 1993), the double
layer from grad S0 (normal part = -h int 1/R^3, in-plane part =
int (y - rho)/R^3) -- and the outer integral over the x-triangle by scipy's
adaptive dblquad to 1e-10.  The closed forms are those of the oracle
(orc_tri_potentials, orc_tri_S0_grad), pinned against adaptive quadrature
and finite differences in tests/test_oracle_cpu.py.

Block layout: B[iy, ix] = int_{T_x} int_{T_y} phi_iy(y) phi_ix(x) k(x, y),
T_y the first face of the pair (the row face fy), T_x the second.

Run: python tests/golden/make_laplace_pairs.py   (about 70 s)
"""

import os
import sys

import numpy as np
from scipy import integrate

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

X = [np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.1, 0.0]), np.array([0.2, 0.9, 0.1])]
# (name, relation, Y vertices): the shared vertex / edge sits first in both
# faces, in the same order, as the Sauter-Schwab maps assume
PAIRS = [
    ("separate", 0, [X[0] + [0.3, 0.2, 1.5], X[1] + [0.1, 0.4, 1.2], X[2] + [0.5, 0.1, 1.4]]),
    ("vertex", 1, [X[0], np.array([-0.8, 0.3, 0.4]), np.array([-0.2, -0.9, 0.5])]),
    ("edge", 2, [X[0], X[1], np.array([0.4, -0.7, 0.6])]),
    ("face", 3, [X[0], X[1], X[2]]),
]


def inner(Y, iy, x, layer):
    """int_{T_y} phi_iy(y) k(x, y) dy in closed form."""
    n = np.cross(Y[1] - Y[0], Y[2] - Y[1])
    a2 = np.linalg.norm(n)
    n = n / a2
    grad = np.cross(n, Y[(iy + 2) % 3] - Y[(iy + 1) % 3]) / a2  # in-plane gradient of phi_iy
    if layer == 1:
        S0, S1, rho = O.tri_potentials(Y[0], Y[1], Y[2], x)
        return (grad @ (rho - Y[(iy + 1) % 3]) * S0 + grad @ S1) / (4 * np.pi)
    S0, G = O.tri_S0_grad(Y[0], Y[1], Y[2], x)
    h = (x - Y[0]) @ n
    rho = x - h * n
    gn = G @ n
    return (-(grad @ (rho - Y[(iy + 1) % 3])) * gn + h * (grad @ (G - gn * n))) / (4 * np.pi)


def block(Xt, Yt, layer, eps=1e-10):
    ax2 = np.linalg.norm(np.cross(Xt[1] - Xt[0], Xt[2] - Xt[0]))
    out = np.zeros((3, 3))
    for iy in range(3):
        for ix in range(3):
            def f(v, u):
                x = Xt[0] + u * (Xt[1] - Xt[0]) + v * (Xt[2] - Xt[0])
                return (1 - u - v, u, v)[ix] * inner(Yt, iy, x, layer)

            val, _ = integrate.dblquad(f, 0, 1, 0, lambda u: 1 - u, epsabs=eps, epsrel=eps)
            out[iy, ix] = val * ax2
    return out


def main():
    data = {}
    for name, rel, Y in PAIRS:
        data[f"{name}_Y"] = np.array(Y)
        data[f"{name}_X"] = np.array(X)
        data[f"{name}_rel"] = np.array(rel)
        data[f"{name}_V"] = block(X, Y, 1)
        # the double layer of a flat face on itself vanishes identically
        data[f"{name}_K"] = np.zeros((3, 3)) if rel == 3 else block(X, Y, 2)
        print(name, "done", flush=True)
    np.savez(os.path.join(HERE, "laplace_pairs.npz"), **data)


if __name__ == "__main__":
    main()
