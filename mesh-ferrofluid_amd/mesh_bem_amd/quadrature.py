"""Quadrature rules on [0,1] and on triangles.

Two families are provided:

* ``get_gaussion_integration_points_and_weights`` -- the reference's 1-D
  Gauss-Legendre table mapped to [0,1] (src/utils/BEM_utils.py:360-588).  The
  reference writes its nodes as decimal literals inside ``torch.Tensor([...])``,
  which creates float32 values that are then cast to the requested dtype, so a
  float64 run integrates with float32-rounded nodes.  This module computes the
  nodes with ``numpy.polynomial.legendre.leggauss`` and applies the same
  float32 rounding, then the same ``(p + 1) / 2`` and ``w / 2`` in the target
  dtype.  Supported orders are the reference's: 1..8 and 12.

* ``radon7`` -- the 7-point degree-5 rule on a triangle (Radon 1948; listed
  by Dunavant 1985 as rule 5): centroid weight 9/40 and two orbits with
  barycentric a = (6 -+ sqrt(15))/21, weights (155 -+ sqrt(15))/1200.

Triangle rules are returned as barycentric tables ``lam[N,3]`` and weights
``w[N]`` normalised so that the physical weight of point n on a triangle of
area A is ``w[n] * 2 * A``; a point is ``lam0*v0 + lam1*v1 + lam2*v2``.
"""

import math
from typing import List, Tuple

import numpy as np
import torch

_SUPPORTED_GL = (1, 2, 3, 4, 5, 6, 7, 8, 12)


def gauss_legendre_f32_nodes(N: int) -> Tuple[np.ndarray, np.ndarray]:
    """Nodes/weights on [-1,1], rounded to float32 (as the reference's literals are)."""
    if N not in _SUPPORTED_GL:
        raise NotImplementedError("Please refer N as 2~8 for gaussion interation")
    x, w = np.polynomial.legendre.leggauss(N)
    return x.astype(np.float32), w.astype(np.float32)


def get_gaussion_integration_points_and_weights(
    N: int,
    device: torch.device = torch.device("cpu"),
    dtype: torch.dtype = torch.float32,
) -> List[torch.Tensor]:
    """1-D Gauss-Legendre points on [0,1] and weights summing to 1.

    Same signature, values and error behaviour as the reference
    (src/utils/BEM_utils.py:360-588): float32-rounded nodes, then
    ``(p + 1) / 2`` and ``w / 2`` evaluated in ``dtype``.
    """
    x32, w32 = gauss_legendre_f32_nodes(N)
    p = torch.from_numpy(x32).to(dtype).to(device)
    w = torch.from_numpy(w32).to(dtype).to(device)
    return [(p + 1.0) / 2.0, w / 2.0]


def collapsed_gl_rule(gp: np.ndarray, gw: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Collapsed (Duffy) tensor Gauss rule on the unit triangle.

    Point (a, b) -> r1 = gp[a], r2 = gp[b] * r1, weight gw[a] * gw[b] * r1, the
    map x = (1 - r1) v0 + (r1 - r2) v1 + r2 v2 of
    src/mesh_BEM/cuda/mesh_MOM_3d.cu:410-420.  Returns barycentric ``lam`` and
    weights ``w`` (physical weight = w * 2A).  Order of points: a major, b minor,
    matching the kernel's ``i / N`` and ``i % N``.
    """
    gp = np.asarray(gp, dtype=np.float64)
    gw = np.asarray(gw, dtype=np.float64)
    n = gp.size
    lam = np.zeros((n * n, 3))
    w = np.zeros(n * n)
    for a in range(n):
        for b in range(n):
            r1 = gp[a]
            r2 = gp[b] * r1
            lam[a * n + b] = (1.0 - r1, r1 - r2, r2)
            w[a * n + b] = gw[a] * gw[b] * r1
    return lam, w


def radon7() -> Tuple[np.ndarray, np.ndarray]:
    """7-point degree-5 triangle rule; weights sum to 1/2 (physical = w * 2A)."""
    s15 = math.sqrt(15.0)
    a1 = (6.0 - s15) / 21.0
    a2 = (6.0 + s15) / 21.0
    w1 = (155.0 - s15) / 1200.0
    w2 = (155.0 + s15) / 1200.0
    lam = [(1.0 / 3, 1.0 / 3, 1.0 / 3)]
    w = [9.0 / 40.0]
    for a, wa in ((a1, w1), (a2, w2)):
        b = 1.0 - 2.0 * a
        lam += [(a, a, b), (a, b, a), (b, a, a)]
        w += [wa, wa, wa]
    return np.array(lam), 0.5 * np.array(w)


def triangle_monomial_integral(a: int, b: int) -> float:
    """Exact integral of x^a y^b over the unit triangle {x, y >= 0, x + y <= 1}."""
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
