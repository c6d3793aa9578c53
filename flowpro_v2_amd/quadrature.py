"""Quadrature rule selection of QuadratureCentral (FlowPro/src/main/java/flowpro/core/
quadrature/QuadratureCentral.java:34-91): the rule of order k is exact to degree
2k - 1; the triangle rules of degree 3 and 7 have negative weights and are replaced
by those of degree 4 and 8.  The rule files are FlowPro's gaussRules data
(BSD licence, gauss_rules/LICENSE.txt).
"""
from __future__ import annotations

import os

import numpy as np

RULES_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gauss_rules")


def _read(path: str, cols: int):
    with open(path) as f:
        tok = f.read().split()
    n = int(tok[0])
    vals = [float(t) for t in tok[1:]]
    x = np.array(vals[: n * cols]).reshape(n, cols)
    w = np.array(vals[n * cols: n * cols + n])
    return x, w


def line_rule(order: int):
    """Gauss-Legendre rule with `order` points on [-1, 1] (QuadratureLine)."""
    x, w = _read(os.path.join(RULES_DIR, "gauss", f"{order}.txt"), 1)
    return x[:, 0].copy(), w


def triangle_rule(order: int):
    """Reference-triangle rule (0,0),(1,0),(0,1): Cartesian points = barycentric (b1, b2)
    (QuadratureTriangle.setCoords)."""
    deg = 2 * order - 1
    bary, w = _read(os.path.join(RULES_DIR, "gaussTri", f"{deg}.txt"), 3)
    if (w < 0).any():
        bary, w = _read(os.path.join(RULES_DIR, "gaussTri", f"{deg + 1}.txt"), 3)
    xy = np.stack([0.0 * bary[:, 0] + 1.0 * bary[:, 1] + 0.0 * bary[:, 2],
                   0.0 * bary[:, 0] + 0.0 * bary[:, 1] + 1.0 * bary[:, 2]], 1)
    return np.ascontiguousarray(xy), w


def square_rule(order: int):
    """QuadratureSquare (FlowPro/src/main/java/flowpro/core/quadrature/QuadratureSquare.java:7-25):
    the tensor product of the Gauss rule of the order on [0, 1]^2, point (i, j) ->
    ((1 + x_i) / 2, (1 + x_j) / 2), weight w_i w_j / 4, i outer."""
    x, w = line_rule(order)
    X = np.array([[(1 + x[i]) / 2, (1 + x[j]) / 2] for i in range(len(x)) for j in range(len(x))])
    W = np.array([w[i] * w[j] / 4 for i in range(len(x)) for j in range(len(x))])
    return np.ascontiguousarray(X), W


def tetra_rule(order: int):
    """Reference-tetrahedron rule of degree 2k - 1 (QuadratureTetra: the file holds the
    (xi, eta, zeta) coordinates directly; the rules have positive weights, no swap)."""
    x, w = _read(os.path.join(RULES_DIR, "gaussTetra", f"{2 * order - 1}.txt"), 3)
    return np.ascontiguousarray(x), w
