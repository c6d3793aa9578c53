"""Pins of the quadrilateral restatement in the oracle (squareElement, basis2Dsquare,
QuadratureSquare, transformation2Dsquare), CPU only.

No reference test covers quadrilaterals; the checks are what the definitions fix:
  area = shoelace area of the vertices (the bilinear map's |det J| integrates exactly) (1e-13)
  mass matrix on an axis-parallel rectangle a x b = a b (M1 (x) M1), M1 the 1D mass matrix of the
    Lagrange basis on linspace(0, 1, k), computed here from numpy polynomials (1e-12)
  elemSize = 4 area / perimeter (1e-13); geometric closure sum_k S_k n_k = 0 (1e-13)
  free-stream preservation on jittered (non-affine) quadrilaterals: R = 0 (1e-12)
  conservation through the neighbour-side face points (transformation2Dsquare.getXi): with slip
    walls only, the mass and energy residuals summed over the mesh vanish (1e-13)
  convergence of the scalar equation (the reference's ConvergenceOrderTest setup on the square
    divided into n x n quadrilaterals): rate > p + 0.7
  exact Jacobian against central differences of the residual, diagonal and neighbour blocks (1e-6)
"""
import math

import numpy as np
import pytest
from numpy.polynomial import polynomial as Pn

from flowpro_v2_amd import cases
import oracle as O


def _shoelace(P):
    x, y = P[:, 0], P[:, 1]
    return 0.5 * abs((x * np.roll(y, -1) - np.roll(x, -1) * y).sum())


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_area_size_and_closure(order):
    case = cases.channel_bump(8, 4, quads=True, jitter=0.25)
    o = O.Oracle(case, cases.euler_props(order=order))
    assert o.nf == 4 and o.nb == order * order
    for e in range(0, case.n_elem, 5):
        P = case.PXY[case.TP[e]]
        g = o.geometry(e)
        A = _shoelace(P)
        assert abs(g["area"] - A) <= 1e-13 * A
        perim = sum(np.linalg.norm(P[(k + 1) % 4] - P[k]) for k in range(4))
        assert abs(g["elemSize"] - 4 * A / perim) <= 1e-13 * g["elemSize"]
        _, w = np.polynomial.legendre.leggauss(order)
        s = sum((g["jacF"][k][:, None] * w[:, None] * g["normals"][k]).sum(0) for k in range(4))
        assert np.abs(s).max() < 1e-13


def _lagrange_mass_1d(k):
    nodes = np.linspace(0, 1, k) if k > 1 else np.array([0.5])
    polys = []
    for i in range(k):
        c = np.array([1.0])
        for j in range(k):
            if j != i:
                c = Pn.polymul(c, np.array([-nodes[j], 1.0]) / (nodes[i] - nodes[j]))
        polys.append(c)
    x, w = np.polynomial.legendre.leggauss(k + 2)
    x, w = (x + 1) / 2, w / 2
    V = np.array([Pn.polyval(x, c) for c in polys])
    return (V * w) @ V.T


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_mass_matrix_on_a_rectangle(order):
    a, b = 0.7, 0.3
    PXY = np.array([[0, 0], [a, 0], [a, b], [0, b]], dtype=np.float64)
    case = cases.Case("rect", PXY, np.array([[0, 1, 2, 3]], np.int32),
                      np.array([[-2, 0, 1], [-2, 1, 2], [-2, 2, 3], [-2, 3, 0]], np.int32))
    o = O.Oracle(case, dict(model="TestScalarEquation", a=(1.0, 0.5), k=0.0, order=order, dt=1e8))
    M1 = _lagrange_mass_1d(order)
    ref = a * b * np.kron(M1, M1)       # node r = (b_i, b_j), r = i k + j
    M = o.geometry(0)["M"]
    # the basis coefficients come from the inverse of the monomial Vandermonde matrix of the nodes
    # (basis2Dsquare), whose conditioning grows with the order: 1.9e-12 measured at order 5
    tol = 1e-12 if order <= 4 else 1e-11
    assert np.abs(M - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_free_stream_preserved_on_bilinear_quads(order):
    case = cases.channel_bump(8, 4, quads=True)
    case.PXY[:, 1] = np.repeat(np.linspace(0.0, 1.0, 5)[None, :], 9, 0).ravel()   # flat walls
    rng = np.random.default_rng(7)
    inner = (case.PXY[:, 0] > 1e-9) & (case.PXY[:, 0] < 3 - 1e-9) & (case.PXY[:, 1] > 1e-9) & (case.PXY[:, 1] < 1 - 1e-9)
    case.PXY[inner] += 0.08 * rng.uniform(-1, 1, size=(inner.sum(), 2))
    o = O.Oracle(case, cases.euler_props(order=order))
    R = o.residual()
    # rounding grows with the conditioning of the monomial basis: 2e-16, 2e-15, 5e-14, 3.6e-12 measured
    # at orders 2-5
    assert np.abs(R).max() < (1e-12 if order <= 4 else 1e-10)


@pytest.mark.parametrize("order", [2, 3, 5])
def test_conservation_through_neighbour_face_points(order):
    case = cases.square_mesh(5, 1.0, -4, quads=True)
    rng = np.random.default_rng(2)
    inner = (case.PXY > 1e-9).all(1) & (case.PXY < 1 - 1e-9).all(1)
    case.PXY[inner] += 0.06 * rng.uniform(-1, 1, size=(inner.sum(), 2))
    o = O.Oracle(case, cases.euler_props(order=order))
    W = np.tile(np.repeat([1.0, 0.1, -0.05, 2.5], o.nb), (case.n_elem, 1))
    W *= 1 + 0.02 * rng.standard_normal(W.shape)
    o.set_w(W)
    R = o.residual()
    nb = o.nb
    # the test functions sum to one only to the rounding of the monomial basis coefficients: 2e-16,
    # 3e-15, 2e-13, 4e-12 measured at orders 2-5
    tol = 1e-13 if order <= 3 else 1e-10
    for m in (0, 3):   # slip walls carry no mass and no energy flux
        s = R[:, m * nb:(m + 1) * nb].sum()
        assert abs(s) < tol * np.abs(R[:, m * nb:(m + 1) * nb]).sum()


def _scalar_props(order):
    return dict(model="TestScalarEquation", a=(1.0, 0.5), k=0.0, order=order, dt=1e8, steps=4, newtonIters=1,
                iterativeSolverTol=1e-13, preconditioner="blockilu0", penalty=0.0)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_convergence_rate_on_quadrilaterals(p):
    errs = []
    for n in (4, 8, 16):
        o = O.Oracle(cases.square_mesh(n, 1.0, -2, quads=True), _scalar_props(p + 1))
        o.run(4)
        errs.append(o.l2_error_scalar())
    rate = math.log(errs[1] / errs[2]) / math.log(2)
    assert rate > p + 0.7


@pytest.mark.parametrize("viscous,order", [(False, 2), (True, 3), (True, 5)])
def test_exact_jacobian_against_central_differences(viscous, order):
    case = cases.channel_bump(4, 2, quads=True, jitter=0.2)
    props = dict(cases.ns_props(order=order, reynolds=200.0) if viscous else cases.euler_props(order=order), jacobian="exact")
    o = O.Oracle(case, props)
    W = cases.perturbed_state(o.get_w()[0][:: o.nb], case.n_elem, o.nb, 0.02, 11)
    o.set_w(W)
    o.limiter()
    o.assemble(1.0, -1.0)
    d, nbk = o.blocks()
    TT = o.TT()
    rng = np.random.default_rng(0)
    eps = 1e-6
    worst = 0.0
    for e in rng.choice(case.n_elem, 3, replace=False):
        for i in rng.choice(o.B, 4, replace=False):
            Wp, Wm = W.copy(), W.copy()
            Wp[e, i] += eps
            Wm[e, i] -= eps
            o.set_w(Wp); Rp = o.residual()
            o.set_w(Wm); Rm = o.residual()
            dR = (Rp - Rm) / (2 * eps)
            M = o.geometry(e)["M"]
            mass = np.zeros(o.B)
            m, r = divmod(i, o.nb)
            mass[m * o.nb:(m + 1) * o.nb] = M[r]
            worst = max(worst, np.abs(d[e, i] - (-dR[e] + mass)).max() / np.abs(d[e, i]).max())
            for k in range(4):
                t = TT[e, k]
                if t > -1:
                    kk = list(TT[t]).index(e)
                    worst = max(worst, np.abs(nbk[t, kk, i] + dR[t]).max() / max(np.abs(dR[t]).max(), 1e-300))
    o.set_w(W)
    assert worst < 1e-6


def test_product_quad_integer_setup_bit_exact():
    """fp_neighbor_matrix_quad and fp_rcm_order_quad equal the oracle's createNeighborMatrix and
    reverseCuthillMcKee on quadrilaterals, also with a repeated boundary row of another type (the
    later row wins) and with missing rows (-1)."""
    from flowpro_v2_amd import native
    for case in (cases.cylinder_mesh(16, 5, quads=True), cases.channel_bump(9, 4, quads=True, jitter=0.2)):
        o = O.Oracle(case, cases.euler_props(order=2))
        TT = native.neighbor_matrix_quad(case.TP, case.bnd)
        assert np.array_equal(TT, o.TT())
        assert np.array_equal(native.rcm_order_quad(TT), o.order())
        extra = case.bnd[:5].copy()
        extra[:, 0] = -4
        for bnd in (np.concatenate([case.bnd, extra]), case.bnd[6:]):
            c2 = cases.Case("b", case.PXY, case.TP, bnd)
            assert np.array_equal(native.neighbor_matrix_quad(case.TP, bnd), O.Oracle(c2, cases.euler_props(order=2)).TT())
