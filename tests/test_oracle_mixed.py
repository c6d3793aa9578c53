"""Pins of the oracle on meshes that mix element types (triangles and squares, elementType.txt)
and element orders (order.txt), CPU only.

No reference test covers mixed meshes; the checks are what the definitions fix:
  a mixed build with one type and one order equals the uniform build bit for bit (residual,
    Jacobian blocks)
  free-stream preservation on a flat channel of jittered triangles and quadrilaterals with
    per-element orders 1-4: R = 0 (1e-12)
  conservation through faces whose two sides have different bases and a face rule of the larger
    order: slip walls only, the mass and energy residuals sum to zero (1e-13)
  the exact Jacobian (square diagonal blocks, rectangular neighbour blocks) against central
    differences of the residual (1e-6), with penalty 0: the reference writes a neighbour block
    from the own element's face integrand (KR), whose interior-penalty term carries the own
    element's c_IP; DGpure.limiter sets c_IP = penalty only on elements of order > 1, so at a face
    between an order-1 element and a higher-order one the KR block is not the derivative of the
    neighbour's residual (by -c_IP (WL - WR) with the two c_IP differing), in the reference and here
  GMRES with the rectangular-block ILU0, ilu0 and blockjacobi against a dense solve of the
    assembled matrix (1e-8; the reference's LinearSolverTest criterion)
"""
import numpy as np
import pytest

from flowpro_v2_amd import cases
import oracle as O


def _mixed(nx=8, ny=4, flat=False, orders=(1, 2, 3, 4), seed=5, jitter=0.2):
    c = cases.mixed_channel(nx, ny, flat=flat, jitter=jitter)
    c.orders = cases.seeded_orders(c.n_elem, orders, seed)
    return c


def test_mixed_build_with_one_type_and_order_equals_uniform():
    u = cases.channel_bump(6, 3, quads=True, jitter=0.2)
    m = cases.Case("m", u.PXY, u.TP, u.bnd, types=np.full(u.n_elem, 4, np.int32),
                   orders=np.full(u.n_elem, 3, np.int32))
    props = cases.ns_props(order=3, reynolds=200.0, jacobian="exact")
    ou, om = O.Oracle(u, props), O.Oracle(m, props)
    W = cases.perturbed_state(ou.get_w()[0][:: ou.nb], u.n_elem, ou.nb, 0.02, 4)
    ou.set_w(W)
    om.set_w(W.ravel())
    assert np.array_equal(ou.residual().ravel(), om.residual())
    for o in (ou, om):
        o.limiter()
        o.assemble(0.1, -1.0)
    d, n = ou.blocks()
    dm, nm = om.blocks_flat()
    assert all(np.array_equal(d[e], dm[e]) for e in range(u.n_elem))
    assert all(np.array_equal(n[e, k], blk) for (e, k), blk in nm.items())


def test_free_stream_preserved_on_mixed_types_and_orders():
    c = _mixed(8, 4, flat=True)
    o = O.Oracle(c, cases.euler_props(order=4))
    assert o.mixed and len(set(o.sizes)) > 2
    assert np.abs(o.residual()).max() < 1e-12


@pytest.mark.parametrize("seed", [1, 2])
def test_conservation_across_faces_of_different_orders(seed):
    base = cases.mixed_channel(6, 4, flat=True, jitter=0.3, seed=seed)
    bnd = base.bnd.copy()
    bnd[:, 0] = -4                      # slip walls everywhere
    c = cases.Case("slip", base.PXY, base.TP, bnd, types=base.types,
                   orders=cases.seeded_orders(base.n_elem, (1, 2, 3, 4), seed))
    o = O.Oracle(c, cases.euler_props(order=4))
    rng = np.random.default_rng(seed)
    W = np.concatenate([np.repeat([1.0, 0.1, -0.05, 2.5], s // 4) for s in o.sizes])
    W *= 1 + 0.02 * rng.standard_normal(W.shape)
    o.set_w(W)
    R = o.residual()
    for m in (0, 3):
        parts = [R[o.off[e] + m * (o.sizes[e] // 4): o.off[e] + (m + 1) * (o.sizes[e] // 4)] for e in range(c.n_elem)]
        s = sum(p.sum() for p in parts)
        assert abs(s) < 1e-13 * sum(np.abs(p).sum() for p in parts)


@pytest.mark.parametrize("viscous", [False, True])
def test_exact_jacobian_against_central_differences(viscous):
    c = _mixed(4, 2, orders=(1, 2, 3))
    props = dict(cases.ns_props(order=3, reynolds=200.0, penalty=0.0) if viscous else cases.euler_props(order=3),
                 jacobian="exact")
    o = O.Oracle(c, props)
    w0 = np.asarray(o.get_w())
    rng = np.random.default_rng(0)
    W = w0 * (1 + 0.02 * rng.standard_normal(w0.shape))
    o.set_w(W)
    o.limiter()
    o.assemble(1.0, -1.0)
    diag, neigh = o.blocks_flat()
    TT = o.TT()
    eps, worst = 1e-6, 0.0
    for e in rng.choice(c.n_elem, 4, replace=False):
        b, nbas = int(o.sizes[e]), int(o.sizes[e]) // 4
        for i in rng.choice(b, 3, replace=False):
            Wp, Wm = W.copy(), W.copy()
            Wp[o.off[e] + i] += eps
            Wm[o.off[e] + i] -= eps
            o.set_w(Wp); Rp = o.residual()
            o.set_w(Wm); Rm = o.residual()
            dR = (Rp - Rm) / (2 * eps)
            mass = np.zeros(b)
            m, r = divmod(i, nbas)
            mass[m * nbas:(m + 1) * nbas] = o.mass(e)[r]
            col = -dR[o.off[e]:o.off[e + 1]] + mass
            worst = max(worst, np.abs(diag[e][i] - col).max() / np.abs(diag[e][i]).max())
            for k, t in enumerate(TT[e][:4]):
                if t > -1 and (t, list(TT[t]).index(e)) in neigh:
                    kk = list(TT[t]).index(e)
                    ref = -dR[o.off[t]:o.off[t + 1]]
                    # scaled by the element's own column: a basis function that vanishes on the
                    # shared edge has an exactly zero neighbour column (central differences: noise)
                    worst = max(worst, np.abs(neigh[(t, kk)][i] - ref).max() / np.abs(col).max())
    o.set_w(W)
    assert worst < 1e-6


@pytest.mark.parametrize("prec", ["blockilu0", "ilu0", "blockjacobi"])
def test_gmres_with_rectangular_blocks_matches_dense_solve(prec):
    c = _mixed(6, 3)
    o = O.Oracle(c, cases.euler_props(order=4, preconditioner=prec, iterativeSolverTol=1e-12, jacobian="exact"))
    w0 = np.asarray(o.get_w())
    W = w0 * (1 + 0.01 * np.random.default_rng(3).standard_normal(w0.shape))
    o.set_w(W)
    o.limiter()
    o.assemble(0.05, -1.0)
    ok, x, it, err = o.linear_solve()
    A = o.global_matrix()
    r, _ = o.rhs()
    xd = np.linalg.solve(A, o.to_dof(r))
    assert ok and np.abs(x - xd).max() / np.abs(xd).max() < 1e-8


def test_product_mixed_integer_setup_bit_exact():
    """fp_neighbor_matrix_mixed and fp_rcm_order_quad (the face graph of 4-slot rows) equal the
    oracle's createNeighborMatrix and reverseCuthillMcKee on triangles and squares together, also
    with a repeated boundary row of another type (the later row wins) and with missing rows (-1)."""
    from flowpro_v2_amd import native
    for seed in (7, 8):
        c = cases.mixed_channel(9, 4, jitter=0.2, seed=seed)
        o = O.Oracle(c, cases.euler_props(order=2))
        TT = native.neighbor_matrix_mixed(c.TP, c.types, c.bnd)
        assert np.array_equal(TT, o.TT())
        assert (TT[c.types == 3, 3] == -1000).all()
        assert np.array_equal(native.rcm_order_quad(TT), o.order())
        extra = c.bnd[:5].copy()
        extra[:, 0] = -4
        for bnd in (np.concatenate([c.bnd, extra]), c.bnd[6:]):
            c2 = cases.Case("b", c.PXY, c.TP, bnd, types=c.types)
            assert np.array_equal(native.neighbor_matrix_mixed(c.TP, c.types, bnd), O.Oracle(c2, cases.euler_props(order=2)).TT())


def test_triangle_orders_with_one_order_equal_uniform():
    """Per-element orders on triangles only (order.txt without elementType.txt), all equal: the
    residual and Jacobian blocks of the uniform triangle build, bit for bit."""
    u = cases.channel_bump(6, 3, jitter=0.2)
    m = cases.Case("m", u.PXY, u.TP, u.bnd, orders=np.full(u.n_elem, 3, np.int32))
    props = cases.ns_props(order=3, reynolds=200.0, jacobian="exact")
    ou, om = O.Oracle(u, props), O.Oracle(m, props)
    assert om.mixed and (om.sizes == ou.B).all()
    W = cases.perturbed_state(ou.get_w()[0][:: ou.nb], u.n_elem, ou.nb, 0.02, 4)
    ou.set_w(W)
    om.set_w(W.ravel())
    assert np.array_equal(ou.residual().ravel(), om.residual())
    for o in (ou, om):
        o.limiter()
        o.assemble(0.1, -1.0)
    d, n = ou.blocks()
    dm, nm = om.blocks_flat()
    assert all(np.array_equal(d[e], dm[e]) for e in range(u.n_elem))
    assert all(np.array_equal(n[e, k], blk) for (e, k), blk in nm.items())


def test_order_five_triangles_pins():
    """Order 5 (15 functions, the degree-9 triangle rule of 19 points, 5-point face rules): a
    uniform order.txt of 5 equals the uniform order-5 build bit for bit, and free-stream
    preservation and conservation hold on triangles with per-element orders 3-5."""
    u = cases.channel_bump(4, 2, jitter=0.2)
    m = cases.Case("m", u.PXY, u.TP, u.bnd, orders=np.full(u.n_elem, 5, np.int32))
    props = cases.euler_props(order=5)
    ou, om = O.Oracle(u, props), O.Oracle(m, props)
    assert ou.nb == 15 and (om.sizes == 60).all()
    W = cases.perturbed_state(ou.get_w()[0][:: ou.nb], u.n_elem, ou.nb, 0.02, 4)
    ou.set_w(W)
    om.set_w(W.ravel())
    assert np.array_equal(ou.residual().ravel(), om.residual())
    b = cases.channel_bump(6, 3, jitter=0.2, seed=2)
    b.PXY[:, 1] = np.repeat(np.linspace(0.0, 1.0, 4)[None, :], 7, 0).ravel()
    c = cases.Case("t", b.PXY, b.TP, b.bnd, orders=cases.seeded_orders(b.n_elem, (3, 4, 5), 2))
    o = O.Oracle(c, props)
    assert np.abs(o.residual()).max() < 1e-12
    bnd = c.bnd.copy()
    bnd[:, 0] = -4
    cs = cases.Case("slip", c.PXY, c.TP, bnd, orders=c.orders)
    o = O.Oracle(cs, props)
    rng = np.random.default_rng(5)
    W = np.concatenate([np.repeat([1.0, 0.1, -0.05, 2.5], s // 4) for s in o.sizes])
    W *= 1 + 0.02 * rng.standard_normal(W.shape)
    o.set_w(W)
    R = o.residual()
    for mq in (0, 3):
        parts = [R[o.off[e] + mq * (o.sizes[e] // 4): o.off[e] + (mq + 1) * (o.sizes[e] // 4)] for e in range(cs.n_elem)]
        assert abs(sum(p.sum() for p in parts)) < 1e-13 * sum(np.abs(p).sum() for p in parts)
