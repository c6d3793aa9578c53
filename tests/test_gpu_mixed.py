"""Parity of the kernels on meshes that mix triangles and squares (elementType.txt) and element
orders 1-4 (order.txt) with the oracle, through the C ABI.

The device pads every element to the largest basis of the mesh (4, 9 or 16 functions) with zero
tables and an identity mass on the padded unknowns; the exports convert back to the reference's
layout (element after element, each [m][nb_e]; blocks B_t x B_e).  Meshes: the channel with the
bump, each cell a square or two triangles at random, with seeded per-element orders.
Tolerances (per-equation relative unless said otherwise):
  residual, BDF right-hand side                     1e-12
  residual norms, rhs norm                          1e-11 (relative)
  exact Jacobian blocks (square diagonal, rectangular neighbour)   1e-10 (max |diff| / max |entry|)
  J v, J^T v                                        1e-10
  GMRES correction (tol 1e-13) vs oracle and dense  1e-8 (max |diff| / max |x|)
  Newton residual history, state                    1e-8
  min elemSize / lambda, element means              1e-13
"""
import math

import numpy as np
import pytest

from flowpro_v2_amd import cases, solver
import oracle as O
from helpers import eq_rel_err

pytestmark = pytest.mark.gpu


def _mixed(nx, ny, orders=(1, 2, 3, 4), seed=5, jitter=0.2):
    """Squares and triangles (mixed_channel) with seeded per-element orders (None: one order)."""
    c = cases.mixed_channel(nx, ny, jitter=jitter, seed=seed)
    if orders is not None:
        c.orders = cases.seeded_orders(c.n_elem, orders, seed)
    return c


def _tri_orders(nx, ny, orders, seed):
    """Triangles only, per-element orders: the channel_bump triangles with an order.txt."""
    b = cases.channel_bump(nx, ny, jitter=0.2, seed=seed)
    return cases.Case("tri_orders", b.PXY, b.TP, b.bnd, orders=cases.seeded_orders(b.n_elem, orders, seed))


MIXED_CONFIGS = [
    ("euler_mixed_1234", lambda: _mixed(8, 4), lambda: cases.euler_props(order=4)),
    ("euler_types_only_q3", lambda: _mixed(8, 4, orders=None), lambda: cases.euler_props(order=3)),
    ("euler_tri_orders_vanleer", lambda: _tri_orders(8, 4, (1, 2, 3), 2),
     lambda: cases.euler_props(order=3, numericalFlux="vanLeer")),
    ("ns_mixed_123_penalty", lambda: _mixed(8, 4, orders=(1, 2, 3), seed=6),
     lambda: cases.ns_props(order=3, reynolds=200.0, penalty=1.0)),
    ("ns_mixed_234", lambda: _mixed(6, 3, orders=(2, 3, 4), seed=8), lambda: cases.ns_props(order=4, reynolds=300.0)),
    ("ae_mixed_12", lambda: _mixed(8, 4, orders=(1, 2), seed=9),
     lambda: cases.euler_props(model="AdiabaticEulers", order=2)),
    # order 5 (15 functions, the degree-9 triangle rule of 19 points) on triangles
    ("euler_tri_orders_345", lambda: _tri_orders(8, 4, (3, 4, 5), 3), lambda: cases.euler_props(order=5)),
    ("ns_tri_orders_45", lambda: _tri_orders(6, 3, (4, 5), 4), lambda: cases.ns_props(order=5, reynolds=300.0)),
]


def _eq_err(a, b, off, sizes, neq):
    """Per-equation relative error over a flat vector in the reference's layout."""
    errs = []
    for m in range(neq):
        pa = np.concatenate([a[off[e] + m * (sizes[e] // neq): off[e] + (m + 1) * (sizes[e] // neq)] for e in range(len(sizes))])
        pb = np.concatenate([b[off[e] + m * (sizes[e] // neq): off[e] + (m + 1) * (sizes[e] // neq)] for e in range(len(sizes))])
        den = np.abs(pb).max()
        errs.append(np.abs(pa - pb).max() / (den if den > 0 else 1.0))
    return max(errs)


def _state(o, seed=3, rel=0.01):
    """The constant initial state times (1 + rel N(0,1)) plus rel max|w0_m| N(0,1) per coefficient,
    so that zero components (rho v) are perturbed too."""
    w0 = np.asarray(o.get_w(), np.float64)
    rng = np.random.default_rng(seed)
    scale = np.concatenate([np.repeat(np.abs(w0[o.off[e]:o.off[e + 1]]).max(), o.sizes[e]) for e in range(o.n_elem)])
    return w0 * (1 + rel * rng.standard_normal(w0.shape)) + rel * scale * rng.standard_normal(w0.shape)


def _setup(case, props, seed=3):
    props = dict(props, jacobian="exact")
    o = O.Oracle(case, props)
    dev = solver.DeviceCase(case, props)
    assert dev.mixed and dev.neq == o.neq and dev.dofs == o.dofs
    assert np.array_equal(dev.sizes, o.sizes) and np.array_equal(dev.off, o.off)
    W = _state(o, seed)
    o.set_w(W)
    o.limiter()
    dev.set_state(W)
    return o, dev, W


def test_mixed_state_round_trip_and_initial_state():
    case = _mixed(6, 3)
    o, dev, W = _setup(case, cases.euler_props(order=4))
    assert np.array_equal(dev.get_state(), W)
    o2 = O.Oracle(case, dict(cases.euler_props(order=4)))
    np.testing.assert_allclose(dev.initial_state(), o2.get_w(), rtol=1e-15, atol=0)
    dev.close()


@pytest.mark.parametrize("name,mk_case,mk_props", MIXED_CONFIGS, ids=[c[0] for c in MIXED_CONFIGS])
def test_mixed_residual_and_jacobian(name, mk_case, mk_props):
    case, props = mk_case(), mk_props()
    o, dev, W = _setup(case, props)
    dt, dto = 0.05, -1.0
    o.assemble(dt, dto)
    norms, rhs = dev.assemble(dt, dto, 1, True)
    R, RHS = dev.residual()
    RHSref, _ = o.rhs()
    assert _eq_err(R, o.residual(), o.off, o.sizes, o.neq) < 1e-12
    assert _eq_err(RHS, RHSref, o.off, o.sizes, o.neq) < 1e-12
    np.testing.assert_allclose(norms, o.residuum_norms(), rtol=1e-11)
    assert math.isclose(rhs, o.rhs_norm(), rel_tol=1e-11)
    d, n = dev.jacobian()
    dref, nref = o.blocks_flat()
    scale = max(np.abs(x).max() for x in dref)
    assert max(np.abs(a - b).max() for a, b in zip(d, dref)) / scale < 1e-10
    assert set(n) == set(nref)
    assert max(np.abs(n[k] - nref[k]).max() for k in nref) / scale < 1e-10
    dev.close()


@pytest.mark.parametrize("name,mk_case,mk_props", MIXED_CONFIGS, ids=[c[0] for c in MIXED_CONFIGS])
def test_mixed_jv_jtv(name, mk_case, mk_props):
    case, props = mk_case(), mk_props()
    o, dev, W = _setup(case, props)
    v = np.random.default_rng(7).standard_normal(o.dofs)
    a1 = 20.0
    for got, ref in ((dev.jv(a1, v), o.jvp(a1, v)), (dev.jtv(a1, v), o.vjp(a1, v))):
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-10
    dev.close()


@pytest.mark.parametrize("prec", ["blockilu0", "ilu0", "blockjacobi"])
def test_mixed_linear_solve(prec):
    case = _mixed(8, 4)
    props = cases.euler_props(order=4, preconditioner=prec, iterativeSolverTol=1e-13)
    o, dev, W = _setup(case, props)
    o.assemble(0.05, -1.0)
    dev.assemble(0.05, -1.0, 1, True)
    ok, xref, _, _ = o.linear_solve()
    rc, it, err = dev.linear_solve()
    x = dev.solution()
    assert ok and rc == 0
    assert np.abs(x - xref).max() / np.abs(xref).max() < 1e-8
    A = o.global_matrix()
    r, _ = o.rhs()
    xd = np.linalg.solve(A, o.to_dof(r))
    assert np.abs(x - xd).max() / np.abs(xd).max() < 1e-8
    dev.close()


@pytest.mark.parametrize("name,mk_case,mk_props,steps", [
    ("euler_mixed", lambda: _mixed(8, 4, orders=(1, 2, 3)),
     lambda: cases.euler_props(order=3, CFL=5.0, newtonIters=2, iterativeSolverTol=1e-10), 6),
    ("ns_mixed_bdf2", lambda: _mixed(8, 4, orders=(2, 3), seed=6),
     lambda: cases.ns_props(order=3, reynolds=200.0, CFL=2.0, timeOrder=2, iterativeSolverTol=1e-10), 5),
], ids=["euler_mixed", "ns_mixed_bdf2"])
def test_mixed_newton_history(name, mk_case, mk_props, steps):
    case, props = mk_case(), dict(mk_props(), jacobian="exact")
    o = O.Oracle(case, props)
    W0 = _state(o, 5)
    o.set_w(W0)
    hist_ref = o.run(steps)
    dev = solver.DeviceCase(case, props)
    dev.set_state(W0)
    hist = [r.residuum for r in solver.GpuImplicitSolver(dev, props).solve(steps)]
    assert len(hist) == len(hist_ref) == steps
    np.testing.assert_allclose(hist, hist_ref, rtol=1e-8)
    Wg, Wr = dev.get_state(), o.get_w()
    assert np.abs(Wg - Wr).max() / np.abs(Wr).max() < 1e-8
    dev.close()


def test_mixed_time_step_and_averages():
    case = _mixed(8, 4)
    o, dev, W = _setup(case, cases.euler_props(order=4))
    o.set_cfl(1.0)
    assert math.isclose(dev.size_over_lambda(), o.time_step(), rel_tol=1e-13)
    avg = dev.averages()
    for e in range(0, case.n_elem, 3):
        M = o.mass(e)
        nb = M.shape[0]
        area = M.sum()   # sum_ab M_ab = int (sum_a phi_a)^2 = |K| for a nodal basis (partition of unity)
        ref = W[o.off[e]:o.off[e + 1]].reshape(o.neq, nb) @ (M @ np.ones(nb) / area)
        assert np.abs(avg[e] - ref).max() <= 1e-13 * np.abs(ref).max()
    dev.close()


def test_mixed_freestream_residual_vanishes():
    case = cases.mixed_channel(8, 4, flat=True, jitter=0.2)
    case.orders = cases.seeded_orders(case.n_elem, (1, 2, 3, 4), 5)
    dev = solver.DeviceCase(case, cases.euler_props(order=4))
    dev.set_state(dev.initial_state())
    dev.assemble(0.05, -1.0, 1, False)
    R, _ = dev.residual()
    assert np.abs(R).max() < 1e-12
    dev.close()


@pytest.mark.parametrize("which", [0, 1])
def test_mixed_functional_and_gradient(which):
    """Drag / lift on the cylinder with per-element orders 2-4: the value against the oracle, the
    exact gradient against the oracle's forward differences (OptimisationToolExport, h = 1e-8) and
    zero away from the wall."""
    c = cases.cylinder_mesh(32, 8)
    case = cases.Case("cyl_orders", c.PXY, c.TP, c.bnd, orders=cases.seeded_orders(c.n_elem, (2, 3, 4), 4))
    o, dev, W = _setup(case, cases.ns_props(order=4, reynolds=300.0))
    val, g = dev.functional(which)
    assert abs(val - o.functional(which)) <= 1e-12
    gfd = o.functional_grad_fd(which)
    assert np.abs(g - gfd).max() <= 1e-6 * np.abs(gfd).max()
    TT = o.TT()
    wall = np.concatenate([np.full(o.sizes[e], (TT[e] == -1).any()) for e in range(case.n_elem)])
    assert np.abs(g[~wall]).max() == 0.0 and np.abs(g[wall]).max() > 0
    dev.close()


def test_mixed_unsupported_features():
    """Mixed meshes have kernels for the NavierStokes and AdiabaticEulers models: the scalar test
    equation is refused."""
    case = _mixed(6, 2)
    with pytest.raises(solver.UnsupportedCase):
        solver.DeviceCase(case, dict(model="TestScalarEquation", a=(1.0, 0.5), k=0.0, order=4, dt=1e8))


def test_mixed_steady_adjoint_matches_dense():
    """The steady adjoint (a1 = 0: (-dR/dW)^T lambda = dI/dW) on a padded mixed mesh: the padded unknowns
    keep a unit diagonal at a1 = 0, so the block-ILU0 preconditioner stays regular; lambda against a dense
    solve of the oracle's unpadded matrix (a mesh of under 150 unknowns, which GMRES(30) x 5 solves
    exactly whatever the conditioning)."""
    case = _mixed(3, 2, orders=(1, 2), seed=13)
    props = cases.ns_props(order=2, reynolds=300.0, iterativeSolverTol=1e-12)
    o, dev, W = _setup(case, props)
    assert dev.dofs < 150
    _, g = dev.functional(0)
    rc, lam, it, err = dev.adjoint_solve(g, a1=0.0)
    assert rc == 0, (it, err)
    o.assemble(float("inf"), -1.0)   # a1 = 1/dt = 0
    A = o.global_matrix()
    lam_ref = np.linalg.solve(A.T, o.to_dof(g))
    assert np.abs(o.to_dof(lam) - lam_ref).max() <= 1e-8 * np.abs(lam_ref).max()
    dev.close()


def test_uniform_order5_triangles_without_order_file():
    """A uniform mesh of order-5 triangles given without order.txt (the uniform-mesh layout of the
    reference: every element 15 functions) runs on the padded mixed-mesh kernels: the residual, the
    Newton history and the state against the oracle's uniform-mesh path."""
    case = cases.channel_bump(6, 3, jitter=0.2)
    props = dict(cases.euler_props(order=5, CFL=2.0, iterativeSolverTol=1e-12), jacobian="exact")
    print("[mixed] uniform order-5 triangles", flush=True)
    o = O.Oracle(case, props)
    assert o.nb == 15
    W0 = cases.perturbed_state(o.get_w()[0][:: o.nb], case.n_elem, o.nb, 0.01, 5)
    o.set_w(W0)
    dev = solver.DeviceCase(case, props)
    assert dev.mixed and dev.dofs == case.n_elem * 15 * 4
    dev.set_state(W0.ravel())
    o.limiter()
    o.assemble(0.05, -1.0)
    dev.assemble(0.05, -1.0, 1, True)
    R, _ = dev.residual()
    Rref = o.residual()
    assert eq_rel_err(np.asarray(R).reshape(Rref.shape), Rref, 15, 4) < 1e-12
    o.set_w(W0)
    dev.set_state(W0.ravel())
    steps = 3
    hist_ref = o.run(steps)
    hist = [r.residuum for r in solver.GpuImplicitSolver(dev, props).solve(steps)]
    np.testing.assert_allclose(hist, hist_ref, rtol=1e-8)
    Wg, Wr = np.asarray(dev.get_state()).reshape(case.n_elem, -1), o.get_w()
    assert np.abs(Wg - Wr).max() / np.abs(Wr).max() < 1e-8
    dev.close()
