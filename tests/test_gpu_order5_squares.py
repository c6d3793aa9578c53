"""Order 5 on squares (25 basis functions, blocks of B = 100 with 4 equations): the quadrilateral
kernels, the padded mixed-mesh kernels (triangles and squares of orders up to 5, every element padded to
25 functions), the two-wavefront block-ILU0 factor (k_ilu_factor_level_2w) and the per-level sweeps
with one 128-thread workgroup per block row, against the oracle through the C ABI.

Tolerances (equation-norm relative unless said otherwise):
  residual, BDF right-hand side                       1e-11 (the oracle's own rounding at order 5: its
                                                      free-stream residual is 3.6e-12, the monomial
                                                      basis coefficients lose digits with the order)
  residual norms, rhs norm                            1e-10 (relative)
  exact Jacobian blocks                               1e-10 (max |diff| / max |entry|)
  J v, J^T v                                          1e-10
  first Krylov direction, GMRES(1) one cycle         FP64 factors 1e-9, FP32 factors 1e-4 (u32 cond(D))
  GMRES correction (tol 1e-13) vs oracle and dense    1e-8
  Newton residual history, state                      1e-8
"""
import math

import numpy as np
import pytest

from flowpro_v2_amd import cases, solver
import oracle as O
from helpers import block_rel_err, eq_rel_err
from test_gpu_mixed import _eq_err, _mixed, _setup as _mixed_setup

pytestmark = pytest.mark.gpu

SQ5_CONFIGS = [
    ("euler_q5", lambda: cases.channel_bump(6, 3, quads=True, jitter=0.2), lambda: cases.euler_props(order=5)),
    ("ns_q5", lambda: cases.cylinder_mesh(16, 3, quads=True), lambda: cases.ns_props(order=5, reynolds=300.0)),
    ("ae_q5", lambda: cases.channel_bump(6, 3, quads=True, jitter=0.2),
     lambda: cases.euler_props(model="AdiabaticEulers", order=5)),
]


def _setup(case, props, seed=3, rel=0.01, precond_bits=64):
    props = dict(props, jacobian="exact")
    o = O.Oracle(case, props)
    dev = solver.DeviceCase(case, props, precond_bits=precond_bits)
    assert (dev.nb, dev.neq) == (o.nb, o.neq) and dev.nb == 25
    W = cases.perturbed_state(dev.initial_state()[0][:: dev.nb], case.n_elem, dev.nb, rel, seed)
    o.set_w(W)
    o.limiter()
    dev.set_state(W)
    return o, dev, W


@pytest.mark.parametrize("name,mk_case,mk_props", SQ5_CONFIGS, ids=[c[0] for c in SQ5_CONFIGS])
def test_order5_squares_residual_and_jacobian(name, mk_case, mk_props):
    case, props = mk_case(), mk_props()
    print(f"[order5] residual and Jacobian {name}", flush=True)
    o, dev, W = _setup(case, props)
    o.assemble(0.05, -1.0)
    norms, rhs = dev.assemble(0.05, -1.0, 1, True)
    R, RHS = dev.residual()
    assert eq_rel_err(R, o.residual(), dev.nb, dev.neq) < 1e-11
    assert eq_rel_err(RHS, o.rhs()[0], dev.nb, dev.neq) < 1e-11
    np.testing.assert_allclose(norms, o.residuum_norms(), rtol=1e-10)
    assert math.isclose(rhs, o.rhs_norm(), rel_tol=1e-10)
    d, n = dev.jacobian()
    dref, nref = o.blocks()
    assert block_rel_err(d, dref) < 1e-10
    assert block_rel_err(n, nref) < 1e-10
    dev.close()


@pytest.mark.parametrize("name,mk_case,mk_props", SQ5_CONFIGS[:2], ids=[c[0] for c in SQ5_CONFIGS[:2]])
def test_order5_squares_jv_jtv(name, mk_case, mk_props):
    case, props = mk_case(), mk_props()
    print(f"[order5] J v, J^T v {name}", flush=True)
    o, dev, W = _setup(case, props)
    v = np.random.default_rng(7).standard_normal((case.n_elem, dev.B))
    a1 = 20.0
    assert block_rel_err(dev.jv(a1, v), o.jvp(a1, v)) < 1e-10
    assert block_rel_err(dev.jtv(a1, v), o.vjp(a1, v)) < 1e-10
    dev.close()


@pytest.mark.parametrize("bits,tol", [(64, 1e-9), (32, 1e-4)])
def test_order5_squares_first_krylov_direction(bits, tol):
    """GMRES(1), one cycle from x0 = 0: x = y1 M^-1 b / |M^-1 b| pins the B = 100 factor kernel and both
    level sweeps against the oracle's sequential BlockILU0."""
    case = cases.cylinder_mesh(16, 3, quads=True)
    props = cases.ns_props(order=5, reynolds=300.0, gmresRestart=1, gmresCycles=1, iterativeSolverTol=1e-14)
    print(f"[order5] first Krylov direction FP{bits}", flush=True)
    o, dev, W = _setup(case, props, precond_bits=bits)
    o.assemble(0.05, -1.0)
    dev.assemble(0.05, -1.0, 1, True)
    _, xref, _, _ = o.linear_solve()
    rc, it, err = dev.linear_solve()
    assert rc in (0, 1) and it == 1
    x = dev.solution()
    assert np.abs(x - xref).max() <= tol * np.abs(xref).max()
    dev.close()


def test_order5_squares_adjoint_preconditioner_first_direction():
    """The per-level transposed sweeps at B = 100: GMRES(1), one cycle, on the adjoint system returns a
    multiple of M^-T b (the oracle's transposed block-ILU0 application)."""
    case = cases.cylinder_mesh(16, 3, quads=True)
    props = cases.ns_props(order=5, reynolds=300.0, gmresRestart=1, gmresCycles=1, iterativeSolverTol=1e-14)
    print("[order5] adjoint first direction", flush=True)
    o, dev, W = _setup(case, props)
    a1 = 20.0
    _, g = dev.functional(0)
    rc, lam, it, err = dev.adjoint_solve(g, a1=a1)
    assert rc in (0, 1) and it == 1
    o.assemble(1.0 / a1, -1.0)
    o.linear_solve()
    r = o.ilu_apply(o.to_dof(g), trans=True)
    x = o.to_dof(lam)
    x, r = x / np.linalg.norm(x), r / np.linalg.norm(r)
    assert np.abs(x - np.sign(x @ r) * r).max() <= 1e-9
    dev.close()


@pytest.mark.parametrize("prec", ["blockilu0", "blockjacobi"])
def test_order5_squares_linear_solve(prec):
    case = cases.channel_bump(6, 3, quads=True, jitter=0.2)
    props = cases.euler_props(order=5, preconditioner=prec, iterativeSolverTol=1e-13)
    print(f"[order5] linear solve {prec}", flush=True)
    o, dev, W = _setup(case, props)
    # dt 0.01: at 0.05 blockJacobi GMRES(30) x 5 stalls near 5e-7 on the oracle as well
    o.assemble(0.01, -1.0)
    dev.assemble(0.01, -1.0, 1, True)
    ok, xref, _, _ = o.linear_solve()
    rc, it, err = dev.linear_solve()
    assert ok and rc == 0
    x = dev.solution()
    assert np.abs(x - xref).max() / np.abs(xref).max() < 1e-8
    xd = np.linalg.solve(o.global_matrix(), o.to_dof(o.rhs()[0]))
    assert np.abs(x - xd).max() / np.abs(xd).max() < 1e-8
    dev.close()


def test_order5_squares_newton_history():
    case = cases.cylinder_mesh(16, 3, quads=True)
    props = dict(cases.ns_props(order=5, reynolds=200.0, CFL=2.0, timeOrder=2, iterativeSolverTol=1e-10), jacobian="exact")
    print("[order5] BDF2 Newton history", flush=True)
    o = O.Oracle(case, props)
    W0 = cases.perturbed_state(o.get_w()[0][:: o.nb], case.n_elem, o.nb, 0.01, 5)
    o.set_w(W0)
    steps = 4
    hist_ref = o.run(steps)
    dev = solver.DeviceCase(case, props)
    dev.set_state(W0)
    hist = [r.residuum for r in solver.GpuImplicitSolver(dev, props).solve(steps)]
    assert len(hist) == len(hist_ref) == steps
    np.testing.assert_allclose(hist, hist_ref, rtol=1e-8)
    Wg, Wr = dev.get_state(), o.get_w()
    assert np.abs(Wg - Wr).max() / np.abs(Wr).max() < 1e-8
    dev.close()


# meshes of squares and triangles with per-element orders up to 5: the padded basis is 25 (B = 100)
MIXED5_CONFIGS = [
    ("euler_mixed_345", lambda: _mixed(6, 3, orders=(3, 4, 5), seed=11), lambda: cases.euler_props(order=5)),
    ("ns_mixed_45", lambda: _mixed(5, 3, orders=(4, 5), seed=12), lambda: cases.ns_props(order=5, reynolds=300.0)),
]


@pytest.mark.parametrize("name,mk_case,mk_props", MIXED5_CONFIGS, ids=[c[0] for c in MIXED5_CONFIGS])
def test_mixed_order5_residual_jacobian_and_products(name, mk_case, mk_props):
    case, props = mk_case(), mk_props()
    print(f"[order5] mixed {name}", flush=True)
    assert ((case.types == 4) & (case.orders == 5)).any()
    o, dev, W = _mixed_setup(case, props)
    assert dev.nb == 25
    o.assemble(0.05, -1.0)
    norms, rhs = dev.assemble(0.05, -1.0, 1, True)
    R, RHS = dev.residual()
    assert _eq_err(R, o.residual(), o.off, o.sizes, o.neq) < 1e-11
    assert _eq_err(RHS, o.rhs()[0], o.off, o.sizes, o.neq) < 1e-11
    np.testing.assert_allclose(norms, o.residuum_norms(), rtol=1e-10)
    d, n = dev.jacobian()
    dref, nref = o.blocks_flat()
    scale = max(np.abs(x).max() for x in dref)
    assert max(np.abs(a - b).max() for a, b in zip(d, dref)) / scale < 1e-10
    assert set(n) == set(nref)
    assert max(np.abs(n[k] - nref[k]).max() for k in nref) / scale < 1e-10
    v = np.random.default_rng(7).standard_normal(o.dofs)
    for got, ref in ((dev.jv(20.0, v), o.jvp(20.0, v)), (dev.jtv(20.0, v), o.vjp(20.0, v))):
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-10
    dev.close()


def test_mixed_order5_linear_solve():
    case = _mixed(6, 3, orders=(3, 4, 5), seed=11)
    props = cases.euler_props(order=5, iterativeSolverTol=1e-13)
    print("[order5] mixed linear solve", flush=True)
    o, dev, W = _mixed_setup(case, props)
    o.assemble(0.05, -1.0)
    dev.assemble(0.05, -1.0, 1, True)
    ok, xref, _, _ = o.linear_solve()
    rc, it, err = dev.linear_solve()
    assert ok and rc == 0
    x = dev.solution()
    assert np.abs(x - xref).max() / np.abs(xref).max() < 1e-8
    xd = np.linalg.solve(o.global_matrix(), o.to_dof(o.rhs()[0]))
    assert np.abs(x - xd).max() / np.abs(xd).max() < 1e-8
    dev.close()
