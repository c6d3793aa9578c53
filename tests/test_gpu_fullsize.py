"""Checks at the benchmark sizes, where index width, level schedules and the dataflow sweeps and
factorisation meet their real shapes: the 200,000-triangle NS p3 cylinder (8.0 M DOFs, 1.28e9
FP64 matrix entries) and the 1.57 M-tetrahedron Kuhn cube (31.5 M DOFs, 3.1e9 entries).

  residual of the 200k case vs the oracle (one evaluation)          1e-12 equation-norm relative
  <u, J v> = <J^T u, v> (matrix-free kernels)                        1e-10 relative
  Jacobian-free GMRES correction vs the assembled-matrix correction  1e-8 (max |diff| / max |x|)
    (the same Krylov solve through J v and through the assembled SpMV: the assembled matrix
    equals the matrix-free derivative at full size)
  true residual |b - A x| / |b| of the FP64 solve, A x by the
    matrix-free J v (independent of the assembled matrix)             1e-6 at GMRES tol 1e-10
  FP32 against FP64 preconditioner: same corrections                  1e-6
"""
import numpy as np
import pytest

from flowpro_v2_amd import cases, solver
import oracle as O
from helpers import eq_rel_err

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


def _state(dev, case, seed):
    return cases.perturbed_state(dev.initial_state()[0][:: dev.nb], case.n_elem, dev.nb, 0.01, seed)


def _linear_checks(dev, a1):
    rng = np.random.default_rng(11)
    shape = dev.get_state().shape
    u, v = rng.standard_normal(shape), rng.standard_normal(shape)
    lhs, rhs = float((u * dev.jv(a1, v)).sum()), float((dev.jtv(a1, u) * v).sum())
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1.0)
    rc, it, err = dev.linear_solve()
    assert rc == 0
    x = dev.solution()
    rc2, it2, err2 = dev.linear_solve_jfnk()
    assert rc2 == 0
    x_mf = dev.solution()
    assert np.abs(x_mf - x).max() <= 1e-8 * np.abs(x).max()
    return x


def test_ns_p3_cylinder_200k():
    case = cases.cylinder_mesh(400, 250)
    props = dict(cases.ns_props(order=4, iterativeSolverTol=1e-10), jacobian="exact")
    dev = solver.DeviceCase(case, props)
    W = _state(dev, case, 7)
    dev.set_state(W)
    dt = 1e-3
    dev.assemble(dt, -1.0, 1, True)
    R, RHS = dev.residual()
    o = O.Oracle(case, props, threads=16)
    o.set_w(W)
    o.limiter()
    assert eq_rel_err(R, o.residual(), dev.nb, dev.neq) < 1e-12
    del o
    x = _linear_checks(dev, 1.0 / dt)
    # true residual of the correction in DOF numbering: b = RHS in DOF order, A x = J v with v the
    # correction in element layout
    order = dev.element_order()
    x_elem = np.empty_like(RHS)
    x_elem[order] = x.reshape(-1, dev.B)
    r = RHS - dev.jv(1.0 / dt, x_elem)
    assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(RHS)
    dev.close()
    # the FP32 preconditioner reaches the same correction
    dev = solver.DeviceCase(case, props, precond_bits=32)
    dev.set_state(W)
    dev.assemble(dt, -1.0, 1, True)
    rc, it, err = dev.linear_solve()
    assert rc == 0
    assert np.abs(dev.solution() - x).max() <= 1e-6 * np.abs(x).max()
    dev.close()


def test_euler_p1_kuhn_cube_1p57m_tets():
    case = cases.kuhn_cube(64, 0.05, 4)
    props = dict(cases.euler3d_props(order=2, iterativeSolverTol=1e-10), jacobian="exact")
    dev = solver.DeviceCase(case, props)
    W = cases.density_pulse_state(case, dev.nb, dev.initial_state()[0][:: dev.nb], amp=0.05, width=0.1,
                                  pressure_amp=0.05)
    dev.set_state(W)
    dt = 2e-3
    dev.assemble(dt, -1.0, 1, True)
    _linear_checks(dev, 1.0 / dt)
    dev.close()
