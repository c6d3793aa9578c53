"""Parity of the tetrahedron kernels (3D inviscid model, orders 1 and 2) with the tetrahedron
oracle on the jittered Kuhn cube (inlet x = 0, outlet x = 1, slip walls elsewhere).

Tolerances:
  residual, BDF right-hand side            1e-12 (equation-norm relative)
  residual norms, rhs norm                 1e-11 (relative)
  exact Jacobian (all blocks, dense)       1e-10 (max |diff| / max |entry|)
  J v, J^T v                               1e-10
  GMRES correction (tol 1e-13) vs dense    1e-8  (max |diff| / max |x|)
  min elemSize / lambda                    1e-13 (relative)
The oracle's dense Jacobian differentiates every residual that depends on a state; the device
writes the off-diagonal blocks from the neighbour-side face integrand (the reference's KR
blocks), so the two agree because the triangle face rules are symmetric under the vertex
permutations that relate the two sides of a face.
"""
import math

import numpy as np
import pytest

from flowpro_v2_amd import cases, solver
import oracle as O
from helpers import block_rel_err, eq_rel_err

pytestmark = pytest.mark.gpu


def _state(case, nb, w0, seed=3):
    W = cases.density_pulse_state(case, nb, w0, amp=0.1, width=0.3)
    return W * (1 + 0.01 * np.random.default_rng(seed).standard_normal(W.shape))


def _setup(order, n=3, precond="blockilu0", tol=1e-2, seed=3):
    case = cases.kuhn_cube(n, 0.05, 4)
    props = cases.euler3d_props(order=order, preconditioner=precond, iterativeSolverTol=tol)
    o = O.OracleTet(case, props)
    dev = solver.DeviceCase(case, props)
    assert (dev.nb, dev.neq) == (o.nb, o.neq)
    W = _state(case, dev.nb, dev.initial_state()[0][:: dev.nb], seed)
    o.set_w(W)
    dev.set_state(W)
    return case, o, dev, W


def _dense_from_blocks(o, d, n):
    """a1 M - dR/dW from the exported blocks (transposed storage: [perturbed][residual])."""
    TT, B = o.TT(), o.B
    A = np.zeros((o.dofs, o.dofs))
    for e in range(o.n_elem):
        A[e * B:(e + 1) * B, e * B:(e + 1) * B] = d[e].T
        for k in range(4):
            t = TT[e, k]
            if t > -1:
                A[e * B:(e + 1) * B, t * B:(t + 1) * B] = n[e, k].T
    return A


@pytest.mark.parametrize("order", [1, 2])
def test_tet_residual_rhs_and_norms(order):
    case, o, dev, W = _setup(order)
    Wo = W * (1 + 0.002 * np.random.default_rng(11).standard_normal(W.shape))
    dev.set_state(W, Wo)
    dt = 0.05
    norms, rhs_norm = dev.assemble(dt, -1.0, 1, False)
    R, RHS = dev.residual()
    Rref = o.residual()
    assert eq_rel_err(R, Rref, dev.nb, dev.neq) < 1e-12
    # BDF1: RHS = R - M (W - Wo) / dt with the oracle's element mass matrices
    RHSref = Rref.copy()
    for e in range(case.n_elem):
        M = o.mass(e)
        dW = ((W[e] - Wo[e]) / dt).reshape(dev.neq, dev.nb)
        RHSref[e] -= (dW @ M.T).ravel()
    assert eq_rel_err(RHS, RHSref, dev.nb, dev.neq) < 1e-12
    # Element.l2NormsSquared summed over the elements: R_m^T M^-1 R_m
    nref, href = np.zeros(dev.neq), 0.0
    for e in range(case.n_elem):
        Mi = np.linalg.inv(o.mass(e))
        r, h = Rref[e].reshape(dev.neq, dev.nb), RHSref[e].reshape(dev.neq, dev.nb)
        nref += np.einsum("mi,ij,mj->m", r, Mi, r)
        href += float(np.einsum("mi,ij,mj->", h, Mi, h))
    np.testing.assert_allclose(norms, np.sqrt(nref), rtol=1e-11)
    assert math.isclose(rhs_norm, math.sqrt(href), rel_tol=1e-11)
    dev.close()


@pytest.mark.parametrize("order", [1, 2])
def test_tet_freestream_residual_vanishes(order):
    case = cases.kuhn_cube(3, 0.05, 4)
    dev = solver.DeviceCase(case, cases.euler3d_props(order=order))
    dev.set_state(dev.initial_state())
    dev.assemble(0.05, -1.0, 1, False)
    R, _ = dev.residual()
    assert np.abs(R).max() <= 1e-12
    dev.close()


@pytest.mark.parametrize("order", [1, 2])
def test_tet_exact_jacobian_blocks(order):
    case, o, dev, W = _setup(order)
    dt = 0.05
    dev.assemble(dt, -1.0, 1, True)
    d, n = dev.jacobian()
    A = _dense_from_blocks(o, d, n)
    Aref = o.jacobian(1.0 / dt)
    assert block_rel_err(A, Aref) < 1e-10
    # boundary faces carry no block
    TT = o.TT()
    assert np.all(n[TT < 0] == 0)
    dev.close()


def test_tet_van_leer_residual_and_jacobian():
    """numericalFlux = vanLeer on the tetrahedra (the 3D split fluxes of NavierStokes.fplus /
    fminus): residual 1e-12, dense Jacobian 1e-10."""
    case = cases.kuhn_cube(3, 0.05, 4)
    props = cases.euler3d_props(order=2, numericalFlux="vanLeer")
    o = O.OracleTet(case, props)
    dev = solver.DeviceCase(case, props)
    W = _state(case, dev.nb, dev.initial_state()[0][:: dev.nb])
    o.set_w(W)
    dev.set_state(W)
    dt = 0.05
    dev.assemble(dt, -1.0, 1, True)
    R, _ = dev.residual()
    assert eq_rel_err(R, o.residual(), dev.nb, dev.neq) < 1e-12
    d, n = dev.jacobian()
    assert block_rel_err(_dense_from_blocks(o, d, n), o.jacobian(1.0 / dt)) < 1e-10
    dev.close()


@pytest.mark.parametrize("order", [1, 2])
def test_tet_jv_jtv(order):
    case, o, dev, W = _setup(order)
    a1 = 20.0
    Aref = o.jacobian(a1)
    rng = np.random.default_rng(7)
    v = rng.standard_normal((case.n_elem, dev.B))
    assert block_rel_err(dev.jv(a1, v), (Aref @ v.ravel()).reshape(v.shape)) < 1e-10
    assert block_rel_err(dev.jtv(a1, v), (Aref.T @ v.ravel()).reshape(v.shape)) < 1e-10
    dev.close()


@pytest.mark.parametrize("precond", ["blockilu0", "blockjacobi"])
def test_tet_linear_solve_matches_dense(precond):
    case, o, dev, W = _setup(2, precond=precond, tol=1e-13)
    dt = 0.05
    dev.assemble(dt, -1.0, 1, True)
    rc, it, err = dev.linear_solve()
    assert rc == 0
    x = dev.solution()
    # DOF numbering of the device is the element order (RCM): map to mesh order
    order = dev.element_order()
    xm = np.zeros((case.n_elem, dev.B))
    xm[order] = x.reshape(case.n_elem, dev.B)
    _, RHS = dev.residual()
    xd = np.linalg.solve(o.jacobian(1.0 / dt), RHS.ravel()).reshape(xm.shape)
    assert np.abs(xm - xd).max() / np.abs(xd).max() < 1e-8
    dev.close()


@pytest.mark.parametrize("bits", [64, 32])
def test_tet_row_factor_kernel_matches_level_kernel(bits, monkeypatch):
    """The register-resident row factor kernel at B = 20 with five blocks per row (P1 tetrahedra)
    against the one-wave level kernel, which the triangle cases pin against the oracle's
    BlockILU0: the first GMRES(1) direction M^-1 b agrees (1e-12 FP64, 1e-6 FP32)."""
    xs = []
    for rows in ("1", "0"):
        monkeypatch.setenv("FLOWPRO_ILU_ROWS", rows)
        case = cases.kuhn_cube(3, 0.05, 4)
        props = cases.euler3d_props(order=2, gmresRestart=1, gmresCycles=1, iterativeSolverTol=1e-14)
        dev = solver.DeviceCase(case, props, precond_bits=bits)
        dev.set_state(_state(case, dev.nb, dev.initial_state()[0][:: dev.nb]))
        dev.assemble(0.05, -1.0, 1, True)
        rc, it, err = dev.linear_solve()
        assert rc in (0, 1) and it == 1
        xs.append(dev.solution())
        dev.close()
    tol = 1e-12 if bits == 64 else 1e-6
    assert np.abs(xs[0] - xs[1]).max() <= tol * np.abs(xs[1]).max()


@pytest.mark.parametrize("n", [3, 8])
def test_tet_fp32_preconditioner_same_solution(n):
    """FP32 block-ILU0 factors and sweeps (the factorisation starting from the FP32 copy the
    Jacobian kernel writes) solve the same FP64 system: at a tight GMRES tolerance the correction
    agrees with the FP64-preconditioned one to the FP32 rounding of the preconditioned operator
    (1e-6)."""
    xs = []
    for bits in (64, 32):
        case = cases.kuhn_cube(n, 0.05, 4)
        props = cases.euler3d_props(order=2, iterativeSolverTol=1e-12)
        dev = solver.DeviceCase(case, props, precond_bits=bits)
        dev.set_state(_state(case, dev.nb, dev.initial_state()[0][:: dev.nb]))
        dev.assemble(0.05, -1.0, 1, True)
        rc, it, err = dev.linear_solve()
        assert rc == 0
        xs.append(dev.solution())
        dev.close()
    assert np.abs(xs[0] - xs[1]).max() <= 1e-6 * np.abs(xs[0]).max()


def test_tet_fp32_preconditioner_same_history():
    """Newton steps with FP32 and with FP64 preconditioner factors reach the same states and the
    same normalised residual history.  The state is a density and pressure pulse, so that every
    equation's residual is of the order of the pulse: the history is the largest per-equation
    ratio to its first-steps maximum, and an equation whose residual sits at round-off (the
    cross-stream momenta of a pure density pulse) turns the FP32 rounding of the GMRES
    correction into an O(1) change of that ratio.  States 1e-7 relative, history 1e-5."""
    case = cases.kuhn_cube(8, 0.05, 4)
    props = cases.euler3d_props(mach=0.8, order=2, CFL=10.0)
    hist, Ws = [], []
    for bits in (64, 32):
        dev = solver.DeviceCase(case, props, precond_bits=bits)
        dev.set_state(cases.density_pulse_state(case, dev.nb, dev.initial_state()[0][:: dev.nb], amp=0.05,
                                                width=0.1, pressure_amp=0.05))
        s = solver.GpuImplicitSolver(dev, props)
        hist.append(np.array([h.residuum for h in s.solve(8)]))
        Ws.append(dev.get_state())
        dev.close()
    print("histories", hist)
    assert hist[0][-1] < 0.9 * hist[0][0]
    assert np.abs(Ws[1] - Ws[0]).max() <= 1e-7 * np.abs(Ws[0]).max()
    assert np.abs(hist[1] / hist[0] - 1).max() < 1e-5


@pytest.mark.parametrize("order", [1, 2])
def test_tet_size_over_lambda(order):
    case, o, dev, W = _setup(order)
    assert math.isclose(dev.size_over_lambda(), o.size_over_lambda(), rel_tol=1e-13)
    dev.close()


def test_tet_positivity_limiter_properties():
    """Element.positivityLimiter: after the update every volume and face point of a limited
    element is admissible and the element mean is unchanged."""
    case = cases.kuhn_cube(3, 0.05, 4)
    props = cases.euler3d_props(order=2)
    dev = solver.DeviceCase(case, props)
    W = _state(case, 4, dev.initial_state()[0][::4])
    # drive the vertex densities of a few elements negative (mean stays positive)
    Wb = W.reshape(case.n_elem, 5, 4).copy()
    bad = np.arange(0, case.n_elem, 17)
    Wb[bad, 0, 0] = -0.3 * Wb[bad, 0, 1]
    Wb = Wb.reshape(W.shape)
    dev.set_state(Wb)
    # a tiny time step makes the Newton correction x ~ dt M^-1 R small, so the planted
    # negative vertex densities survive W + x and the limiter acts on those elements
    dev.assemble(1e-6, -1.0, 1, True)
    dev.linear_solve()
    x = dev.solution()
    n_lim = dev.update()
    Wn = dev.get_state().reshape(case.n_elem, 5, 4)
    assert n_lim >= len(bad)
    # barycentric points of the volume and face rules: the P1 basis values are the barycentric
    # coordinates, so the point values are convex combinations of the vertex values
    from flowpro_v2_amd import quadrature
    vx, _ = quadrature.tetra_rule(2)
    fx, _ = quadrature.triangle_rule(2)
    faces = [(0, 2, 1), (0, 1, 3), (1, 2, 3), (2, 0, 3)]
    XI = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    pts = [p for p in vx]
    for f in faces:
        for u, v in fx:
            pts.append(XI[f[0]] + u * (XI[f[1]] - XI[f[0]]) + v * (XI[f[2]] - XI[f[0]]))
    pts = np.array(pts)
    phi = np.stack([1 - pts.sum(1), pts[:, 0], pts[:, 1], pts[:, 2]], 1)   # (npts, 4)
    Wx = np.einsum("pj,emj->epm", phi, Wn)
    rho, mom, E = Wx[..., 0], Wx[..., 1:4], Wx[..., 4]
    ok = (rho > 1e-10) & (E - (mom ** 2).sum(-1) / (2 * rho) > 1e-10)
    assert ok.all()
    # the update is W + x then the limiter; the element means (equal weights 1/4 for P1) of
    # W + x are kept
    order = dev.element_order()
    xm = np.zeros((case.n_elem, 20))
    xm[order] = x.reshape(case.n_elem, 20)
    Wu = (Wb + xm).reshape(case.n_elem, 5, 4)
    assert np.abs(Wn.mean(-1) - Wu.mean(-1)).max() <= 1e-11 * np.abs(Wu).max()
    dev.close()


def test_tet_newton_steps_converge():
    """The device Newton loop (BDF1, blockilu0 GMRES) on the density pulse: the residual of the
    pulse decays and every step is accepted."""
    case = cases.kuhn_cube(4, 0.05, 4)
    props = cases.euler3d_props(order=2, CFL=5.0)
    dev = solver.DeviceCase(case, props)
    W = _state(case, 4, dev.initial_state()[0][::4])
    dev.set_state(W)
    s = solver.GpuImplicitSolver(dev, props)
    hist = s.solve(8)
    res = [h.residuum for h in hist]
    assert all(np.isfinite(res))
    assert res[-1] < res[0]
    dev.close()
