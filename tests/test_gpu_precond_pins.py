"""Preconditioner pins on the device against the oracle's sequential preconditioners.

GMRES(1) with one cycle from x0 = 0 returns x = y1 v1 with v1 = M^-1 b / |M^-1 b|, so x pins
the preconditioner application M^-1 b itself.  Covered here:
  * blockJacobi (blockJacobi.java:74-90 with Mat.lsolveT, API/Mat.java:1307-1329): the first
    direction, and restarted solves whose later applies see the diagonal blocks eliminated in
    place by the first; the iteration count must be the oracle's (a kernel that refactored a
    clean copy on every apply converges differently);
  * block-ILU0 at B = 36 and 64 (quadrilaterals of order 3 and 4) and at B = 64 on meshes mixing
    triangles and squares (padded to 16 functions), with both factor kernels and FP64 / FP32
    factors;
  * the Spalart-Allmaras plate system (B = 30) with FP32 factors.
Tolerances (max |x - x_ref| / max |x_ref|): FP64 factors 1e-10; FP32 factors as stated per case,
the FP32 rounding of the factors (u = 6e-8) amplified by the conditioning of the blocks.
"""
import numpy as np
import pytest

from flowpro_v2_amd import cases, solver
import oracle as O
from test_gpu_mixed import _mixed, _state

pytestmark = pytest.mark.gpu


def _pair(case, props, bits=64, mixed=False, seed=3):
    props = dict(props, jacobian="exact")
    o = O.Oracle(case, props)
    dev = solver.DeviceCase(case, props, precond_bits=bits)
    if mixed:
        W = _state(o, seed)
    else:
        W = cases.perturbed_state(o.get_w()[0][:: o.nb], case.n_elem, o.nb, 0.01, seed)
    o.set_w(W)
    o.limiter()
    dev.set_state(W)
    o.assemble(0.05, -1.0)
    dev.assemble(0.05, -1.0, 1, True)
    return o, dev


def _rel(x, xref):
    return float(np.abs(x - xref).max() / np.abs(xref).max())


# ---------------------------------------------------------------- blockJacobi
def test_block_jacobi_first_direction():
    case = cases.channel_bump(16, 4)
    o, dev = _pair(case, cases.euler_props(order=3, preconditioner="blockjacobi", gmresRestart=1, gmresCycles=1,
                                           iterativeSolverTol=1e-14))
    _, xref, it_ref, _ = o.linear_solve()
    rc, it, _ = dev.linear_solve()
    assert it == it_ref == 1 and rc in (0, 1)
    assert _rel(dev.solution(), xref) < 1e-10
    dev.close()


@pytest.mark.parametrize("restart,cycles,tol", [(3, 5, 1e-3), (30, 5, 1e-2), (2, 5, 1e-4)])
def test_block_jacobi_restarted_solve_iterations(restart, cycles, tol):
    """Every apply after the first is the U^-1-only back substitution of the blocks the first
    apply eliminated in place; the oracle's iteration count and correction are reproduced."""
    case = cases.channel_bump(16, 4)
    o, dev = _pair(case, cases.ns_props(order=2, reynolds=200.0, preconditioner="blockjacobi", gmresRestart=restart,
                                        gmresCycles=cycles, iterativeSolverTol=tol), seed=5)
    ok, xref, it_ref, err_ref = o.linear_solve()
    rc, it, err = dev.linear_solve()
    assert it == it_ref and it > 1
    assert (rc == 0) == bool(ok)
    assert abs(err - err_ref) <= 1e-8 * err_ref
    assert _rel(dev.solution(), xref) < 1e-8
    dev.close()


# ---------------------------------------------------------------- block-ILU0 at B = 36, 64
PIN_CASES = [
    ("quad_B36", lambda: cases.channel_bump(12, 4, quads=True, jitter=0.2), lambda **kw: cases.euler_props(order=3, **kw),
     False, 1e-5),
    ("quad_B64", lambda: cases.channel_bump(8, 3, quads=True, jitter=0.2), lambda **kw: cases.euler_props(order=4, **kw),
     False, 1e-5),
    ("mixed_B64", lambda: _mixed(8, 4), lambda **kw: cases.euler_props(order=4, **kw), True, 1e-5),
    ("mixed_ns_B64", lambda: _mixed(6, 3, orders=(2, 3, 4), seed=8),
     lambda **kw: cases.ns_props(order=4, reynolds=300.0, **kw), True, 1e-5),
]


@pytest.mark.parametrize("rows_kernel", ["1", "0"], ids=["row_kernel", "level_kernel"])
@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("name,mk_case,mk_props,mixed,tol32", PIN_CASES, ids=[c[0] for c in PIN_CASES])
def test_block_ilu0_first_direction_large_blocks(name, mk_case, mk_props, mixed, tol32, bits, rows_kernel, monkeypatch):
    monkeypatch.setenv("FLOWPRO_ILU_ROWS", rows_kernel)
    case = mk_case()
    o, dev = _pair(case, mk_props(gmresRestart=1, gmresCycles=1, iterativeSolverTol=1e-14), bits=bits, mixed=mixed)
    _, xref, it_ref, _ = o.linear_solve()
    rc, it, _ = dev.linear_solve()
    assert it == it_ref == 1 and rc in (0, 1)
    err = _rel(dev.solution(), xref)
    print(f"{name} bits={bits} rows={rows_kernel}: {err:.3e}")
    assert err < (1e-10 if bits == 64 else tol32)
    dev.close()


# ---------------------------------------------------------------- Spalart-Allmaras, FP32 factors
def test_sa_fp32_first_direction():
    """SA p2 on the plate channel (B = 30, cond(A) ~ 1e8, diagonal blocks up to cond ~5e3).
    Tolerance 2e-2: every FP32 block inverse carries u32 cond(D) ~ 3e-4, compounded along the
    chains of the factorisation (measured 6.3e-3 with the reference's scaled pivoting)."""
    case = cases.plate_channel(12, 5)
    o, dev = _pair(case, cases.sa_props(order=3, reynolds=1e5, gmresRestart=1, gmresCycles=1, iterativeSolverTol=1e-14),
                   bits=32)
    _, xref, it_ref, _ = o.linear_solve()
    rc, it, _ = dev.linear_solve()
    err = _rel(dev.solution(), xref)
    print(f"sa fp32 first direction: {err:.3e}")
    assert it == it_ref == 1 and err < 2e-2
    dev.close()


def test_sa_fp32_newton_history_against_fp64_oracle():
    """SA p2, BDF2, 6 steps from a perturbed state with FP32 block-ILU0 factors and FP64
    residual, Jacobian and GMRES (tol 1e-10), against the all-FP64 oracle.  With left
    preconditioning GMRES minimises |M^-1 (b - A x)|, and M^-1 applied in FP32 bounds the
    attainable accuracy of each correction near u32 cond(A) in the ill-conditioned directions
    (cond(A) ~ 1e8 here), so the states drift apart by O(1e-5) per step: tolerances 1e-4 on the
    residual history and 5e-4 on the state (measured 1.0e-5 and 6.6e-5)."""
    case = cases.plate_channel(12, 5)
    props = dict(cases.sa_props(order=3, reynolds=1e5, CFL=2.0, iterativeSolverTol=1e-10, timeOrder=2),
                 jacobian="exact")
    o = O.Oracle(case, props)
    W0 = cases.perturbed_state(o.get_w()[0][:: o.nb], case.n_elem, o.nb, 0.01, 21)
    o.set_w(W0)
    hist_ref = o.run(6)
    dev = solver.DeviceCase(case, props, precond_bits=32)
    dev.set_state(W0)
    recs = solver.GpuImplicitSolver(dev, props).solve(6)
    hist = np.array([r.residuum for r in recs])
    herr = float(np.abs(hist - hist_ref).max() / np.abs(hist_ref).max())
    Wg, Wr = dev.get_state(), o.get_w()
    werr = float(np.abs(Wg - Wr).max() / np.abs(Wr).max())
    print(f"sa fp32 history {herr:.3e} state {werr:.3e} gmres {[r.gmres_iters for r in recs]}")
    assert len(recs) == 6 and herr < 1e-4 and werr < 5e-4
    dev.close()
