"""The GMRES of the device path (bookkeeping on the device, Arnoldi steps enqueued in chunks, optionally
replayed from HIP graphs: FLOWPRO_GMRES_GRAPH=1) against the same solver with one control-block read per step
and no graphs.

The chunk graphs are replayed many times within one run, with the host's own reads and other launches
between the replays; every Newton step's correction, GMRES iteration count and residual history must be
bit-identical to the eager run's (same kernels, same order), for each preconditioner, and for restarted
cycles (restart 5) as well as single cycles.
"""
import numpy as np
import pytest

from flowpro_v2_amd import cases, solver

pytestmark = pytest.mark.gpu


def _run(monkeypatch, graph, chunk, props, steps, bits=64):
    monkeypatch.setenv("FLOWPRO_GMRES_GRAPH", "1" if graph else "0")
    monkeypatch.setenv("FLOWPRO_GMRES_CHUNK", str(chunk))
    case = cases.cylinder_mesh(16, 6)
    dev = solver.DeviceCase(case, props, precond_bits=bits)
    dev.set_state(dev.initial_state())
    s = solver.GpuImplicitSolver(dev, props)
    recs = s.solve(steps)
    out = (np.array([r.residuum for r in recs]), [r.gmres_iters for r in recs], dev.get_state(), dev.solution())
    dev.close()
    return out


@pytest.mark.parametrize("prec", ["blockilu0", "blockjacobi", "none"])
@pytest.mark.parametrize("restart", [30, 5])
def test_graph_chunks_bit_identical_to_eager(monkeypatch, prec, restart):
    props = cases.ns_props(order=3, reynolds=100.0, CFL=5.0, iterativeSolverTol=1e-6, timeOrder=2,
                           preconditioner=prec, gmresRestart=restart)
    print(f"[gmres] {prec} restart {restart}", flush=True)
    ref = _run(monkeypatch, False, 1, props, 6)
    for chunk in (4, 3):
        got = _run(monkeypatch, True, chunk, props, 6)
        assert np.array_equal(got[0], ref[0])
        assert got[1] == ref[1]
        assert np.array_equal(got[2], ref[2])
        assert np.array_equal(got[3], ref[3])
    # more than one cycle was exercised with the short restart (without a preconditioner the solves run
    # out of cycles: the steps are repeated at a smaller CFL and record no iteration count)
    its = [it for step in ref[1] for it in step]
    if restart == 5 and prec != "none":
        assert max(its) > 5


def test_graph_chunks_fp32_factors(monkeypatch):
    props = cases.ns_props(order=3, reynolds=100.0, CFL=5.0, iterativeSolverTol=1e-6, timeOrder=2)
    ref = _run(monkeypatch, False, 1, props, 5, bits=32)
    got = _run(monkeypatch, True, 4, props, 5, bits=32)
    assert np.array_equal(got[0], ref[0]) and got[1] == ref[1]
    assert np.array_equal(got[2], ref[2])


def _run_env(monkeypatch, env, props, steps, bits=64, case=None):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    case = case or cases.cylinder_mesh(16, 6)
    dev = solver.DeviceCase(case, props, precond_bits=bits)
    dev.set_state(dev.initial_state())
    recs = solver.GpuImplicitSolver(dev, props).solve(steps)
    out = (np.array([r.residuum for r in recs]), [r.gmres_iters for r in recs], dev.get_state(), dev.solution())
    dev.close()
    return out


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("restart", [30, 5])
def test_spmv_folded_into_forward_sweep_bit_identical(monkeypatch, bits, restart):
    """The Arnoldi step's A v formed row by row inside the forward dataflow sweep of M^-1
    (FLOWPRO_SPMV_SWEEP=1) against the product launch followed by the sweep (the default): the same fused
    multiply-adds in the same order, so every residual, iteration count, correction and state is
    bit-identical, for FP64 and FP32 factors."""
    props = cases.ns_props(order=3, reynolds=100.0, CFL=5.0, iterativeSolverTol=1e-6, timeOrder=2, gmresRestart=restart)
    print(f"[gmres] folded product FP{bits} restart {restart}", flush=True)
    ref = _run_env(monkeypatch, {"FLOWPRO_SPMV_SWEEP": "0"}, props, 5, bits)
    got = _run_env(monkeypatch, {"FLOWPRO_SPMV_SWEEP": "1"}, props, 5, bits)
    assert np.array_equal(got[0], ref[0]) and got[1] == ref[1]
    assert np.array_equal(got[2], ref[2])
    assert np.array_equal(got[3], ref[3])


@pytest.mark.parametrize("prec", ["blockilu0", "none"])
def test_lds_triangular_solve_bit_identical(monkeypatch, prec):
    """The GMRES triangular solve with its working copy in LDS and the rows spread over a wavefront (the
    default) against the one-thread form on global memory (FLOWPRO_GMRES_LSOLVE_ONE_THREAD=1): the same
    operations per entry in the same order, so every history, correction and state is bit-identical,
    over restarted (restart 5) cycles."""
    props = cases.ns_props(order=3, reynolds=100.0, CFL=5.0, iterativeSolverTol=1e-6, timeOrder=2, gmresRestart=5,
                           preconditioner=prec)
    ref = _run_env(monkeypatch, {"FLOWPRO_GMRES_LSOLVE_ONE_THREAD": "1"}, props, 5)
    got = _run_env(monkeypatch, {"FLOWPRO_GMRES_LSOLVE_ONE_THREAD": "0"}, props, 5)
    assert np.array_equal(got[0], ref[0]) and got[1] == ref[1]
    assert np.array_equal(got[2], ref[2])
    assert np.array_equal(got[3], ref[3])
