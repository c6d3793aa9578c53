"""A device run written to a simulation directory in the Java tool chain's formats and continued
from it (continueComputation), plus the element means of W.txt against the oracle's mass matrix.

  W.txt means vs sum_j W_j Is_j with Is = M 1 / area from the oracle  1e-14 (relative)
  continuing from We.txt/state.txt reproduces the uninterrupted run    1e-12 (residual history)
"""
import os

import numpy as np
import pytest

from flowpro_v2_amd import cases, results_io, solver
import oracle as O

pytestmark = pytest.mark.gpu


def test_save_and_continue(tmp_path):
    case = cases.channel_bump(16, 4)
    props = cases.gamm_props(order=3, saveRate=6, iterativeSolverTol=1e-10)
    # uninterrupted: 9 steps
    dev = solver.DeviceCase(case, props)
    W0 = cases.perturbed_state(dev.initial_state()[0][:: dev.nb], case.n_elem, dev.nb, 0.05, 2)
    dev.set_state(W0)
    ref = [h.residuum for h in solver.GpuImplicitSolver(dev, props).solve(9)]
    dev.close()
    # 6 steps (saved at step 6, after the 5 normalisation steps State.save needs), then a new
    # context continues from the files
    path = str(tmp_path)
    dev = solver.DeviceCase(case, props)
    dev.set_state(W0)
    s1 = solver.GpuImplicitSolver(dev, props, simulation_path=path)
    first = [h.residuum for h in s1.solve(6)]
    for f in ("state.txt", "W.txt", "We.txt", "artificialViscosity.txt", "referenceValues.txt", "residuum.txt"):
        assert os.path.exists(os.path.join(path, f)), f
    lines = open(os.path.join(path, "residuum.txt")).read().splitlines()
    assert len(lines) == 6 and len(lines[0].split()) == 3 + dev.neq
    # W.txt holds the element means
    o = O.Oracle(case, props)
    means = np.array(results_io.load_matrix(os.path.join(path, "W.txt")))
    We = np.array(results_io.load_matrix(os.path.join(path, "We.txt")))
    ref_means = np.array([We[e].reshape(dev.neq, dev.nb) @ (o.geometry(e)["M"] @ np.ones(dev.nb) / o.geometry(e)["area"])
                          for e in range(case.n_elem)])
    assert np.abs(means - ref_means).max() <= 1e-14 * np.abs(ref_means).max()
    dev.close()
    dev2 = solver.DeviceCase(case, props)
    s2 = solver.GpuImplicitSolver(dev2, props, simulation_path=path)
    s2.continue_from()
    assert s2.state.steps == 6
    # the continued run restarts BDF from a single level, as the reference does (dto unknown); with
    # BDF1 and the CFL kept, the next steps match the uninterrupted run
    rest = [h.residuum for h in s2.solve(3)]
    np.testing.assert_allclose(first + rest, ref, rtol=1e-12)
    dev2.close()


def test_save_and_continue_mixed_orders(tmp_path):
    """A mesh of squares and triangles with per-element orders: We.txt holds one row per element of
    its own length (nEqs nBasis_e), and continuing from the files reproduces the uninterrupted run."""
    case = cases.mixed_channel(8, 4, jitter=0.2)
    case.orders = cases.seeded_orders(case.n_elem, (1, 2, 3), 6)
    props = cases.euler_props(order=3, CFL=5.0, saveRate=6, iterativeSolverTol=1e-10)
    dev = solver.DeviceCase(case, props)
    W0 = dev.initial_state() * (1 + 0.01 * np.random.default_rng(2).standard_normal(dev.dofs))
    dev.set_state(W0)
    ref = [h.residuum for h in solver.GpuImplicitSolver(dev, props).solve(9)]
    dev.close()
    path = str(tmp_path)
    dev = solver.DeviceCase(case, props)
    dev.set_state(W0)
    first = [h.residuum for h in solver.GpuImplicitSolver(dev, props, simulation_path=path).solve(6)]
    rows = results_io.load_matrix(os.path.join(path, "We.txt"))
    assert [len(r) for r in rows] == list(dev.sizes)
    assert np.array_equal(np.concatenate(rows), dev.get_state())
    dev.close()
    dev2 = solver.DeviceCase(case, props)
    s2 = solver.GpuImplicitSolver(dev2, props, simulation_path=path)
    s2.continue_from()
    rest = [h.residuum for h in s2.solve(3)]
    np.testing.assert_allclose(first + rest, ref, rtol=1e-12)
    dev2.close()
