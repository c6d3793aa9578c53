"""The mixed-precision path converged to a steady state: the cylinder at Re = 40 (steady flow), NS
p=2, CFL varied by the residual up to 1e3, solved until the relative residual (State.residuum) is
below 1e-8 with the FP32 and with the FP64 preconditioner; both converged states lie within 1e-6
(max |diff| / max |W|) of the oracle's converged state (FP64 throughout, the same criterion).  The
FP32 factors only change the linear solves' inner accuracy, which the FP64 Newton residual absorbs."""
import numpy as np
import pytest

from flowpro_v2_amd import cases, solver
import oracle as O

pytestmark = pytest.mark.gpu


def test_mixed_precision_converged_state_matches_fp64_oracle():
    case = cases.cylinder_mesh(24, 6)
    props = dict(cases.ns_props(order=3, reynolds=40.0, CFL=-1.0, maxCFL=1e3, residuum=1e-8), jacobian="exact")
    o = O.Oracle(case, props, threads=8)
    hist = o.run(400)
    assert hist[-1] <= 1e-8
    Wref = np.asarray(o.get_w())
    for bits in (32, 64):
        dev = solver.DeviceCase(case, props, precond_bits=bits)
        dev.set_state(dev.initial_state())
        recs = solver.GpuImplicitSolver(dev, props).solve(400)
        assert recs[-1].residuum <= 1e-8
        W = dev.get_state()
        assert np.abs(W - Wref).max() / np.abs(Wref).max() < 1e-6, bits
        dev.close()
