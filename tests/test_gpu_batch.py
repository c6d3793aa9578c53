"""Independent cases solved concurrently (flowpro_v2_amd.batch): every instance's residual history
and final state equal those of the same instance solved alone, bit for bit (each context runs the
same arithmetic on its own stream; concurrency only interleaves the instances' kernels)."""
import numpy as np
import pytest

from flowpro_v2_amd import batch, cases, solver

pytestmark = pytest.mark.gpu


def _instances():
    out = batch.seeded_ns_instances(n=3, seed=4, min_elems=600, max_elems=1500, orders=(2, 3, 4))
    out.append((cases.channel_bump(12, 4, quads=True, jitter=0.2), cases.euler_props(order=3, CFL=5.0)))
    out += batch.seeded_ns_instances(n=1, seed=5, min_elems=600, max_elems=1200, orders=(2, 3, 4), per_element_orders=True)
    return out


@pytest.mark.parametrize("bits", [64, 32])
def test_batched_equals_alone(bits):
    inst = _instances()
    b = batch.BatchedSolver(inst, precond_bits=bits)
    hists = b.solve(4)
    states = [it.dev.get_state() for it in b.instances]
    lifts = b.functionals(1)
    b.close()
    for i, (case, props) in enumerate(inst):
        dev = solver.DeviceCase(case, props, precond_bits=bits)
        dev.set_state(dev.initial_state())
        h = solver.GpuImplicitSolver(dev, props).solve(4)
        assert [r.residuum for r in h] == [r.residuum for r in hists[i]]
        assert [r.gmres_iters for r in h] == [r.gmres_iters for r in hists[i]]
        assert np.array_equal(dev.get_state(), states[i])
        if i != 3:   # the quadrilateral instance has no force coefficients
            val, g = dev.functional(1, gradient=True)
            assert val == lifts[i][0] and np.array_equal(g, lifts[i][1])
        dev.close()
