"""fp_time_step_noalloc (the allocation-free variant of fp_time_step, its scratch in the
library's static device array) against fp_time_step through the C ABI: bit-identical minima
(the same kernels and reduction order), on two calls in a row, the second on a smaller case
that reuses the array the first one left behind; and the size limit is rejected cleanly."""
import ctypes as C

import pytest

from flowpro_v2_amd import cases, native, solver

pytestmark = pytest.mark.gpu


def _both(dev):
    L = native.lib()
    a, b = C.c_double(), C.c_double()
    assert L.fp_time_step(dev.h, C.byref(a)) == native.FP_OK
    assert L.fp_time_step_noalloc(dev.h, C.byref(b)) == native.FP_OK
    return a.value, b.value


@pytest.mark.parametrize("first,second", [
    (lambda: (cases.cylinder_mesh(64, 20), cases.ns_props(order=3)), lambda: (cases.channel_bump(12, 4), cases.euler_props(order=2))),
    (lambda: (cases.channel_bump(24, 8, quads=True, jitter=0.2), cases.euler_props(order=3)),
     lambda: (cases.kuhn_cube(2), cases.euler3d_props(order=2))),
], ids=["tri_then_small_tri", "quad_then_small_tet"])
def test_noalloc_time_step_matches(first, second):
    for mk, seed in ((first, 3), (second, 5)):
        case, props = mk()
        dev = solver.DeviceCase(case, props)
        W = dev.initial_state()
        if dev.neq > 1 and not dev.mixed:
            W = cases.perturbed_state(W[0][:: dev.nb], case.n_elem, dev.nb, 0.05, seed)
        dev.set_state(W)
        a, b = _both(dev)
        assert a == b and a > 0
        dev.close()


def test_noalloc_time_step_twice_same_context():
    case = cases.channel_bump(16, 4)
    dev = solver.DeviceCase(case, cases.euler_props(order=2))
    dev.set_state(cases.perturbed_state(dev.initial_state()[0][:: dev.nb], case.n_elem, dev.nb, 0.05, 1))
    r1 = _both(dev)
    r2 = _both(dev)
    assert r1[0] == r1[1] == r2[0] == r2[1]
    dev.close()
