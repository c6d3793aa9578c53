"""The launch-shape autotuner of the block-ILU0 factor kernel (fp_autotune_report): every shape is
timed on the case's own matrix at the first factorisation, the fastest is kept, and the choice does
not change a bit of the result (the shapes compute every entry with the same operations in the
same order): the GMRES correction with tuning equals the one with the default shape
(FLOWPRO_AUTOTUNE=0) exactly, in FP64 and FP32 factors."""
import os

import numpy as np
import pytest

from flowpro_v2_amd import cases, solver

pytestmark = pytest.mark.gpu


def _solve(case, props, bits, tune):
    old = os.environ.get("FLOWPRO_AUTOTUNE")
    os.environ["FLOWPRO_AUTOTUNE"] = "1" if tune else "0"
    try:
        dev = solver.DeviceCase(case, props, precond_bits=bits)
        assert dev.autotune_report()[0] == -1          # nothing tuned before the first factorisation
        W = cases.perturbed_state(dev.initial_state()[0][:: dev.nb], case.n_elem, dev.nb, 0.01, 3)
        dev.set_state(W)
        dev.assemble(0.05, -1.0, 1, True)
        rc, it, err = dev.linear_solve()
        x = dev.solution()
        rep = dev.autotune_report()
        dev.assemble(0.05, -1.0, 1, True)              # a second factorisation keeps the choice
        dev.linear_solve()
        assert dev.autotune_report()[0] == rep[0]
        dev.close()
        return rc, it, x, rep
    finally:
        if old is None:
            os.environ.pop("FLOWPRO_AUTOTUNE", None)
        else:
            os.environ["FLOWPRO_AUTOTUNE"] = old


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("name,mk_case,mk_props", [
    ("ns_p3_tri_B40", lambda: cases.cylinder_mesh(64, 20), lambda: cases.ns_props(order=4, reynolds=500.0)),
    ("euler_p1_tet_B20", lambda: cases.kuhn_cube(4, jitter=0.05, seed=1), lambda: cases.euler3d_props(order=2)),
], ids=["ns_p3_tri_B40", "euler_p1_tet_B20"])
def test_autotuned_factor_is_bit_identical_and_fastest(name, mk_case, mk_props, bits):
    case, props = mk_case(), dict(mk_props(), iterativeSolverTol=1e-10)
    rc1, it1, x1, (v1, shapes, ms1) = _solve(case, props, bits, True)
    rc0, it0, x0, (v0, _, ms0) = _solve(case, props, bits, False)
    assert rc1 == rc0 == 0 and it1 == it0
    assert np.array_equal(x1, x0)
    assert len(shapes) == len(ms1) >= 4 and all(nw in (4, 8, 16) for nw, _ in shapes)
    assert all(t > 0 for t in ms1) and v1 == int(np.argmin(ms1))
    assert all(t == 0 for t in ms0) and v0 == (0 if x0.size // case.n_elem == 20 else 1)


@pytest.mark.parametrize("bits", [64, 32])
def test_every_launch_shape_gives_the_same_factor(monkeypatch, bits):
    """Each launch shape of the factor kernel forced in turn (FLOWPRO_ILU_VARIANT): the GMRES correction,
    and so the factor and both sweeps, is bit-identical to the default shape's, for every shape the
    autotuner may pick (4, 8 and 16 task waves; redundant and split inversions)."""
    case = cases.cylinder_mesh(32, 10)
    props = dict(cases.ns_props(order=4, reynolds=500.0), iterativeSolverTol=1e-10)
    monkeypatch.setenv("FLOWPRO_AUTOTUNE", "0")
    ref = _solve(case, props, bits, False)
    nv = len(ref[3][1])
    assert nv >= 9
    for v in range(nv):
        monkeypatch.setenv("FLOWPRO_ILU_VARIANT", str(v))
        dev = solver.DeviceCase(case, props, precond_bits=bits)
        dev.set_state(cases.perturbed_state(dev.initial_state()[0][:: dev.nb], case.n_elem, dev.nb, 0.01, 3))
        dev.assemble(0.05, -1.0, 1, True)
        rc, it, _ = dev.linear_solve()
        x = dev.solution()
        assert dev.autotune_report()[0] == v
        dev.close()
        assert rc == ref[0] and it == ref[1]
        assert np.array_equal(x, ref[2]), v
    monkeypatch.delenv("FLOWPRO_ILU_VARIANT")
