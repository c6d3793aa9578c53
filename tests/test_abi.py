"""The C ABI: the library loads, exports every declared entry point, and its
integer setup (neighbour matrix, RCM numbering) is bit-exact with the oracle.
No device memory is touched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from flowpro_v2_amd import cases, native
import oracle as O

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "flowpro_hip.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fp_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_the_binding():
    assert declared() == sorted(native.EXPORTS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


@pytest.mark.parametrize("mk", [lambda: cases.square_mesh(5, 1.0, -2), lambda: cases.channel_bump(64, 16),
                                lambda: cases.cylinder_mesh(40, 12)])
def test_integer_setup_bit_exact(mk):
    c = mk()
    o = O.Oracle(c, cases.euler_props(order=1))
    TT = native.neighbor_matrix(c.TP, c.bnd)
    assert (TT == o.TT()).all()
    assert (native.rcm_order(TT) == o.order()).all()


def test_bad_arguments_are_rejected():
    assert native.lib().fp_create(None, None) == native.FP_EINVAL
    assert native.lib().fp_destroy(None) == native.FP_EINVAL
