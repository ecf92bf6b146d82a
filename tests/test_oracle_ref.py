"""The oracle's avg / diff / diff2 (oracle/ns_oracle.c) against the
reference's own avg_old / diff_old / diff2_old (FD_functions.f90:8-219),
built with flang from the reference's source into oracle/_ref/
(`make -C oracle ref`, run by __graft_entry__.build() where the reference is
present): bitwise equal on random arrays, every axis and code.  The tests
skip where that library is not built."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libfd_ref.so")

_lib = None


def ref_lib():
    global _lib
    if _lib is None:
        if not os.path.exists(REF_LIB):
            pytest.skip("oracle/_ref/libfd_ref.so is not built (the reference's source is not present)")
        _lib = ctypes.CDLL(REF_LIB)
    return _lib


def _call(name, A, out_shape, *ints):
    B = np.zeros(out_shape, dtype=np.float64, order="F")
    D = ctypes.POINTER(ctypes.c_double)
    rc = getattr(ref_lib(), name)(A.ctypes.data_as(D), *map(ctypes.c_int, A.shape), *map(ctypes.c_int, ints),
                                  B.ctypes.data_as(D), *map(ctypes.c_int, out_shape))
    assert rc == 0, f"{name}: the reference's result has another shape than {out_shape}"
    return B


SHAPES = [(7, 6, 5), (2, 9, 4), (12, 3, 8)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("bc", [1, 2, 4])
def test_avg_and_diff2_equal_reference_old(oracle, shape, k, bc):
    A = np.asfortranarray(np.random.default_rng(hash((shape, k, bc)) % 2 ** 32).standard_normal(shape))
    mine = oracle.avg(A, k, bc)
    assert np.array_equal(_call("fdref_avg_old", A, mine.shape, k, bc), mine)
    if shape[k - 1] > 2:
        mine = oracle.diff2(A, k, bc)
        assert np.array_equal(_call("fdref_diff2_old", A, mine.shape, k, bc), mine)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("k", [1, 2, 3])
def test_diff_equals_reference_old(oracle, shape, k):
    A = np.asfortranarray(np.random.default_rng(7 * k + shape[0]).standard_normal(shape))
    mine = oracle.diff(A, k)
    assert np.array_equal(_call("fdref_diff_old", A, mine.shape, k), mine)
