"""The C-ABI library: it loads, exports every function include/libcc_mi355x.h declares,
its host-only legacy entries (trant3 family, tranmd_23, TRANMD_SMP) match the reference
tables, and
its argument / state checks return the documented status codes.  No GPU needed."""

import ctypes
import re

import numpy as np
import pytest

import libcc_amd
from libcc_amd import lib
from oracle import port


def declared_functions():
    text = open(libcc_amd.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)
    skip = {"if", "sizeof", "defined"}
    return sorted(set(n for n in names if n not in skip))


def test_header_declares_reference_names():
    names = declared_functions()
    for n in ("cc_triples_acc_", "cc_triples_", "common_cc_init_", "ijk_tuple_", "trant3_1_", "trant3_3_",
              "trant3_4_", "trant3_5_", "trant3_acc_", "trant3_smp_", "tranmd_23_"):
        assert n in names


def test_library_exports_every_declared_function():
    L = lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def _i64(v):
    return ctypes.byref(ctypes.c_int64(v))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("name,op", [("trant3_1_", 1), ("trant3_3_", 3), ("trant3_4_", 4), ("trant3_5_", 5)])
def test_trant3_named(golden, name, op):
    for n in (2, 3):
        c = np.asfortranarray(np.arange(1, n ** 3 + 1, dtype=float).reshape(n, n, n, order="F"))
        getattr(lib(), name)(_i64(n), _dp(c))
        assert lib().ccx_last_status() == 0
        idx = {1: 0, 3: 1, 4: 2, 5: 3}[op]
        assert np.array_equal(c.ravel(order="F"), golden[f"cube_{n}"][idx])


@pytest.mark.parametrize("fn", ["trant3_acc_", "trant3_smp_"])
def test_trant3_by_id(golden, fn):
    for n in (2, 3):
        for op in range(1, 6):
            c = np.asfortranarray(np.arange(1, n ** 3 + 1, dtype=float).reshape(n, n, n, order="F"))
            getattr(lib(), fn)(_dp(c), _i64(n), _i64(op))
            assert np.array_equal(c.ravel(order="F"), golden[f"cube_{n}"][3 + op])


def test_tranmd_23(golden):
    n = 3
    c = np.asfortranarray(np.arange(1, n ** 3 + 1, dtype=float).reshape(n, n, n, order="F"))
    lib().tranmd_23_(_dp(c), _i64(n), _i64(n), _i64(n), _i64(1))
    assert np.array_equal(c.ravel(order="F"), golden["cube_3"][9])
    # 4-index form: swap 2 and 3 for every trailing index
    rng = np.random.default_rng(1)
    a = np.asfortranarray(rng.normal(size=(2, 4, 4, 3)))
    want = np.einsum("ikjl->ijkl", a).copy()
    lib().tranmd_23_(_dp(a), _i64(2), _i64(4), _i64(4), _i64(3))
    assert np.array_equal(a, want)


def _tranmd_cases():
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import TRANMD
    return TRANMD


@pytest.mark.parametrize("shape,code", _tranmd_cases(), ids=lambda v: str(v))
def test_tranmd_smp_matches_reference(golden, shape, code):
    """Every TRANMD_SMP code, on the 3^4 array and on shapes where only the permuted extents
    agree, against the reference routine's own output (golden vectors)."""
    a = np.arange(1, np.prod(shape) + 1, dtype=float)
    lib().tranmd_smp_(_dp(a), *[_i64(v) for v in shape], _i64(code))
    assert lib().ccx_last_status() == 0
    assert np.array_equal(a, golden["tranmd_%d_%d_%d_%d_%d" % (tuple(shape) + (code,))])


def test_tranmd_smp_rejects_unequal_extents():
    a = np.arange(1, 2 * 3 * 4 * 5 + 1, dtype=float)
    before = a.copy()
    lib().tranmd_smp_(_dp(a), _i64(2), _i64(3), _i64(4), _i64(5), _i64(12))
    assert lib().ccx_last_status() == 1
    assert np.array_equal(a, before)


def test_async_and_queue_variants_equal_plain():
    rng = np.random.default_rng(5)
    for plain, asy in (("trant3_1_", "trant3_1_async_"), ("trant3_4_", "trant3_4_async_")):
        c = np.asfortranarray(rng.normal(size=(5, 5, 5)))
        d = c.copy(order="F")
        getattr(lib(), plain)(_i64(5), _dp(c))
        getattr(lib(), asy)(_i64(3), _i64(5), _dp(d))
        assert lib().ccx_last_status() == 0
        assert np.array_equal(c, d)
    a = np.asfortranarray(rng.normal(size=(2, 4, 4, 3)))
    b = a.copy(order="F")
    lib().tranmd_23_(_dp(a), _i64(2), _i64(4), _i64(4), _i64(3))
    lib().tranmd_23_acc_(_i64(1), _dp(b), _i64(2), _i64(4), _i64(4), _i64(3))
    assert np.array_equal(a, b)


def test_gpu_trio_order_errors():
    """ijk_gpu_driver_ / ijk_gpu_finalize_ before ijk_gpu_init_ report a state error."""
    L = lib()
    z = np.zeros(8)
    L.ijk_gpu_finalize_(_i64(2), _i64(2), _dp(z))
    assert L.ccx_last_status() == 4
    L.ijk_gpu_driver_(_i64(2), _i64(2), _i64(1), _i64(1), _i64(1), _dp(z), _dp(z), _dp(z), _dp(z), _dp(z), _dp(z),
                      _dp(z))
    assert L.ccx_last_status() == 4


def test_trant3_random_cube_matches_port():
    rng = np.random.default_rng(2)
    for op in range(1, 6):
        c = np.asfortranarray(rng.normal(size=(7, 7, 7)))
        want = port.cube_op(c, op)
        lib().trant3_acc_(_dp(c), _i64(7), _i64(op))
        assert np.array_equal(c, want)


def test_argument_errors():
    L = lib()
    ctx = ctypes.c_void_p()
    assert L.ccx_create(0, 5, ctypes.byref(ctx)) == 1
    assert L.ccx_create(3, 0, ctypes.byref(ctx)) == 1
    assert L.ccx_create(3, 5, None) == 1
    assert L.ccx_destroy(None) == 1
    assert L.ccx_energy(None, 0, None, None, None, 0) == 1
    c = np.zeros(8)
    L.trant3_acc_(_dp(c), _i64(2), _i64(9))
    assert L.ccx_last_status() == 1
    L.common_cc_init_(_i64(0), _i64(5))
    assert L.ccx_last_status() == 1
    for s in range(8):
        assert L.ccx_error_string(s)


def test_legacy_entry_without_registered_vvvo_reports_state():
    L = lib()
    L.common_cc_init_(_i64(3), _i64(5))
    assert L.ccx_last_status() == 0
    z = [np.zeros(200) for _ in range(9)]
    e1, e2 = ctypes.c_double(), ctypes.c_double()
    L.cc_triples_acc_(*[_dp(a) for a in z], _i64(987654), ctypes.byref(e1), ctypes.byref(e2))
    assert L.ccx_last_status() == 4


def test_no_gpu_is_reported_not_crashed():
    n = ctypes.c_int(0)
    have_gpu = lib().ccx_hip_device_count(ctypes.byref(n)) == 0 and n.value > 0
    if have_gpu:
        pytest.skip("a GPU is present")
    ctx = ctypes.c_void_p()
    assert lib().ccx_create(3, 5, ctypes.byref(ctx)) == 2


def test_host_input_shapes_checked():
    """The Python mirror turns logical-shape arrays and 1-D Fortran buffers into the same
    buffer and rejects a wrong size before the library can read past it."""
    from libcc_amd.triples import _fortran_buffer_np
    a = np.arange(24.0).reshape(2, 3, 4, order="F")
    assert np.array_equal(_fortran_buffer_np(a, (2, 3, 4), "x").ravel(order="K"), np.arange(24.0))
    c = np.ascontiguousarray(a)
    assert np.array_equal(_fortran_buffer_np(c, (2, 3, 4), "x").ravel(order="F"), a.ravel(order="F"))
    assert np.array_equal(_fortran_buffer_np(a.ravel(order="F"), (2, 3, 4), "x"), np.arange(24.0))
    for wrong in (np.zeros((2, 3, 5)), np.zeros(23), np.zeros((4, 3, 2))):
        with pytest.raises(ValueError):
            _fortran_buffer_np(wrong, (2, 3, 4), "x")
