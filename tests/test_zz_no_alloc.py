"""GPU check of the allocation-free sort (sbs_util_sort_pairs_noalloc): device-resident pairs sorted
in place on the caller's stream with the library's per-GPU scratch array, against the allocating
entry point (sbs_util_sort_pairs, host arrays in and out), twice in a row, the second call with a
smaller input (the scratch array keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest

import sbs_amd as S
from sbs_amd import _lib

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _hip():
    # the HIP runtime the library links (same soname, so the same loaded instance): device buffers
    # for the test's inputs, which the entry point sorts in place
    import os

    lib = _lib.lib()  # loads libsbs_gpu.so and with it the runtime
    hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so.7"))
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    for f in (hip.hipMalloc, hip.hipMemcpy, hip.hipFree, hip.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return lib, hip


def _ok(rc, what):
    assert rc == 0, f"{what} returned {rc}"


def _noalloc_sort(keys, vals, bits):
    lib, hip = _hip()
    n = len(keys)
    dk, dv = ctypes.c_void_p(), ctypes.c_void_p()
    _ok(hip.hipMalloc(ctypes.byref(dk), max(4 * n, 4)), "hipMalloc")
    _ok(hip.hipMalloc(ctypes.byref(dv), max(4 * n, 4)), "hipMalloc")
    try:
        k, v = np.ascontiguousarray(keys, np.uint32), np.ascontiguousarray(vals, np.uint32)
        _ok(hip.hipMemcpy(dk, k.ctypes.data, 4 * n, 1), "hipMemcpy")  # host to device
        _ok(hip.hipMemcpy(dv, v.ctypes.data, 4 * n, 1), "hipMemcpy")
        _lib.check(lib.sbs_util_sort_pairs_noalloc(0, dk, dv, n, bits, None))
        _ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        ok, ov = np.empty(n, np.uint32), np.empty(n, np.uint32)
        _ok(hip.hipMemcpy(ok.ctypes.data, dk, 4 * n, 2), "hipMemcpy")  # device to host
        _ok(hip.hipMemcpy(ov.ctypes.data, dv, 4 * n, 2), "hipMemcpy")
        return ok, ov
    finally:
        hip.hipFree(dk)
        hip.hipFree(dv)


def test_noalloc_sort_matches_the_allocating_sort_twice():
    rng = np.random.default_rng(21)
    for n, bits in ((1 << 20, 24), (77_777, 32)):  # the maximum size, then a smaller input
        say(f"allocation-free sort n={n} bits={bits}")
        keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        vals = np.arange(n, dtype=np.uint32)
        rk, rv = S.sort_pairs(keys, vals, bits)
        ok, ov = _noalloc_sort(keys, vals, bits)
        assert np.array_equal(ok, rk) and np.array_equal(ov, rv)
        # stable: equal keys (low bits) keep their input order
        mask = np.uint32((1 << bits) - 1) if bits < 32 else np.uint32(0xFFFFFFFF)
        lo = ok & mask
        assert np.all(lo[1:] >= lo[:-1])


def test_noalloc_sort_rejects_inputs_beyond_its_scratch():
    lib, hip = _hip()
    n = (1 << 20) + 1
    d = ctypes.c_void_p()
    _ok(hip.hipMalloc(ctypes.byref(d), 4 * n), "hipMalloc")
    try:
        with pytest.raises(_lib.SbsError):
            _lib.check(lib.sbs_util_sort_pairs_noalloc(0, d, d, n, 32, None))
    finally:
        hip.hipFree(d)


def test_noalloc_sort_edge_cases():
    # empty input (null pointers allowed), one pair, zero key bits: nothing to do and no error;
    # key bits beyond 32 and a device that does not exist: errors, before any launch
    lib, hip = _hip()
    _lib.check(lib.sbs_util_sort_pairs_noalloc(0, None, None, 0, 32, None))
    ok, ov = _noalloc_sort(np.array([7], np.uint32), np.array([3], np.uint32), 32)
    assert ok.tolist() == [7] and ov.tolist() == [3]
    keys = np.array([5, 1, 4], np.uint32)
    ok, ov = _noalloc_sort(keys, np.arange(3, dtype=np.uint32), 0)
    assert ok.tolist() == [5, 1, 4] and ov.tolist() == [0, 1, 2]
    with pytest.raises(_lib.SbsError):
        _lib.check(lib.sbs_util_sort_pairs_noalloc(0, None, None, 0, 33, None))
    with pytest.raises(_lib.SbsError):
        _lib.check(lib.sbs_util_sort_pairs_noalloc(-1, None, None, 0, 8, None))
