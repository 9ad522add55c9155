"""A reference-style caller of libgqf on the GPU (tests/drop_in/dropin.hip):
the reference's wrapper calls (include/gqf_wrapper.cuh:36-86) and its
point_bulk_insert<<<(n - 1) / 32 + 1, 32>>> launch
(include/point_wrapper.cuh:63), compiled unchanged against include/gqf.h
and include/gqf_device.h, plus an application kernel calling point_insert.
The filters they build must hold the oracle's images.  The library is built
on the CPU by build(); every test prints its case first."""
import ctypes
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "drop_in", "libdropin.so")


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dropin(gpu):
    from cqf_amd import gqf
    gqf.load_library()  # the caller links the same libgqf.so
    if not os.path.exists(LIB):
        pytest.fail("tests/drop_in/libdropin.so not built (run build())")
    L = ctypes.CDLL(LIB)
    vp, u64 = ctypes.c_void_p, ctypes.c_uint64
    for fn in ("dropin_gqf_init", "dropin_point_init"):
        getattr(L, fn).argtypes = [ctypes.c_int]
        getattr(L, fn).restype = vp
    for fn in ("dropin_gqf_bulk_insert", "dropin_gqf_bulk_delete", "dropin_point_bulk_insert"):
        getattr(L, fn).argtypes = [vp, u64]
    for fn in ("dropin_gqf_bulk_get", "dropin_point_bulk_get"):
        getattr(L, fn).argtypes = [vp, u64]
        getattr(L, fn).restype = u64
    L.dropin_count_kmers.argtypes = [vp, vp, u64]
    return L


def test_point_bulk_insert_kernel_and_user_kernel(dropin, oracle):
    from cqf_amd.gqf import GQF, to_device_u64
    q = 16
    k1 = oracle.splitmix64(61, 20000)
    k2 = np.concatenate([oracle.splitmix64(62, 15000), k1[:3000]])  # repeats: counts of 2
    _say(f"case drop-in point wrappers q={q}: point_bulk_insert<<<>>> {k1.size} keys, user kernel {k2.size}")
    h = dropin.dropin_point_init(q)
    assert h
    f = GQF(q, _handle=ctypes.c_void_p(h))  # a view on the caller's filter: export only
    try:
        d1 = to_device_u64(k1)
        dropin.dropin_point_bulk_insert(d1.data_ptr(), k1.size)
        o = oracle.OracleQF(q)
        for k in k1:
            o.insert(int(k))
        assert np.array_equal(f.export_image(), o.image())
        d2 = to_device_u64(k2)
        dropin.dropin_count_kmers(ctypes.c_void_p(h), d2.data_ptr(), k2.size)
        for k in k2:
            o.insert(int(k))
        assert np.array_equal(f.export_image(), o.image())
        allk = np.concatenate([k1, k2])
        assert dropin.dropin_point_bulk_get(to_device_u64(allk).data_ptr(), allk.size) == 0
        assert f.status() == 0 and f.layout_errors() == 0
        o.close()
    finally:
        f.handle = None  # the caller destroys its own filter
        dropin.dropin_point_destroy()


def test_gqf_wrapper_bulk_calls(dropin, oracle):
    from cqf_amd.gqf import GQF, to_device_u64
    q = 18
    keys = oracle.splitmix64(63, int(0.9 * (1 << q)))
    _say(f"case drop-in gqf wrappers q={q}: bulk insert/get/delete of {keys.size} keys")
    h = dropin.dropin_gqf_init(q)
    assert h
    f = GQF(q, _handle=ctypes.c_void_p(h))
    try:
        dk = to_device_u64(keys)
        dropin.dropin_gqf_bulk_insert(dk.data_ptr(), keys.size)
        o = oracle.OracleQF(q)
        o.bulk_insert(keys.copy())
        assert np.array_equal(f.export_image(), o.image())
        # the misses count of the inserted keys is 0 (src/test.cu:478-484)
        assert dropin.dropin_gqf_bulk_get(to_device_u64(keys).data_ptr(), keys.size) == 0
        dk = to_device_u64(keys)
        dropin.dropin_gqf_bulk_delete(dk.data_ptr(), keys.size)
        torch.cuda.synchronize()
        assert not f.export_image().any()
        # and all of them miss after the full delete (src/gqf_verify.cu:439)
        assert dropin.dropin_gqf_bulk_get(to_device_u64(keys).data_ptr(), keys.size) == keys.size
        o.close()
    finally:
        f.handle = None
        dropin.dropin_gqf_destroy()
