"""The point operations of include/gqf_device.h, compiled for the host, against
the CPU oracle (no GPU needed).

The same C++ code (insert1, insert1_if_not_exists, count) runs on the GPU
inside point_insert / point_insert_not_exists / point_query; here it runs on
a table in host memory, so each step's image can be compared byte for byte
with the oracle's sequential restatement of the reference host path
(qf_insert / qf_insert_not_exists / qf_count_key_value)."""
import ctypes

import numpy as np
import pytest

from tests import spec_model as S

KEY_IS_HASH = 0x08


@pytest.fixture(scope="module")
def ph():
    from tests.point_host import build as B
    L = ctypes.CDLL(B.build())
    vp, u64 = ctypes.c_void_p, ctypes.c_uint64
    L.ph_table_bytes.argtypes = [ctypes.c_int]
    L.ph_table_bytes.restype = u64
    L.ph_insert1.argtypes = [vp, ctypes.c_int, u64]
    L.ph_insert1.restype = ctypes.c_int
    L.ph_insert_not_exists.argtypes = [vp, ctypes.c_int, u64, ctypes.POINTER(ctypes.c_uint8)]
    L.ph_insert_not_exists.restype = ctypes.c_int
    L.ph_count.argtypes = [vp, ctypes.c_int, u64]
    L.ph_count.restype = u64
    L.ph_export.argtypes = [vp, ctypes.c_int, vp]
    L.ph_used_mismatches.argtypes = [vp, ctypes.c_int]
    L.ph_used_mismatches.restype = u64
    return L


class HostTable:
    def __init__(self, ph, nbits, oracle):
        self.ph, self.nbits = ph, nbits
        self.buf = np.zeros(ph.ph_table_bytes(nbits), dtype=np.uint8)
        self.nblocks = oracle.sizes(nbits)["nblocks"]

    def _p(self):
        return self.buf.ctypes.data_as(ctypes.c_void_p)

    def insert(self, h):
        return self.ph.ph_insert1(self._p(), self.nbits, int(h))

    def insert_not_exists(self, h):
        v = ctypes.c_uint8(77)
        r = self.ph.ph_insert_not_exists(self._p(), self.nbits, int(h), ctypes.byref(v))
        return r, v.value

    def count(self, h):
        return int(self.ph.ph_count(self._p(), self.nbits, int(h)))

    def image(self):
        # every image read also checks the device layout's used-slot masks
        assert self.ph.ph_used_mismatches(self._p(), self.nbits) == 0
        out = np.zeros(self.nblocks * 81, dtype=np.uint8)
        self.ph.ph_export(self._p(), self.nbits, out.ctypes.data_as(ctypes.c_void_p))
        return out


def _hashes(oracle, nbits, seed, n):
    return oracle.hash_many(oracle.splitmix64(seed, n), (1 << (nbits + 8)) - 1)


@pytest.mark.parametrize("nbits,load,seed", [(8, 0.9, 1), (10, 0.95, 2), (12, 0.9, 3)])
def test_insert1_matches_oracle(ph, oracle, nbits, load, seed):
    hs = _hashes(oracle, nbits, seed, int(load * (1 << nbits)))
    t = HostTable(ph, nbits, oracle)
    o = oracle.OracleQF(nbits)
    for i, h in enumerate(hs):
        assert t.insert(h) == 0
        assert o.insert(int(h), 1, KEY_IS_HASH) >= 0
        if i % 97 == 0:
            assert np.array_equal(t.image(), o.image()), i
    assert np.array_equal(t.image(), o.image())
    for h in hs[:500]:
        assert t.count(h) == o.count(int(h), KEY_IS_HASH)
    o.close()


def test_insert1_counters_match_oracle(ph, oracle):
    """Repeated hashes walk every counter length change (1, 2, 3, digits,
    leading 0) on neighbouring quotients that push each other."""
    nbits = 9
    rng = np.random.default_rng(7)
    univ = _hashes(oracle, nbits, 5, 60)
    # cluster some on adjacent quotients and remainders 0, 1, 2, 255
    extra = np.array([(40 << 8) | r for r in (0, 1, 2, 3, 255)] + [(41 << 8) | r for r in (0, 2, 254)],
                     dtype=np.uint64)
    univ = np.concatenate([univ, extra])
    draws = univ[rng.zipf(1.3, size=6000) % univ.size]
    t = HostTable(ph, nbits, oracle)
    o = oracle.OracleQF(nbits)
    for i, h in enumerate(draws):
        assert t.insert(h) == 0
        o.insert(int(h), 1, KEY_IS_HASH)
        if i % 251 == 0:
            assert np.array_equal(t.image(), o.image()), i
    assert np.array_equal(t.image(), o.image())
    ms = {}
    for h in draws:
        ms[int(h)] = ms.get(int(h), 0) + 1
    assert bytes(t.image()) == S.build_image(nbits, ms)
    for h, c in ms.items():
        assert t.count(h) == c
    o.close()


def test_insert_not_exists_matches_oracle_with_counters(ph, oracle):
    """insert1_if_not_exists searches the run slot by slot; on runs holding
    multi-slot counters its result depends on that quirk, which the host
    build of the device code must reproduce exactly."""
    nbits = 10
    rng = np.random.default_rng(11)
    base = _hashes(oracle, nbits, 9, 400)
    t = HostTable(ph, nbits, oracle)
    o = oracle.OracleQF(nbits)
    # counted items first (counts up to a few thousand)
    for h in base[:80]:
        for _ in range(int(rng.choice([1, 2, 3, 5, 40, 300]))):
            t.insert(h)
            o.insert(int(h), 1, KEY_IS_HASH)
    assert np.array_equal(t.image(), o.image())
    # then not-exists inserts: new remainders on the same quotients, repeats
    probes = np.concatenate([base, base[:200], (base[:80] & ~np.uint64(0xFF)) | np.uint64(7),
                             (base[:80] & ~np.uint64(0xFF)) | np.uint64(250)])
    for i, h in enumerate(probes):
        got = t.insert_not_exists(h)
        want = o.insert_not_exists(int(h), KEY_IS_HASH)
        assert got[0] == want[0], (i, got, want)
        if got[0] == 1:
            assert got[1] == 0
        if i % 37 == 0:
            assert np.array_equal(t.image(), o.image()), i
    assert np.array_equal(t.image(), o.image())
    o.close()


def test_full_table_reports_no_space_and_keeps_image(ph, oracle):
    nbits = 7
    t = HostTable(ph, nbits, oracle)
    o = oracle.OracleQF(nbits)
    hs = _hashes(oracle, nbits, 13, 4 * (1 << nbits))
    saw_full = False
    for h in hs:
        before = t.image().copy()
        r = t.insert(h)
        ro = o.insert(int(h), 1, KEY_IS_HASH)
        assert (r == -1) == (ro == -1)
        if r == -1:
            saw_full = True
            assert np.array_equal(t.image(), before)
        assert np.array_equal(t.image(), o.image())
    assert saw_full
    o.close()
