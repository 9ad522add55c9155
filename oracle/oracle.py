"""ctypes binding of the CPU oracle (oracle/cqf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the cqf_amd package.  The oracle is a
C restatement of the reference GQF host path (PASSIONLab/cqf src/gqf.cu,
src/hashutil.cu); see the citations in cqf_oracle.c.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

KEY_IS_HASH = 0x08
NO_LOCK = 0x01


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class _OrcQF(ctypes.Structure):
    _fields_ = [
        ("nslots", ctypes.c_uint64),
        ("xnslots", ctypes.c_uint64),
        ("nblocks", ctypes.c_uint64),
        ("key_bits", ctypes.c_uint64),
        ("key_remainder_bits", ctypes.c_uint64),
        ("bits_per_slot", ctypes.c_uint64),
        ("range", ctypes.c_uint64),
        ("num_locks", ctypes.c_uint64),
        ("total_size_in_bytes", ctypes.c_uint64),
        ("image", ctypes.POINTER(ctypes.c_uint8)),
        ("overflow_events", ctypes.c_uint64),
    ]


_lib = None
_u64p = ctypes.POINTER(ctypes.c_uint64)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(_OrcQF)
        L.orc_init.argtypes = [P, ctypes.c_int]
        L.orc_free.argtypes = [P]
        L.orc_reset.argtypes = [P]
        L.orc_sizes.argtypes = [ctypes.c_int, _u64p]
        L.orc_sizes.restype = ctypes.c_uint64
        L.orc_hash_64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_hash_64.restype = ctypes.c_uint64
        L.orc_hash_64i.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_hash_64i.restype = ctypes.c_uint64
        L.orc_hash_many.argtypes = [_u64p, _u64p, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_encode_counter.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint8)]
        L.orc_encode_counter.restype = ctypes.c_int
        L.orc_decode_counter_run.argtypes = [ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint64,
                                             ctypes.c_uint64, _u64p, _u64p]
        L.orc_decode_counter_run.restype = ctypes.c_uint64
        L.orc_qf_insert.argtypes = [P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint8]
        L.orc_qf_insert.restype = ctypes.c_int
        L.orc_qf_remove.argtypes = [P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint8]
        L.orc_qf_remove.restype = ctypes.c_int
        L.orc_qf_count_key_value.argtypes = [P, ctypes.c_uint64, ctypes.c_uint8]
        L.orc_qf_count_key_value.restype = ctypes.c_uint64
        L.orc_qf_insert_not_exists.argtypes = [P, ctypes.c_uint64, ctypes.c_uint8,
                                               ctypes.POINTER(ctypes.c_uint8)]
        L.orc_qf_insert_not_exists.restype = ctypes.c_int
        L.orc_run_end.argtypes = [P, ctypes.c_uint64]
        L.orc_run_end.restype = ctypes.c_uint64
        for fn in ("orc_bulk_insert", "orc_bulk_insert_reduce", "orc_bulk_delete"):
            getattr(L, fn).argtypes = [P, _u64p, ctypes.c_uint64, ctypes.c_uint8]
            getattr(L, fn).restype = ctypes.c_uint64
        L.orc_bulk_get.argtypes = [P, _u64p, ctypes.c_uint64, _u64p]
        L.orc_bulk_get_misses.argtypes = [P, _u64p, ctypes.c_uint64]
        L.orc_bulk_get_misses.restype = ctypes.c_uint64
        L.orc_build_canonical.argtypes = [P, _u64p, _u64p, ctypes.c_uint64]
        L.orc_time_inserts.argtypes = [P, _u64p, ctypes.c_uint64]
        L.orc_time_inserts.restype = ctypes.c_double
        L.orc_time_queries.argtypes = [P, _u64p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        L.orc_time_queries.restype = ctypes.c_double
        L.orc_build_canonical.restype = ctypes.c_int
        L.orc_enumerate.argtypes = [P, _u64p, _u64p, ctypes.c_uint64]
        L.orc_enumerate.restype = ctypes.c_uint64
        L.orc_cluster_census.argtypes = [P, _u64p, ctypes.c_uint64]
        L.orc_cluster_census.restype = ctypes.c_uint64
        L.orc_splitmix64_fill.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _u64p]
        L.orc_time_insert_query.argtypes = [ctypes.c_int, _u64p, ctypes.c_uint64, _u64p,
                                            ctypes.c_uint64, _u64p,
                                            ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double)]
        L.orc_time_insert_query.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(_u64p)


def sizes(nbits):
    out = (ctypes.c_uint64 * 8)()
    lib().orc_sizes(nbits, out)
    keys = ("nslots", "xnslots", "nblocks", "key_bits", "key_remainder_bits", "range",
            "num_locks", "total_size_in_bytes")
    return dict(zip(keys, [int(v) for v in out]))


def hash_64(key, mask):
    return int(lib().orc_hash_64(key, mask))


def hash_64i(key, mask):
    return int(lib().orc_hash_64i(key, mask))


def hash_many(keys, mask):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.empty_like(keys)
    lib().orc_hash_many(_p(keys), _p(out), keys.size, mask)
    return out


def encode_counter(rem, count):
    buf = (ctypes.c_uint8 * 67)()
    n = lib().orc_encode_counter(rem, count, buf)
    return bytes(buf[:n])


def decode_counter_run(run, index=0):
    """decode_counter over a stand-alone run (last slot = run end)."""
    arr = (ctypes.c_uint8 * max(1, len(run)))(*run)
    r = ctypes.c_uint64()
    c = ctypes.c_uint64()
    end = lib().orc_decode_counter_run(arr, len(run), index, ctypes.byref(r), ctypes.byref(c))
    return int(end), int(r.value), int(c.value)


def splitmix64(seed, n, skip=0):
    out = np.empty(n, dtype=np.uint64)
    if n:
        lib().orc_splitmix64_fill(seed, skip, n, _p(out))
    return out


class OracleQF:
    """One CQF table image on the host, driven with reference semantics."""

    def __init__(self, nbits):
        self.nbits = nbits
        self._s = _OrcQF()
        if lib().orc_init(ctypes.byref(self._s), nbits) != 0:
            raise MemoryError("oracle init failed")

    def close(self):
        if self._s.image:
            lib().orc_free(ctypes.byref(self._s))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # sizing
    @property
    def nslots(self):
        return int(self._s.nslots)

    @property
    def xnslots(self):
        return int(self._s.xnslots)

    @property
    def nblocks(self):
        return int(self._s.nblocks)

    @property
    def range(self):
        return int(self._s.range)

    @property
    def mask(self):
        return int(self._s.range) - 1

    @property
    def overflow_events(self):
        return int(self._s.overflow_events)

    def reset(self):
        lib().orc_reset(ctypes.byref(self._s))

    def image(self):
        n = int(self._s.total_size_in_bytes)
        return np.ctypeslib.as_array(self._s.image, shape=(n,)).copy()

    def set_image(self, img):
        n = int(self._s.total_size_in_bytes)
        img = np.ascontiguousarray(img, dtype=np.uint8)
        assert img.size == n
        ctypes.memmove(self._s.image, img.ctypes.data, n)

    # per-key
    def insert(self, key, count=1, flags=0):
        return lib().orc_qf_insert(ctypes.byref(self._s), key, count, flags)

    def remove(self, key, count=1, flags=0):
        return lib().orc_qf_remove(ctypes.byref(self._s), key, count, flags)

    def count(self, key, flags=0):
        return int(lib().orc_qf_count_key_value(ctypes.byref(self._s), key, flags))

    def insert_not_exists(self, key, flags=0):
        rv = ctypes.c_uint8(77)
        ret = lib().orc_qf_insert_not_exists(ctypes.byref(self._s), key, flags, ctypes.byref(rv))
        return ret, int(rv.value)

    def run_end(self, h):
        return int(lib().orc_run_end(ctypes.byref(self._s), h))

    # bulk (sequential, reference semantics); keys arrays are overwritten
    # with ascending hashes like the reference's bulk mutations.
    def bulk_insert(self, keys, flags=0):
        return int(lib().orc_bulk_insert(ctypes.byref(self._s), _p(keys), keys.size, flags))

    def bulk_insert_reduce(self, keys, flags=0):
        return int(lib().orc_bulk_insert_reduce(ctypes.byref(self._s), _p(keys), keys.size, flags))

    def bulk_delete(self, keys, flags=0):
        return int(lib().orc_bulk_delete(ctypes.byref(self._s), _p(keys), keys.size, flags))

    def bulk_get(self, vals):
        vals = np.ascontiguousarray(vals, dtype=np.uint64)
        out = np.empty_like(vals)
        lib().orc_bulk_get(ctypes.byref(self._s), _p(vals), vals.size, _p(out))
        return out

    def bulk_get_misses(self, vals):
        vals = np.ascontiguousarray(vals, dtype=np.uint64)
        return int(lib().orc_bulk_get_misses(ctypes.byref(self._s), _p(vals), vals.size))

    def time_inserts(self, keys):
        """Seconds of one sequential qf_insert per key (keys hashed)."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        return float(lib().orc_time_inserts(ctypes.byref(self._s), _p(keys), keys.size))

    def time_queries(self, keys):
        """(seconds, keys found) of one qf_count_key_value per key."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        hits = ctypes.c_uint64()
        t = float(lib().orc_time_queries(ctypes.byref(self._s), _p(keys), keys.size, ctypes.byref(hits)))
        return t, int(hits.value)

    def build_canonical(self, hashes, counts):
        hashes = np.ascontiguousarray(hashes, dtype=np.uint64)
        counts = np.ascontiguousarray(counts, dtype=np.uint64)
        return lib().orc_build_canonical(ctypes.byref(self._s), _p(hashes), _p(counts), hashes.size)

    def enumerate(self, cap=None):
        if cap is None:
            cap = self.xnslots
        h = np.empty(cap, dtype=np.uint64)
        c = np.empty(cap, dtype=np.uint64)
        m = int(lib().orc_enumerate(ctypes.byref(self._s), _p(h), _p(c), cap))
        return h[:m], c[:m]


    def cluster_census(self):
        """Cluster lengths in slot order (the reference's census loop)."""
        cap = self.xnslots
        out = np.empty(cap, dtype=np.uint64)
        m = int(lib().orc_cluster_census(ctypes.byref(self._s), _p(out), cap))
        return out[:m]


def canonical_image(nbits, hashes, counts):
    """81-byte block image fixed by the multiset {hash: count}."""
    o = OracleQF(nbits)
    rc = o.build_canonical(hashes, counts)
    img = o.image()
    o.close()
    if rc != 0:
        raise ValueError("multiset does not fit the table")
    return img


def time_insert_query(nbits, keys, qkeys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    qkeys = np.ascontiguousarray(qkeys, dtype=np.uint64)
    hits = ctypes.c_uint64()
    ti = ctypes.c_double()
    tq = ctypes.c_double()
    lib().orc_time_insert_query(nbits, _p(keys), keys.size, _p(qkeys), qkeys.size,
                                ctypes.byref(hits), ctypes.byref(ti), ctypes.byref(tq))
    return ti.value, tq.value, int(hits.value)
