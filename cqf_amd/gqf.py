"""Python mirror of the GQF bulk API over libgqf's C ABI.

The names and argument meaning follow the reference API
(PASSIONLab/cqf include/gqf.cuh:87-140): `bulk_insert`, `bulk_insert_reduce`
and `bulk_delete` overwrite the key tensor with the ascending hashes,
`bulk_get` leaves `vals` unchanged and fills `returns`.  Device memory is
held in torch tensors (dtype int64 carries the uint64 bit patterns); torch
is only the allocator and stream owner here -- every operation runs in the
HIP kernels of libgqf.so, and importing this module fails loudly when that
library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgqf.so")

QF_NO_LOCK = 0x01
QF_TRY_ONCE_LOCK = 0x02
QF_WAIT_FOR_LOCK = 0x04
QF_KEY_IS_HASH = 0x08

QF_STATUS_NO_SPACE = 0x1
QF_STATUS_CLUSTER_OVERFLOW = 0x2
QF_STATUS_BAD_HASH = 0x4
QF_STATUS_POINT_FULL = 0x8

# every extern "C" entry point declared in include/gqf.h
EXPORTS = [
    "qf_malloc_device", "qf_destroy_device",
    "bulk_insert", "bulk_insert_reduce", "bulk_delete", "bulk_get", "bulk_get_misses_wrapper",
    "point_get_wrapper", "point_bulk_insert_wrapper",
    "point_insert_async", "point_insert_not_exists_async", "point_query_async",
    "qf_get_info", "qf_get_status", "qf_clear_status",
    "qf_export_image", "qf_import_image", "qf_export_image_device", "qf_import_image_device",
    "qf_reserve", "qf_set_tuning", "qf_reset_device", "bulk_insert_async", "bulk_insert_reduce_async", "bulk_delete_async",
    "bulk_get_async", "qf_last_error", "gqf_profile_enable", "gqf_profile_read", "gqf_fill_splitmix64",
    "gqf_group_create", "gqf_group_destroy", "gqf_group_bulk_insert", "gqf_group_bulk_get",
    "gqf_group_set_graphs", "gqf_autotune", "gqf_autotune_set", "gqf_group_autotune",
    "gqf_group_set_workgroups_per_cu", "gqf_group_get_workgroups_per_cu", "gqf_group_set_wait_bound",
    "qf_batch_pieces", "bulk_insert_noalloc", "qf_check_layout",
    "qf_server_create", "qf_server_add_filter", "qf_server_submit", "qf_server_run", "qf_server_wait",
    "qf_server_request_times", "qf_server_get_stats", "qf_server_filter", "qf_server_destroy",
    "qf_malloc_device_kv", "qf_enumerate", "qf_get_stats", "qf_cluster_census", "qf_serialize",
    "qf_deserialize_device", "qf_resize_device", "qf_merge_device", "qf_multi_merge_device",
    "qf_bulk_mutate_seq", "qf_bulk_get_seq_adjust", "qf_set_wide_wait_bound",
]

QF_SEQ_MAX_REQUESTS = 256

QF_ENUM_HASH = 0
QF_ENUM_KEY = 1


class GqfInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "nslots", "xnslots", "nblocks", "key_bits", "key_remainder_bits", "bits_per_slot",
        "range", "num_locks", "nregions", "total_size_in_bytes", "device_bytes", "region_blocks",
        "sort_digit_bits", "value_bits")]


_lib = None
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64


def load_library(path=LIB_PATH):
    """Load libgqf.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libgqf.so not built at {path}: run `python -m cqf_amd.build`")
    L = ctypes.CDLL(path)
    L.qf_malloc_device.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_bool]
    L.qf_destroy_device.argtypes = [_vp]
    for fn in ("bulk_insert", "bulk_insert_reduce", "bulk_delete"):
        getattr(L, fn).argtypes = [_vp, _u64, _vp, ctypes.c_uint8]
    for fn in ("bulk_insert_async", "bulk_insert_reduce_async", "bulk_delete_async"):
        getattr(L, fn).argtypes = [_vp, _u64, _vp, ctypes.c_uint8, _vp]
    L.bulk_get.argtypes = [_vp, _u64, _vp, _vp]
    L.bulk_get_async.argtypes = [_vp, _u64, _vp, _vp, _vp]
    L.bulk_get_misses_wrapper.argtypes = [_vp, _vp, _u64]
    L.bulk_get_misses_wrapper.restype = _u64
    L.point_get_wrapper.argtypes = [_vp, _vp, _u64]
    L.point_get_wrapper.restype = _u64
    L.point_bulk_insert_wrapper.argtypes = [_vp, _vp, _u64]
    L.point_insert_async.argtypes = [_vp, _u64, _vp, ctypes.c_uint8, _vp, _vp]
    L.point_insert_not_exists_async.argtypes = [_vp, _u64, _vp, ctypes.c_uint8, _vp, _vp]
    L.point_query_async.argtypes = [_vp, _u64, _vp, ctypes.c_uint8, _vp, ctypes.c_int, _vp]
    L.qf_get_info.argtypes = [_vp, ctypes.POINTER(GqfInfo)]
    L.qf_get_info.restype = ctypes.c_int
    L.qf_get_status.argtypes = [_vp]
    L.qf_get_status.restype = ctypes.c_uint32
    L.qf_clear_status.argtypes = [_vp]
    L.qf_export_image.argtypes = [_vp, _vp, _u64]
    L.qf_export_image.restype = ctypes.c_int
    L.qf_import_image.argtypes = [_vp, _vp, _u64]
    L.qf_import_image.restype = ctypes.c_int
    L.qf_export_image_device.argtypes = [_vp, _vp, _u64, _vp]
    L.qf_export_image_device.restype = ctypes.c_int
    L.qf_import_image_device.argtypes = [_vp, _vp, _u64, _vp]
    L.qf_import_image_device.restype = ctypes.c_int
    L.qf_reset_device.argtypes = [_vp, _vp]
    L.qf_reset_device.restype = ctypes.c_int
    L.qf_reserve.argtypes = [_vp, _u64]
    L.qf_reserve.restype = ctypes.c_int
    L.qf_set_tuning.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32]
    L.qf_set_tuning.restype = ctypes.c_int
    L.qf_last_error.restype = ctypes.c_int
    L.gqf_profile_enable.argtypes = [ctypes.c_int]
    L.gqf_profile_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_u64)]
    L.gqf_profile_read.restype = ctypes.c_int
    L.gqf_fill_splitmix64.argtypes = [_vp, _u64, _u64, _u64, _vp]
    L.gqf_group_create.argtypes = [ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_uint32]
    L.gqf_group_create.restype = ctypes.c_int
    L.gqf_group_destroy.argtypes = [_vp]
    L.gqf_group_bulk_insert.argtypes = [_vp, _vp, ctypes.POINTER(_u64), ctypes.c_uint8, _vp]
    L.gqf_group_bulk_insert.restype = ctypes.c_int
    L.gqf_group_bulk_get.argtypes = [_vp, _vp, ctypes.POINTER(_u64), _vp, _vp]
    L.gqf_group_bulk_get.restype = ctypes.c_int
    L.gqf_group_set_graphs.argtypes = [_vp, ctypes.c_int]
    L.gqf_group_set_graphs.restype = ctypes.c_int
    L.gqf_autotune.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_uint32),
                               ctypes.POINTER(ctypes.c_uint32)]
    L.gqf_autotune.restype = ctypes.c_int
    L.gqf_autotune_set.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32]
    L.gqf_autotune_set.restype = ctypes.c_int
    L.gqf_group_autotune.argtypes = [_vp, ctypes.POINTER(_u64)]
    L.gqf_group_autotune.restype = ctypes.c_int
    L.gqf_group_set_workgroups_per_cu.argtypes = [_vp, ctypes.c_uint32]
    L.gqf_group_set_workgroups_per_cu.restype = ctypes.c_int
    L.gqf_group_get_workgroups_per_cu.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint32)]
    L.gqf_group_get_workgroups_per_cu.restype = ctypes.c_int
    L.qf_set_wide_wait_bound.argtypes = [ctypes.c_uint32]
    L.qf_set_wide_wait_bound.restype = ctypes.c_uint32
    L.gqf_group_set_wait_bound.argtypes = [_vp, ctypes.c_uint32]
    L.gqf_group_set_wait_bound.restype = ctypes.c_int
    L.qf_check_layout.argtypes = [_vp, ctypes.POINTER(_u64)]
    L.qf_check_layout.restype = ctypes.c_int
    L.bulk_insert_noalloc.argtypes = [_vp, _u64, _vp, ctypes.c_uint8, _vp]
    L.bulk_insert_noalloc.restype = ctypes.c_int
    L.qf_batch_pieces.argtypes = [_u64, _u64, ctypes.POINTER(_u64), ctypes.c_uint32]
    L.qf_batch_pieces.restype = ctypes.c_uint32
    L.qf_malloc_device_kv.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int]
    L.qf_enumerate.argtypes = [_vp, _u64, _u64, ctypes.c_int, _vp, _vp, _vp, _u64, ctypes.POINTER(_u64), _vp]
    L.qf_enumerate.restype = ctypes.c_int
    L.qf_get_stats.argtypes = [_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64), ctypes.POINTER(_u64)]
    L.qf_get_stats.restype = ctypes.c_int
    L.qf_cluster_census.argtypes = [_vp, _vp, _u64, _vp, ctypes.c_uint32, ctypes.POINTER(_u64),
                                    ctypes.POINTER(_u64)]
    L.qf_cluster_census.restype = ctypes.c_int
    L.qf_serialize.argtypes = [_vp, ctypes.c_char_p]
    L.qf_serialize.restype = _u64
    L.qf_deserialize_device.argtypes = [ctypes.POINTER(_vp), ctypes.c_char_p]
    L.qf_deserialize_device.restype = _u64
    L.qf_resize_device.argtypes = [_vp, ctypes.c_int]
    L.qf_resize_device.restype = ctypes.c_int64
    L.qf_merge_device.argtypes = [_vp, _vp, _vp]
    L.qf_merge_device.restype = ctypes.c_int64
    L.qf_multi_merge_device.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, _vp]
    _u32p = ctypes.POINTER(ctypes.c_uint32)
    L.qf_bulk_mutate_seq.argtypes = [_vp, _u64, _vp, ctypes.c_uint32, _u32p, _u32p, _vp]
    L.qf_bulk_mutate_seq.restype = ctypes.c_int
    L.qf_bulk_get_seq_adjust.argtypes = [_vp, _u64, _vp, ctypes.c_uint32, _u32p, _u32p, _vp, _u64, _vp, _vp]
    L.qf_bulk_get_seq_adjust.restype = ctypes.c_int
    L.qf_multi_merge_device.restype = ctypes.c_int64
    _lib = L
    return L


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(stream.cuda_stream)


def to_device_u64(arr, device="cuda"):
    """numpy uint64 -> torch int64 device tensor holding the same bits."""
    import torch
    a = np.ascontiguousarray(arr, dtype=np.uint64).view(np.int64)
    return torch.from_numpy(a.copy()).to(device)


def to_numpy_u64(t):
    return t.detach().cpu().numpy().view(np.uint64)


class GQF:
    """One device-resident counting quotient filter (2^nbits slots of 8
    bits; value_bits > 0 gives the layout of a grown filter, see resize)."""

    def __init__(self, nbits, bulk_config=True, value_bits=0, _handle=None):
        self._lib = load_library()
        if _handle is None:
            h = ctypes.c_void_p()
            if value_bits:
                self._lib.qf_malloc_device_kv(ctypes.byref(h), nbits, value_bits)
            else:
                self._lib.qf_malloc_device(ctypes.byref(h), nbits, bulk_config)
            if not h.value:
                raise RuntimeError(f"qf_malloc_device failed for nbits={nbits} value_bits={value_bits}")
        else:
            h = _handle
        self.handle = h
        self.refresh_info()
        self.nbits = self.info["nslots"].bit_length() - 1

    @classmethod
    def load(cls, path):
        """A new filter from a file written by save() (qf_deserialize_device)."""
        h = ctypes.c_void_p()
        n = load_library().qf_deserialize_device(ctypes.byref(h), os.fsencode(path))
        if not n or not h.value:
            raise RuntimeError(f"qf_deserialize_device failed for {path}")
        return cls(0, _handle=h)

    # -- lifecycle
    def close(self):
        if self.handle is not None and self.handle.value:
            self._lib.qf_destroy_device(self.handle)
        self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def range(self):
        return self.info["range"]

    @property
    def nbytes(self):
        return self.info["total_size_in_bytes"]

    # -- bulk API (tensors are int64 device tensors)
    def bulk_insert(self, keys, flags=0, stream=None):
        self._lib.bulk_insert_async(self.handle, keys.numel(), _ptr(keys), flags, _stream_ptr(stream))

    def bulk_insert_noalloc(self, keys, flags=0, stream=None):
        """bulk_insert through the entry point that allocates nothing;
        returns its status (0, or < 0 with nothing launched)."""
        return int(self._lib.bulk_insert_noalloc(self.handle, keys.numel(), _ptr(keys), flags,
                                                 _stream_ptr(stream)))

    def bulk_insert_reduce(self, keys, flags=0, stream=None):
        self._lib.bulk_insert_reduce_async(self.handle, keys.numel(), _ptr(keys), flags, _stream_ptr(stream))

    def bulk_delete(self, keys, flags=0, stream=None):
        self._lib.bulk_delete_async(self.handle, keys.numel(), _ptr(keys), flags, _stream_ptr(stream))

    def bulk_get(self, vals, returns=None, stream=None):
        import torch
        if returns is None:
            returns = torch.empty_like(vals)
        self._lib.bulk_get_async(self.handle, vals.numel(), _ptr(vals), _ptr(returns), _stream_ptr(stream))
        return returns

    # -- sequenced batches: requests of one filter in arrival order
    @staticmethod
    def _u32_array(vals):
        return (ctypes.c_uint32 * len(vals))(*vals)

    def mutate_seq(self, keys, ends, tags, stream=None):
        """Inserts and deletes of several requests as one rebuild: request r
        owns keys[ends[r-1]:ends[r]] and has tags[r] = seq << 1 | is_delete
        (seq ascending); the result is the requests applied one at a time.
        keys is left holding the tagged hashes sorted by hash."""
        rc = self._lib.qf_bulk_mutate_seq(self.handle, keys.numel(), _ptr(keys), len(ends), self._u32_array(ends),
                                          self._u32_array(tags), _stream_ptr(stream))
        if rc != 0:
            raise ValueError(f"qf_bulk_mutate_seq rc={rc}")

    def get_seq_adjust(self, keys, ends, seqs, counts, mut_sorted, stream=None):
        """counts (bulk_get of keys before mutate_seq) -> the counts after the
        mutation requests whose sequence numbers precede each query's
        request (seqs[r] of query request r owning keys[ends[r-1]:ends[r]])."""
        rc = self._lib.qf_bulk_get_seq_adjust(self.handle, keys.numel(), _ptr(keys), len(ends), self._u32_array(ends),
                                              self._u32_array(seqs), _ptr(counts), mut_sorted.numel(),
                                              _ptr(mut_sorted), _stream_ptr(stream))
        if rc != 0:
            raise ValueError(f"qf_bulk_get_seq_adjust rc={rc}")
        return counts

    def bulk_get_misses(self, vals):
        return int(self._lib.bulk_get_misses_wrapper(self.handle, _ptr(vals), vals.numel()))

    def point_get_misses(self, keys):
        return int(self._lib.point_get_wrapper(self.handle, _ptr(keys), keys.numel()))

    def point_bulk_insert(self, keys):
        self._lib.point_bulk_insert_wrapper(self.handle, _ptr(keys), keys.numel())

    # -- device point API over an array of keys (one lane per key)
    def point_insert(self, keys, flags=0, stream=None):
        """point_insert per key; returns an int32 tensor of qf_returns."""
        import torch
        rets = torch.empty(keys.numel(), dtype=torch.int32, device=keys.device)
        self._lib.point_insert_async(self.handle, keys.numel(), _ptr(keys), flags, _ptr(rets), _stream_ptr(stream))
        return rets

    def point_insert_not_exists(self, keys, flags=0, stream=None):
        import torch
        rets = torch.empty(keys.numel(), dtype=torch.int32, device=keys.device)
        self._lib.point_insert_not_exists_async(self.handle, keys.numel(), _ptr(keys), flags, _ptr(rets),
                                                _stream_ptr(stream))
        return rets

    def point_query(self, keys, flags=0, concurrent=False, stream=None):
        import torch
        counts = torch.empty(keys.numel(), dtype=torch.int64, device=keys.device)
        self._lib.point_query_async(self.handle, keys.numel(), _ptr(keys), flags, _ptr(counts),
                                    1 if concurrent else 0, _stream_ptr(stream))
        return counts

    def reset(self, stream=None):
        return self._lib.qf_reset_device(self.handle, _stream_ptr(stream))

    def reserve(self, max_batch):
        return self._lib.qf_reserve(self.handle, max_batch)

    def set_tuning(self, region_blocks=0, sort_digit_bits=0):
        rc = self._lib.qf_set_tuning(self.handle, region_blocks, sort_digit_bits)
        if rc != 0:
            raise ValueError(f"qf_set_tuning({region_blocks}, {sort_digit_bits}) rc={rc}")
        self.refresh_info()

    def refresh_info(self):
        info = GqfInfo()
        self._lib.qf_get_info(self.handle, ctypes.byref(info))
        self.info = {n: int(getattr(info, n)) for n, _ in GqfInfo._fields_}

    # -- whole-table operations
    def enumerate(self, q_lo=0, q_hi=None, form=QF_ENUM_HASH, values=False, stream=None):
        """(hashes or keys, counts[, values]) int64 device tensors of the
        items of quotients [q_lo, q_hi), in ascending hash order."""
        import torch
        if q_hi is None:
            q_hi = self.info["nslots"]
        n = _u64()
        rc = self._lib.qf_enumerate(self.handle, q_lo, q_hi, form, None, None, None, 0, ctypes.byref(n),
                                    _stream_ptr(stream))
        if rc != 0:
            raise RuntimeError(f"qf_enumerate rc={rc}")
        m = int(n.value)
        h = torch.empty(m, dtype=torch.int64, device="cuda")
        c = torch.empty(m, dtype=torch.int64, device="cuda")
        v = torch.empty(m if values else 0, dtype=torch.int64, device="cuda")
        if m:
            rc = self._lib.qf_enumerate(self.handle, q_lo, q_hi, form, _ptr(h), _ptr(c),
                                        _ptr(v) if values else None, m, ctypes.byref(n), _stream_ptr(stream))
            if rc != 0 or int(n.value) != m:
                raise RuntimeError(f"qf_enumerate rc={rc}")
        return (h, c, v) if values else (h, c)

    def stats(self):
        """{nelts, ndistinct_elts, noccupied_slots} computed from the table."""
        a, b, c = _u64(), _u64(), _u64()
        rc = self._lib.qf_get_stats(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        if rc != 0:
            raise RuntimeError(f"qf_get_stats rc={rc}")
        return {"nelts": int(a.value), "ndistinct_elts": int(b.value), "noccupied_slots": int(c.value)}

    def cluster_census(self, nbins=0, lengths=True):
        """(lengths int64 device tensor or None, histogram or None, max length)."""
        import torch
        nc, mx = _u64(), _u64()
        rc = self._lib.qf_cluster_census(self.handle, None, 0, None, 0, ctypes.byref(nc), ctypes.byref(mx))
        if rc != 0:
            raise RuntimeError(f"qf_cluster_census rc={rc}")
        lens = torch.empty(int(nc.value), dtype=torch.int64, device="cuda") if lengths else None
        hist = torch.zeros(nbins, dtype=torch.int64, device="cuda") if nbins else None
        if lengths or nbins:
            rc = self._lib.qf_cluster_census(self.handle, _ptr(lens) if lengths else None,
                                             lens.numel() if lengths else 0, _ptr(hist) if nbins else None,
                                             nbins, ctypes.byref(nc), ctypes.byref(mx))
            if rc != 0:
                raise RuntimeError(f"qf_cluster_census rc={rc}")
        return lens, hist, int(mx.value)

    def save(self, path):
        n = self._lib.qf_serialize(self.handle, os.fsencode(path))
        if not n:
            raise RuntimeError(f"qf_serialize failed for {path}")
        return int(n)

    def resize(self, new_nbits):
        """Grow to 2^new_nbits slots keeping the hash; returns items moved."""
        m = int(self._lib.qf_resize_device(self.handle, new_nbits))
        if m < 0:
            raise RuntimeError(f"qf_resize_device({new_nbits}) rc={m}")
        self.refresh_info()
        self.nbits = new_nbits
        return m

    def merge_from(self, *sources):
        """self += every item of the sources (qf_multi_merge_device)."""
        arr = (_vp * len(sources))(*[f.handle for f in sources])
        m = int(self._lib.qf_multi_merge_device(arr, len(sources), self.handle))
        if m < 0:
            raise RuntimeError(f"qf_multi_merge_device rc={m}")
        return m

    # -- status and image
    def status(self):
        return int(self._lib.qf_get_status(self.handle))

    def clear_status(self):
        self._lib.qf_clear_status(self.handle)

    def layout_errors(self):
        """Blocks whose stored used-slot mask disagrees with their runs
        (0 for a consistent table; synchronises)."""
        bad = _u64()
        rc = self._lib.qf_check_layout(self.handle, ctypes.byref(bad))
        if rc != 0:
            raise RuntimeError(f"qf_check_layout rc={rc}")
        return int(bad.value)

    def export_image(self):
        buf = np.empty(self.nbytes, dtype=np.uint8)
        rc = self._lib.qf_export_image(self.handle, buf.ctypes.data_as(ctypes.c_void_p), buf.size)
        if rc != 0:
            raise RuntimeError(f"qf_export_image rc={rc}")
        return buf

    def import_image(self, img):
        img = np.ascontiguousarray(img, dtype=np.uint8)
        rc = self._lib.qf_import_image(self.handle, img.ctypes.data_as(ctypes.c_void_p), img.size)
        if rc != 0:
            raise RuntimeError(f"qf_import_image rc={rc}")

    def export_image_device(self, out=None, stream=None):
        """The 81-byte-block image into a uint8 device tensor (stream-ordered)."""
        import torch
        if out is None:
            out = torch.empty(self.nbytes, dtype=torch.uint8, device="cuda")
        rc = self._lib.qf_export_image_device(self.handle, _ptr(out), out.numel(), _stream_ptr(stream))
        if rc != 0:
            raise RuntimeError(f"qf_export_image_device rc={rc}")
        return out

    def import_image_device(self, img, stream=None):
        rc = self._lib.qf_import_image_device(self.handle, _ptr(img), img.numel(), _stream_ptr(stream))
        if rc != 0:
            raise RuntimeError(f"qf_import_image_device rc={rc}")


class GQFGroup:
    """Several GQF filters served by shared launches (one device key array
    cut into one segment per filter)."""

    def __init__(self, filters, graphs=True):
        self._lib = load_library()
        self.filters = list(filters)
        arr = (_vp * len(self.filters))(*[f.handle for f in self.filters])
        h = ctypes.c_void_p()
        rc = self._lib.gqf_group_create(ctypes.byref(h), arr, len(self.filters))
        if rc != 0:
            raise RuntimeError(f"gqf_group_create rc={rc}")
        self.handle = h
        self._lib.gqf_group_set_graphs(h, 1 if graphs else 0)

    @staticmethod
    def _offsets(offsets):
        return (_u64 * len(offsets))(*[int(x) for x in offsets])

    def bulk_insert(self, keys, offsets, flags=0, stream=None):
        rc = self._lib.gqf_group_bulk_insert(self.handle, _ptr(keys), self._offsets(offsets), flags,
                                             _stream_ptr(stream))
        if rc != 0:
            raise RuntimeError(f"gqf_group_bulk_insert rc={rc}")

    def bulk_get(self, vals, offsets, returns=None, stream=None):
        import torch
        if returns is None:
            returns = torch.empty_like(vals)
        rc = self._lib.gqf_group_bulk_get(self.handle, _ptr(vals), self._offsets(offsets), _ptr(returns),
                                          _stream_ptr(stream))
        if rc != 0:
            raise RuntimeError(f"gqf_group_bulk_get rc={rc}")
        return returns

    def autotune(self, offsets):
        rc = self._lib.gqf_group_autotune(self.handle, self._offsets(offsets))
        if rc != 0:
            raise RuntimeError(f"gqf_group_autotune rc={rc}")
        for f in self.filters:
            f.refresh_info()

    @property
    def workgroups_per_cu(self):
        """Workgroups per CU of the group's persistent rebuild."""
        v = ctypes.c_uint32()
        rc = self._lib.gqf_group_get_workgroups_per_cu(self.handle, ctypes.byref(v))
        if rc != 0:
            raise RuntimeError(f"gqf_group_get_workgroups_per_cu rc={rc}")
        return v.value

    @workgroups_per_cu.setter
    def workgroups_per_cu(self, per_cu):
        rc = self._lib.gqf_group_set_workgroups_per_cu(self.handle, int(per_cu))
        if rc != 0:
            raise ValueError(f"gqf_group_set_workgroups_per_cu({per_cu}) rc={rc}")

    def set_wait_bound(self, max_polls):
        """Bound on an odd region's neighbour wait in the persistent rebuild
        (regions whose wait runs out are redone by the wide pass)."""
        rc = self._lib.gqf_group_set_wait_bound(self.handle, int(max_polls))
        if rc != 0:
            raise RuntimeError(f"gqf_group_set_wait_bound rc={rc}")

    def close(self):
        if self.handle is not None and self.handle.value:
            self._lib.gqf_group_destroy(self.handle)
        self.handle = None


def batch_pieces(n, max_piece):
    """Start indices (p + 1 of them) of the pieces a bulk mutation of n keys
    is cut into, pieces of at most max_piece keys."""
    L = load_library()
    p = int(L.qf_batch_pieces(n, max_piece, None, 0))
    arr = (_u64 * (p + 1))()
    L.qf_batch_pieces(n, max_piece, arr, p + 1)
    return [int(x) for x in arr]


def autotune(nbits, load):
    """(region_blocks, sort_digit_bits) the autotuner picks for a table of
    2^nbits slots loaded to `load` in one batch."""
    r, d = ctypes.c_uint32(), ctypes.c_uint32()
    rc = load_library().gqf_autotune(nbits, load, ctypes.byref(r), ctypes.byref(d))
    if rc != 0:
        raise RuntimeError(f"gqf_autotune rc={rc}")
    return r.value, d.value


def set_wide_wait_bound(max_polls):
    """Bound on the polls of each wait between the wide pass's classes
    (process-wide, default 2^24); returns the previous bound.  A wait that
    runs out leaves its class's regions unchanged and sets the filter's
    QF_STATUS_CLUSTER_OVERFLOW: a hook for tests of that path."""
    return int(load_library().qf_set_wide_wait_bound(int(max_polls)))


def fill_splitmix64(out, seed, skip=0, stream=None):
    """out (int64 device tensor) <- splitmix64 outputs skip.. of `seed`."""
    load_library().gqf_fill_splitmix64(_ptr(out), out.numel(), seed, skip, _stream_ptr(stream))
    return out


def profile_enable(on=True):
    load_library().gqf_profile_enable(1 if on else 0)


def profile_read(name):
    ms = ctypes.c_double()
    cnt = _u64()
    load_library().gqf_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
    return ms.value, int(cnt.value)


def last_error():
    return int(load_library().qf_last_error())
