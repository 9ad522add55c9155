"""Serving loop: insert / query / delete requests of many sizes against many
resident filters, with continuous batching -- the Python face of the native
loop in libgqf (qf_server_*, cqf_amd/csrc/gqf_server.hip).

No counterpart in the reference (it exposes only per-call bulk APIs).  The
loop runs on a C++ thread inside the library.  Each request is submitted
with its arrival time; the loop releases it into its filter's FIFO at that
time and, each round, retires finished batches and launches, for every
filter with queued requests and fewer than `max_inflight` batches on the
device, the head of its FIFO -- requests of every kind, in arrival order,
up to `max_batch_keys` keys and 256 requests per side -- as ONE sequenced
batch on the filter's stream: a gather of the get keys and ONE bulk_get
(the counts before the batch's mutations); a gather of the insert / delete
keys and ONE qf_bulk_mutate_seq (one sort and one rebuild, each hash's
items applied in request order); when a get follows a mutation, ONE
qf_bulk_get_seq_adjust (each count becomes the count after exactly the
mutations that arrived before it); and a scatter of the counts to each
get's output.  Filters are spread over `n_streams` streams, so batches of
different filters overlap on the device while requests keep arriving.

Per-filter order: a filter's batches go out in FIFO order on one stream,
and within a batch every get's count is the count after exactly the
mutations that arrived before it -- the same result as the requests applied
one by one in arrival order ("alone").

Admission: a filter is created on its first request if its device memory
fits the HBM budget; otherwise the least recently used idle filter is
evicted (its 81-byte image exported to host memory) until it fits.  An
evicted filter is re-admitted by importing its image.

The plain mode of the same loop is the baseline: every request its own
bulk call, in arrival order, on one stream, every filter resident.
"""
import ctypes

import numpy as np
import torch

from . import gqf as G

OPS = ("get", "insert", "delete")
OPCODE = {"get": 0, "insert": 1, "delete": 2}


class Request:
    __slots__ = ("rid", "fid", "op", "keys", "t_arrival", "t_launch", "t_done", "result", "_sid")

    def __init__(self, rid, fid, op, keys, t_arrival):
        self.rid, self.fid, self.op, self.keys, self.t_arrival = rid, fid, op, keys, t_arrival
        self.t_launch = self.t_done = None
        self.result = None
        self._sid = -1


def filter_bytes(nbits, max_batch_keys):
    """Device bytes a resident filter takes in the loop's budget: table,
    point locks, bulk workspace for batches of max_batch_keys (sort buffer,
    histograms, run-length buffers, region tables) and the loop's three
    batch buffers (as gqf_server.hip counts them)."""
    nslots = 1 << nbits
    xn = nslots + int(10 * (nslots ** 0.5))
    nblocks = (xn + 63) // 64
    table = nblocks * 128 + 4096
    locks = (xn // 8192 + 2) * 64 * 4
    ws = max_batch_keys * 8 * 4 + 1024 * ((max_batch_keys + 4095) // 4096) * 4
    regions = (nblocks // 16 + 2) * 12
    return table + locks + ws + regions + 3 * max_batch_keys * 8


class _Stats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_uint64), ("evictions", ctypes.c_uint64), ("admissions", ctypes.c_uint64)]


def _lib():
    L = G.load_library()
    if not getattr(L, "_server_bound", False):
        vp, u64 = ctypes.c_void_p, ctypes.c_uint64
        L.qf_server_create.argtypes = [ctypes.POINTER(vp), u64, ctypes.c_uint32, u64, ctypes.c_uint32, ctypes.c_int]
        L.qf_server_create.restype = ctypes.c_int
        L.qf_server_add_filter.argtypes = [vp, ctypes.c_uint32, ctypes.c_int]
        L.qf_server_add_filter.restype = ctypes.c_int
        L.qf_server_submit.argtypes = [vp, u64, ctypes.c_uint32, ctypes.c_int, vp, u64, vp]
        L.qf_server_submit.restype = ctypes.c_int64
        L.qf_server_run.argtypes = [vp]
        L.qf_server_run.restype = ctypes.c_int
        L.qf_server_wait.argtypes = [vp, ctypes.POINTER(u64)]
        L.qf_server_wait.restype = ctypes.c_int
        L.qf_server_request_times.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                              ctypes.POINTER(u64)]
        L.qf_server_request_times.restype = ctypes.c_int
        L.qf_server_get_stats.argtypes = [vp, ctypes.POINTER(_Stats)]
        L.qf_server_get_stats.restype = ctypes.c_int
        L.qf_server_filter.argtypes = [vp, ctypes.c_uint32]
        L.qf_server_filter.restype = vp
        L.qf_server_destroy.argtypes = [vp]
        L._server_bound = True
    return L


class Server:
    """The native serving loop (qf_server_*).  add_filter() every filter,
    then run() a list of Requests sorted by arrival time (seconds from the
    start); each get's result is a device tensor of counts."""

    def __init__(self, budget_bytes, n_streams=4, max_batch_keys=1 << 22, max_inflight=2, plain=False):
        self._lib = _lib()
        h = ctypes.c_void_p()
        rc = self._lib.qf_server_create(ctypes.byref(h), int(budget_bytes), int(n_streams), int(max_batch_keys),
                                        int(max_inflight), 1 if plain else 0)
        if rc != 0:
            raise RuntimeError(f"qf_server_create rc={rc}")
        self.handle = h
        self.nbits = {}

    def add_filter(self, fid, nbits):
        rc = self._lib.qf_server_add_filter(self.handle, int(fid), int(nbits))
        if rc != 0:
            raise RuntimeError(f"qf_server_add_filter({fid}, {nbits}) rc={rc}")
        self.nbits[fid] = nbits

    def run(self, requests):
        """Serves `requests` in real time on the loop thread; returns the
        wall time (s) until the last one completed.  Sets each request's
        t_launch / t_done (s from the start) and, for gets, result."""
        L = self._lib
        for r in requests:
            if r.op == "get":
                r.result = torch.empty_like(r.keys)
        torch.cuda.synchronize()  # keys and outputs exist before the clock starts
        for r in requests:
            sid = L.qf_server_submit(self.handle, int(r.t_arrival * 1e9), int(r.fid), OPCODE[r.op],
                                     r.keys.data_ptr(), r.keys.numel(),
                                     r.result.data_ptr() if r.op == "get" else None)
            if sid < 0:
                raise RuntimeError(f"qf_server_submit rc={sid} for request {r.rid}")
            r._sid = sid
        if L.qf_server_run(self.handle) != 0:
            raise RuntimeError("qf_server_run failed")
        wall = ctypes.c_uint64()
        rc = L.qf_server_wait(self.handle, ctypes.byref(wall))
        if rc != 0:
            raise RuntimeError(f"qf_server_wait rc={rc}")
        ta, tl, td = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        for r in requests:
            L.qf_server_request_times(self.handle, r._sid, ctypes.byref(ta), ctypes.byref(tl), ctypes.byref(td))
            r.t_launch, r.t_done = tl.value * 1e-9, td.value * 1e-9
        return wall.value * 1e-9

    def _stats(self):
        st = _Stats()
        self._lib.qf_server_get_stats(self.handle, ctypes.byref(st))
        return st

    @property
    def batches(self):
        return int(self._stats().batches)

    @property
    def evictions(self):
        return int(self._stats().evictions)

    @property
    def admissions(self):
        return int(self._stats().admissions)

    def filter(self, fid):
        """The resident filter fid as a GQF view (None if evicted); the
        loop owns it."""
        h = self._lib.qf_server_filter(self.handle, int(fid))
        return G.GQF(0, _handle=ctypes.c_void_p(h)) if h else None

    def close(self):
        if self.handle is not None and self.handle.value:
            self._lib.qf_server_destroy(self.handle)
        self.handle = None


def statuses(server):
    """QF_STATUS_* word and used-slot-mask errors of every resident filter
    (synchronises)."""
    out = {}
    for fid in server.nbits:
        v = server.filter(fid)
        if v is not None:
            out[fid] = v.status() | (0x100 if v.layout_errors() else 0)
            v.handle = None  # a view: the loop destroys its filters
    return out


def run_plain(filters_nbits, requests, max_batch_keys=1 << 22):
    """The baseline: every request as its own bulk call, in arrival order,
    on one stream (filters all resident), on the same native loop.  Returns
    the wall time (s) until the last request completed."""
    srv = Server(budget_bytes=1 << 62, n_streams=1, max_batch_keys=max_batch_keys, plain=True)
    try:
        for fid, nb in filters_nbits.items():
            srv.add_filter(fid, nb)
        return srv.run(requests)
    finally:
        srv.close()


def replay_alone(filters_nbits, requests):
    """Reference results: each filter's requests applied one at a time in
    arrival order, synchronously.  Returns {rid: counts (numpy)} for gets."""
    qs = {fid: G.GQF(nb) for fid, nb in filters_nbits.items()}
    out = {}
    for r in requests:
        q = qs[r.fid]
        k = r.keys.clone()
        if r.op == "get":
            out[r.rid] = q.bulk_get(k).cpu().numpy()
        elif r.op == "insert":
            q.bulk_insert(k)
        else:
            q.bulk_delete(k)
    torch.cuda.synchronize()
    for q in qs.values():
        q.close()
    return out


def latency_stats(requests):
    lat = np.array([(r.t_done - r.t_arrival) * 1e3 for r in requests if r.t_done is not None])
    return {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3)}
