"""Serving loop: insert / query / delete requests of many sizes against many
resident filters, with continuous batching.

No counterpart in the reference (it exposes only per-call bulk APIs).  The
loop keeps every admitted filter resident in HBM and, each round:

  * admits arrived requests into a FIFO per filter;
  * for every filter whose previous batch has left the device (at most
    `max_inflight` batches in flight per filter), takes the requests at the
    head of its FIFO -- of every kind, in arrival order, up to
    `max_batch_keys` keys and 256 requests of each side -- and launches
    them as one sequenced batch on the filter's stream: ONE bulk_get of all
    the batch's query keys, taken before the batch's mutations; ONE
    sequenced mutation (qf_bulk_mutate_seq) applying all its inserts and
    deletes in request order with one sort and one rebuild; and, when a
    query follows a mutation in the batch, ONE correction
    (qf_bulk_get_seq_adjust) that replays on each query's count the
    mutations of its hash that arrived before it;
  * filters are spread over `n_streams` HIP streams, so batches of
    different filters overlap on the device while requests keep arriving.

Per-filter order: a filter's batches are issued in FIFO order on a single
stream, and within a batch every query's count is the count after exactly
the mutations that arrived before it -- the same result as the requests
applied one by one in arrival order ("alone").

Admission: a filter is created on its first request if its table and
workspace fit the HBM budget; otherwise the least recently used idle filter
is evicted (its 81-byte image exported to host memory, device memory freed)
until it fits.  An evicted filter is re-admitted by importing its image.
"""
import collections
import time

import numpy as np
import torch

from . import gqf as G

OPS = ("get", "insert", "delete")


class Request:
    __slots__ = ("rid", "fid", "op", "keys", "t_arrival", "t_launch", "t_done", "result")

    def __init__(self, rid, fid, op, keys, t_arrival):
        self.rid, self.fid, self.op, self.keys, self.t_arrival = rid, fid, op, keys, t_arrival
        self.t_launch = self.t_done = None
        self.result = None


class _Filter:
    def __init__(self, fid, nbits):
        self.fid, self.nbits = fid, nbits
        self.q = None            # resident GQF, or None
        self.image = None        # host image while evicted
        self.fifo = collections.deque()
        self.inflight = collections.deque()  # (event, [requests], keep-alive tensors)
        self.last_use = 0.0
        self.stream = None


def filter_bytes(nbits, max_batch_keys):
    """Device bytes a resident filter takes: table + point locks + bulk
    workspace for batches of max_batch_keys (sort buffer, histograms,
    run-length buffers, region tables)."""
    nslots = 1 << nbits
    xn = nslots + int(10 * (nslots ** 0.5))
    nblocks = (xn + 63) // 64
    table = nblocks * 128 + 4096
    locks = (xn // 8192 + 2) * 64 * 4
    ws = max_batch_keys * 8 * 4 + 1024 * ((max_batch_keys + 4095) // 4096) * 4
    regions = (nblocks // 16 + 2) * 12
    return table + locks + ws + regions


class Server:
    def __init__(self, budget_bytes, n_streams=4, max_batch_keys=1 << 22, max_inflight=2):
        self.budget = budget_bytes
        self.used = 0
        self.max_batch = max_batch_keys
        self.max_inflight = max_inflight
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]
        self.filters = {}
        self.evictions = 0
        self.admissions = 0
        self.batches = 0

    # -- filters and admission
    def add_filter(self, fid, nbits):
        f = _Filter(fid, nbits)
        f.stream = self.streams[len(self.filters) % len(self.streams)]
        self.filters[fid] = f

    def _evict_one(self, need):
        idle = [f for f in self.filters.values() if f.q is not None and not f.fifo and not f.inflight]
        if not idle:
            return False
        v = min(idle, key=lambda f: f.last_use)
        torch.cuda.current_stream().wait_stream(v.stream)
        v.image = v.q.export_image()
        v.q.close()
        v.q = None
        self.used -= filter_bytes(v.nbits, self.max_batch)
        self.evictions += 1
        return True

    def _admit(self, f):
        need = filter_bytes(f.nbits, self.max_batch)
        while self.used + need > self.budget:
            if not self._evict_one(need):
                return False
        f.q = G.GQF(f.nbits)
        f.q.reserve(self.max_batch)
        if f.image is not None:
            f.q.import_image(f.image)
            f.image = None
        self.used += need
        self.admissions += 1
        return True

    # -- requests
    def submit(self, req):
        self.filters[req.fid].fifo.append(req)

    def _retire(self, now):
        done = 0
        for f in self.filters.values():
            while f.inflight and f.inflight[0][0].query():
                _, reqs, _keep = f.inflight.popleft()
                for r in reqs:
                    r.t_done = now
                done += len(reqs)
        return done

    def _take(self, f):
        """The FIFO prefix of one sequenced batch: (gets, mutations), each a
        list of (sequence number, request) in arrival order."""
        gets, muts = [], []
        n_get = n_mut = 0
        seq = 0
        while f.fifo:
            r = f.fifo[0]
            k = r.keys.numel()
            if r.op == "get":
                if len(gets) == G.QF_SEQ_MAX_REQUESTS or (gets and n_get + k > self.max_batch):
                    break
                gets.append((seq, r))
                n_get += k
            else:
                if len(muts) == G.QF_SEQ_MAX_REQUESTS or (muts and n_mut + k > self.max_batch):
                    break
                muts.append((seq, r))
                n_mut += k
            f.fifo.popleft()
            seq += 1
        return gets, muts

    @staticmethod
    def _concat(side):
        keys = [r.keys for _, r in side]
        batch = torch.cat(keys) if len(keys) > 1 else keys[0].clone()
        ends = np.cumsum([k.numel() for k in keys]).tolist()
        return batch, ends

    def _launch(self, f, now):
        gets, muts = self._take(f)
        reqs = [r for _, r in sorted(gets + muts, key=lambda x: x[0])]
        s = f.stream
        with torch.cuda.stream(s):
            keep = []
            if gets:
                gk, g_ends = self._concat(gets)
                out = f.q.bulk_get(gk, stream=s)  # the counts before this batch's mutations
                keep += [gk, out]
            if muts:
                mk, m_ends = self._concat(muts)
                f.q.mutate_seq(mk, m_ends, [(q << 1) | (r.op == "delete") for q, r in muts], s)
                keep.append(mk)
                if gets and gets[-1][0] > muts[0][0]:  # a query follows a mutation
                    f.q.get_seq_adjust(gk, g_ends, [q for q, _ in gets], out, mk, s)
            if gets:
                o = 0
                for _, r in gets:
                    r.result = out[o:o + r.keys.numel()]
                    o += r.keys.numel()
            ev = torch.cuda.Event()
            ev.record(s)
        for r in reqs:
            r.t_launch = now
        f.inflight.append((ev, reqs, keep))
        f.last_use = now
        self.batches += 1

    def pump(self, now):
        """One scheduling round; returns the number of requests retired."""
        done = self._retire(now)
        for f in self.filters.values():
            if not f.fifo or len(f.inflight) >= self.max_inflight:
                continue
            if f.q is None and not self._admit(f):
                continue
            self._launch(f, now)
        return done

    def idle(self):
        return all(not f.fifo and not f.inflight for f in self.filters.values())

    def run(self, requests):
        """Serve `requests` (sorted by t_arrival, seconds from the start) in
        real time; returns the wall time until the last one completed."""
        t0 = time.perf_counter()
        i = 0
        while i < len(requests) or not self.idle():
            now = time.perf_counter() - t0
            while i < len(requests) and requests[i].t_arrival <= now:
                self.submit(requests[i])
                i += 1
            self.pump(now)
        return time.perf_counter() - t0

    def close(self):
        torch.cuda.synchronize()
        for f in self.filters.values():
            if f.q is not None:
                f.q.close()
                f.q = None


def statuses(server):
    """QF_STATUS_* word of every resident filter (synchronises)."""
    return {fid: f.q.status() for fid, f in server.filters.items() if f.q is not None}


def run_plain(filters_nbits, requests, max_batch_keys=1 << 22):
    """The baseline: every request as its own bulk call, in arrival order,
    on one stream (filters all resident).  Returns (wall time, latencies)."""
    qs = {fid: G.GQF(nb) for fid, nb in filters_nbits.items()}
    for q in qs.values():
        q.reserve(max_batch_keys)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pending = collections.deque()
    i = 0
    while i < len(requests) or pending:
        now = time.perf_counter() - t0
        while pending and pending[0][0].query():
            _, r = pending.popleft()
            r.t_done = now
        if i < len(requests) and requests[i].t_arrival <= now:
            r = requests[i]
            i += 1
            q = qs[r.fid]
            with torch.cuda.stream(s):
                k = r.keys.clone()
                if r.op == "get":
                    r.result = q.bulk_get(k, stream=s)
                elif r.op == "insert":
                    q.bulk_insert(k, 0, s)
                else:
                    q.bulk_delete(k, 0, s)
                ev = torch.cuda.Event()
                ev.record(s)
            r.t_launch = now
            pending.append((ev, r))
    wall = time.perf_counter() - t0
    torch.cuda.synchronize()
    for q in qs.values():
        q.close()
    return wall


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
