// gqf_server.hip -- the serving loop: many resident filters, requests of
// any size and kind arriving over time, continuous batching on a native
// thread.  No counterpart in the reference (it offers per-call bulk APIs).
//
// Requests are submitted with an arrival time; the loop thread releases
// each into its filter's FIFO at that time and, every round:
//   * retires batches whose completion event has fired (their requests are
//     done: the time is recorded);
//   * for every filter with queued requests and fewer than max_inflight
//     batches on the device, admits the filter if it is not resident
//     (evicting least recently used idle filters to pinned host copies of
//     their device blocks until it fits the HBM budget) and launches the head of its FIFO -- requests of
//     every kind, in arrival order, up to max_batch_keys keys and
//     QF_SEQ_MAX_REQUESTS requests per side -- as ONE sequenced batch on
//     the filter's stream:
//       gather the get keys, bulk_get (counts before the batch's mutations),
//       gather the insert / delete keys, qf_bulk_mutate_seq (one sort and
//       one rebuild, each hash's items applied in request order),
//       qf_bulk_get_seq_adjust when a get follows a mutation (each count
//       becomes the count after exactly the mutations that arrived before
//       its request), scatter the counts to each get's output array;
//   * filters are spread over n_streams streams, so batches of different
//     filters overlap on the device while requests keep arriving.
// A filter's batches go out in FIFO order on one stream, so every get sees
// exactly the mutations that arrived before it: the results equal the
// requests applied one by one in arrival order.
//
// The plain mode is the baseline it is measured against: every request is
// its own bulk call, in arrival order, on one stream, all filters resident.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/gqf.h"
#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {

struct CopyDesc {  // one request's keys or counts: n words from src to dst
  const uint64_t* src;
  uint64_t* dst;
  uint64_t n;
};

// one workgroup per descriptor (blockIdx.y), threads striding over its words
__global__ void __launch_bounds__(256) gqf_server_copy(const CopyDesc* __restrict__ d) {
  const CopyDesc c = d[blockIdx.y];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += (uint64_t)gridDim.x * blockDim.x)
    c.dst[i] = c.src[i];
}

struct Req {
  uint32_t fid;
  int op;
  const uint64_t* keys;
  uint64_t n;
  uint64_t* counts;
  uint64_t t_arrival, t_launch = 0, t_done = 0;
  bool done = false;
};

struct Batch {
  hipEvent_t ev;
  std::vector<int64_t> rids;
  int slot;  // descriptor slot it used
};

struct SFilter {
  uint32_t fid = 0;
  int nbits = 0;
  QF* qf = nullptr;
  uint8_t* host_blocks = nullptr;  // pinned copy of the device blocks while evicted
  uint64_t host_bytes = 0;
  uint32_t host_status = 0;
  bool evicted = false;
  std::deque<int64_t> fifo;
  std::deque<Batch> inflight;
  uint64_t last_use = 0;
  hipStream_t stream = nullptr;
  uint64_t* gbuf = nullptr;  // concatenated get keys
  uint64_t* mbuf = nullptr;  // concatenated mutation keys (left sorted and tagged)
  uint64_t* cbuf = nullptr;  // counts of the concatenated gets
  CopyDesc* desc = nullptr;  // pinned: [max_inflight][3][QF_SEQ_MAX_REQUESTS] (get keys, mutation keys, counts)
  std::vector<bool> slot_busy;
};

constexpr int kMaxReq = QF_SEQ_MAX_REQUESTS;

}  // namespace

struct qf_server {
  uint64_t budget = 0, used = 0, max_batch = 0;
  uint32_t max_inflight = 2;
  int plain = 0;
  std::vector<hipStream_t> streams;
  std::map<uint32_t, SFilter> filters;
  std::vector<Req> reqs;
  size_t next_release = 0;  // requests are released in submission order (arrival times must not decrease)
  std::mutex mu;
  std::condition_variable cv;
  std::thread th;
  std::atomic<bool> stop{false};
  bool running = false;
  uint64_t n_done = 0;
  std::chrono::steady_clock::time_point t0;
  qf_server_stats stats{};
  std::vector<hipEvent_t> event_pool;
  // plain mode
  uint64_t* staging = nullptr;
  int err = 0;

  uint64_t now_ns() const {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0)
        .count();
  }
};

namespace {

void check(qf_server* s, hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (!s->err) s->err = (int)e;
  fprintf(stderr, "libgqf server: %s failed: %s\n", what, hipGetErrorString(e));
}

hipEvent_t take_event(qf_server* s) {
  if (!s->event_pool.empty()) {
    hipEvent_t e = s->event_pool.back();
    s->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  check(s, hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  return e;
}

// device bytes a resident filter holds: table, point locks, bulk workspace
// for batches of max_batch keys, and the loop's three batch buffers
uint64_t filter_bytes(int nbits, uint64_t max_batch) {
  const uint64_t nslots = 1ULL << nbits;
  const uint64_t xn = nslots + (uint64_t)(10 * sqrt((double)nslots));
  const uint64_t nblocks = (xn + 63) / 64;
  const uint64_t table = gqf_table_bytes(nblocks);
  const uint64_t locks = (xn / GQF_REF_LOCK_SLOTS + 2) * GQF_REF_LOCK_DIST * 4;
  const uint64_t ws = max_batch * 8 * 4 + 1024 * ((max_batch + 4095) / 4096) * 4;
  const uint64_t regions = (nblocks / GQF_REGION_BLOCKS_MIN + 2) * 12;
  return table + locks + ws + regions + 3 * max_batch * 8;
}

void release_buffers(SFilter& f) {
  if (f.gbuf) (void)hipFree(f.gbuf);
  if (f.mbuf) (void)hipFree(f.mbuf);
  if (f.cbuf) (void)hipFree(f.cbuf);
  f.gbuf = f.mbuf = f.cbuf = nullptr;
}

bool evict_one(qf_server* s) {
  SFilter* v = nullptr;
  for (auto& kv : s->filters) {
    SFilter& f = kv.second;
    if (f.qf && f.fifo.empty() && f.inflight.empty() && (!v || f.last_use < v->last_use)) v = &f;
  }
  if (!v) return false;
  // the device blocks as they are (no format change) into a pinned host
  // buffer kept for the filter's next eviction, and the status word
  QF d;
  if (!gqf_filter_host(v->qf, &d)) return false;
  const uint64_t bytes = d.nblocks * GQF_BLOCK_BYTES;
  if (v->host_bytes < bytes) {
    if (v->host_blocks) (void)hipHostFree(v->host_blocks);
    v->host_blocks = nullptr;
    v->host_bytes = 0;
    check(s, hipHostMalloc((void**)&v->host_blocks, bytes, hipHostMallocDefault), "hipHostMalloc");
    if (!v->host_blocks) return false;
    v->host_bytes = bytes;
  }
  check(s, hipMemcpyAsync(v->host_blocks, d.blocks, bytes, hipMemcpyDeviceToHost, v->stream), "hipMemcpyAsync");
  check(s, hipMemcpyAsync(&v->host_status, d.status, sizeof(uint32_t), hipMemcpyDeviceToHost, v->stream),
        "hipMemcpyAsync");
  check(s, hipStreamSynchronize(v->stream), "hipStreamSynchronize");
  v->evicted = true;
  qf_destroy_device(v->qf);
  v->qf = nullptr;
  release_buffers(*v);
  s->used -= filter_bytes(v->nbits, s->max_batch);
  s->stats.evictions++;
  return true;
}

bool admit(qf_server* s, SFilter& f) {
  const uint64_t need = filter_bytes(f.nbits, s->max_batch);
  while (s->used + need > s->budget)
    if (!evict_one(s)) return false;
  qf_malloc_device(&f.qf, f.nbits, true);
  if (!f.qf) return false;
  qf_reserve(f.qf, s->max_batch);
  check(s, hipMalloc(&f.gbuf, s->max_batch * 8), "hipMalloc");
  check(s, hipMalloc(&f.mbuf, s->max_batch * 8), "hipMalloc");
  check(s, hipMalloc(&f.cbuf, s->max_batch * 8), "hipMalloc");
  if (f.evicted) {
    QF d;
    if (!gqf_filter_host(f.qf, &d) || d.nblocks * GQF_BLOCK_BYTES > f.host_bytes) {
      s->err = -4;
    } else {
      check(s, hipMemcpyAsync(d.blocks, f.host_blocks, d.nblocks * GQF_BLOCK_BYTES, hipMemcpyHostToDevice, f.stream),
            "hipMemcpyAsync");
      check(s, hipMemcpyAsync(d.status, &f.host_status, sizeof(uint32_t), hipMemcpyHostToDevice, f.stream),
            "hipMemcpyAsync");
      // the host words are read by the copies: wait before they can change
      check(s, hipStreamSynchronize(f.stream), "hipStreamSynchronize");
    }
    f.evicted = false;
  }
  s->used += need;
  s->stats.admissions++;
  return true;
}

void launch_copies(qf_server* s, SFilter& f, CopyDesc* d, uint32_t nd, uint64_t longest) {
  if (nd == 0) return;
  uint64_t gx = (longest + 255) / 256;
  gx = gx < 1 ? 1 : (gx > 256 ? 256 : gx);
  hipLaunchKernelGGL(gqf_server_copy, dim3((uint32_t)gx, nd), dim3(256), 0, f.stream, (const CopyDesc*)d);
  check(s, hipGetLastError(), "gqf_server_copy");
}

// one sequenced batch from the head of f's FIFO
void launch_batch(qf_server* s, SFilter& f, uint64_t now) {
  int slot = 0;
  while (slot < (int)s->max_inflight && f.slot_busy[slot]) slot++;
  if (slot == (int)s->max_inflight) return;
  // the three descriptor tables of this slot are read by kernels still in
  // flight until the batch's event fires, so each use has its own
  CopyDesc* dg = f.desc + (size_t)slot * 3 * kMaxReq;  // get keys in
  CopyDesc* dm = dg + kMaxReq;                         // mutation keys in
  CopyDesc* dc = dm + kMaxReq;                         // counts out
  std::vector<int64_t> rids;
  uint32_t ng = 0, nm = 0, g_end[kMaxReq], g_seq[kMaxReq], m_end[kMaxReq], m_tag[kMaxReq];
  uint64_t n_get = 0, n_mut = 0, long_g = 0, long_m = 0;
  uint32_t seq = 0, first_mut_seq = ~0u, last_get_seq = 0;
  while (!f.fifo.empty()) {
    const int64_t rid = f.fifo.front();
    const Req& r = s->reqs[(size_t)rid];
    if (r.op == QF_SERVER_GET) {
      if (ng == kMaxReq || (ng && n_get + r.n > s->max_batch)) break;
      dg[ng] = CopyDesc{r.keys, f.gbuf + n_get, r.n};
      n_get += r.n;
      g_end[ng] = (uint32_t)n_get;
      g_seq[ng] = seq;
      last_get_seq = seq;
      long_g = std::max(long_g, r.n);
      ng++;
    } else {
      if (nm == kMaxReq || (nm && n_mut + r.n > s->max_batch)) break;
      dm[nm] = CopyDesc{r.keys, f.mbuf + n_mut, r.n};
      n_mut += r.n;
      m_end[nm] = (uint32_t)n_mut;
      m_tag[nm] = (seq << 1) | (r.op == QF_SERVER_DELETE ? 1u : 0u);
      if (first_mut_seq == ~0u) first_mut_seq = seq;
      long_m = std::max(long_m, r.n);
      nm++;
    }
    rids.push_back(rid);
    f.fifo.pop_front();
    seq++;
  }
  if (rids.empty()) return;
  const hipStream_t st = f.stream;
  if (ng) {
    launch_copies(s, f, dg, ng, long_g);
    bulk_get_async(f.qf, n_get, f.gbuf, f.cbuf, st);
  }
  if (nm) {
    launch_copies(s, f, dm, nm, long_m);
    if (qf_bulk_mutate_seq(f.qf, n_mut, f.mbuf, nm, m_end, m_tag, st) != 0) s->err = -5;
    if (ng && last_get_seq > first_mut_seq &&
        qf_bulk_get_seq_adjust(f.qf, n_get, f.gbuf, ng, g_end, g_seq, f.cbuf, n_mut, f.mbuf, st) != 0)
      s->err = -5;
  }
  if (ng) {  // the counts back to each get's output array
    uint64_t off = 0;
    uint32_t k = 0;
    for (int64_t rid : rids) {
      const Req& r = s->reqs[(size_t)rid];
      if (r.op != QF_SERVER_GET) continue;
      dc[k] = CopyDesc{f.cbuf + off, r.counts, r.n};
      off += r.n;
      k++;
    }
    launch_copies(s, f, dc, ng, long_g);
  }
  Batch b;
  b.ev = take_event(s);
  b.rids = std::move(rids);
  b.slot = slot;
  check(s, hipEventRecord(b.ev, st), "hipEventRecord");
  for (int64_t rid : b.rids) s->reqs[(size_t)rid].t_launch = now;
  f.slot_busy[slot] = true;
  f.inflight.push_back(std::move(b));
  f.last_use = now;
  s->stats.batches++;
}

// plain mode: every request its own bulk call, in arrival order, on one stream
void launch_plain(qf_server* s, SFilter& f, int64_t rid, uint64_t now) {
  Req& r = s->reqs[(size_t)rid];
  const hipStream_t st = s->streams[0];
  if (r.op == QF_SERVER_GET) {
    bulk_get_async(f.qf, r.n, (uint64_t*)r.keys, r.counts, st);
  } else {
    // the bulk mutations leave their keys sorted: work on a copy
    check(s, hipMemcpyAsync(s->staging, r.keys, r.n * 8, hipMemcpyDeviceToDevice, st), "hipMemcpyAsync");
    if (r.op == QF_SERVER_INSERT) bulk_insert_async(f.qf, r.n, s->staging, 0, st);
    else bulk_delete_async(f.qf, r.n, s->staging, 0, st);
  }
  Batch b;
  b.ev = take_event(s);
  b.rids.push_back(rid);
  b.slot = 0;
  check(s, hipEventRecord(b.ev, st), "hipEventRecord");
  r.t_launch = now;
  f.inflight.push_back(std::move(b));
  s->stats.batches++;
}

void loop(qf_server* s) {
  std::deque<int64_t> plain_queue;  // arrival order across filters
  for (;;) {
    const uint64_t now = s->now_ns();
    bool stop_now = false;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      while (s->next_release < s->reqs.size() && s->reqs[s->next_release].t_arrival <= now) {
        const int64_t rid = (int64_t)s->next_release++;
        if (s->plain) plain_queue.push_back(rid);
        else s->filters[s->reqs[(size_t)rid].fid].fifo.push_back(rid);
      }
      stop_now = s->stop.load();
    }
    // retire
    uint64_t retired = 0;
    for (auto& kv : s->filters) {
      SFilter& f = kv.second;
      while (!f.inflight.empty()) {
        const hipError_t q = hipEventQuery(f.inflight.front().ev);
        if (q == hipErrorNotReady) break;
        check(s, q, "hipEventQuery");
        Batch& b = f.inflight.front();
        const uint64_t t = s->now_ns();
        for (int64_t rid : b.rids) {
          s->reqs[(size_t)rid].t_done = t;
          s->reqs[(size_t)rid].done = true;
        }
        retired += b.rids.size();
        if (!s->plain) f.slot_busy[b.slot] = false;
        s->event_pool.push_back(b.ev);
        f.inflight.pop_front();
      }
    }
    if (retired) {
      std::lock_guard<std::mutex> lk(s->mu);
      s->n_done += retired;
      s->cv.notify_all();
    }
    // launch
    if (s->plain) {
      while (!plain_queue.empty()) {
        const int64_t rid = plain_queue.front();
        plain_queue.pop_front();
        SFilter& f = s->filters[s->reqs[(size_t)rid].fid];
        launch_plain(s, f, rid, now);
      }
    } else {
      for (auto& kv : s->filters) {
        SFilter& f = kv.second;
        if (f.fifo.empty() || f.inflight.size() >= s->max_inflight) continue;
        if (!f.qf && !admit(s, f)) continue;
        launch_batch(s, f, now);
      }
    }
    if (stop_now) {
      bool idle = plain_queue.empty();
      for (auto& kv : s->filters) idle = idle && kv.second.fifo.empty() && kv.second.inflight.empty();
      std::lock_guard<std::mutex> lk(s->mu);
      if (idle && s->next_release == s->reqs.size()) break;
    }
    if (!retired) std::this_thread::yield();
  }
}

}  // namespace

extern "C" int qf_server_create(qf_server** out, uint64_t budget_bytes, uint32_t n_streams, uint64_t max_batch_keys,
                                uint32_t max_inflight, int plain) {
  if (!out || n_streams == 0 || max_batch_keys == 0 || max_batch_keys > 0xFFFFFFFFULL || max_inflight == 0 ||
      max_inflight > 8)
    return -2;
  qf_server* s = new qf_server();
  s->budget = budget_bytes;
  s->max_batch = max_batch_keys;
  s->max_inflight = max_inflight;
  s->plain = plain ? 1 : 0;
  const uint32_t ns = s->plain ? 1 : n_streams;
  for (uint32_t k = 0; k < ns; k++) {
    hipStream_t st = nullptr;
    check(s, hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    s->streams.push_back(st);
  }
  if (s->plain) check(s, hipMalloc(&s->staging, max_batch_keys * 8), "hipMalloc");
  if (s->err) {
    qf_server_destroy(s);
    return -3;
  }
  *out = s;
  return 0;
}

extern "C" int qf_server_add_filter(qf_server* s, uint32_t fid, int nbits) {
  if (!s || s->running || s->filters.count(fid) || nbits < 6 || nbits > 40) return -2;
  SFilter& f = s->filters[fid];
  f.fid = fid;
  f.nbits = nbits;
  f.stream = s->streams[s->filters.size() % s->streams.size()];
  f.slot_busy.assign(s->max_inflight, false);
  if (hipHostMalloc((void**)&f.desc, sizeof(CopyDesc) * s->max_inflight * 3 * kMaxReq, hipHostMallocMapped) !=
      hipSuccess)
    return -3;
  if (s->plain) {  // the baseline keeps every filter resident
    qf_malloc_device(&f.qf, nbits, true);
    if (!f.qf) return -3;
    qf_reserve(f.qf, s->max_batch);
  }
  return 0;
}

extern "C" int64_t qf_server_submit(qf_server* s, uint64_t t_arrival_ns, uint32_t fid, int op, const uint64_t* keys,
                                    uint64_t n, uint64_t* counts) {
  if (!s || (op != QF_SERVER_GET && op != QF_SERVER_INSERT && op != QF_SERVER_DELETE) || n == 0 ||
      n > s->max_batch || !keys || (op == QF_SERVER_GET && !counts))
    return -2;
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->filters.count(fid)) return -2;
  if (!s->reqs.empty() && t_arrival_ns < s->reqs.back().t_arrival) return -2;
  Req r;
  r.fid = fid;
  r.op = op;
  r.keys = keys;
  r.n = n;
  r.counts = counts;
  r.t_arrival = t_arrival_ns;
  s->reqs.push_back(r);
  return (int64_t)s->reqs.size() - 1;
}

extern "C" int qf_server_run(qf_server* s) {
  if (!s || s->running) return -2;
  s->stop = false;
  s->running = true;
  s->n_done = 0;
  s->t0 = std::chrono::steady_clock::now();
  s->th = std::thread(loop, s);
  return 0;
}

extern "C" int qf_server_wait(qf_server* s, uint64_t* wall_ns) {
  if (!s || !s->running) return -2;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [&] { return s->n_done >= s->reqs.size(); });
    s->stop = true;
  }
  s->th.join();
  s->running = false;
  uint64_t last = 0;
  for (const Req& r : s->reqs) last = std::max(last, r.t_done);
  if (wall_ns) *wall_ns = last;
  return s->err ? -3 : 0;
}

extern "C" int qf_server_request_times(qf_server* s, int64_t rid, uint64_t* t_arrival_ns, uint64_t* t_launch_ns,
                                       uint64_t* t_done_ns) {
  if (!s || rid < 0 || (size_t)rid >= s->reqs.size()) return -2;
  const Req& r = s->reqs[(size_t)rid];
  if (t_arrival_ns) *t_arrival_ns = r.t_arrival;
  if (t_launch_ns) *t_launch_ns = r.t_launch;
  if (t_done_ns) *t_done_ns = r.t_done;
  return r.done ? 1 : 0;
}

extern "C" int qf_server_get_stats(qf_server* s, qf_server_stats* out) {
  if (!s || !out) return -2;
  *out = s->stats;
  return 0;
}

extern "C" QF* qf_server_filter(qf_server* s, uint32_t fid) {
  if (!s || s->running) return nullptr;
  auto it = s->filters.find(fid);
  return it == s->filters.end() ? nullptr : it->second.qf;
}

extern "C" void qf_server_destroy(qf_server* s) {
  if (!s) return;
  if (s->running) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
      s->next_release = s->reqs.size();  // drop what has not arrived
    }
    s->th.join();
  }
  (void)hipDeviceSynchronize();
  for (auto& kv : s->filters) {
    SFilter& f = kv.second;
    if (f.qf) qf_destroy_device(f.qf);
    release_buffers(f);
    if (f.desc) (void)hipHostFree(f.desc);
    if (f.host_blocks) (void)hipHostFree(f.host_blocks);
    for (Batch& b : f.inflight) (void)hipEventDestroy(b.ev);
  }
  for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
  for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
  if (s->staging) (void)hipFree(s->staging);
  delete s;
}
