// gqf_group.hip -- many filters driven by shared launches, and the
// autotuner of the per-filter knobs.
//
// No counterpart in the reference (it inserts one filter per call, each a
// chain of host-synchronised launches).  A group holds existing filters.  A
// group insert takes one device array of keys cut into one segment per
// filter and runs, on one stream:
//   tag      hash every key with its filter's mask and put the filter index
//            above the hash bits: (f << S) | hash;
//   sort     one radix sort of all tagged keys (filters come out in
//            segment order, each sorted by hash);
//   rebuild  ONE persistent launch over every filter's regions: workgroup
//            w takes work items w, w + G, w + 2G, ... (even regions of
//            every filter first, then odd ones that wait for their even
//            neighbours; gqf_rebuild.hip), plus a wide-redo tail that
//            exits at once when nothing overflowed;
//   untag    each segment is left holding its filter's ascending hashes,
//            as bulk_insert leaves its keys.
// The sequence is captured once into a hipGraph per shape (key pointer,
// segment sizes, flags, tuning) and replayed.  A group get is one launch.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/gqf.h"
#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {

#define GCHECK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) fprintf(stderr, "libgqf: %s failed: %s\n", #call, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
void dalloc(T** p, size_t n) {
  *p = nullptr;
  if (hipMalloc((void**)p, (n ? n : 1) * sizeof(T)) != hipSuccess) *p = nullptr;
}
template <typename T>
void dfree(T*& p) {
  if (p) GCHECK(hipFree(p));
  p = nullptr;
}

__device__ inline uint32_t seg_of(const uint64_t* __restrict__ seg, uint32_t nf, uint64_t i) {
  uint32_t lo = 0, hi = nf;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (seg[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

}  // namespace

// keys[i] <- (f << S) | hash_f(keys[i]); a pre-hashed key at or past its
// filter's range becomes (f << S) | (2^S - 1), past every region, and is
// flagged (S is one above the widest filter then).
extern "C" __global__ void __launch_bounds__(256)
gqf_group_tag(gqf_group_dev g, uint64_t* __restrict__ keys, uint64_t n, int is_hash) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t f = seg_of(g.seg, g.nf, i);
    const QF& d = g.descs[f];
    uint64_t h = keys[i];
    if (!is_hash) {
      h = gqf_key_hash(d, h);
    } else if (h > gqf_mask(d.key_bits)) {
      h = (1ULL << g.tag_shift) - 1;
      atomicOr(d.status, GQF_STATUS_BAD_HASH);
    } else {
      h <<= d.value_bits;
    }
    keys[i] = ((uint64_t)f << g.tag_shift) | h;
  }
}

extern "C" __global__ void __launch_bounds__(256)
gqf_group_untag(uint64_t* __restrict__ keys, uint64_t n, int tag_shift) {
  const uint64_t m = (1ULL << tag_shift) - 1;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    keys[i] &= m;
}

// next epoch of the done flags (a kernel, so a captured graph replays it)
extern "C" __global__ void gqf_group_reset(uint32_t* counters) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counters[1] = counters[1] + 1;
}

namespace {

inline uint32_t flat_grid(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  const uint64_t cap = 256ULL * 64;
  return (uint32_t)(g < cap ? (g ? g : 1) : cap);
}

// per-segment sorts run on this many side streams, forked from the call's
// stream and joined back (captured into the group's graph with it)
constexpr int kSortBranches = 4;

struct Group {
  std::vector<QF*> handles;
  uint32_t nf = 0;
  std::mutex mu;
  // device tables
  QF* d_descs = nullptr;
  uint64_t* d_tab = nullptr;  // seg | even_base | odd_base | reg_base, each nf + 1
  uint32_t* d_counters = nullptr;
  uint32_t* d_done = nullptr;
  uint8_t* d_wide = nullptr;
  uint64_t flag_cap = 0;
  // uploaded table contents (to skip unchanged uploads)
  std::vector<QF> h_descs;
  std::vector<uint64_t> h_tab;
  // sort workspace
  uint64_t* tmp = nullptr;
  uint32_t* hist = nullptr;
  uint32_t* totals = nullptr;
  uint64_t cap = 0;
  // side streams of the per-segment sorts, each with its own two-level
  // workspace (histogram for the largest segment, totals, slow flags)
  hipStream_t side[kSortBranches] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kSortBranches] = {};
  uint32_t* bhist[kSortBranches] = {};
  uint32_t* btotals[kSortBranches] = {};
  uint32_t* bslow[kSortBranches] = {};
  uint64_t bcap = 0;
  uint32_t per_cu = 0;  // workgroups per CU of the persistent rebuild (0: default)
  uint32_t max_polls = GQF_GROUP_MAX_POLLS;  // bound of an odd region's neighbour wait
  // graph cache
  bool use_graph = true;
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> graph_key;
};

std::mutex g_groups_mu;
std::map<gqf_group*, Group*> g_groups;

Group* glookup(gqf_group* h) {
  std::lock_guard<std::mutex> lk(g_groups_mu);
  auto it = g_groups.find(h);
  return it == g_groups.end() ? nullptr : it->second;
}

// Refreshes the device tables from the filters' current descriptors and
// the given segments; fills `dev`.  Returns false on a bad argument.
bool refresh(Group* g, const uint64_t* offsets, bool is_hash, gqf_group_dev* dev, uint64_t* n_out) {
  const uint32_t nf = g->nf;
  std::vector<QF> descs(nf);
  std::vector<uint64_t> tab(4 * (size_t)(nf + 1), 0);
  uint64_t* seg = tab.data();
  uint64_t* ev = seg + (nf + 1);
  uint64_t* od = ev + (nf + 1);
  uint64_t* rb = od + (nf + 1);
  uint64_t maxbits = 0;
  for (uint32_t f = 0; f < nf; f++) {
    if (!gqf_filter_host(g->handles[f], &descs[f])) return false;
    if (offsets && offsets[f + 1] < offsets[f]) return false;
    seg[f + 1] = offsets ? offsets[f + 1] : 0;
    const uint64_t nr = descs[f].nregions;
    ev[f + 1] = ev[f] + (nr + 1) / 2;
    od[f + 1] = od[f] + nr / 2;
    rb[f + 1] = rb[f] + nr;
    const uint64_t slot_bits = descs[f].key_bits + descs[f].value_bits;  // log2 of the slot-hash range
    if (slot_bits > maxbits) maxbits = slot_bits;
  }
  if (offsets && offsets[0] != 0) return false;
  const int shift = (int)maxbits + (is_hash ? 1 : 0);
  uint32_t fbits = 0;
  while ((1ULL << fbits) < nf) fbits++;
  if (shift + (int)fbits > 64) return false;
  if (rb[nf] + 1 > g->flag_cap) {
    GCHECK(hipDeviceSynchronize());
    dfree(g->d_done);
    dfree(g->d_wide);
    g->flag_cap = rb[nf] + 1;
    dalloc(&g->d_done, g->flag_cap);
    dalloc(&g->d_wide, g->flag_cap);
    if (!g->d_done || !g->d_wide) return false;
    // done flags hold epochs; the epoch counter only grows, so 0 never matches
    GCHECK(hipMemset(g->d_done, 0, g->flag_cap * sizeof(uint32_t)));
    GCHECK(hipMemset(g->d_wide, 0, g->flag_cap));
  }
  const bool descs_changed = g->h_descs.size() != nf || memcmp(descs.data(), g->h_descs.data(), sizeof(QF) * nf) != 0;
  const bool tab_changed = tab != g->h_tab;
  if (descs_changed || tab_changed) {
    // earlier calls may still read the tables on another stream
    GCHECK(hipDeviceSynchronize());
    if (descs_changed) GCHECK(hipMemcpy(g->d_descs, descs.data(), sizeof(QF) * nf, hipMemcpyHostToDevice));
    if (tab_changed) GCHECK(hipMemcpy(g->d_tab, tab.data(), sizeof(uint64_t) * tab.size(), hipMemcpyHostToDevice));
    g->h_descs = descs;
    g->h_tab = tab;
  }
  dev->descs = g->d_descs;
  dev->seg = g->d_tab;
  dev->even_base = g->d_tab + (nf + 1);
  dev->odd_base = g->d_tab + 2 * (nf + 1);
  dev->reg_base = g->d_tab + 3 * (nf + 1);
  dev->nf = nf;
  dev->tag_shift = shift;
  dev->counters = g->d_counters;
  dev->done = g->d_done;
  dev->wide = g->d_wide;
  dev->max_polls = g->max_polls;
  *n_out = seg[nf];
  return true;
}

void ensure_sort_space(Group* g, uint64_t n) {
  if (n <= g->cap) return;
  GCHECK(hipDeviceSynchronize());
  dfree(g->tmp);
  dfree(g->hist);
  g->cap = n < 4096 ? 4096 : n;
  dalloc(&g->tmp, g->cap);
  dalloc(&g->hist, gqf_sort_hist_words(g->cap));
}

// Keys to be hashed whose every segment fits the two-level sort (at most
// 2^24 keys, 18..42 hash bits) are sorted segment by segment, hashing on
// load, its filter's tag ORed in as the bucket sort writes each key: equal
// (filter, hash) values are equal keys, so the unstable two-level sort
// gives the same array as one LSD sort of the tagged keys (4 passes over
// 40 bits).  The 256-filter k-mer group's insert call (7.3e8 keys) went
// from 147 ms to 104 ms, to 98 ms without the separate tag pass, and to
// 92 ms with the segments' sorts spread over kSortBranches side streams.
bool per_segment_sort(const Group* g, int is_hash) {
  if (is_hash) return false;
  for (uint32_t f = 0; f < g->nf; f++) {
    const uint64_t m = g->h_tab[f + 1] - g->h_tab[f];
    const int bits = (int)(g->h_descs[f].key_bits + g->h_descs[f].value_bits);
    if (m > (1ULL << 24) || bits < 18 || bits > 42) return false;
  }
  return true;
}

// the side streams' workspaces, sized for the largest segment (outside
// any capture: it may allocate)
bool ensure_branch_space(Group* g) {
  uint64_t big = 0;
  for (uint32_t f = 0; f < g->nf; f++) big = std::max(big, g->h_tab[f + 1] - g->h_tab[f]);
  if (big <= g->bcap && g->bhist[0]) return true;
  GCHECK(hipDeviceSynchronize());
  for (int k = 0; k < kSortBranches; k++) {
    dfree(g->bhist[k]);
    dalloc(&g->bhist[k], gqf_sort_hist_words(big));
    if (!g->bhist[k]) return false;
  }
  g->bcap = big;
  return true;
}

void enqueue_insert(Group* g, const gqf_group_dev& dev, uint64_t* keys, uint64_t n, int is_hash, int mode,
                    hipStream_t s) {
  if (per_segment_sort(g, is_hash)) {
    // the segments' sorts are independent: greedy by size over the side
    // streams, so the many short launches of small segments overlap
    GCHECK(hipEventRecord(g->fork, s));
    for (int k = 0; k < kSortBranches; k++) GCHECK(hipStreamWaitEvent(g->side[k], g->fork, 0));
    uint64_t load[kSortBranches] = {};
    for (uint32_t f = 0; f < g->nf; f++) {
      const uint64_t a = g->h_tab[f], m = g->h_tab[f + 1] - a;
      if (m == 0) continue;
      int k = 0;
      for (int j = 1; j < kSortBranches; j++)
        if (load[j] < load[k]) k = j;
      load[k] += m;
      const QF& d = g->h_descs[f];
      gqf_launch_sort_two_level(keys + a, g->tmp + a, m, gqf_mask(d.key_bits), (int)d.value_bits,
                                (int)(d.key_bits + d.value_bits), g->bhist[k], g->btotals[k], g->bslow[k],
                                g->side[k], nullptr, (uint64_t)f << dev.tag_shift);
    }
    for (int k = 0; k < kSortBranches; k++) {
      GCHECK(hipEventRecord(g->join[k], g->side[k]));
      GCHECK(hipStreamWaitEvent(s, g->join[k], 0));
    }
  } else {
    hipLaunchKernelGGL(gqf_group_tag, dim3(flat_grid(n)), dim3(256), 0, s, dev, keys, n, is_hash);
    uint32_t fbits = 0;
    while ((1ULL << fbits) < dev.nf) fbits++;
    gqf_launch_sort(keys, g->tmp, n, 0, 0, 0, dev.tag_shift + (int)fbits, 0, g->hist, g->totals, s);
  }
  hipLaunchKernelGGL(gqf_group_reset, dim3(1), dim3(64), 0, s, g->d_counters);
  gqf_launch_group_rebuild(dev, keys, mode, gqf_group_rebuild_grid(g->per_cu), s);
  hipLaunchKernelGGL(gqf_group_untag, dim3(flat_grid(n)), dim3(256), 0, s, keys, n, dev.tag_shift);
}

}  // namespace

extern "C" int gqf_group_create(gqf_group** out, QF* const* filters, uint32_t nfilters) {
  if (!out) return -1;
  *out = nullptr;
  if (!filters || nfilters == 0) return -2;
  Group* g = new Group();
  g->nf = nfilters;
  for (uint32_t f = 0; f < nfilters; f++) {
    QF tmp;
    if (!gqf_filter_host(filters[f], &tmp)) {
      delete g;
      return -2;
    }
    g->handles.push_back(filters[f]);
  }
  dalloc(&g->d_descs, nfilters);
  dalloc(&g->d_tab, 4 * (size_t)(nfilters + 1));
  dalloc(&g->d_counters, 4);
  dalloc(&g->totals, 1024);
  bool ok = true;
  for (int k = 0; k < kSortBranches; k++) {
    dalloc(&g->btotals[k], 1024);
    dalloc(&g->bslow[k], GQF_SORT_MAX_BUCKETS);
    ok = ok && g->btotals[k] && g->bslow[k];
    if (g->bslow[k]) GCHECK(hipMemset(g->bslow[k], 0, GQF_SORT_MAX_BUCKETS * sizeof(uint32_t)));
    GCHECK(hipStreamCreateWithFlags(&g->side[k], hipStreamNonBlocking));
    GCHECK(hipEventCreateWithFlags(&g->join[k], hipEventDisableTiming));
  }
  GCHECK(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
  if (!g->d_descs || !g->d_tab || !g->d_counters || !g->totals || !ok) {
    delete g;
    return -3;
  }
  GCHECK(hipMemset(g->d_counters, 0, 4 * sizeof(uint32_t)));
  GCHECK(hipStreamCreateWithFlags(&g->cap_stream, hipStreamNonBlocking));
  gqf_group* h = (gqf_group*)g;
  {
    std::lock_guard<std::mutex> lk(g_groups_mu);
    g_groups[h] = g;
  }
  *out = h;
  return 0;
}

extern "C" void gqf_group_destroy(gqf_group* h) {
  Group* g = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_groups_mu);
    auto it = g_groups.find(h);
    if (it == g_groups.end()) return;
    g = it->second;
    g_groups.erase(it);
  }
  GCHECK(hipDeviceSynchronize());
  if (g->exec) GCHECK(hipGraphExecDestroy(g->exec));
  if (g->cap_stream) GCHECK(hipStreamDestroy(g->cap_stream));
  dfree(g->d_descs);
  dfree(g->d_tab);
  dfree(g->d_counters);
  dfree(g->d_done);
  dfree(g->d_wide);
  dfree(g->tmp);
  dfree(g->hist);
  dfree(g->totals);
  for (int k = 0; k < kSortBranches; k++) {
    if (g->side[k]) GCHECK(hipStreamDestroy(g->side[k]));
    if (g->join[k]) GCHECK(hipEventDestroy(g->join[k]));
    dfree(g->bhist[k]);
    dfree(g->btotals[k]);
    dfree(g->bslow[k]);
  }
  if (g->fork) GCHECK(hipEventDestroy(g->fork));
  delete g;
}

extern "C" int gqf_group_set_graphs(gqf_group* h, int on) {
  Group* g = glookup(h);
  if (!g) return -1;
  std::lock_guard<std::mutex> lk(g->mu);
  g->use_graph = on != 0;
  return 0;
}

extern "C" int gqf_group_set_wait_bound(gqf_group* h, uint32_t max_polls) {
  Group* g = glookup(h);
  if (!g) return -1;
  std::lock_guard<std::mutex> lk(g->mu);
  g->max_polls = max_polls;
  return 0;
}

extern "C" int gqf_group_bulk_insert(gqf_group* h, uint64_t* keys, const uint64_t* offsets, uint8_t flags,
                                     void* stream) {
  Group* g = glookup(h);
  if (!g) return -1;
  std::lock_guard<std::mutex> lk(g->mu);
  const int is_hash = (flags & QF_KEY_IS_HASH) ? 1 : 0;
  gqf_group_dev dev;
  uint64_t n = 0;
  if (!refresh(g, offsets, is_hash, &dev, &n)) return -2;
  if (n == 0) return 0;
  if (n > GQF_MAX_PIECE) {  // one sort over all segments: 32-bit digit counts
    fprintf(stderr, "libgqf: group insert of %llu keys (at most 2^30 per call)\n", (unsigned long long)n);
    return -2;
  }
  ensure_sort_space(g, n);
  if (!g->tmp || !g->hist) return -3;
  if (per_segment_sort(g, is_hash) && !ensure_branch_space(g)) return -3;
  hipStream_t s = (hipStream_t)stream;
  (void)gqf_group_rebuild_grid(g->per_cu);  // occupancy query outside any capture
  if (!g->use_graph) {
    enqueue_insert(g, dev, keys, n, is_hash, 0, s);
    return hipGetLastError() == hipSuccess ? 0 : -3;
  }
  std::vector<uint64_t> key = g->h_tab;
  key.push_back((uint64_t)(uintptr_t)keys);
  key.push_back((uint64_t)is_hash);
  key.push_back((uint64_t)(uintptr_t)g->tmp);
  key.push_back((uint64_t)(uintptr_t)g->bhist[0]);
  key.push_back((uint64_t)(uintptr_t)g->d_done);
  key.push_back((uint64_t)g->per_cu);
  key.push_back((uint64_t)g->max_polls);
  for (const QF& d : g->h_descs) key.push_back(d.region_blocks);
  if (!g->exec || key != g->graph_key) {
    if (g->exec) {
      GCHECK(hipGraphExecDestroy(g->exec));
      g->exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    GCHECK(hipStreamBeginCapture(g->cap_stream, hipStreamCaptureModeThreadLocal));
    enqueue_insert(g, dev, keys, n, is_hash, 0, g->cap_stream);
    GCHECK(hipStreamEndCapture(g->cap_stream, &graph));
    if (!graph) return -3;
    hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    GCHECK(hipGraphDestroy(graph));
    if (e != hipSuccess) {
      g->exec = nullptr;
      return -3;
    }
    g->graph_key = key;
  }
  return hipGraphLaunch(g->exec, s) == hipSuccess ? 0 : -3;
}

extern "C" int gqf_group_bulk_get(gqf_group* h, const uint64_t* vals, const uint64_t* offsets, uint64_t* returns,
                                  void* stream) {
  Group* g = glookup(h);
  if (!g) return -1;
  std::lock_guard<std::mutex> lk(g->mu);
  gqf_group_dev dev;
  uint64_t n = 0;
  if (!refresh(g, offsets, 0, &dev, &n)) return -2;
  gqf_launch_group_get(dev, vals, n, returns, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// ---------------------------------------------------------------------------
// Autotuner: for a table of 2^nbits slots loaded to `load` in one batch,
// times bulk_insert for each candidate region size and sort digit on a
// scratch filter (synthetic splitmix64 keys) and keeps the fastest.
// Results are cached per (nbits, load rounded to 0.05).

namespace {
std::mutex g_tune_mu;
std::map<std::pair<int, int>, std::pair<uint32_t, uint32_t>> g_tune_cache;
}  // namespace

extern "C" int gqf_autotune(int nbits, double load, uint32_t* region_blocks, uint32_t* sort_digit_bits) {
  if (nbits < 8 || nbits > 32 || !(load > 0) || load > 1.0) return -2;
  const int bucket = (int)lround(load * 20.0);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find({nbits, bucket});
    if (it != g_tune_cache.end()) {
      if (region_blocks) *region_blocks = it->second.first;
      if (sort_digit_bits) *sort_digit_bits = it->second.second;
      return 0;
    }
  }
  QF* qf = nullptr;
  qf_malloc_device(&qf, nbits, true);
  if (!qf) return -3;
  const uint64_t n = (uint64_t)((double)(1ULL << nbits) * bucket / 20.0);
  uint64_t* src = nullptr;
  uint64_t* work = nullptr;
  dalloc(&src, n);
  dalloc(&work, n);
  hipEvent_t a, b;
  GCHECK(hipEventCreate(&a));
  GCHECK(hipEventCreate(&b));
  gqf_fill_splitmix64(src, n, 0x7A11u + (uint64_t)nbits, 0, nullptr);
  qf_reserve(qf, n);
  static const uint32_t kRegions[] = {24, 32, 40, 48, 56};
  static const uint32_t kDigits[] = {0, 8};
  float best = 1e30f;
  uint32_t best_r = GQF_REGION_BLOCKS, best_d = 0;
  for (uint32_t d : kDigits) {
    for (uint32_t r : kRegions) {
      qf_set_tuning(qf, r, d);
      float ms_sum = 0;
      for (int rep = 0; rep < 2; rep++) {
        qf_reset_device(qf, nullptr);
        GCHECK(hipMemcpy(work, src, n * sizeof(uint64_t), hipMemcpyDeviceToDevice));
        GCHECK(hipEventRecord(a, nullptr));
        bulk_insert_async(qf, n, work, 0, nullptr);
        GCHECK(hipEventRecord(b, nullptr));
        GCHECK(hipEventSynchronize(b));
        float ms = 0;
        GCHECK(hipEventElapsedTime(&ms, a, b));
        if (rep) ms_sum += ms;  // first run warms up
      }
      if (qf_get_status(qf) != 0) {
        qf_clear_status(qf);
        continue;
      }
      if (ms_sum < best) {
        best = ms_sum;
        best_r = r;
        best_d = d;
      }
    }
  }
  GCHECK(hipEventDestroy(a));
  GCHECK(hipEventDestroy(b));
  dfree(src);
  dfree(work);
  qf_destroy_device(qf);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[{nbits, bucket}] = {best_r, best_d};
  }
  if (region_blocks) *region_blocks = best_r;
  if (sort_digit_bits) *sort_digit_bits = best_d;
  return 0;
}

extern "C" int gqf_autotune_set(int nbits, double load, uint32_t region_blocks, uint32_t sort_digit_bits) {
  if (nbits < 8 || nbits > 32 || !(load > 0) || load > 1.0) return -2;
  if (region_blocks < GQF_REGION_BLOCKS_MIN || region_blocks > GQF_REGION_BLOCKS_MAX) return -2;
  if (sort_digit_bits && (sort_digit_bits < 6 || sort_digit_bits > 10)) return -2;  // as qf_set_tuning
  const int bucket = (int)lround(load * 20.0);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[{nbits, bucket}] = {region_blocks, sort_digit_bits};
  return 0;
}

namespace {

// Workgroups per CU of the group's persistent rebuild: times the group
// insert of a batch cut like `offsets` into scratch filters of the same
// sizes and region size, for each candidate up to what the GPU keeps
// resident, and returns the fastest (0 if nothing could be timed).
uint32_t tune_per_cu(Group* g, const uint64_t* offsets) {
  const uint64_t n = offsets[g->nf];
  const uint32_t mx = gqf_group_rebuild_max_per_cu();
  if (n == 0) return 0;
  std::vector<QF*> scratch(g->nf, nullptr);
  bool ok = true;
  for (uint32_t f = 0; f < g->nf && ok; f++) {
    QF d;
    ok = gqf_filter_host(g->handles[f], &d);
    if (!ok) break;
    qf_malloc_device_kv(&scratch[f], (int)(d.key_bits + d.value_bits) - 8, (int)d.value_bits);
    ok = scratch[f] && qf_set_tuning(scratch[f], d.region_blocks, 0) == 0;
  }
  gqf_group* sh = nullptr;
  uint64_t* src = nullptr;
  uint64_t* work = nullptr;
  if (ok) ok = gqf_group_create(&sh, scratch.data(), g->nf) == 0;
  if (ok) {
    dalloc(&src, n);
    dalloc(&work, n);
    ok = src && work;
  }
  uint32_t best_pc = 0;
  if (ok) {
    Group* sg = glookup(sh);
    sg->use_graph = false;
    gqf_fill_splitmix64(src, n, 0x9C0Fu + n, 0, nullptr);
    hipEvent_t a, b;
    GCHECK(hipEventCreate(&a));
    GCHECK(hipEventCreate(&b));
    static const uint32_t kPerCu[] = {4, 6, 8, 10, 12, 16};
    float best = 1e30f;
    for (uint32_t pc : kPerCu) {
      if (pc > mx) break;
      sg->per_cu = pc;
      float ms_sum = 0;
      bool full = false;
      for (int rep = 0; rep < 2; rep++) {
        for (QF* q : scratch) qf_reset_device(q, nullptr);
        GCHECK(hipMemcpy(work, src, n * sizeof(uint64_t), hipMemcpyDeviceToDevice));
        GCHECK(hipEventRecord(a, nullptr));
        const int rc = gqf_group_bulk_insert(sh, work, offsets, 0, nullptr);
        GCHECK(hipEventRecord(b, nullptr));
        GCHECK(hipEventSynchronize(b));
        float ms = 0;
        GCHECK(hipEventElapsedTime(&ms, a, b));
        if (rc != 0) full = true;
        if (rep) ms_sum += ms;  // the first run warms up
      }
      for (QF* q : scratch) {
        if (qf_get_status(q) != 0) full = true;
        qf_clear_status(q);
      }
      if (!full && ms_sum < best) {
        best = ms_sum;
        best_pc = pc;
      }
    }
    GCHECK(hipEventDestroy(a));
    GCHECK(hipEventDestroy(b));
  }
  if (sh) gqf_group_destroy(sh);
  dfree(src);
  dfree(work);
  for (QF* q : scratch)
    if (q) qf_destroy_device(q);
  return best_pc;
}

}  // namespace

// Tunes every filter of the group for the load its segment brings
// (offsets as for gqf_group_bulk_insert), then the group's workgroups per
// CU for that batch shape; synchronises.
extern "C" int gqf_group_autotune(gqf_group* h, const uint64_t* offsets) {
  Group* g = glookup(h);
  if (!g || !offsets) return -1;
  if (offsets[0] != 0) return -2;
  for (uint32_t f = 0; f < g->nf; f++) {
    QF d;
    if (!gqf_filter_host(g->handles[f], &d)) return -2;
    if (offsets[f + 1] < offsets[f]) return -2;
    double load = (double)(offsets[f + 1] - offsets[f]) / (double)d.nslots;
    if (load <= 0) continue;
    if (load > 1.0) load = 1.0;
    uint32_t r = 0, dg = 0;
    int rc = gqf_autotune((int)(d.key_bits + d.value_bits) - 8, load, &r, &dg);
    if (rc != 0) return rc;
    if (qf_set_tuning(g->handles[f], r, dg) != 0) return -3;
  }
  const uint32_t pc = tune_per_cu(g, offsets);
  if (pc) {
    std::lock_guard<std::mutex> lk(g->mu);
    g->per_cu = pc;
  }
  return 0;
}

// Workgroups per CU the group's persistent rebuild launches with: 0 sets
// the default; above what the GPU keeps resident -> -2.
extern "C" int gqf_group_set_workgroups_per_cu(gqf_group* h, uint32_t per_cu) {
  Group* g = glookup(h);
  if (!g) return -1;
  if (per_cu > gqf_group_rebuild_max_per_cu()) return -2;
  std::lock_guard<std::mutex> lk(g->mu);
  g->per_cu = per_cu;
  return 0;
}

extern "C" int gqf_group_get_workgroups_per_cu(gqf_group* h, uint32_t* per_cu) {
  Group* g = glookup(h);
  if (!g || !per_cu) return -1;
  std::lock_guard<std::mutex> lk(g->mu);
  const uint32_t mx = gqf_group_rebuild_max_per_cu();
  const uint32_t pc = g->per_cu ? g->per_cu : GQF_GROUP_PER_CU;
  *per_cu = pc < mx ? pc : mx;
  return 0;
}
