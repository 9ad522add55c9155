// gqf_api.hip -- host side of libgqf: lifecycle, workspace and the C ABI.
//
// The handle returned by qf_malloc_device is a device pointer to a
// `quotient_filter` descriptor (as the reference's QF* is,
// src/gqf.cu:2212-2217).  The host keeps a registry from that pointer to
// the filter's host-side state: sizes, the workspace of the bulk pipeline
// and the device status word.
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gqf.h"
#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {

constexpr int kModeInsert = 0;
constexpr int kModeDelete = 1;

std::mutex g_mu;
int g_last_error = 0;

void note_error(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_last_error) g_last_error = (int)e;
  fprintf(stderr, "libgqf: %s failed: %s\n", what, hipGetErrorString(e));
}

#define GQF_CHECK(call) note_error((call), #call)

struct Workspace {
  uint64_t cap = 0;            // keys the buffers below are sized for
  uint64_t* tmp = nullptr;     // sort ping-pong, cap keys
  uint32_t* hist = nullptr;    // [digit][tile]
  uint32_t* totals = nullptr;  // 1024 digit totals
  uint32_t* slow_buckets = nullptr;  // flags of the two-level sort, kept zero by its kernels
  uint32_t* tile_heads = nullptr;
  uint64_t* ukeys = nullptr;   // run-length reduce outputs, cap keys each
  uint64_t* upos = nullptr;
  uint64_t* ucounts = nullptr;
  uint64_t* m_dev = nullptr;
  uint32_t* region_start = nullptr;  // nregions + 1
  uint32_t* region_list = nullptr;   // non-empty regions of each parity
  uint32_t* list_count = nullptr;    // 2 counts
  uint8_t* wide_flags = nullptr;     // nregions, kept zero by the kernels
  uint32_t* wide_count = nullptr;    // GQF_WIDE_WORDS: count, flagged regions, finished workgroups
  unsigned long long* misses = nullptr;
  // counting reduce (allocated on first use): hash table, counts by
  // compaction index, scan tile sums, control words
  uint64_t* agg_stage = nullptr;   // GQF_AGG_BUCKETS x GQF_AGG_BUCKET_CAP staged (hash, count) entries
  uint64_t* agg_by_idx = nullptr;
  uint64_t* agg_tiles = nullptr;
  uint64_t* agg_hot = nullptr;
  uint32_t* agg_ctrl = nullptr;    // 4 control words, then the bucket fills
};

struct Filter {
  QF host;           // host mirror of the device descriptor
  QF* dev = nullptr; // the public handle
  int sort_digit = 0;  // radix digit bits of the bulk sort (0: chosen from the key bits)
  Workspace ws;
  std::mutex mu;     // serialises workspace growth
};

std::unordered_map<QF*, Filter*> g_filters;

Filter* lookup(QF* qf) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_filters.find(qf);
  if (it == g_filters.end()) {
    fprintf(stderr, "libgqf: unknown QF handle %p\n", (void*)qf);
    return nullptr;
  }
  return it->second;
}

template <typename T>
void dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  GQF_CHECK(hipMalloc((void**)p, count * sizeof(T)));
}

template <typename T>
void dev_free(T*& p) {
  if (p) GQF_CHECK(hipFree((void*)p));
  p = nullptr;
}

void free_workspace(Workspace& w) {
  dev_free(w.tmp);
  dev_free(w.hist);
  dev_free(w.tile_heads);
  dev_free(w.ukeys);
  dev_free(w.upos);
  dev_free(w.ucounts);
  w.cap = 0;
}

// the per-table part of a workspace: region tables of the rebuild
void init_region_workspace(Workspace& w, uint64_t nblocks) {
  if (!w.totals) {
    dev_alloc(&w.totals, 1024);
    dev_alloc(&w.slow_buckets, GQF_SORT_MAX_BUCKETS);
    GQF_CHECK(hipMemset(w.slow_buckets, 0, GQF_SORT_MAX_BUCKETS * sizeof(uint32_t)));
    dev_alloc(&w.m_dev, 1);
    // sized for the smallest tunable region, so retuning never reallocates
    const uint64_t maxreg = (nblocks + GQF_REGION_BLOCKS_MIN - 1) / GQF_REGION_BLOCKS_MIN;
    dev_alloc(&w.region_start, maxreg + 2);
    dev_alloc(&w.region_list, 2 * ((maxreg + 1) / 2 + 1));
    dev_alloc(&w.list_count, 2);
    dev_alloc(&w.wide_flags, maxreg + 1);
    GQF_CHECK(hipMemset(w.wide_flags, 0, maxreg + 1));
    dev_alloc(&w.wide_count, GQF_WIDE_WORDS);
    GQF_CHECK(hipMemset(w.wide_count, 0, GQF_WIDE_WORDS * sizeof(uint32_t)));
    dev_alloc(&w.misses, 1);
  }
}

// grow the per-key buffers to hold n keys (synchronises when it allocates)
void ensure_workspace(Filter* f, uint64_t n) {
  std::lock_guard<std::mutex> lk(f->mu);
  Workspace& w = f->ws;
  init_region_workspace(w, f->host.nblocks);
  if (n <= w.cap) return;
  GQF_CHECK(hipDeviceSynchronize());
  free_workspace(w);
  uint64_t cap = n < 4096 ? 4096 : n;
  dev_alloc(&w.tmp, cap);
  dev_alloc(&w.hist, gqf_sort_hist_words(cap));
  dev_alloc(&w.tile_heads, (cap + 4095) / 4096 + 1);
  dev_alloc(&w.ukeys, cap);
  dev_alloc(&w.upos, cap);
  dev_alloc(&w.ucounts, cap);
  w.cap = cap;
}

// Counting reduce of a duplicate-heavy batch of hashed keys: the hashes are
// counted per tile, staged by bucket and summed per bucket, and only the
// distinct ones sorted (gqf_agg.hip), which leaves `keys` and the (hash,
// count) pairs exactly as the sort and the run-length reduce would.
// Batches of at least 2^22 keys whose sample of 4096 holds at most 2048
// distinct keys try it; a batch whose distinct hashes do not fit (a staging
// bucket, a bucket's LDS table or the 2^22 output entries) falls back to the
// sort.  Reads the control words back, so it synchronises the stream twice
// (the reference's bulk_insert_reduce synchronises as well,
// src/gqf.cu:3247-3250, and reads reduce_by_key's end back, :3273).
// Returns false when not taken (nothing in `keys` changed); else *m_out =
// distinct hashes.  A stream under graph capture keeps the sort path (no
// host reads there).
constexpr uint64_t kAggMinKeys = 1ULL << 22;

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return s && hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

bool reduce_by_counting(Filter* f, uint64_t n, uint64_t* keys, int bits, hipStream_t s, uint64_t* m_out) {
  Workspace& w = f->ws;
  const int width = gqf_sort_width(bits, 0);
  if (n < kAggMinKeys || bits > 40 || width + GQF_AGG_LOG_CAP > 64 || stream_capturing(s)) return false;
  const uint64_t cap = 1ULL << GQF_AGG_LOG_CAP;
  {
    std::lock_guard<std::mutex> lk(f->mu);
    if (!w.agg_stage) {
      dev_alloc(&w.agg_stage, (uint64_t)GQF_AGG_BUCKETS * GQF_AGG_BUCKET_CAP);
      dev_alloc(&w.agg_by_idx, cap);
      dev_alloc(&w.agg_tiles, 4096);
      dev_alloc(&w.agg_hot, GQF_AGG_HOT + 1);
      dev_alloc(&w.agg_ctrl, 4 + GQF_AGG_BUCKETS);
      if (!w.agg_stage || !w.agg_by_idx || !w.agg_tiles || !w.agg_hot || !w.agg_ctrl) return false;
    }
  }
  const uint64_t mask = gqf_mask(f->host.key_bits);
  const int vb = (int)f->host.value_bits;
  uint32_t* fill = w.agg_ctrl + 4;
  uint32_t ctrl[4] = {0, 0, 0, 0};
  gqf_launch_agg_sample(keys, n, mask, w.agg_ctrl, fill, w.agg_hot, s);
  GQF_CHECK(hipMemcpyAsync(ctrl, w.agg_ctrl, sizeof(ctrl), hipMemcpyDeviceToHost, s));
  GQF_CHECK(hipStreamSynchronize(s));
  if (ctrl[0] > GQF_AGG_SAMPLE / 2) return false;
  gqf_launch_agg_tiles(keys, n, mask, vb, w.agg_hot, w.agg_stage, fill, w.agg_ctrl, s);
  // comp goes to upos (free until the finish's prefix sums overwrite it)
  gqf_launch_agg_merge(w.agg_stage, fill, width, w.upos, w.agg_by_idx, std::min(cap, w.cap), w.agg_ctrl, s);
  GQF_CHECK(hipMemcpyAsync(ctrl, w.agg_ctrl, sizeof(ctrl), hipMemcpyDeviceToHost, s));
  GQF_CHECK(hipStreamSynchronize(s));
  const uint64_t m = ctrl[1];
  if (ctrl[2] != 0 || m > cap || m == 0) return false;
  gqf_launch_agg_finish(keys, n, m, width, w.upos, w.agg_by_idx, w.tmp, w.hist, w.totals, w.agg_tiles, w.tile_heads,
                        w.ukeys, w.ucounts, w.upos, s);
  *m_out = m;
  return true;
}

void run_mutation(Filter* f, Workspace& w, uint64_t n, uint64_t* keys, uint8_t flags, int mode, bool reduce,
                  hipStream_t s);

void mutate(QF* qf, uint64_t n, uint64_t* keys, uint8_t flags, int mode, bool reduce, hipStream_t s) {
  if (n == 0) return;  // the reference's grid size underflows here
  Filter* f = lookup(qf);
  if (!f) return;
  // Region starts, sort histograms and reduce counts are 32-bit: a batch of
  // more than GQF_MAX_PIECE keys is applied as consecutive pieces.  The
  // table ends up the same (it depends on the multiset only; a bulk_delete
  // floors each count at 0 in either order), and each piece of `keys` is
  // left holding its own ascending hashes.
  if (n > GQF_MAX_PIECE) {
    uint64_t start[64];
    const uint32_t np = gqf_batch_pieces(n, GQF_MAX_PIECE, start, 64);
    if (np + 1 > 64) {  // over 6.7e10 keys: more than any device array here holds
      fprintf(stderr, "libgqf: batch of %llu keys is too large\n", (unsigned long long)n);
      return;
    }
    for (uint32_t k = 0; k < np; k++) mutate(qf, start[k + 1] - start[k], keys + start[k], flags, mode, reduce, s);
    return;
  }
  ensure_workspace(f, n);
  run_mutation(f, f->ws, n, keys, flags, mode, reduce, s);
}

// The bulk pipeline of one batch (n <= GQF_MAX_PIECE) on a workspace sized
// for it; allocates nothing itself (the counting reduce allocates its table
// on first use, so reduce batches take workspaces that may allocate).
void run_mutation(Filter* f, Workspace& w, uint64_t n, uint64_t* keys, uint8_t flags, int mode, bool reduce,
                  hipStream_t s) {
  const bool hash_it = !(flags & QF_KEY_IS_HASH);
  const int vb = (int)f->host.value_bits;
  const int bits = hash_it ? (int)(f->host.key_bits + f->host.value_bits) : 64;
  uint64_t m = 0;
  if (reduce && hash_it && reduce_by_counting(f, n, keys, bits, s, &m)) {
    gqf_launch_region_bounds(f->dev, f->host.nregions, w.ukeys, m, nullptr, w.region_start, w.region_list,
                             w.list_count, ~0ULL, s);
    gqf_launch_rebuild(f->dev, f->host.nregions, f->host.region_blocks, w.ukeys, w.ucounts, m, w.region_start,
                       w.region_list, w.list_count, mode, w.wide_flags, w.wide_count, s);
    GQF_CHECK(hipGetLastError());
    return;
  }
  // the two-level sort also writes the region starts of a plain batch
  // (reduce batches index their distinct keys instead)
  const bool two_level = hash_it && gqf_sort_two_level_fits(n, bits);
  if (two_level) {
    gqf_region_args ra;
    ra.start = w.region_start;
    ra.step = f->host.region_blocks << 14;
    ra.nregions = f->host.nregions;
    ra.range = f->host.range;
    gqf_launch_sort_two_level(keys, w.tmp, n, gqf_mask(f->host.key_bits), vb, bits, w.hist, w.totals,
                              w.slow_buckets, s, reduce ? nullptr : &ra);
  } else
    gqf_launch_sort(keys, w.tmp, n, hash_it ? 1 : (vb ? 2 : 0), gqf_mask(f->host.key_bits), vb, bits, f->sort_digit,
                    w.hist, w.totals, s);
  // A sparse insert batch -- at most one key per point-lock region and one
  // per two rebuild regions on average -- would make the rebuild rewrite a
  // whole window for about one key each; the point path inserts each key
  // where it belongs instead (the image is the same: it depends on the
  // multiset only).  Hashed keys of a filter without value bits, so the
  // sorted keys are the hashes point_insert takes.
  if (mode == kModeInsert && hash_it && vb == 0 && n <= f->host.nregions / 2 && n <= f->host.num_locks) {
    gqf_launch_sparse_insert(f->dev, keys, n, s);
    GQF_CHECK(hipGetLastError());
    return;
  }
  const uint64_t* items = keys;
  const uint64_t* counts = nullptr;
  const uint64_t* n_dev = nullptr;
  if (reduce) {
    gqf_launch_rle(keys, n, w.tile_heads, w.ukeys, w.upos, w.ucounts, w.m_dev, s);
    items = w.ukeys;
    counts = w.ucounts;
    n_dev = w.m_dev;
  }
  if (two_level && !reduce)
    gqf_launch_region_lists(f->dev, f->host.nregions, n, w.region_start, w.region_list, w.list_count, s);
  else
    gqf_launch_region_bounds(f->dev, f->host.nregions, items, n, n_dev, w.region_start, w.region_list,
                             w.list_count, ~0ULL, s);
  gqf_launch_rebuild(f->dev, f->host.nregions, f->host.region_blocks, items, counts, n, w.region_start,
                     w.region_list, w.list_count, mode, w.wide_flags, w.wide_count, s);
  GQF_CHECK(hipGetLastError());
}

uint64_t count_misses(QF* qf, uint64_t* vals, uint64_t n) {
  if (n == 0) return 0;
  Filter* f = lookup(qf);
  if (!f) return 0;
  ensure_workspace(f, 0);
  GQF_CHECK(hipMemsetAsync(f->ws.misses, 0, sizeof(unsigned long long), 0));
  gqf_launch_misses(f->dev, vals, n, f->ws.misses, 0);
  unsigned long long h = 0;
  GQF_CHECK(hipMemcpy(&h, f->ws.misses, sizeof(h), hipMemcpyDeviceToHost));
  return (uint64_t)h;
}

inline uint64_t load_le64(const uint8_t* p) {
  uint64_t v = 0;
  for (int k = 7; k >= 0; k--) v = (v << 8) | p[k];
  return v;
}

inline uint32_t select64_host(uint64_t w, uint32_t r) {  // position of the r-th set bit (0-based)
  for (uint32_t k = 0; k < r; k++) w &= w - 1;
  return (uint32_t)__builtin_ctzll(w);
}

// The reference's 81-byte image -> the device layout, with each block's
// exact offset recomputed from the bit vectors: the runs of the quotients
// below 64 b end at the P-th run end, P = occupied quotients below 64 b.
// Rejects an image whose offset bytes disagree with that (min(offset, 255))
// or whose run ends and occupied quotients do not pair up.
// `out` receives the whole table allocation (gqf_table_bytes), slack zeroed.
bool image_to_device_layout(const uint8_t* img, uint64_t nblocks, uint64_t xnslots, std::vector<uint8_t>& out) {
  out.assign(gqf_table_bytes(nblocks), 0);
  uint64_t occ_before = 0;   // occupied quotients < 64 b
  uint64_t cb = 0;           // run-end cursor block
  uint64_t ends_before = 0;  // run ends in blocks < cb
  uint64_t total_ends = 0;
  for (uint64_t b = 0; b < nblocks; b++) {
    const uint8_t* src = img + b * GQF_REF_BLOCK_BYTES;
    const uint64_t occ = load_le64(src + 1);
    int64_t x = (int64_t)(b * 64) - 1;
    if (occ_before > 0) {
      while (cb < nblocks &&
             ends_before + (uint64_t)__builtin_popcountll(load_le64(img + cb * GQF_REF_BLOCK_BYTES + 9)) < occ_before) {
        ends_before += (uint64_t)__builtin_popcountll(load_le64(img + cb * GQF_REF_BLOCK_BYTES + 9));
        cb++;
      }
      if (cb == nblocks) return false;
      const uint64_t w = load_le64(img + cb * GQF_REF_BLOCK_BYTES + 9);
      const int64_t end = (int64_t)(cb * 64 + select64_host(w, (uint32_t)(occ_before - ends_before - 1)));
      if (end > x) x = end;
    }
    const uint64_t off = (uint64_t)(x - ((int64_t)(b * 64) - 1));
    if (src[0] != (off < GQF_OFFSET_SAT ? off : GQF_OFFSET_SAT)) return false;
    uint8_t* dst = out.data() + b * GQF_BLOCK_BYTES;
    memcpy(dst, src + 1, 16);  // occupieds, runends (little endian both sides)
    const uint32_t off32 = (uint32_t)off;
    memcpy(dst + 16, &off32, 4);
    const uint64_t lim = xnslots - b * 64;
    const uint64_t used = gqf_used_mask(occ, load_le64(src + 9), off32, lim < 64 ? (uint32_t)lim : 64u);
    memcpy(dst + 24, &used, 8);
    memcpy(dst + GQF_SLOT_BASE, src + 17, 64);
    occ_before += (uint64_t)__builtin_popcountll(occ);
    total_ends += (uint64_t)__builtin_popcountll(load_le64(src + 9));
  }
  return total_ends == occ_before;
}

}  // namespace

// ---------------------------------------------------------------------------
// lifecycle

namespace {

// Sizes of a table of 2^nbits slots and `vb` value bits, and its device
// storage (blocks zeroed, point locks); status and descriptor are the
// caller's.  Sizing arithmetic of qf_init / qf_malloc_device
// (src/gqf.cu:1939-1996, 2128-2140): 8-bit slots, key_bits = nbits + 8 - vb,
// double-precision sqrt.
bool alloc_table(QF& h, int nbits, int vb) {
  memset(&h, 0, sizeof(h));
  h.nslots = 1ULL << nbits;
  h.xnslots = h.nslots + (uint64_t)(10 * sqrt((double)h.nslots));
  h.nblocks = (h.xnslots + GQF_SLOTS_PER_BLOCK - 1) / GQF_SLOTS_PER_BLOCK;
  h.value_bits = (uint64_t)vb;
  h.key_bits = (uint64_t)nbits + 8 - h.value_bits;
  h.key_remainder_bits = 8 - h.value_bits;
  h.bits_per_slot = 8;
  h.range = h.nslots << 8;
  h.num_locks = h.xnslots / GQF_REF_LOCK_SLOTS + 2;
  h.region_blocks = GQF_REGION_BLOCKS;
  h.nregions = (h.nblocks + h.region_blocks - 1) / h.region_blocks;
  const size_t bytes = gqf_table_bytes(h.nblocks);
  dev_alloc(&h.blocks, bytes);
  if (!h.blocks) return false;
  GQF_CHECK(hipMemset(h.blocks, 0, bytes));  // qf_malloc memsets the image (src/gqf.cu:2081)
  // the reference allocates the point-API locks only when !bulk_config
  // (src/gqf.cu:2166-2190); they are always allocated here (8 MB at 2^28
  // slots) so a point call on a bulk-configured filter cannot fault
  const size_t nl = (size_t)h.num_locks * GQF_REF_LOCK_DIST;
  dev_alloc(&h.locks, nl);
  if (!h.locks) {
    dev_free(h.blocks);
    return false;
  }
  GQF_CHECK(hipMemset(h.locks, 0, nl * sizeof(uint32_t)));
  return true;
}

Filter* create_filter(int nbits, int vb) {
  if (nbits < 6 || nbits > 40 || vb < 0 || vb > 6) {
    fprintf(stderr, "libgqf: nbits %d / value bits %d out of range\n", nbits, vb);
    return nullptr;
  }
  Filter* f = new Filter();
  if (!alloc_table(f->host, nbits, vb)) {
    delete f;
    return nullptr;
  }
  dev_alloc(&f->host.status, 1);
  GQF_CHECK(hipMemset(f->host.status, 0, sizeof(uint32_t)));
  dev_alloc(&f->dev, 1);
  GQF_CHECK(hipMemcpy(f->dev, &f->host, sizeof(QF), hipMemcpyHostToDevice));
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_filters[f->dev] = f;
  }
  return f;
}

void free_all_workspace(Workspace& w) {
  free_workspace(w);
  dev_free(w.totals);
  dev_free(w.slow_buckets);
  dev_free(w.m_dev);
  dev_free(w.region_start);
  dev_free(w.region_list);
  dev_free(w.list_count);
  dev_free(w.wide_flags);
  dev_free(w.wide_count);
  dev_free(w.misses);
  dev_free(w.agg_stage);
  dev_free(w.agg_by_idx);
  dev_free(w.agg_tiles);
  dev_free(w.agg_hot);
  dev_free(w.agg_ctrl);
}

}  // namespace

extern "C" void qf_malloc_device(QF** qf, int nbits, bool bulk_config) {
  (void)bulk_config;  // locks are always allocated, see alloc_table
  Filter* f = create_filter(nbits, 0);
  *qf = f ? f->dev : nullptr;
}

extern "C" void qf_malloc_device_kv(QF** qf, int nbits, int value_bits) {
  Filter* f = create_filter(nbits, value_bits);
  *qf = f ? f->dev : nullptr;
}

extern "C" void qf_destroy_device(QF* qf) {
  Filter* f = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_filters.find(qf);
    if (it == g_filters.end()) return;
    f = it->second;
    g_filters.erase(it);
  }
  GQF_CHECK(hipDeviceSynchronize());
  free_all_workspace(f->ws);
  dev_free(f->host.blocks);
  dev_free(f->host.status);
  dev_free(f->host.locks);
  dev_free(f->dev);
  delete f;
}

// ---------------------------------------------------------------------------
// bulk API

extern "C" void bulk_insert(QF* qf, uint64_t nvals, uint64_t* keys, uint8_t flags) {
  mutate(qf, nvals, keys, flags, kModeInsert, false, 0);
}

extern "C" void bulk_insert_reduce(QF* qf, uint64_t nvals, uint64_t* keys, uint8_t flags) {
  mutate(qf, nvals, keys, flags, kModeInsert, true, 0);
}

extern "C" void bulk_delete(QF* qf, uint64_t nvals, uint64_t* keys, uint8_t flags) {
  mutate(qf, nvals, keys, flags, kModeDelete, false, 0);
}

extern "C" void bulk_get(QF* qf, uint64_t nvals, uint64_t* vals, uint64_t* returns) {
  bulk_get_async(qf, nvals, vals, returns, nullptr);
}

extern "C" uint64_t bulk_get_misses_wrapper(QF* qf, uint64_t* vals, uint64_t nvals) {
  return count_misses(qf, vals, nvals);
}

// point_query of every key without locks, number not found
// (point_bulk_get + point_get_wrapper, src/gqf.cu:3112-3150); synchronises.
extern "C" uint64_t point_get_wrapper(QF* qf, uint64_t* hashes, uint64_t nitems) {
  if (nitems == 0) return 0;
  Filter* f = lookup(qf);
  if (!f) return 0;
  ensure_workspace(f, 0);
  GQF_CHECK(hipMemsetAsync(f->ws.misses, 0, sizeof(unsigned long long), 0));
  gqf_launch_point_query(f->dev, hashes, nitems, 0, nullptr, 0, f->ws.misses, 0);
  unsigned long long h = 0;
  GQF_CHECK(hipMemcpy(&h, f->ws.misses, sizeof(h), hipMemcpyDeviceToHost));
  return (uint64_t)h;
}

// One locked point_insert per key (the point_bulk_insert kernel,
// src/gqf.cu:3153-3162); synchronises.
extern "C" void point_bulk_insert_wrapper(QF* qf, uint64_t* hashes, uint64_t nitems) {
  if (nitems == 0 || !lookup(qf)) return;
  gqf_launch_point_insert(qf, hashes, nitems, 0, nullptr, 0, 0);
  GQF_CHECK(hipDeviceSynchronize());
}

extern "C" void point_insert_async(QF* qf, uint64_t nvals, uint64_t* keys, uint8_t flags, int32_t* rets,
                                   void* stream) {
  if (nvals == 0 || !lookup(qf)) return;
  gqf_launch_point_insert(qf, keys, nvals, flags, rets, 0, (hipStream_t)stream);
}

extern "C" void point_insert_not_exists_async(QF* qf, uint64_t nvals, uint64_t* keys, uint8_t flags,
                                              int32_t* rets, void* stream) {
  if (nvals == 0 || !lookup(qf)) return;
  gqf_launch_point_insert(qf, keys, nvals, flags, rets, 1, (hipStream_t)stream);
}

extern "C" void point_query_async(QF* qf, uint64_t nvals, uint64_t* keys, uint8_t flags, uint64_t* counts,
                                  int concurrent, void* stream) {
  if (nvals == 0 || !lookup(qf)) return;
  gqf_launch_point_query(qf, keys, nvals, flags, counts, concurrent, nullptr, (hipStream_t)stream);
}

extern "C" void bulk_insert_async(QF* qf, uint64_t nvals, uint64_t* keys, uint8_t flags, void* stream) {
  mutate(qf, nvals, keys, flags, kModeInsert, false, (hipStream_t)stream);
}

extern "C" void bulk_insert_reduce_async(QF* qf, uint64_t nvals, uint64_t* keys, uint8_t flags, void* stream) {
  mutate(qf, nvals, keys, flags, kModeInsert, true, (hipStream_t)stream);
}

extern "C" void bulk_delete_async(QF* qf, uint64_t nvals, uint64_t* keys, uint8_t flags, void* stream) {
  mutate(qf, nvals, keys, flags, kModeDelete, false, (hipStream_t)stream);
}

extern "C" void bulk_get_async(QF* qf, uint64_t nvals, uint64_t* vals, uint64_t* returns, void* stream) {
  if (nvals == 0) return;
  Filter* f = lookup(qf);
  if (!f) return;
  gqf_launch_bulk_get(f->dev, vals, nvals, returns, (hipStream_t)stream);
  GQF_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// bulk_insert without allocation
//
// For callers that cannot allocate device memory (code under graph capture,
// tight serving loops, harnesses that forbid hipMalloc): the workspace of
// the pipeline -- sort buffer and histograms, region table and lists, wide
// flags -- is carved out of one file-scope device array instead of the
// filter's own workspace.  Each GPU holds one copy of the array for the
// life of the process; every call on that GPU shares it, so calls that
// overlap on one GPU race (issue them on one stream, or synchronise).
namespace {

constexpr uint64_t kNoAllocMaxKeys = 1ULL << 20;       // keys per call
constexpr uint64_t kNoAllocMaxRegions = 1200000;       // rebuild regions of the filter (2^30 slots at 16 blocks)
constexpr uint64_t kNoAllocAlign = 256;

constexpr uint64_t na_round(uint64_t b) { return (b + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign; }
// sub-buffers, in order
constexpr uint64_t kNaTmp = na_round(kNoAllocMaxKeys * 8);
constexpr uint64_t kNaHist = na_round(1024ULL * (kNoAllocMaxKeys / 4096 + 64) * 4);  // gqf_sort_hist_words(max) words
constexpr uint64_t kNaTotals = na_round(1024 * 4);
constexpr uint64_t kNaSlow = na_round(GQF_SORT_MAX_BUCKETS * 4);
constexpr uint64_t kNaStart = na_round((kNoAllocMaxRegions + 2) * 4);
constexpr uint64_t kNaList = na_round(2 * ((kNoAllocMaxRegions + 1) / 2 + 1) * 4);
constexpr uint64_t kNaCount = na_round(2 * 4);
constexpr uint64_t kNaWideFlags = na_round(kNoAllocMaxRegions + 1);
constexpr uint64_t kNaWideCount = na_round(GQF_WIDE_WORDS * 4);
constexpr uint64_t kNoAllocBytes =
    kNaTmp + kNaHist + kNaTotals + kNaSlow + kNaStart + kNaList + kNaCount + kNaWideFlags + kNaWideCount;

// The scratch array.  In host code `scratch` is only a stand-in for the
// symbol: its device address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up once per
// device, with that device current); nullptr on failure
void* scratch_address() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      note_error(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

__global__ void gqf_zero_bytes(uint8_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

}  // namespace

extern "C" int bulk_insert_noalloc(QF* qf, uint64_t nvals, uint64_t* keys, uint8_t flags, void* stream) {
  if (nvals == 0) return 0;
  Filter* f = lookup(qf);
  if (!f) return -1;
  if (nvals > kNoAllocMaxKeys) {
    fprintf(stderr, "libgqf: bulk_insert_noalloc takes at most %llu keys per call (got %llu)\n",
            (unsigned long long)kNoAllocMaxKeys, (unsigned long long)nvals);
    return -2;
  }
  if (f->host.nregions > kNoAllocMaxRegions) {
    fprintf(stderr, "libgqf: bulk_insert_noalloc takes filters of at most %llu rebuild regions (this one has %llu)\n",
            (unsigned long long)kNoAllocMaxRegions, (unsigned long long)f->host.nregions);
    return -2;
  }
  if (gqf_sort_hist_words(nvals) * 4 > kNaHist) return -2;  // cannot happen for nvals <= kNoAllocMaxKeys
  uint8_t* base = (uint8_t*)scratch_address();
  if (!base) return -3;
  Workspace w;  // a view into scratch: nothing here is freed
  uint8_t* p = base;
  w.cap = kNoAllocMaxKeys;
  w.tmp = (uint64_t*)p;                 p += kNaTmp;
  w.hist = (uint32_t*)p;                p += kNaHist;
  w.totals = (uint32_t*)p;              p += kNaTotals;
  w.slow_buckets = (uint32_t*)p;        p += kNaSlow;
  w.region_start = (uint32_t*)p;        p += kNaStart;
  w.region_list = (uint32_t*)p;         p += kNaList;
  w.list_count = (uint32_t*)p;          p += kNaCount;
  w.wide_flags = (uint8_t*)p;           p += kNaWideFlags;
  w.wide_count = (uint32_t*)p;
  const hipStream_t s = (hipStream_t)stream;
  // the flags the kernels expect at zero: the last call on this GPU left
  // them so, but clear them anyway (a fill kernel, not a memset node, so
  // the call can be captured into a graph)
  hipLaunchKernelGGL(gqf_zero_bytes, dim3(64), dim3(256), 0, s, (uint8_t*)w.slow_buckets, (uint64_t)kNaSlow);
  hipLaunchKernelGGL(gqf_zero_bytes, dim3(64), dim3(256), 0, s, w.wide_flags, f->host.nregions + 1);
  hipLaunchKernelGGL(gqf_zero_bytes, dim3(1), dim3(256), 0, s, (uint8_t*)w.wide_count,
                     (uint64_t)GQF_WIDE_WORDS * 4);
  gqf_prof_suppress(true);  // per-kernel timing would create events
  run_mutation(f, w, nvals, keys, flags, kModeInsert, false, s);
  gqf_prof_suppress(false);
  const hipError_t e = hipGetLastError();
  note_error(e, "bulk_insert_noalloc");
  return e == hipSuccess ? 0 : -3;
}

// ---------------------------------------------------------------------------
// sequenced batches

namespace {

// request table of a sequenced call; `tagged`: req_tag carries seq << 1 | delete
bool seq_requests(uint64_t n, uint32_t nreq, const uint32_t* req_end, const uint32_t* req_tag, bool tagged,
                  gqf_seq_reqs& out) {
  if (nreq == 0 || nreq > QF_SEQ_MAX_REQUESTS || !req_end || !req_tag || n > 0xFFFFFFFFULL) {
    fprintf(stderr, "libgqf: sequenced call with %u requests (1..%d) or %llu keys\n", nreq, QF_SEQ_MAX_REQUESTS,
            (unsigned long long)n);
    return false;
  }
  uint32_t prev_end = 0;
  int64_t prev_seq = -1;
  for (uint32_t r = 0; r < nreq; r++) {
    const uint32_t seq = tagged ? req_tag[r] >> 1 : req_tag[r];
    if (req_end[r] < prev_end || (int64_t)seq <= prev_seq || seq > 0x7FFFu) {
      fprintf(stderr, "libgqf: sequenced request %u: end %u after %u, sequence number %u after %lld\n", r,
              req_end[r], prev_end, seq, (long long)prev_seq);
      return false;
    }
    prev_end = req_end[r];
    prev_seq = seq;
  }
  if ((uint64_t)prev_end != n) {
    fprintf(stderr, "libgqf: sequenced requests end at %u, not at n = %llu\n", prev_end, (unsigned long long)n);
    return false;
  }
  memset(&out, 0, sizeof(out));
  out.nreq = nreq;
  for (uint32_t r = 0; r < nreq; r++) {
    out.end[r] = req_end[r];
    out.tag[r] = req_tag[r];
  }
  return true;
}

}  // namespace

extern "C" int qf_bulk_mutate_seq(QF* qf, uint64_t n, uint64_t* keys, uint32_t nreq, const uint32_t* req_end,
                                  const uint32_t* req_tag, void* stream) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  if (f->host.key_bits + f->host.value_bits > 40) {
    fprintf(stderr, "libgqf: sequenced batches take filters of at most 2^32 slots\n");
    return -2;
  }
  gqf_seq_reqs reqs;
  if (!seq_requests(n, nreq, req_end, req_tag, true, reqs)) return -3;
  if (n == 0) return 0;
  ensure_workspace(f, n);
  Workspace& w = f->ws;
  const hipStream_t s = (hipStream_t)stream;
  gqf_launch_seq_tag(f->dev, keys, n, reqs, s);
  // hash bits only: the tags above bit 48 ride along, and the passes are
  // stable, so each hash's items stay in request order
  gqf_launch_sort(keys, w.tmp, n, 0, 0, 0, (int)(f->host.key_bits + f->host.value_bits), f->sort_digit, w.hist,
                  w.totals, s);
  gqf_launch_region_bounds(f->dev, f->host.nregions, keys, n, nullptr, w.region_start, w.region_list, w.list_count,
                           GQF_SEQ_HASH_MASK, s);
  gqf_launch_rebuild(f->dev, f->host.nregions, f->host.region_blocks, keys, nullptr, n, w.region_start,
                     w.region_list, w.list_count, GQF_REBUILD_SEQ, w.wide_flags, w.wide_count, s);
  GQF_CHECK(hipGetLastError());
  return 0;
}

extern "C" int qf_bulk_get_seq_adjust(QF* qf, uint64_t n, const uint64_t* keys, uint32_t nreq, const uint32_t* req_end,
                                      const uint32_t* req_seq, uint64_t* counts, uint64_t n_mut, const uint64_t* mut,
                                      void* stream) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  gqf_seq_reqs reqs;
  if (!seq_requests(n, nreq, req_end, req_seq, false, reqs)) return -3;
  gqf_launch_seq_adjust(f->dev, keys, n, reqs, counts, mut, n_mut, (hipStream_t)stream);
  GQF_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// additions

extern "C" int qf_get_info(QF* qf, gqf_info* out) {
  Filter* f = lookup(qf);
  if (!f || !out) return -1;
  const QF& h = f->host;
  out->nslots = h.nslots;
  out->xnslots = h.xnslots;
  out->nblocks = h.nblocks;
  out->key_bits = h.key_bits;
  out->key_remainder_bits = h.key_remainder_bits;
  out->bits_per_slot = h.bits_per_slot;
  out->range = 1ULL << h.key_bits;  // the reference's range: nslots << key_remainder_bits
  out->num_locks = h.num_locks;
  out->nregions = h.nregions;
  out->region_blocks = h.region_blocks;
  out->sort_digit_bits = (uint64_t)f->sort_digit;
  out->value_bits = h.value_bits;
  out->total_size_in_bytes = h.nblocks * GQF_REF_BLOCK_BYTES;  // the reference's image size
  out->device_bytes = gqf_table_bytes(h.nblocks);
  return 0;
}

extern "C" uint32_t qf_get_status(QF* qf) {
  Filter* f = lookup(qf);
  if (!f) return 0;
  uint32_t s = 0;
  GQF_CHECK(hipDeviceSynchronize());
  GQF_CHECK(hipMemcpy(&s, f->host.status, sizeof(s), hipMemcpyDeviceToHost));
  return s;
}

extern "C" void qf_clear_status(QF* qf) {
  Filter* f = lookup(qf);
  if (!f) return;
  GQF_CHECK(hipDeviceSynchronize());
  GQF_CHECK(hipMemset(f->host.status, 0, sizeof(uint32_t)));
}

// The image these four calls exchange is the reference's packed 81-byte
// block array (include/gqf_int.cuh:38-57), nblocks * 81 bytes: the device
// layout is converted on the way out and rebuilt (exact offsets) on the way in.

extern "C" int qf_export_image(QF* qf, void* host_dst, uint64_t nbytes) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  const uint64_t sz = f->host.nblocks * GQF_REF_BLOCK_BYTES;
  if (nbytes < sz) return -2;
  GQF_CHECK(hipDeviceSynchronize());
  uint8_t* tmp = nullptr;
  hipError_t e = hipMalloc(&tmp, sz + 8);
  if (e == hipSuccess) {
    gqf_launch_export(f->dev, f->host.nblocks, tmp, nullptr);
    e = hipMemcpy(host_dst, tmp, sz, hipMemcpyDeviceToHost);
    GQF_CHECK(hipFree(tmp));
  }
  note_error(e, "qf_export_image");
  return e == hipSuccess ? 0 : -3;
}

extern "C" int qf_import_image(QF* qf, const void* host_src, uint64_t nbytes) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  const uint64_t sz = f->host.nblocks * GQF_REF_BLOCK_BYTES;
  if (nbytes != sz) return -2;
  std::vector<uint8_t> dev_img;
  if (!image_to_device_layout((const uint8_t*)host_src, f->host.nblocks, f->host.xnslots, dev_img)) {
    fprintf(stderr, "libgqf: qf_import_image: inconsistent block image (offsets / run ends)\n");
    return -4;
  }
  GQF_CHECK(hipDeviceSynchronize());
  hipError_t e = hipMemcpy(f->host.blocks, dev_img.data(), dev_img.size(), hipMemcpyHostToDevice);
  note_error(e, "qf_import_image");
  return e == hipSuccess ? 0 : -3;
}

// stream-ordered: the conversion kernel runs on `stream`
extern "C" int qf_export_image_device(QF* qf, void* dev_dst, uint64_t nbytes, void* stream) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  const uint64_t sz = f->host.nblocks * GQF_REF_BLOCK_BYTES;
  if (nbytes < sz) return -2;
  gqf_launch_export(f->dev, f->host.nblocks, (uint8_t*)dev_dst, (hipStream_t)stream);
  hipError_t e = hipGetLastError();
  note_error(e, "qf_export_image_device");
  return e == hipSuccess ? 0 : -3;
}

// synchronises `stream`: the offsets are rebuilt on the host
extern "C" int qf_import_image_device(QF* qf, const void* dev_src, uint64_t nbytes, void* stream) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  const uint64_t sz = f->host.nblocks * GQF_REF_BLOCK_BYTES;
  if (nbytes != sz) return -2;
  std::vector<uint8_t> img(sz);
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (e == hipSuccess) e = hipMemcpy(img.data(), dev_src, sz, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    note_error(e, "qf_import_image_device");
    return -3;
  }
  return qf_import_image(qf, img.data(), sz);
}

// Empties the table on the device (stream-ordered), as qf_reset empties the
// host image (src/gqf.cu:2283-2299).
extern "C" int qf_reset_device(QF* qf, void* stream) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  const size_t bytes = gqf_table_bytes(f->host.nblocks);
  hipError_t e = hipMemsetAsync(f->host.blocks, 0, bytes, (hipStream_t)stream);
  note_error(e, "qf_reset_device");
  return e == hipSuccess ? 0 : -3;
}

// Tuning knobs (0 keeps the current value): blocks per rebuild region and
// radix digit bits of the bulk sort.  Takes effect for later calls on the
// stream; synchronises.
extern "C" uint32_t qf_set_wide_wait_bound(uint32_t max_polls) { return gqf_wide_wait_bound_set(max_polls); }

extern "C" int qf_set_tuning(QF* qf, uint32_t region_blocks, uint32_t sort_digit_bits) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  if (region_blocks && (region_blocks < GQF_REGION_BLOCKS_MIN || region_blocks > GQF_REGION_BLOCKS_MAX)) return -2;
  if (sort_digit_bits && (sort_digit_bits < 6 || sort_digit_bits > 10)) return -2;
  GQF_CHECK(hipDeviceSynchronize());
  if (sort_digit_bits) f->sort_digit = (int)sort_digit_bits;
  if (region_blocks) {
    f->host.region_blocks = region_blocks;
    f->host.nregions = (f->host.nblocks + region_blocks - 1) / region_blocks;
    hipError_t e = hipMemcpy(f->dev, &f->host, sizeof(QF), hipMemcpyHostToDevice);
    note_error(e, "qf_set_tuning");
    if (e != hipSuccess) return -3;
  }
  return 0;
}

extern "C" int qf_reserve(QF* qf, uint64_t max_batch) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  ensure_workspace(f, max_batch);
  return f->ws.cap >= max_batch ? 0 : -2;
}

extern "C" uint32_t qf_batch_pieces(uint64_t n, uint64_t max_piece, uint64_t* starts, uint32_t cap) {
  return gqf_batch_pieces(n, max_piece, starts, cap);
}

extern "C" int qf_last_error(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  int e = g_last_error;
  g_last_error = 0;
  return e;
}

// ---------------------------------------------------------------------------
// per-kernel event timing

namespace {
struct ProfRec {
  std::string name;
  hipEvent_t a = nullptr, b = nullptr;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRec> g_prof;
std::unordered_map<std::string, size_t> g_prof_open;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t take_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  GQF_CHECK(hipEventCreate(&e));
  return e;
}

void recycle_all() {
  for (auto& r : g_prof) {
    if (r.a) g_event_pool.push_back(r.a);
    if (r.b) g_event_pool.push_back(r.b);
  }
  g_prof.clear();
  g_prof_open.clear();
}
}  // namespace

static bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return s && hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// set around calls that must not create events or grow the record vectors
static thread_local bool g_prof_suppressed = false;
void gqf_prof_suppress(bool on) { g_prof_suppressed = on; }

void gqf_prof_begin(const char* name, hipStream_t s) {
  if (g_prof_suppressed) return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (!g_prof_on || capturing(s)) return;
  ProfRec r;
  r.name = name;
  r.a = take_event();
  GQF_CHECK(hipEventRecord(r.a, s));
  g_prof_open[r.name] = g_prof.size();
  g_prof.push_back(r);
}

void gqf_prof_end(const char* name, hipStream_t s) {
  if (g_prof_suppressed) return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (!g_prof_on || capturing(s)) return;
  auto it = g_prof_open.find(name);
  if (it == g_prof_open.end()) return;
  ProfRec& r = g_prof[it->second];
  r.b = take_event();
  GQF_CHECK(hipEventRecord(r.b, s));
  g_prof_open.erase(it);
}

bool gqf_filter_host(QF* handle, QF* out) {
  Filter* f = lookup(handle);
  if (!f) return false;
  *out = f->host;
  return true;
}

// Turns per-kernel event timing on (clearing earlier records) or off.
extern "C" void gqf_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  recycle_all();
  g_prof_on = on != 0;
}

// Sum of the completed launches of kernel `name` since the last enable:
// total milliseconds and launch count.  Synchronises on the events.
extern "C" int gqf_profile_read(const char* name, double* total_ms, uint64_t* launches) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  double t = 0;
  uint64_t c = 0;
  for (auto& r : g_prof) {
    if (r.name != name || !r.b) continue;
    GQF_CHECK(hipEventSynchronize(r.b));
    float ms = 0;
    GQF_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
    t += ms;
    c++;
  }
  if (total_ms) *total_ms = t;
  if (launches) *launches = c;
  return 0;
}

// Fills dev_out[0..n) with splitmix64 outputs skip .. skip+n-1 of `seed`.
extern "C" void gqf_fill_splitmix64(uint64_t* dev_out, uint64_t n, uint64_t seed, uint64_t skip, void* stream) {
  gqf_launch_splitmix64(dev_out, n, seed, skip, (hipStream_t)stream);
  GQF_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// whole-table operations: enumeration, statistics, census, files, resize,
// merge

namespace {

// device scratch of one count pass over nblocks blocks
struct EnumScratch {
  uint64_t nwg = 0;
  uint32_t* wg_items = nullptr;
  uint64_t* wg_base = nullptr;           // nwg + 1
  unsigned long long* stats = nullptr;   // items, sum of counts, used slots, spare
  // marks of item starts for the write pass: one word per block from the
  // first enumerated block to the end of the table
  uint64_t nwords = 0;
  unsigned long long* starts = nullptr;
  unsigned long long* max_word = nullptr;
  uint32_t* local_prefix = nullptr;
  uint32_t* group_sums = nullptr;
  uint64_t* group_base = nullptr;
  bool init(uint64_t nblocks) {
    nwg = gqf_enum_wg_count(nblocks);
    dev_alloc(&wg_items, nwg);
    dev_alloc(&wg_base, nwg + 1);
    dev_alloc(&stats, 4);
    return wg_items && wg_base && stats;
  }
  bool init_marks(uint64_t words, hipStream_t s) {
    nwords = words;
    const uint64_t ng = gqf_enum_wg_count(words);
    dev_alloc(&starts, words);
    dev_alloc(&max_word, 1);
    dev_alloc(&local_prefix, words);
    dev_alloc(&group_sums, ng);
    dev_alloc(&group_base, ng + 1);
    if (!starts || !max_word || !local_prefix || !group_sums || !group_base) return false;
    GQF_CHECK(hipMemsetAsync(starts, 0, words * sizeof(unsigned long long), s));
    return true;
  }
  ~EnumScratch() {
    dev_free(wg_items);
    dev_free(wg_base);
    dev_free(stats);
    dev_free(starts);
    dev_free(max_word);
    dev_free(local_prefix);
    dev_free(group_sums);
    dev_free(group_base);
  }
};

uint64_t quotient_blocks(const QF& h) { return (h.nslots + GQF_SLOTS_PER_BLOCK - 1) / GQF_SLOTS_PER_BLOCK; }

// Counted insert of every item of `src` into `dst` (dst_dev / dst, with the
// region tables of dst_ws): the items are enumerated in chunks of whole
// 256-block groups, each chunk written as ascending unique slot hashes of
// dst with their counts and merged into dst by the rebuild.  Returns the
// number of items moved, or < 0.  Synchronises.
int64_t move_items(const QF* src_dev, const QF& src, const QF* dst_dev, const QF& dst, Workspace& dst_ws,
                   hipStream_t s) {
  const uint64_t nb = quotient_blocks(src);
  EnumScratch sc;
  if (!sc.init(nb)) return -3;
  GQF_CHECK(hipMemsetAsync(sc.stats, 0, 4 * sizeof(unsigned long long), s));
  gqf_launch_enum_count(src_dev, 0, nb, 0, src.nslots, sc.wg_items, sc.wg_base, sc.stats, nullptr, nullptr, s);
  std::vector<uint64_t> base(sc.nwg + 1);
  GQF_CHECK(hipMemcpyAsync(base.data(), sc.wg_base, base.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  GQF_CHECK(hipStreamSynchronize(s));
  const uint64_t total = base[sc.nwg];
  if (total == 0) return 0;
  uint64_t cap = total < (1ULL << 24) ? total : (1ULL << 24);
  for (uint64_t u = 0; u < sc.nwg; u++)
    if (base[u + 1] - base[u] > cap) cap = base[u + 1] - base[u];
  uint64_t* items = nullptr;
  uint64_t* counts = nullptr;
  dev_alloc(&items, cap);
  dev_alloc(&counts, cap);
  if (!items || !counts) {
    dev_free(items);
    dev_free(counts);
    return -3;
  }
  init_region_workspace(dst_ws, dst.nblocks);
  if (!sc.init_marks(src.nblocks, s)) {
    dev_free(items);
    dev_free(counts);
    return -3;
  }
  for (uint64_t u = 0; u < sc.nwg;) {
    uint64_t v = u + 1;
    while (v < sc.nwg && base[v + 1] - base[u] <= cap) v++;
    const uint64_t n = base[v] - base[u];
    if (n) {
      const uint64_t b_lo = u * GQF_ENUM_WG_BLOCKS, b_hi = v * GQF_ENUM_WG_BLOCKS < nb ? v * GQF_ENUM_WG_BLOCKS : nb;
      const uint64_t q_hi = b_hi * 64 < src.nslots ? b_hi * 64 : src.nslots;
      GQF_CHECK(hipMemsetAsync(sc.max_word, 0, sizeof(unsigned long long), s));
      unsigned long long* marks = sc.starts + b_lo;  // word 0 of this pass is block b_lo
      gqf_launch_enum_count(src_dev, b_lo, b_hi, b_lo * 64, q_hi, sc.wg_items, sc.wg_base, sc.stats, marks,
                            sc.max_word, s);
      // the write pass covers the marked words only: up to the last one
      unsigned long long last = 0;
      GQF_CHECK(hipMemcpyAsync(&last, sc.max_word, sizeof(last), hipMemcpyDeviceToHost, s));
      GQF_CHECK(hipStreamSynchronize(s));
      gqf_launch_enum_emit(src_dev, b_lo, (uint64_t)last + 1, marks, sc.max_word, sc.local_prefix, sc.group_sums,
                           sc.group_base, GQF_ENUM_OUT_MOVE, (int)src.value_bits, (int)dst.value_bits,
                           gqf_mask(src.key_bits), items, counts, nullptr, cap, s);
      gqf_launch_region_bounds(dst_dev, dst.nregions, items, n, nullptr, dst_ws.region_start, dst_ws.region_list,
                               dst_ws.list_count, ~0ULL, s);
      gqf_launch_rebuild(dst_dev, dst.nregions, dst.region_blocks, items, counts, n, dst_ws.region_start, dst_ws.region_list,
                         dst_ws.list_count, kModeInsert, dst_ws.wide_flags, dst_ws.wide_count, s);
    }
    u = v;
  }
  GQF_CHECK(hipStreamSynchronize(s));
  dev_free(items);
  dev_free(counts);
  return (int64_t)total;
}

int log2_exact(uint64_t v) {
  int k = 0;
  while ((1ULL << k) < v) k++;
  return (1ULL << k) == v ? k : -1;
}

// the reference's qfmetadata (include/gqf_int.cuh:96-113) as laid out by
// an x86-64 compiler: the enum is an int, 4 bytes of padding after `seed`
struct qf_file_meta {
  uint64_t magic_endian_number;
  int32_t hash_mode;
  uint32_t reserved;
  uint64_t total_size_in_bytes;
  uint32_t seed;
  uint32_t pad;
  uint64_t nslots, xnslots, key_bits, value_bits, key_remainder_bits, bits_per_slot, range, nblocks;
  uint64_t nelts, ndistinct_elts, noccupied_slots;
};
static_assert(sizeof(qf_file_meta) == 120, "qfmetadata is 120 bytes");
static_assert(offsetof(qf_file_meta, nslots) == 32, "qfmetadata field offsets");
constexpr uint64_t kQfMagic = 1018874902021329732ULL;  // MAGIC_NUMBER, include/gqf_int.cuh:23

}  // namespace

extern "C" int qf_enumerate(QF* qf, uint64_t q_lo, uint64_t q_hi, int form, uint64_t* hashes, uint64_t* counts,
                            uint64_t* values, uint64_t cap, uint64_t* n_items, void* stream) {
  if (n_items) *n_items = 0;
  Filter* f = lookup(qf);
  if (!f) return -1;
  if (form != QF_ENUM_HASH && form != QF_ENUM_KEY) return -2;
  const QF& h = f->host;
  hipStream_t s = (hipStream_t)stream;
  if (q_hi > h.nslots) q_hi = h.nslots;
  if (q_lo >= q_hi) return 0;
  const uint64_t b_lo = q_lo / 64, b_hi = (q_hi + 63) / 64;
  EnumScratch sc;
  const bool write = cap && (hashes || counts || values);
  if (!sc.init(b_hi - b_lo) || (write && !sc.init_marks(h.nblocks - b_lo, s))) return -3;
  GQF_CHECK(hipMemsetAsync(sc.stats, 0, 4 * sizeof(unsigned long long), s));
  if (write) GQF_CHECK(hipMemsetAsync(sc.max_word, 0, sizeof(unsigned long long), s));
  gqf_launch_enum_count(f->dev, b_lo, b_hi, q_lo, q_hi, sc.wg_items, sc.wg_base, sc.stats, sc.starts, sc.max_word,
                        s);
  unsigned long long last = 0;
  if (write) {
    GQF_CHECK(hipMemcpyAsync(&last, sc.max_word, sizeof(last), hipMemcpyDeviceToHost, s));
    GQF_CHECK(hipStreamSynchronize(s));
  }
  if (write)
    gqf_launch_enum_emit(f->dev, b_lo, (uint64_t)last + 1, sc.starts, sc.max_word, sc.local_prefix, sc.group_sums,
                         sc.group_base, form == QF_ENUM_KEY ? GQF_ENUM_OUT_KEY : GQF_ENUM_OUT_HASH,
                         (int)h.value_bits, 0, gqf_mask(h.key_bits), hashes, counts, values, cap, s);
  uint64_t total = 0;
  hipError_t e = hipMemcpyAsync(&total, sc.wg_base + sc.nwg, sizeof(total), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  note_error(e, "qf_enumerate");
  if (e != hipSuccess) return -3;
  if (n_items) *n_items = total;
  return 0;
}

extern "C" int qf_get_stats(QF* qf, uint64_t* nelts, uint64_t* ndistinct_elts, uint64_t* noccupied_slots) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  const uint64_t nb = quotient_blocks(f->host);
  EnumScratch sc;
  if (!sc.init(nb)) return -3;
  GQF_CHECK(hipDeviceSynchronize());
  GQF_CHECK(hipMemset(sc.stats, 0, 4 * sizeof(unsigned long long)));
  gqf_launch_enum_count(f->dev, 0, nb, 0, f->host.nslots, sc.wg_items, sc.wg_base, sc.stats, nullptr, nullptr,
                        nullptr);
  unsigned long long st[4] = {0, 0, 0, 0};
  hipError_t e = hipMemcpy(st, sc.stats, sizeof(st), hipMemcpyDeviceToHost);
  note_error(e, "qf_get_stats");
  if (e != hipSuccess) return -3;
  if (ndistinct_elts) *ndistinct_elts = st[0];
  if (nelts) *nelts = st[1];
  if (noccupied_slots) *noccupied_slots = st[2];
  return 0;
}

extern "C" int qf_check_layout(QF* qf, uint64_t* bad_blocks) {
  Filter* f = lookup(qf);
  if (!f || !bad_blocks) return -1;
  unsigned long long* bad = nullptr;
  dev_alloc(&bad, 1);
  if (!bad) return -3;
  GQF_CHECK(hipDeviceSynchronize());
  GQF_CHECK(hipMemset(bad, 0, sizeof(unsigned long long)));
  gqf_launch_check_used(f->dev, f->host.nblocks, bad, nullptr);
  unsigned long long h = 0;
  const hipError_t e = hipMemcpy(&h, bad, sizeof(h), hipMemcpyDeviceToHost);
  note_error(e, "qf_check_layout");
  dev_free(bad);
  *bad_blocks = (uint64_t)h;
  return e == hipSuccess ? 0 : -3;
}

extern "C" int qf_cluster_census(QF* qf, uint64_t* lengths, uint64_t cap, uint64_t* hist, uint32_t nbins,
                                 uint64_t* num_clusters, uint64_t* max_len) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  if (hist && (nbins == 0 || nbins > 4096)) return -2;
  const uint64_t nb = f->host.nblocks;
  const uint64_t nwg = gqf_enum_wg_count(nb);
  uint64_t* used = nullptr;
  uint32_t* wg_starts = nullptr;
  uint64_t* wg_base = nullptr;
  unsigned long long* mx = nullptr;
  dev_alloc(&used, nb);
  dev_alloc(&wg_starts, nwg);
  dev_alloc(&wg_base, nwg + 1);
  dev_alloc(&mx, 1);
  int rc = -3;
  if (used && wg_starts && wg_base && mx) {
    GQF_CHECK(hipDeviceSynchronize());
    GQF_CHECK(hipMemset(mx, 0, sizeof(unsigned long long)));
    gqf_launch_census(f->dev, nb, used, wg_starts, wg_base, lengths, lengths ? cap : 0,
                      (unsigned long long*)hist, hist ? nbins : 0, mx, nullptr);
    uint64_t nc = 0;
    unsigned long long m = 0;
    hipError_t e = hipMemcpy(&nc, wg_base + nwg, sizeof(nc), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&m, mx, sizeof(m), hipMemcpyDeviceToHost);
    note_error(e, "qf_cluster_census");
    if (e == hipSuccess) {
      if (num_clusters) *num_clusters = nc;
      if (max_len) *max_len = (uint64_t)m;
      rc = 0;
    }
  }
  dev_free(used);
  dev_free(wg_starts);
  dev_free(wg_base);
  dev_free(mx);
  return rc;
}

extern "C" uint64_t qf_serialize(QF* qf, const char* filename) {
  Filter* f = lookup(qf);
  if (!f || !filename) return 0;
  const QF& h = f->host;
  qf_file_meta m;
  memset(&m, 0, sizeof(m));
  m.magic_endian_number = kQfMagic;
  m.hash_mode = QF_HASH_INVERTIBLE;
  m.total_size_in_bytes = h.nblocks * GQF_REF_BLOCK_BYTES;
  m.seed = 0;
  m.nslots = h.nslots;
  m.xnslots = h.xnslots;
  m.key_bits = h.key_bits;
  m.value_bits = h.value_bits;
  m.key_remainder_bits = h.key_remainder_bits;
  m.bits_per_slot = h.bits_per_slot;
  m.range = 1ULL << h.key_bits;
  m.nblocks = h.nblocks;
  if (qf_get_stats(qf, &m.nelts, &m.ndistinct_elts, &m.noccupied_slots) != 0) return 0;
  std::vector<uint8_t> img(m.total_size_in_bytes);
  if (qf_export_image(qf, img.data(), img.size()) != 0) return 0;
  FILE* fp = fopen(filename, "wb");
  if (!fp) {
    fprintf(stderr, "libgqf: qf_serialize: cannot open %s\n", filename);
    return 0;
  }
  bool ok = fwrite(&m, sizeof(m), 1, fp) == 1 && fwrite(img.data(), 1, img.size(), fp) == img.size();
  ok = (fclose(fp) == 0) && ok;
  return ok ? sizeof(m) + img.size() : 0;
}

extern "C" uint64_t qf_deserialize_device(QF** qf, const char* filename) {
  *qf = nullptr;
  if (!filename) return 0;
  FILE* fp = fopen(filename, "rb");
  if (!fp) {
    fprintf(stderr, "libgqf: qf_deserialize_device: cannot open %s\n", filename);
    return 0;
  }
  qf_file_meta m;
  bool ok = fread(&m, sizeof(m), 1, fp) == 1;
  const int nbits = ok ? log2_exact(m.nslots) : -1;
  ok = ok && m.magic_endian_number == kQfMagic && m.hash_mode == QF_HASH_INVERTIBLE && nbits >= 6 && nbits <= 40 &&
       m.bits_per_slot == 8 && m.value_bits <= 6 && m.key_remainder_bits + m.value_bits == 8 &&
       m.key_bits == (uint64_t)nbits + m.key_remainder_bits &&
       m.range == (m.nslots << m.key_remainder_bits) &&
       m.xnslots == m.nslots + (uint64_t)(10 * sqrt((double)m.nslots)) && m.nblocks == (m.xnslots + 63) / 64 &&
       m.total_size_in_bytes == m.nblocks * GQF_REF_BLOCK_BYTES;
  std::vector<uint8_t> img;
  if (ok) {
    img.resize(m.total_size_in_bytes);
    ok = fread(img.data(), 1, img.size(), fp) == img.size();
  }
  fclose(fp);
  if (!ok) {
    fprintf(stderr, "libgqf: qf_deserialize_device: %s is not a CQF image this library reads\n", filename);
    return 0;
  }
  Filter* f = create_filter(nbits, (int)m.value_bits);
  if (!f) return 0;
  if (qf_import_image(f->dev, img.data(), img.size()) != 0) {
    qf_destroy_device(f->dev);
    return 0;
  }
  *qf = f->dev;
  return sizeof(m) + img.size();
}

extern "C" int64_t qf_resize_device(QF* qf, int new_nbits) {
  Filter* f = lookup(qf);
  if (!f) return -1;
  const QF old = f->host;
  const int nbits = log2_exact(old.nslots);
  if (new_nbits < nbits || new_nbits > 40) return -2;
  const int vb = (int)old.value_bits + (new_nbits - nbits);
  if (vb > 6) return -2;
  GQF_CHECK(hipDeviceSynchronize());
  QF nh;
  if (!alloc_table(nh, new_nbits, vb)) return -3;
  nh.status = old.status;
  QF* ndev = nullptr;
  dev_alloc(&ndev, 1);
  Workspace nws;
  int64_t moved = -3;
  if (ndev && hipMemcpy(ndev, &nh, sizeof(QF), hipMemcpyHostToDevice) == hipSuccess)
    moved = move_items(f->dev, old, ndev, nh, nws, nullptr);
  if (moved < 0) {
    free_all_workspace(nws);
    dev_free(nh.blocks);
    dev_free(nh.locks);
    dev_free(ndev);
    return moved;
  }
  // the handle's descriptor now describes the new table
  std::lock_guard<std::mutex> lk(f->mu);
  free_all_workspace(f->ws);
  f->ws = nws;
  QF gone = f->host;
  f->host = nh;
  GQF_CHECK(hipMemcpy(f->dev, &f->host, sizeof(QF), hipMemcpyHostToDevice));
  dev_free(gone.blocks);
  dev_free(gone.locks);
  dev_free(ndev);
  return moved;
}

extern "C" int64_t qf_multi_merge_device(QF* const* qfs, int nqf, QF* qfr) {
  Filter* r = lookup(qfr);
  if (!r || nqf < 0 || (nqf && !qfs)) return -1;
  std::vector<Filter*> src;
  for (int i = 0; i < nqf; i++) {
    Filter* f = lookup(qfs[i]);
    if (!f) return -1;
    if (f == r || f->host.key_bits != r->host.key_bits || f->host.value_bits > r->host.value_bits) return -2;
    src.push_back(f);
  }
  GQF_CHECK(hipDeviceSynchronize());
  int64_t total = 0;
  for (Filter* f : src) {
    std::lock_guard<std::mutex> lk(r->mu);
    const int64_t m = move_items(f->dev, f->host, r->dev, r->host, r->ws, nullptr);
    if (m < 0) return m;
    total += m;
  }
  return total;
}

extern "C" int64_t qf_merge_device(QF* qfa, QF* qfb, QF* qfc) {
  QF* const two[2] = {qfa, qfb};
  return qf_multi_merge_device(two, 2, qfc);
}
