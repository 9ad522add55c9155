// gqf_api.hip -- host side of libgqf: lifecycle, workspace and the C ABI.
//
// The handle returned by qf_malloc_device is a device pointer to a
// `quotient_filter` descriptor (as the reference's QF* is,
// src/gqf.cu:2212-2217).  The host keeps a registry from that pointer to
// the filter's host-side state: sizes, the workspace of the bulk pipeline
// and the device status word.
#include <math.h>
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
  uint32_t* tile_heads = nullptr;
  uint64_t* ukeys = nullptr;   // run-length reduce outputs, cap keys each
  uint64_t* upos = nullptr;
  uint64_t* ucounts = nullptr;
  uint64_t* m_dev = nullptr;
  uint32_t* region_start = nullptr;  // nregions + 1
  uint32_t* region_list = nullptr;   // non-empty regions of each parity
  uint32_t* list_count = nullptr;    // 2 counts
  uint8_t* wide_flags = nullptr;     // nregions, kept zero by the kernels
  uint32_t* wide_count = nullptr;
  unsigned long long* misses = nullptr;
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

// grow the per-key buffers to hold n keys (synchronises when it allocates)
void ensure_workspace(Filter* f, uint64_t n) {
  std::lock_guard<std::mutex> lk(f->mu);
  Workspace& w = f->ws;
  if (!w.totals) {
    dev_alloc(&w.totals, 1024);
    dev_alloc(&w.m_dev, 1);
    // sized for the smallest tunable region, so retuning never reallocates
    const uint64_t maxreg = (f->host.nblocks + GQF_REGION_BLOCKS_MIN - 1) / GQF_REGION_BLOCKS_MIN;
    dev_alloc(&w.region_start, maxreg + 2);
    dev_alloc(&w.region_list, 2 * ((maxreg + 1) / 2 + 1));
    dev_alloc(&w.list_count, 2);
    dev_alloc(&w.wide_flags, maxreg + 1);
    GQF_CHECK(hipMemset(w.wide_flags, 0, maxreg + 1));
    dev_alloc(&w.wide_count, 1);
    GQF_CHECK(hipMemset(w.wide_count, 0, sizeof(uint32_t)));
    dev_alloc(&w.misses, 1);
  }
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

void mutate(QF* qf, uint64_t n, uint64_t* keys, uint8_t flags, int mode, bool reduce, hipStream_t s) {
  if (n == 0) return;  // the reference's grid size underflows here
  Filter* f = lookup(qf);
  if (!f) return;
  ensure_workspace(f, n);
  Workspace& w = f->ws;
  const bool hash_it = !(flags & QF_KEY_IS_HASH);
  const int bits = hash_it ? (int)f->host.key_bits : 64;
  gqf_launch_sort(keys, w.tmp, n, hash_it ? 1 : 0, f->host.range - 1, bits, f->sort_digit, w.hist, w.totals, s);
  const uint64_t* items = keys;
  const uint64_t* counts = nullptr;
  const uint64_t* n_dev = nullptr;
  if (reduce) {
    gqf_launch_rle(keys, n, w.tile_heads, w.ukeys, w.upos, w.ucounts, w.m_dev, s);
    items = w.ukeys;
    counts = w.ucounts;
    n_dev = w.m_dev;
  }
  gqf_launch_region_bounds(f->dev, f->host.nregions, items, n, n_dev, w.region_start, w.region_list, w.list_count,
                           s);
  gqf_launch_rebuild(f->dev, f->host.nregions, items, counts, n, w.region_start, w.region_list, w.list_count, mode,
                     w.wide_flags, w.wide_count, s);
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
bool image_to_device_layout(const uint8_t* img, uint64_t nblocks, std::vector<uint8_t>& out) {
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
    memcpy(dst + GQF_SLOT_BASE, src + 17, 64);
    occ_before += (uint64_t)__builtin_popcountll(occ);
    total_ends += (uint64_t)__builtin_popcountll(load_le64(src + 9));
  }
  return total_ends == occ_before;
}

}  // namespace

// ---------------------------------------------------------------------------
// lifecycle

extern "C" void qf_malloc_device(QF** qf, int nbits, bool bulk_config) {
  *qf = nullptr;
  if (nbits < 6 || nbits > 40) {
    fprintf(stderr, "libgqf: nbits %d out of range\n", nbits);
    return;
  }
  Filter* f = new Filter();
  QF& h = f->host;
  memset(&h, 0, sizeof(h));
  // sizing arithmetic of qf_init / qf_malloc_device (src/gqf.cu:1939-1996,
  // 2128-2140): r = 8 remainder bits, 0 value bits, double-precision sqrt
  h.nslots = 1ULL << nbits;
  h.xnslots = h.nslots + (uint64_t)(10 * sqrt((double)h.nslots));
  h.nblocks = (h.xnslots + GQF_SLOTS_PER_BLOCK - 1) / GQF_SLOTS_PER_BLOCK;
  h.key_bits = (uint64_t)nbits + 8;
  h.key_remainder_bits = 8;
  h.bits_per_slot = 8;
  h.range = h.nslots << h.key_remainder_bits;
  h.num_locks = h.xnslots / GQF_REF_LOCK_SLOTS + 2;
  h.region_blocks = GQF_REGION_BLOCKS;
  h.nregions = (h.nblocks + h.region_blocks - 1) / h.region_blocks;
  const size_t bytes = gqf_table_bytes(h.nblocks);
  dev_alloc(&h.blocks, bytes);
  GQF_CHECK(hipMemset(h.blocks, 0, bytes));  // qf_malloc memsets the image (src/gqf.cu:2081)
  dev_alloc(&h.status, 1);
  GQF_CHECK(hipMemset(h.status, 0, sizeof(uint32_t)));
  // the reference allocates the point-API locks only when !bulk_config
  // (src/gqf.cu:2166-2190); they are always allocated here (8 MB at 2^28
  // slots) so a point call on a bulk-configured filter cannot fault
  (void)bulk_config;
  {
    const size_t nl = (size_t)h.num_locks * GQF_REF_LOCK_DIST;
    dev_alloc(&h.locks, nl);
    GQF_CHECK(hipMemset(h.locks, 0, nl * sizeof(uint32_t)));
  }
  dev_alloc(&f->dev, 1);
  GQF_CHECK(hipMemcpy(f->dev, &h, sizeof(QF), hipMemcpyHostToDevice));
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_filters[f->dev] = f;
  }
  *qf = f->dev;
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
  free_workspace(f->ws);
  dev_free(f->ws.totals);
  dev_free(f->ws.m_dev);
  dev_free(f->ws.region_start);
  dev_free(f->ws.region_list);
  dev_free(f->ws.list_count);
  dev_free(f->ws.wide_flags);
  dev_free(f->ws.wide_count);
  dev_free(f->ws.misses);
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
  out->range = h.range;
  out->num_locks = h.num_locks;
  out->nregions = h.nregions;
  out->region_blocks = h.region_blocks;
  out->sort_digit_bits = (uint64_t)f->sort_digit;
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
  if (!image_to_device_layout((const uint8_t*)host_src, f->host.nblocks, dev_img)) {
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

void gqf_prof_begin(const char* name, hipStream_t s) {
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
