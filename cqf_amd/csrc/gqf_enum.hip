// gqf_enum.hip -- whole-table passes over a filter: enumeration of its items
// in ascending hash order, the item / count / used-slot statistics, and the
// cluster census.
//
// Enumeration replaces the reference's host iterator
// (qf_iterator_from_position src/gqf.cu:3655-3700, qfi_next :3814-3840,
// qfi_get_hash :3807, qfi_get_key :3791-3805).  Table order is hash order:
// runs sit in quotient order and a run's counters in remainder order, so
// the iterator's sequence is the table's counters read left to right.  Here
// one thread owns one 64-slot block and decodes the runs of the block's
// occupied quotients (which may start or end in later blocks); a count
// pass, a scan of per-workgroup totals and a write pass put every item at
// its rank.
//
// The census replaces src/cluster_length_test.cu:37-120, which walks the
// table with find_first_empty_slot (src/gqf.cu:728-771) from one thread:
// a cluster is a maximal stretch of used slots.  Here each block's used
// slots become a 64-bit mask; cluster starts are the mask's rising edges,
// and each start finds its cluster's end by scanning the masks forward.
#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {

constexpr int kThreads = GQF_ENUM_WG_BLOCKS;  // one thread per table block

// Slot reader for the runs of block b: slots of blocks b and b + 1 come
// from a per-thread LDS row (loaded with 16-byte vector loads), later ones
// from the table.
struct slot_row {
  const QF* f;
  uint64_t base;        // first slot of block b
  const uint8_t* row;   // 128 bytes in LDS
  __device__ uint8_t operator()(uint64_t s) const {
    const uint64_t rel = s - base;
    return rel < 128 ? row[rel] : gqf_point::slot(*f, s);
  }
};

// slots of blocks b and b + 1 into a 33-word LDS row (the odd stride keeps
// the rows of a wave in distinct banks)
__device__ inline void load_row(const QF& f, uint64_t b, uint32_t* row) {
  const GQF_GLOBAL uint4* p = gqf_block_ptr(f, b);
  const bool next = b + 1 < f.nblocks;
  const GQF_GLOBAL uint4* p2 = gqf_block_ptr(f, next ? b + 1 : b);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    uint4 v = k < 4 ? p[4 + k] : p2[k];
    if (k >= 4 && !next) v = make_uint4(0, 0, 0, 0);
    row[4 * k] = v.x;
    row[4 * k + 1] = v.y;
    row[4 * k + 2] = v.z;
    row[4 * k + 3] = v.w;
  }
}

// Visits the counters of the runs of block b's occupied quotients in
// [q_lo, q_hi), in table order: item(storage hash, count, first slot).
// Returns the slots those runs use, or -1 if a run has no run end (a
// damaged table).
template <typename Item>
__device__ inline int64_t walk_items(const QF& f, uint64_t b, uint64_t q_lo, uint64_t q_hi, const slot_row& sl,
                                     Item item) {
  uint64_t occ = gqf_point::occ_word(f, b);
  int64_t prev = gqf_point::end_before_block(f, b);
  int64_t used = 0;
  // run-end word of block cb, reloaded only when a run starts or ends in
  // another block (a load per run made a chain of ~38 dependent loads)
  uint64_t cb = ~0ULL, cw = 0;
  while (occ) {
    const uint64_t q = b * 64 + (uint64_t)__builtin_ctzll(occ);
    occ &= occ - 1;
    const uint64_t start = (int64_t)q > prev + 1 ? q : (uint64_t)(prev + 1);
    // the run ends at the first run end at or after start
    uint64_t bj = start >> 6;
    if (bj >= f.nblocks) return -1;
    if (bj != cb) {
      cb = bj;
      cw = gqf_point::run_word(f, bj);
    }
    uint64_t w = cw & (~0ULL << (start & 63));
    while (!w) {
      if (++bj >= f.nblocks) return -1;
      cb = bj;
      cw = gqf_point::run_word(f, bj);
      w = cw;
    }
    const int64_t e = (int64_t)(bj * 64) + __builtin_ctzll(w);
    prev = e;
    if (q < q_lo || q >= q_hi) continue;
    const uint32_t len = (uint32_t)(e - (int64_t)start + 1);
    used += len;
    auto at = [&](uint32_t k) -> uint64_t { return sl(start + k); };
    uint32_t idx = 0;
    while (idx < len) {
      uint64_t rem, cnt;
      const uint32_t last = gqf_decode_counter(at, len, idx, &rem, &cnt);
      item((q << 8) | rem, cnt, start + idx);
      idx = last + 1;
    }
  }
  return used;
}

// workgroup exclusive scan of one u32 per thread; returns the prefix and
// writes the workgroup total to *total
__device__ inline uint32_t wg_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl = wave_scan_add(v);
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; k++) {
    if (k < w) before += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *total = tot;
  return before + incl - v;
}

__device__ inline uint64_t wg_sum_u64(uint64_t v) {
  __shared__ uint64_t wsum[kThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if (lane == 0) wsum[w] = v;
  __syncthreads();
  uint64_t t = 0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; k++) t += wsum[k];
  __syncthreads();
  return t;
}

}  // namespace

// Output form of an item (storage hash S = (hash << value_bits) | value):
//   GQF_ENUM_OUT_HASH  S >> value_bits (qfi_get_hash's key), value to `values`
//   GQF_ENUM_OUT_KEY   hash_64i of that (qfi_get_key's key)
//   GQF_ENUM_OUT_MOVE  the storage hash of the same (hash, value) in a filter
//                      with `dst_vb` value bits (value truncated to them)
struct gqf_out_form {
  int mode;
  int vb;       // source value bits
  int dst_vb;   // MOVE: destination value bits
  uint64_t key_mask;
};

static __device__ inline uint64_t form_hash(const gqf_out_form& o, uint64_t s) {
  const uint64_t h = s >> o.vb;
  if (o.mode == GQF_ENUM_OUT_KEY) return gqf_hash64_inv(h, o.key_mask);
  if (o.mode == GQF_ENUM_OUT_MOVE) return (h << o.dst_vb) | (s & gqf_mask((uint64_t)min(o.vb, o.dst_vb)));
  return h;
}

// Pass 1: items per workgroup; statistics of the whole range into stats[0..2]
// (items, sum of counts, used slots) and damaged-table flag into status.
// With `starts` set, also the bit of every item's first slot, in words
// indexed from block b_lo, and the highest word marked in *max_word.
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_enum_count(const QF* __restrict__ qfp, uint64_t b_lo, uint64_t b_hi, uint64_t q_lo, uint64_t q_hi,
               uint32_t* __restrict__ wg_items, unsigned long long* __restrict__ stats,
               unsigned long long* __restrict__ starts, unsigned long long* __restrict__ max_word) {
  __shared__ uint32_t rows[kThreads * 33];
  const QF& f = *qfp;
  const uint64_t b = b_lo + (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  uint32_t items = 0;
  uint64_t sum = 0, used = 0;
  if (b < b_hi) {
    uint32_t* row = rows + threadIdx.x * 33;
    load_row(f, b, row);
    const slot_row sl{&f, b * 64, (const uint8_t*)row};
    uint64_t cw = ~0ULL, cb = 0;  // word being marked and its pending bits
    const int64_t u = walk_items(f, b, q_lo, q_hi, sl, [&](uint64_t, uint64_t c, uint64_t pos) {
      items++;
      sum += c;
      if (starts) {
        const uint64_t wd = (pos >> 6) - b_lo;
        if (wd != cw) {
          if (cb) atomicOr(&starts[cw], (unsigned long long)cb);
          cw = wd;
          cb = 0;
        }
        cb |= 1ULL << (pos & 63);
      }
    });
    if (cb) {
      atomicOr(&starts[cw], (unsigned long long)cb);
      atomicMax(max_word, (unsigned long long)cw);
    }
    if (u < 0) atomicOr(f.status, GQF_STATUS_CLUSTER_OVERFLOW);
    else used = (uint64_t)u;
  }
  uint32_t tot;
  (void)wg_exclusive_scan(items, &tot);
  const uint64_t s_sum = wg_sum_u64(sum);
  const uint64_t s_used = wg_sum_u64(used);
  if (threadIdx.x == 0) {
    wg_items[blockIdx.x] = tot;
    if (tot) atomicAdd(&stats[0], (unsigned long long)tot);
    if (s_sum) atomicAdd(&stats[1], (unsigned long long)s_sum);
    if (s_used) atomicAdd(&stats[2], (unsigned long long)s_used);
  }
}

// Exclusive scan of n u32 counts into u64 bases (one workgroup); the total
// goes to base[n].
extern "C" __global__ void __launch_bounds__(1024) gqf_enum_scan(const uint32_t* __restrict__ cnt, uint32_t n,
                                                      uint64_t* __restrict__ base) {
  __shared__ uint64_t wsum[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t carry = 0;
  for (uint32_t i0 = 0; i0 < n; i0 += 1024) {
    const uint32_t i = i0 + threadIdx.x;
    const uint64_t v = i < n ? cnt[i] : 0;
    uint64_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t t = __shfl_up(incl, d, 64);
      if (lane >= d) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint64_t before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      if (k < w) before += wsum[k];
      tot += wsum[k];
    }
    if (i < n) base[i] = carry + before + incl - v;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) base[n] = carry;
}

// Pass 2a: per word of marks, the number of marks in earlier words of its
// 256-word group (local_prefix) and each group's total (group_sums).
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_enum_words(const unsigned long long* __restrict__ starts, uint64_t nwords, uint32_t* __restrict__ local_prefix,
               uint32_t* __restrict__ group_sums) {
  const uint64_t w = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  const uint32_t c = w < nwords ? (uint32_t)__builtin_popcountll(starts[w]) : 0;
  uint32_t tot;
  const uint32_t before = wg_exclusive_scan(c, &tot);
  if (w < nwords) local_prefix[w] = before;
  if (threadIdx.x == 0) group_sums[blockIdx.x] = tot;
}

// Quotient whose run holds used slot p.  Past block b's offset the runs of
// quotients below 64 b have ended, so it is block b's k-th occupied
// quotient, k = run ends of block b in [offset, p).  Inside the offset the
// slot belongs to an earlier block b2: the last one whose own runs start at
// or before p, and k counts the run ends in [64 b2 + offset(b2), p).
// Returns ~0 for a table whose bits do not pair up.
static __device__ inline uint64_t quotient_of_slot(const QF& f, uint64_t p) {
  const uint64_t b = p >> 6;
  const uint64_t i = p & 63;
  const uint64_t off = gqf_point::offset(f, b);
  if (i >= off) {
    const int k = __builtin_popcountll(gqf_point::run_word(f, b) & gqf_mask(i) & ~gqf_mask(off));
    const int sel = gqf_select64(gqf_point::occ_word(f, b), k);
    return sel < 64 ? b * 64 + (uint64_t)sel : ~0ULL;
  }
  uint64_t b2 = b - 1;  // b > 0: block 0 has offset 0
  while (b2 > 0 && b2 * 64 + gqf_point::offset(f, b2) > p) b2--;
  const uint64_t x = b2 * 64 + gqf_point::offset(f, b2);
  uint64_t k = 0;
  for (uint64_t j = x >> 6; j <= b; j++) {
    uint64_t wd = gqf_point::run_word(f, j);
    if (j == (x >> 6)) wd &= ~gqf_mask(x & 63);
    if (j == b) wd &= gqf_mask(i);
    k += (uint64_t)__builtin_popcountll(wd);
  }
  for (uint64_t j = b2; j < b; j++) {
    const uint64_t o = gqf_point::occ_word(f, j);
    const uint64_t c = (uint64_t)__builtin_popcountll(o);
    if (k < c) return j * 64 + (uint64_t)gqf_select64(o, (int)k);
    k -= c;
  }
  return ~0ULL;
}

// Pass 2b: one lane per slot of the marked words (wave per word): a lane
// whose slot starts a counter decodes it and writes the item at its rank,
// so a wave's items go out as one contiguous stretch.  Each word is cleared
// after it is read, leaving the marks zero for the next use.
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_enum_emit(const QF* __restrict__ qfp, uint64_t w_lo, uint64_t nwords,
              const unsigned long long* __restrict__ max_word, unsigned long long* __restrict__ starts,
              const uint32_t* __restrict__ local_prefix, const uint64_t* __restrict__ group_base, gqf_out_form o,
              uint64_t* __restrict__ hashes, uint64_t* __restrict__ counts, uint64_t* __restrict__ values,
              uint64_t cap) {
  const QF& f = *qfp;
  // the wave's word (block) index, made wave-uniform so that the block's
  // header words below are scalar loads shared by the 64 lanes
  const uint64_t rel = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)));
  const uint64_t i = threadIdx.x & 63;
  if (rel >= nwords || rel > *max_word) return;
  const uint64_t bits = starts[rel];
  if (!bits) return;
  if (i == 0) starts[rel] = 0;
  if (!((bits >> i) & 1)) return;
  const uint64_t rank = group_base[rel / kThreads] + local_prefix[rel] + (uint64_t)__builtin_popcountll(bits & gqf_mask(i));
  if (rank >= cap) return;
  const uint64_t b = w_lo + rel;
  const uint64_t p = b * 64 + i;
  const uint64_t off = gqf_point::offset(f, b);
  const uint64_t runw = gqf_point::run_word(f, b);
  uint64_t q;
  if (i >= off) {
    // past the offset: block b's k-th occupied quotient, k = its run ends in [offset, p)
    const int k = __builtin_popcountll(runw & gqf_mask(i) & ~gqf_mask(off));
    const int sel = gqf_select64(gqf_point::occ_word(f, b), k);
    q = sel < 64 ? b * 64 + (uint64_t)sel : ~0ULL;
  } else {
    q = quotient_of_slot(f, p);
  }
  // the counter's run ends at the first run end at or after p
  const uint64_t wend = runw & (~0ULL << i);
  const int64_t e = wend ? (int64_t)(b * 64) + __builtin_ctzll(wend) : gqf_point::runend_from(f, p, 0);
  if (q == ~0ULL || e < (int64_t)p) {
    atomicOr(f.status, GQF_STATUS_CLUSTER_OVERFLOW);
    return;
  }
  auto at = [&](uint32_t k) -> uint64_t { return gqf_point::slot(f, p + k); };
  uint64_t rem, cnt;
  gqf_decode_counter(at, (uint32_t)(e - (int64_t)p + 1), 0, &rem, &cnt);
  const uint64_t sh = (q << 8) | rem;
  if (hashes) hashes[rank] = form_hash(o, sh);
  if (counts) counts[rank] = cnt;
  if (values) values[rank] = sh & gqf_mask((uint64_t)o.vb);
}

// ---------------------------------------------------------------------------
// census

// used-slot mask of block b, slots >= xnslots excluded.  Slot 64 b + i is
// used iff the runs of earlier blocks reach it (i < offset) or some quotient
// of block b at or below it has not ended its run before it
// (offset_lower_bound != 0, src/gqf.cu:681-705).
static __device__ inline uint64_t used_mask(const QF& f, uint64_t b) {
  const gqf_header h = gqf_load_header(f, b);
  const uint64_t lim = f.xnslots - b * 64;
  return gqf_used_mask(h.occupieds, h.runends, h.offset, lim < 64 ? (uint32_t)lim : 64u);
}

// whether slot 64 b - 1 is used: earlier runs reach into block b, or a run
// ends exactly there
static __device__ inline uint64_t prev_slot_used(const QF& f, uint64_t b) {
  return b > 0 && (gqf_point::offset(f, b) > 0 || (gqf_load_runends(f, b - 1) >> 63)) ? 1 : 0;
}

// Blocks whose stored used-slot mask (bytes 24..31) differs from the one
// their header implies: one thread per block, one atomic per mismatch.
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_check_used(const QF* __restrict__ qfp, unsigned long long* __restrict__ bad) {
  const QF& f = *qfp;
  const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= f.nblocks) return;
  const uint64_t stored = ((const GQF_GLOBAL uint64_t*)gqf_block_ptr(f, b))[3];
  if (stored != used_mask(f, b)) atomicAdd(bad, 1ULL);
}

extern "C" __global__ void __launch_bounds__(kThreads)
gqf_census_mask(const QF* __restrict__ qfp, uint64_t* __restrict__ used, uint32_t* __restrict__ wg_starts) {
  const QF& f = *qfp;
  const uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  uint32_t starts = 0;
  if (b < f.nblocks) {
    const uint64_t u = used_mask(f, b);
    used[b] = u;
    const uint64_t prev = prev_slot_used(f, b);
    starts = (uint32_t)__builtin_popcountll(u & ~((u << 1) | prev));
  }
  uint32_t tot;
  (void)wg_exclusive_scan(starts, &tot);
  if (threadIdx.x == 0) wg_starts[blockIdx.x] = tot;
}

extern "C" __global__ void __launch_bounds__(kThreads)
gqf_census_write(const QF* __restrict__ qfp, const uint64_t* __restrict__ used, const uint64_t* __restrict__ wg_base,
                 uint64_t* __restrict__ lengths, uint64_t cap, unsigned long long* __restrict__ hist, uint32_t nbins,
                 unsigned long long* __restrict__ max_len) {
  extern __shared__ uint32_t lhist[];
  const QF& f = *qfp;
  const uint64_t nb = f.nblocks;
  for (uint32_t i = threadIdx.x; i < nbins; i += kThreads) lhist[i] = 0;
  __syncthreads();
  const uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  uint64_t starts = 0;
  if (b < nb) {
    const uint64_t u = used[b];
    starts = u & ~((u << 1) | prev_slot_used(f, b));
  }
  uint32_t tot;
  uint64_t r = wg_base[blockIdx.x] + wg_exclusive_scan((uint32_t)__builtin_popcountll(starts), &tot);
  uint64_t my_max = 0;
  while (starts) {
    const int i = __builtin_ctzll(starts);
    starts &= starts - 1;
    // first unused slot after 64 b + i
    uint64_t cb = b;
    uint64_t free_bits = ~used[cb] & (~0ULL << i);
    while (!free_bits && ++cb < nb) free_bits = ~used[cb];
    const uint64_t end = cb < nb ? cb * 64 + (uint64_t)__builtin_ctzll(free_bits) : nb * 64;
    const uint64_t len = end - (b * 64 + (uint64_t)i);
    if (lengths && r < cap) lengths[r] = len;
    r++;
    if (nbins) atomicAdd(&lhist[len < nbins ? len : nbins - 1], 1u);
    if (len > my_max) my_max = len;
  }
  __syncthreads();
  if (hist)
    for (uint32_t i = threadIdx.x; i < nbins; i += kThreads)
      if (lhist[i]) atomicAdd(&hist[i], (unsigned long long)lhist[i]);
  if (my_max) atomicMax(max_len, (unsigned long long)my_max);
}

uint64_t gqf_enum_wg_count(uint64_t nblocks) { return (nblocks + kThreads - 1) / kThreads; }

void gqf_launch_enum_count(const QF* qf_dev, uint64_t b_lo, uint64_t b_hi, uint64_t q_lo, uint64_t q_hi,
                           uint32_t* wg_items, uint64_t* wg_base, unsigned long long* stats,
                           unsigned long long* starts, unsigned long long* max_word, hipStream_t s) {
  const uint64_t nwg = gqf_enum_wg_count(b_hi - b_lo);
  if (nwg == 0) return;
  gqf_prof_begin("enum_count", s);
  hipLaunchKernelGGL(gqf_enum_count, dim3((uint32_t)nwg), dim3(kThreads), 0, s, qf_dev, b_lo, b_hi, q_lo, q_hi,
                     wg_items, stats, starts, max_word);
  gqf_prof_end("enum_count", s);
  hipLaunchKernelGGL(gqf_enum_scan, dim3(1), dim3(1024), 0, s, (const uint32_t*)wg_items, (uint32_t)nwg, wg_base);
}

void gqf_launch_enum_emit(const QF* qf_dev, uint64_t w_lo, uint64_t nwords, unsigned long long* starts,
                          const unsigned long long* max_word, uint32_t* local_prefix, uint32_t* group_sums,
                          uint64_t* group_base, int mode, int vb, int dst_vb, uint64_t key_mask, uint64_t* hashes,
                          uint64_t* counts, uint64_t* values, uint64_t cap, hipStream_t s) {
  const uint64_t ngroups = gqf_enum_wg_count(nwords);
  if (ngroups == 0) return;
  hipLaunchKernelGGL(gqf_enum_words, dim3((uint32_t)ngroups), dim3(kThreads), 0, s,
                     (const unsigned long long*)starts, nwords, local_prefix, group_sums);
  hipLaunchKernelGGL(gqf_enum_scan, dim3(1), dim3(1024), 0, s, (const uint32_t*)group_sums, (uint32_t)ngroups,
                     group_base);
  const gqf_out_form o{mode, vb, dst_vb, key_mask};
  const uint64_t wgs = (nwords + kThreads / 64 - 1) / (kThreads / 64);
  gqf_prof_begin("enum_emit", s);
  hipLaunchKernelGGL(gqf_enum_emit, dim3((uint32_t)wgs), dim3(kThreads), 0, s, qf_dev, w_lo, nwords, max_word,
                     starts, (const uint32_t*)local_prefix, (const uint64_t*)group_base, o, hashes, counts, values,
                     cap);
  gqf_prof_end("enum_emit", s);
}

void gqf_launch_check_used(const QF* qf_dev, uint64_t nblocks, unsigned long long* bad, hipStream_t s) {
  if (nblocks == 0) return;
  hipLaunchKernelGGL(gqf_check_used, dim3((uint32_t)((nblocks + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                     qf_dev, bad);
}

void gqf_launch_census(const QF* qf_dev, uint64_t nblocks, uint64_t* used, uint32_t* wg_starts, uint64_t* wg_base,
                       uint64_t* lengths, uint64_t cap, unsigned long long* hist, uint32_t nbins,
                       unsigned long long* max_len, hipStream_t s) {
  const uint64_t nwg = gqf_enum_wg_count(nblocks);
  if (nwg == 0) return;
  gqf_prof_begin("census_mask", s);
  hipLaunchKernelGGL(gqf_census_mask, dim3((uint32_t)nwg), dim3(kThreads), 0, s, qf_dev, used, wg_starts);
  gqf_prof_end("census_mask", s);
  hipLaunchKernelGGL(gqf_enum_scan, dim3(1), dim3(1024), 0, s, (const uint32_t*)wg_starts, (uint32_t)nwg, wg_base);
  gqf_prof_begin("census_write", s);
  hipLaunchKernelGGL(gqf_census_write, dim3((uint32_t)nwg), dim3(kThreads), nbins * sizeof(uint32_t), s, qf_dev,
                     (const uint64_t*)used, (const uint64_t*)wg_base, lengths, cap, hist, nbins, max_len);
  gqf_prof_end("census_write", s);
}
