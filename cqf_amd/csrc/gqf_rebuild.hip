// gqf_rebuild.hip -- bulk mutation of the CQF image by region merge-rebuild.
//
// Replaces the reference's even-odd scheduler in which one *thread* owns an
// 8192-slot region and calls qf_insert / qf_remove once per key
// (src/gqf.cu:2759-2906, 3184-3363).  Here one wave64 owns a 64-block region
// (4096 quotients, one lane per block) and applies ALL of the region's keys
// in one pass:
//
//   1. the region's blocks plus a window of the next region are staged from
//      HBM into LDS with coalesced 16-byte loads and split into
//      slots / occupieds / runends / offsets;
//   2. each lane walks its block's runs (old layout from the stored offsets),
//      merges the block's sorted (hash, count) items counter by counter and
//      gets the new run lengths;
//   3. a wave-wide max-plus scan places every run at
//      max(quotient, end of previous run + 1);
//   4. runs of the next region's first blocks are shifted (right on insert,
//      left on delete) until the layout meets its old position again;
//   5. the new bytes are written back with coalesced stores.
//
// The result is the canonical image the reference's sequential inserts and
// removes produce (counters in encode_counter form, offsets saturated at 255),
// because a CQF image is a function of the multiset of hashes alone.
//
// Regions are processed in two phases (even, then odd), so a region may push
// into the next region's first blocks without racing the wave that owns it.
// A region whose push is not absorbed inside its window is flagged and
// redone by a single-workgroup pass with a 4x wider window, in ascending
// region order, after the phase.
#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {

constexpr int kWinFast = 80;   // 64 region blocks + 16 push blocks
constexpr int kWinWide = 256;  // fallback window
constexpr int64_t kNegInf = INT64_MIN / 4;

enum RebuildMode : int { kInsert = 0, kDelete = 1 };

enum RebuildResult : int { kOk = 0, kNeedWide = 1, kNoSpace = 2 };

template <int W>
struct Smem {
  uint8_t slots_in[W * 64];
  uint8_t slots_out[W * 64];
  uint64_t occ_in[W];
  uint64_t run_in[W];
  uint64_t occ_out[W];
  uint64_t run_out[W];
  int64_t xold[W + 1];   // end of runs with quotient < block start (>= start-1), old layout
  int64_t xnew[W + 1];   // same, new layout
  int64_t agg_a[W];      // per-block max-plus aggregate of the new run lengths
  int64_t agg_b[W];
  uint32_t off_in[W];
  uint32_t blk_cnt[64];
  uint32_t blk_start[64];
  int32_t first_invalid;
  int32_t bstar;
  int32_t result;
};

// ---------------------------------------------------------------------------
// global-image helpers (used when a stored offset is saturated)

__device__ int64_t next_runend_global(const QF& qf, uint64_t pos) {
  uint64_t b = pos >> 6;
  uint32_t bit = (uint32_t)(pos & 63);
  while (b < qf.nblocks) {
    uint64_t w = gqf_load_header(qf.blocks, b).runends >> bit;
    if (w) return (int64_t)(b * 64 + bit + __builtin_ctzll(w));
    b++;
    bit = 0;
  }
  return -1;
}

// End of the runs with quotient < 64*b, at least 64*b - 1: the quantity the
// block offset encodes (src/gqf.cu:619-630).  Walks back to a block with an
// unsaturated offset, then forward over the runs in between.
__device__ int64_t true_x_global(const QF& qf, uint64_t b) {
  uint64_t b0 = b;
  gqf_header h = gqf_load_header(qf.blocks, b0);
  while (b0 > 0 && h.offset == GQF_OFFSET_SAT) {
    b0--;
    h = gqf_load_header(qf.blocks, b0);
  }
  int64_t x = (int64_t)(b0 * 64) - 1 + (int64_t)h.offset;
  for (uint64_t bb = b0; bb < b; bb++) {
    uint64_t occ = gqf_load_header(qf.blocks, bb).occupieds;
    while (occ) {
      int64_t q = (int64_t)(bb * 64 + __builtin_ctzll(occ));
      occ &= occ - 1;
      int64_t st = q > x + 1 ? q : x + 1;
      int64_t en = next_runend_global(qf, (uint64_t)st);
      if (en < 0) return (int64_t)(qf.nblocks * 64);
      x = en;
    }
  }
  int64_t lo = (int64_t)(b * 64) - 1;
  return x > lo ? x : lo;
}

// ---------------------------------------------------------------------------
// wave helpers

__device__ inline int lane_id() { return (int)__lane_id(); }

__device__ inline void compose(int64_t& a, int64_t& b, int64_t a2, int64_t b2) {
  // (a, b) then (a2, b2): x -> max(a2, max(a, x + b) + b2)
  int64_t na = a + b2;
  a = na > a2 ? na : a2;
  b = b + b2;
}

// ---------------------------------------------------------------------------
// counter emission in slot order (same bytes as encode_counter,
// src/gqf.cu:1216-1272)

template <typename Put>
__device__ inline uint32_t emit_counter(uint64_t rem, uint64_t count, Put put) {
  if (count == 0) return 0;
  if (count == 1) { put(0, rem); return 1; }
  if (count == 2) { put(0, rem); put(1, rem); return 2; }
  if (count == 3) {
    put(0, rem);  // r 0 r, or 0 0 0 for r == 0
    put(1, 0);
    put(2, rem);
    return 3;
  }
  gqf_digits g;
  gqf_counter_digits(rem, count, &g);
  uint32_t k = 0;
  put(k++, rem);
  if (rem && g.pad0) put(k++, 0);
  for (int i = 0; i < g.n; i++) put(k++, g.d[i]);
  if (rem == 0) put(k++, 0);
  put(k++, rem);
  return k;
}

// ---------------------------------------------------------------------------
// One block's walk: old runs (from the old layout) merged with the block's
// items.  Pass 1 (WRITE = false) returns the max-plus aggregate of the new
// run lengths and the new occupieds word; pass 2 (WRITE = true) lays the
// runs out from x = xnew[j] into slots_out / run_out.

template <int W, bool WRITE>
__device__ bool walk_block(const QF& qf, Smem<W>& sm, int j, int64_t win_base, int nwin,
                           const uint64_t* __restrict__ items, const uint64_t* __restrict__ counts,
                           uint64_t it, uint64_t ie, int mode,
                           int64_t& agg_a, int64_t& agg_b, uint64_t& newocc) {
  const int64_t blk_base = win_base + 64 * j;
  int64_t xo = sm.xold[j];
  int64_t xn = WRITE ? sm.xnew[j] : 0;
  uint64_t occ = sm.occ_in[j];
  agg_a = kNegInf;
  agg_b = 0;
  newocc = 0;

  uint64_t cur_h = it < ie ? items[it] : 0;
  for (;;) {
    int qo = occ ? __builtin_ctzll(occ) : 64;
    int qi = it < ie ? (int)((cur_h >> 8) & 63) : 64;
    int qrel = qo < qi ? qo : qi;
    if (qrel == 64) break;
    int64_t qa = blk_base + qrel;

    // old run of qa
    int64_t st_o = 0, en_o = -1;
    if (qo == qrel) {
      occ &= occ - 1;
      st_o = qa > xo + 1 ? qa : xo + 1;
      int64_t rel = st_o - win_base;
      int bj = (int)(rel >> 6);
      if (bj >= nwin) return false;
      uint64_t w = sm.run_in[bj] >> (rel & 63);
      int64_t pos = rel;
      while (!w) {
        bj++;
        if (bj >= nwin) return false;
        w = sm.run_in[bj];
        pos = (int64_t)bj * 64;
      }
      en_o = win_base + pos + __builtin_ctzll(w);
      xo = en_o;
    }
    const uint32_t old_len = (uint32_t)(en_o - st_o + 1);  // 0 if no old run

    int64_t st_n = 0;
    if (WRITE) st_n = qa > xn + 1 ? qa : xn + 1;
    uint32_t len = 0;

    if (qi != qrel) {
      // untouched run: same bytes at the new position
      len = old_len;
      if (WRITE) {
        const int64_t src = st_o - win_base, dst = st_n - win_base;
        for (uint32_t k = 0; k < len; k++) sm.slots_out[dst + k] = sm.slots_in[src + k];
      }
    } else {
      // two-pointer merge of the old counters of qa (ascending remainders)
      // with the items whose quotient is qa (ascending hashes)
      const int64_t src = st_o - win_base;
      const int64_t dst = WRITE ? st_n - win_base : 0;
      auto old_slot = [&](uint32_t k) -> uint64_t { return sm.slots_in[src + k]; };
      auto put = [&](uint32_t k, uint64_t v) {
        if (WRITE) sm.slots_out[dst + len + k] = (uint8_t)v;
      };
      uint32_t oi = 0, o_end = 0;
      uint64_t o_rem = 0, o_cnt = 0;
      bool have_old = old_len > 0;
      if (have_old) o_end = gqf_decode_counter(old_slot, old_len, 0, &o_rem, &o_cnt);
      uint64_t n_rem = 0, n_cnt = 0;
      bool have_new = false;
      auto next_new = [&]() {
        have_new = it < ie && (int64_t)(cur_h >> 8) == qa;
        if (!have_new) return;
        const uint64_t h0 = cur_h;
        n_rem = h0 & 0xFF;
        n_cnt = 0;
        while (it < ie && cur_h == h0) {
          n_cnt += counts ? counts[it] : 1ULL;
          it++;
          cur_h = it < ie ? items[it] : 0;
        }
      };
      auto next_old = [&]() {
        oi = o_end + 1;
        have_old = oi < old_len;
        if (have_old) o_end = gqf_decode_counter(old_slot, old_len, oi, &o_rem, &o_cnt);
      };
      next_new();
      while (have_old || have_new) {
        if (have_old && (!have_new || o_rem < n_rem)) {
          const uint32_t clen = o_end - oi + 1;  // unchanged counter: copy its bytes
          if (WRITE)
            for (uint32_t k = 0; k < clen; k++) sm.slots_out[dst + len + k] = sm.slots_in[src + oi + k];
          len += clen;
          next_old();
        } else if (!have_old || n_rem < o_rem) {
          if (mode == kInsert) len += emit_counter(n_rem, n_cnt, put);  // delete of an absent key: no-op
          next_new();
        } else {
          const uint64_t c = mode == kInsert ? o_cnt + n_cnt : (o_cnt > n_cnt ? o_cnt - n_cnt : 0);
          len += emit_counter(n_rem, c, put);
          next_old();
          next_new();
        }
      }
    }

    if (len) {
      newocc |= 1ULL << qrel;
      int64_t a2 = qa + (int64_t)len - 1;
      compose(agg_a, agg_b, a2, (int64_t)len);
      if (WRITE) {
        int64_t en_n = st_n + (int64_t)len - 1;
        int64_t rel = en_n - win_base;
        atomicOr((unsigned long long*)&sm.run_out[rel >> 6], 1ULL << (rel & 63));
        xn = en_n;
      }
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// Region rebuild.  All 64 lanes of the (single-wave) workgroup call it.

template <int W>
__device__ int rebuild_region(const QF& qf, uint64_t r, const uint64_t* __restrict__ items,
                              const uint64_t* __restrict__ counts, uint64_t s, uint64_t e, int mode,
                              Smem<W>& sm) {
  const int lane = lane_id();
  const uint64_t B0 = r * GQF_REGION_BLOCKS;
  const int nwin = (int)min((uint64_t)W, qf.nblocks - B0);
  const int nreg = (int)min((uint64_t)GQF_REGION_BLOCKS, qf.nblocks - B0);
  const int64_t win_base = (int64_t)(B0 * 64);
  const bool table_end_in_window = B0 + (uint64_t)nwin >= qf.nblocks;

  // ---- 1. stage the window into LDS (coalesced 16-byte loads) ----
  {
    const uint8_t* src = qf.blocks + B0 * GQF_BLOCK_BYTES;  // 16-byte aligned: 5184 r
    const uint32_t nbytes = (uint32_t)nwin * GQF_BLOCK_BYTES;
    const uint32_t nchunks = (nbytes + 15) / 16;
    for (uint32_t c = lane; c < nchunks; c += 64) {
      uint4 v = ((const uint4*)src)[c];
      uint32_t words[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 16; k++) {
        uint32_t o = c * 16 + k;
        if (o >= nbytes) break;
        uint8_t byte = (uint8_t)(words[k >> 2] >> (8 * (k & 3)));
        uint32_t j = o / GQF_BLOCK_BYTES;
        uint32_t f = o - j * GQF_BLOCK_BYTES;
        if (f == 0) sm.off_in[j] = byte;
        else if (f < 9) ((uint8_t*)sm.occ_in)[j * 8 + (f - 1)] = byte;
        else if (f < 17) ((uint8_t*)sm.run_in)[j * 8 + (f - 9)] = byte;
        else sm.slots_in[j * 64 + (f - 17)] = byte;
      }
    }
    for (int j = lane; j < 64; j += 64) { sm.blk_cnt[j] = 0; }
    if (lane == 0) { sm.first_invalid = nwin; sm.bstar = -1; sm.result = kOk; }
  }
  __syncthreads();

  // ---- 2. per-block item ranges (items are sorted) ----
  for (uint64_t i = s + lane; i < e; i += 64) {
    uint32_t j = (uint32_t)((items[i] >> 14) - B0);
    atomicAdd(&sm.blk_cnt[j], 1u);
  }
  __syncthreads();
  {
    uint32_t c = sm.blk_cnt[lane];
    uint32_t incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t t = __shfl_up(incl, d, 64);
      if (lane >= d) incl += t;
    }
    sm.blk_start[lane] = incl - c;
  }

  // ---- 3. old end-before for every window block (true offsets) ----
  for (int j = lane; j < nwin; j += 64) {
    uint32_t o = sm.off_in[j];
    int64_t bstart = win_base + 64 * j;
    sm.xold[j] = o < GQF_OFFSET_SAT ? bstart - 1 + (int64_t)o : true_x_global(qf, B0 + j);
  }
  __syncthreads();

  // ---- 4. pass 1: new run lengths per block ----
  for (int j = lane; j < nwin; j += 64) {
    uint64_t it = 0, ie = 0;
    if (j < nreg) {
      it = s + sm.blk_start[j];
      ie = it + sm.blk_cnt[j];
    }
    int64_t a, b;
    uint64_t occn;
    bool ok = walk_block<W, false>(qf, sm, j, win_base, nwin, items, counts, it, ie, mode, a, b, occn);
    sm.agg_a[j] = a;
    sm.agg_b[j] = b;
    sm.occ_out[j] = occn;
    if (!ok) atomicMin(&sm.first_invalid, j);
  }
  __syncthreads();

  // ---- 5. max-plus scan over blocks -> xnew ----
  {
    const int64_t e_in = sm.xold[0];
    int64_t ca = kNegInf, cb = 0;  // carry over previous chunks
    for (int base = 0; base < nwin; base += 64) {
      int j = base + lane;
      int64_t a = kNegInf, b = 0;
      if (j < nwin) { a = sm.agg_a[j]; b = sm.agg_b[j]; }
      // inclusive scan within the wave
      int64_t ia = a, ib = b;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        int64_t pa = __shfl_up(ia, d, 64);
        int64_t pb = __shfl_up(ib, d, 64);
        if (lane >= d) {
          int64_t na = pa + ib;
          ia = na > ia ? na : ia;
          ib = pb + ib;
        }
      }
      // exclusive = inclusive of lane-1, composed after the carry
      int64_t ea = __shfl_up(ia, 1, 64), eb = __shfl_up(ib, 1, 64);
      if (lane == 0) { ea = kNegInf; eb = 0; }
      int64_t pa = ca, pb = cb;
      compose(pa, pb, ea, eb);
      if (j < nwin) {
        int64_t x = pa > e_in + pb ? pa : e_in + pb;
        int64_t lo = win_base + 64 * j - 1;
        sm.xnew[j] = x > lo ? x : lo;
      }
      int64_t ta = __shfl(ia, 63, 64), tb = __shfl(ib, 63, 64);
      compose(ca, cb, ta, tb);
    }
    if (lane == 0) {
      int64_t x = ca > e_in + cb ? ca : e_in + cb;
      int64_t lo = win_base + 64 * nwin - 1;
      sm.xnew[nwin] = x > lo ? x : lo;
    }
  }
  __syncthreads();

  // ---- 6. where does the new layout meet the old one? ----
  {
    const int first_invalid = sm.first_invalid;
    int cand = nwin;  // nwin = "no meeting block inside the window"
    for (int base = GQF_REGION_BLOCKS; base < nwin; base += 64) {
      int j = base + lane;
      bool ok = j < nwin && j <= first_invalid && sm.xnew[j] == sm.xold[j] &&
                sm.xold[j] - (win_base + 64 * j - 1) < 64;
      uint64_t m = __ballot(ok);
      if (m) { cand = base + __builtin_ctzll(m); break; }
    }
    if (lane == 0) {
      int res = kOk;
      int bstar = cand;
      if (cand == nwin) {
        if (table_end_in_window && first_invalid == nwin) {
          // the table ends inside the window: everything up to the end is rewritten
          if (sm.xnew[nwin] >= (int64_t)qf.xnslots) res = kNoSpace;
        } else {
          res = kNeedWide;
        }
      } else {
        // runs placed before the meeting block must end inside the table
        if (sm.xnew[cand] >= (int64_t)qf.xnslots) res = kNoSpace;
      }
      // the spilled-in runs of the previous region must fit the window
      if (sm.xold[0] - win_base + 1 > (int64_t)nwin * 64) res = kNeedWide;
      sm.bstar = bstar;
      sm.result = res;
    }
  }
  __syncthreads();
  const int res = sm.result;
  if (res != kOk) return res;
  const int bstar = sm.bstar;                 // blocks [0, bstar) re-laid, bstar keeps its suffix
  const int nwrite = bstar < nwin ? bstar + 1 : nwin;

  // ---- 7. pass 2: lay out the runs ----
  for (int i = lane; i < nwrite * 64; i += 64) sm.slots_out[i] = 0;
  for (int j = lane; j < nwrite; j += 64) sm.run_out[j] = 0;
  __syncthreads();
  {
    // slots of the previous region's runs spilling into this one are kept
    const int64_t e_in_rel = sm.xold[0] - win_base;  // last spilled slot (relative), may be -1
    for (int64_t i = lane; i <= e_in_rel; i += 64) {
      sm.slots_out[i] = sm.slots_in[i];
      if ((sm.run_in[i >> 6] >> (i & 63)) & 1ULL)
        atomicOr((unsigned long long*)&sm.run_out[i >> 6], 1ULL << (i & 63));
    }
    // the meeting block keeps everything after the new end-before
    if (bstar < nwin) {
      const int64_t from = sm.xnew[bstar] - win_base + 1;
      const int64_t to = (int64_t)(bstar + 1) * 64;
      for (int64_t i = from + lane; i < to; i += 64) {
        sm.slots_out[i] = sm.slots_in[i];
        if ((sm.run_in[i >> 6] >> (i & 63)) & 1ULL)
          atomicOr((unsigned long long*)&sm.run_out[i >> 6], 1ULL << (i & 63));
      }
    }
  }
  __syncthreads();
  const int nrelay = bstar < nwin ? bstar : nwin;
  for (int j = lane; j < nrelay; j += 64) {
    uint64_t it = 0, ie = 0;
    if (j < nreg) {
      it = s + sm.blk_start[j];
      ie = it + sm.blk_cnt[j];
    }
    int64_t a, b;
    uint64_t occn;
    walk_block<W, true>(qf, sm, j, win_base, nwin, items, counts, it, ie, mode, a, b, occn);
  }
  __syncthreads();

  // ---- 8. store blocks [0, nwrite) (coalesced 16-byte chunks) ----
  {
    uint8_t* dst = qf.blocks + B0 * GQF_BLOCK_BYTES;
    const uint32_t nbytes = (uint32_t)nwrite * GQF_BLOCK_BYTES;
    const uint32_t nchunks = (nbytes + 15) / 16;
    for (uint32_t c = lane; c < nchunks; c += 64) {
      uint32_t words[4] = {0, 0, 0, 0};
      uint32_t valid = 0;
#pragma unroll
      for (int k = 0; k < 16; k++) {
        uint32_t o = c * 16 + k;
        if (o >= nbytes) break;
        uint32_t j = o / GQF_BLOCK_BYTES;
        uint32_t f = o - j * GQF_BLOCK_BYTES;
        uint32_t byte;
        if (f == 0) {
          if ((int)j == bstar) {
            byte = sm.off_in[j];
          } else {
            int64_t off = sm.xnew[j] - (win_base + 64 * (int64_t)j - 1);
            byte = off > GQF_OFFSET_SAT ? GQF_OFFSET_SAT : (uint32_t)off;
          }
        } else if (f < 9) {
          uint64_t w = (int)j == bstar ? sm.occ_in[j] : sm.occ_out[j];
          byte = (uint32_t)(w >> (8 * (f - 1))) & 0xFFu;
        } else if (f < 17) {
          byte = (uint32_t)(sm.run_out[j] >> (8 * (f - 9))) & 0xFFu;
        } else {
          byte = sm.slots_out[j * 64 + (f - 17)];
        }
        words[k >> 2] |= byte << (8 * (k & 3));
        valid = k + 1;
      }
      if (valid == 16) {
        ((uint4*)dst)[c] = make_uint4(words[0], words[1], words[2], words[3]);
      } else {
        for (uint32_t k = 0; k < valid; k++) dst[c * 16 + k] = (uint8_t)(words[k >> 2] >> (8 * (k & 3)));
      }
    }
  }
  return kOk;
}

}  // namespace

// ---------------------------------------------------------------------------
// kernels

extern "C" __global__ void __launch_bounds__(64)
gqf_rebuild_phase(const QF* __restrict__ qfp, const uint64_t* __restrict__ items,
                  const uint64_t* __restrict__ counts, const uint32_t* __restrict__ region_start,
                  int parity, int mode, uint8_t* __restrict__ wide_flags,
                  uint32_t* __restrict__ wide_count) {
  __shared__ Smem<kWinFast> sm;
  const QF qf = *qfp;
  const uint64_t r = 2ULL * blockIdx.x + (uint64_t)parity;
  if (r >= qf.nregions) return;
  const uint64_t s = region_start[r], e = region_start[r + 1];
  if (s == e) return;
  int res = rebuild_region<kWinFast>(qf, r, items, counts, s, e, mode, sm);
  if (lane_id() == 0) {
    if (res == kNeedWide) {
      wide_flags[r] = 1;
      atomicAdd(wide_count, 1u);
    } else if (res == kNoSpace) {
      atomicOr(qf.status, GQF_STATUS_NO_SPACE);
    }
  }
}

// Redoes flagged regions of one parity in ascending order with the wide
// window.  One workgroup; exits at once when nothing was flagged.
extern "C" __global__ void __launch_bounds__(64)
gqf_rebuild_wide(const QF* __restrict__ qfp, const uint64_t* __restrict__ items,
                 const uint64_t* __restrict__ counts, const uint32_t* __restrict__ region_start,
                 int parity, int mode, uint8_t* __restrict__ wide_flags,
                 uint32_t* __restrict__ wide_count) {
  __shared__ Smem<kWinWide> sm;
  if (*(volatile uint32_t*)wide_count == 0) return;
  const QF qf = *qfp;
  const int lane = lane_id();
  for (uint64_t base = (uint64_t)parity; base < qf.nregions; base += 128) {
    uint64_t r = base + 2ULL * lane;
    bool f = r < qf.nregions && wide_flags[r];
    uint64_t m = __ballot(f);
    while (m) {
      int l = __builtin_ctzll(m);
      m &= m - 1;
      uint64_t rr = base + 2ULL * l;
      const uint64_t s = region_start[rr], e = region_start[rr + 1];
      int res = rebuild_region<kWinWide>(qf, rr, items, counts, s, e, mode, sm);
      __syncthreads();
      if (lane == 0) {
        wide_flags[rr] = 0;
        if (res == kNeedWide) atomicOr(qf.status, GQF_STATUS_CLUSTER_OVERFLOW);
        else if (res == kNoSpace) atomicOr(qf.status, GQF_STATUS_NO_SPACE);
      }
      __syncthreads();
    }
  }
  if (lane == 0) *wide_count = 0;
}

// region_start[r] = first index i with items[i] >> 20 >= r, for r in
// [0, nregions]; items at or above `range` fall past the last region.
extern "C" __global__ void gqf_region_bounds(const QF* __restrict__ qfp, const uint64_t* __restrict__ items,
                                             uint64_t n_host, const uint64_t* __restrict__ n_dev,
                                             uint32_t* __restrict__ region_start) {
  const QF qf = *qfp;
  uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > qf.nregions) return;
  const uint64_t n = n_dev ? *n_dev : n_host;
  uint64_t lo = 0, hi = n;
  if (r == qf.nregions) {
    // first item at or above the table's hash range
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (items[mid] < qf.range) lo = mid + 1; else hi = mid;
    }
    region_start[r] = (uint32_t)lo;
    if (lo < n) atomicOr(qf.status, GQF_STATUS_BAD_HASH);
    return;
  }
  const uint64_t bound = r << GQF_REGION_SHIFT;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (items[mid] < bound) lo = mid + 1; else hi = mid;
  }
  region_start[r] = (uint32_t)lo;
}

// ---------------------------------------------------------------------------
// launchers

void gqf_launch_region_bounds(const QF* qf_dev, uint64_t nregions, const uint64_t* items, uint64_t n,
                              const uint64_t* n_dev, uint32_t* region_start, hipStream_t s) {
  const uint32_t g = (uint32_t)((nregions + 1 + 255) / 256);
  hipLaunchKernelGGL(gqf_region_bounds, dim3(g), dim3(256), 0, s, qf_dev, items, n, n_dev, region_start);
}

void gqf_launch_rebuild(const QF* qf_dev, uint64_t nregions, const uint64_t* items, const uint64_t* counts,
                        const uint32_t* region_start, int mode, uint8_t* wide_flags, uint32_t* wide_count,
                        hipStream_t s) {
  const uint32_t g = (uint32_t)((nregions + 1) / 2);
  for (int parity = 0; parity < 2; parity++) {
    hipLaunchKernelGGL(gqf_rebuild_phase, dim3(g), dim3(64), 0, s, qf_dev, items, counts, region_start, parity,
                       mode, wide_flags, wide_count);
    hipLaunchKernelGGL(gqf_rebuild_wide, dim3(1), dim3(64), 0, s, qf_dev, items, counts, region_start, parity, mode,
                       wide_flags, wide_count);
  }
}
