// gqf_rebuild.hip -- bulk mutation of the CQF image by region merge-rebuild.
//
// Replaces the reference's even-odd scheduler in which one *thread* owns an
// 8192-slot region and calls qf_insert / qf_remove once per key
// (src/gqf.cu:2759-2906, 3184-3363).  Here one wave64 owns a region of 48
// blocks plus a 16-block push window into the next region, one lane per
// block, and applies ALL of the region's keys in one pass:
//
//   1. lane j loads window block j's 128-byte line with 16-byte loads: the
//      header stays in registers, the 64 slots go to LDS row j; the
//      region's sorted items are cached in LDS as 20-bit relative hashes;
//   2. each lane walks its block's runs in the old layout (starting from the
//      stored offset), merges the block's items counter by counter, and
//      reduces the new run lengths to a max-plus aggregate (a, b): the
//      block maps "end of the runs before it" x to max(a, x + b);
//   3. a wave scan of the aggregates places every run at
//      max(quotient, end of previous run + 1) -- the canonical CQF layout;
//   4. the push into the next region's first blocks (a pure shift of their
//      runs) ends at the first block whose end-before is unchanged;
//   5. the new slots are assembled in LDS rows and every rewritten block's
//      line is stored with 16-byte writes (header, exact offset, slots).
//
// The result is the image the reference's sequential inserts and removes
// produce (counters in encode_counter form; the exported offset byte is the
// exact offset saturated at 255), because a CQF image is a function of the
// multiset of hashes alone.
//
// Regions run in two phases (even, then odd) so a push never races the
// wave that owns the next region.  A region whose push is not absorbed in
// its window is flagged and redone by a single-workgroup pass with a wide
// window, in ascending region order, after the phase.
#include <stdlib.h>

#include <atomic>
#include <type_traits>
#include <utility>

#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {

constexpr int kWinFast = 64;          // region blocks + push blocks
constexpr int kWinMid = 128;          // first fallback window (80 push blocks)
constexpr int kWinWide = 256;         // last fallback window (208 push blocks)
constexpr int kWideClassGrid = 256;  // workgroups of a class launch of the wide pass
constexpr int kItemCap = 256;         // items cached in LDS per region (fast path): 192 +- 14 at 2^24 keys into 2^28 slots
constexpr int32_t kNegInf = INT32_MIN / 4;

enum RebuildMode : int { kInsert = 0, kDelete = 1, kSeq = 2 };
enum RebuildResult : int { kOk = 0, kNeedWide = 1, kNoSpace = 2 };

// Run lengths of the listed quotients: one byte in the fast window (a
// longer run sends the region to the wider windows, which keep 16 bits)
template <int W>
using RunLen = std::conditional_t<W == kWinFast, uint8_t, uint16_t>;
template <int W>
constexpr uint32_t kRunLenMax = W == kWinFast ? 255u : 65535u;

template <int W>
struct Smem {
  uint32_t sin[W * 17];   // old slots: slot p at byte sin_off(p), rows of 68 bytes (bank-conflict free)
  uint32_t sout[W * 17];  // new slots, same rows (word 16 of a row, its pad, takes junk writes)
  uint64_t run_in[W];
  // arrays whose lifetimes do not overlap share space
  union {
    uint64_t occ_in[W];    // old occupieds, until the quotient table is built
    uint64_t run_out[W];   // new run ends, from pass 2 on
  };
  union {
    uint64_t tmask[W];         // quotients of block j touched by the items, until
                               // the walks take them into registers (fast path)
    int16_t tsn[kItemCap];     // new run start of each listed quotient, from pass 2 on
  };
  int32_t xold[W + 1];  // end of the runs of quotients < 64 j (window-relative, >= 64 j - 1)
  int32_t xnew[W + 1];
  union {
    uint32_t items[kItemCap];  // fast path: hash - first hash of the region
    struct {
      uint32_t cnt[64];          // general path (items not cached): items per region block
      uint32_t start[64];        // and the index of each block's first item
    } blk;
  };
  // one entry per distinct quotient among the items (fast path); the
  // quotient itself is items[tit0[t]] >> 8
  uint8_t tit0[kItemCap];    // first item of the quotient (< kItemCap = 256)
  int16_t tst[kItemCap];     // old run start
  RunLen<W> tol[kItemCap];   // old run length (0: quotient was empty)
  RunLen<W> tnl[kItemCap];   // new run length
  uint32_t nq;
  int32_t first_invalid;
  int32_t xend;
  int32_t bstar;
  int32_t result;
};

__device__ inline int lane_id() { return (int)__lane_id(); }

// byte offset of window-relative slot p in the slot rows (sin / sout): one
// 68-byte row per block, so lanes reading "their" block at the same column
// hit different LDS banks
__device__ inline uint32_t sin_off(int32_t p) {
  return (uint32_t)p + ((uint32_t)(p >> 6) << 2);  // = 68 (p >> 6) + (p & 63), for p >= 0
}

__device__ inline void compose(int32_t& a, int32_t& b, int32_t a2, int32_t b2) {
  // (a, b) then (a2, b2): x -> max(a2, max(a, x + b) + b2)
  int32_t na = a + b2;
  a = na > a2 ? na : a2;
  b = b + b2;
}

// Inclusive compose scan over the 64 lanes, on the same DPP steps as
// wave_scan_add (gqf_internal.h); lanes a shift reaches past their row's
// start read the identity map (kNegInf, 0).  All lanes must be active.
template <int kCtrl, int kRows>
__device__ __forceinline__ void scan_step_compose(int32_t& a, int32_t& b) {
  const int32_t pa = __builtin_amdgcn_update_dpp(kNegInf, a, kCtrl, kRows, 0xf, false);
  const int32_t pb = __builtin_amdgcn_update_dpp(0, b, kCtrl, kRows, 0xf, false);
  const int32_t na = pa + b;  // compose(pa, pb, a, b)
  a = na > a ? na : a;
  b = pb + b;
}
__device__ __forceinline__ void wave_scan_compose(int32_t& a, int32_t& b) {
  scan_step_compose<0x111, 0xf>(a, b);
  scan_step_compose<0x112, 0xf>(a, b);
  scan_step_compose<0x114, 0xf>(a, b);
  scan_step_compose<0x118, 0xf>(a, b);
  scan_step_compose<0x142, 0xa>(a, b);
  scan_step_compose<0x143, 0xc>(a, b);
}
// the value of the lane below (lane 0: fill); wave_shr:1
__device__ __forceinline__ int32_t wave_shr1(int32_t x, int32_t fill) {
  return __builtin_amdgcn_update_dpp(fill, x, 0x138, 0xf, 0xf, false);
}


// The region's items: cached relative hashes in LDS, or (when the region
// holds more than kItemCap) read from global memory.  Each item changes its
// hash's count by a delta: +1 (insert), -1 (delete), +count (reduce mode,
// counts array), or +1 / -1 by the item's tag bit 63 (sequenced mode, whose
// items carry a request tag above bit GQF_SEQ_SHIFT that hmask strips).
// Deltas other than +-1 are cached in a separate LDS array that only the
// counting kernels declare, so the plain insert and delete kernels keep
// their LDS small.
enum ItemKind : int { kAdd = 0, kSub = 1, kCounts = 2, kTagged = 3 };

struct Items {
  const uint64_t* __restrict__ g;
  const uint64_t* __restrict__ gc;
  uint32_t* lcnt;  // LDS cache of the deltas (cached, kinds kCounts and kTagged)
  uint64_t s;      // first item index of the region
  uint64_t base;   // first hash of the region
  uint64_t hmask;  // hash bits of an item
  int kind;
  bool cached;
  __device__ inline uint32_t hash(const uint32_t* lds, uint32_t i) const {
    return cached ? lds[i] : (uint32_t)((g[s + i] & hmask) - base);
  }
  __device__ inline int64_t delta(uint32_t i) const {
    if (kind == kAdd) return 1;
    if (kind == kSub) return -1;
    if (kind == kCounts) return cached ? (int64_t)lcnt[i] : (int64_t)gc[s + i];
    return cached ? (int64_t)(int32_t)lcnt[i] : ((g[s + i] >> 63) ? -1 : 1);
  }
  // the value cached in lcnt for item i (global index s + i)
  __device__ inline uint32_t cache_delta(uint64_t i) const {
    return kind == kCounts ? (uint32_t)gc[i] : ((g[i] >> 63) ? 0xFFFFFFFFu : 1u);
  }
};

// The straight-line insert merge (one or two items with distinct hashes on
// a run of count-1 counters, at most 8 slots once merged; see
// merge_quotient), split in two so a caller can keep what the reading half
// finds in registers until the writing half.  lean_read reads the old run
// as aligned words with the row's pad word dropped, checks it, and returns
// the merged run little-endian in lo / hi (byte k: slot k of the new run)
// with its length; lean_write ORs those bytes into the (cleared) new rows,
// at most three words.  A row's 64 slots end on a word boundary, so the pad
// word of the row holding slot st is word 16 - (st & 63) / 4 counted from
// the word holding st.
// The 8 old slots from st (whatever follows a shorter run included), little
// endian in x0 / x1: the words holding them read aligned, the row's pad
// word dropped, and funnel-shifted to st's byte.
template <int W>
__device__ __forceinline__ void read_slots8(const Smem<W>& sm, int32_t st, uint32_t& x0, uint32_t& x1) {
  const uint32_t a = sin_off(st);
  const uint32_t* wp = sm.sin + (a >> 2);
  const uint32_t w0 = wp[0], w1 = wp[1], w2 = wp[2], w3 = wp[3];
  const uint32_t pad = 16u - ((uint32_t)(st & 63) >> 2);
  const uint32_t v1 = pad == 1 ? w2 : w1;
  const uint32_t v2 = pad <= 2 ? w3 : w2;
  const uint32_t sh = (a & 3u) * 8u;
  x0 = __builtin_amdgcn_alignbit(v1, w0, sh);
  x1 = __builtin_amdgcn_alignbit(v2, v1, sh);
}

template <int W>
__device__ __forceinline__ bool lean_read(const Smem<W>& sm, const Items& src, uint32_t it0, uint32_t it1,
                                          int32_t st_o, uint32_t old_len, uint32_t h_first, uint32_t& lo,
                                          uint32_t& hi, uint32_t& len) {
  const uint32_t n = it1 - it0;
  len = old_len + n;
  const uint32_t r0 = h_first & 0xFFu;
  const uint32_t r1 = n == 2 ? (src.hash(sm.items, it0 + 1) & 0xFFu) : 0x200u;
  uint32_t x0, x1;
  read_slots8(sm, st_o, x0, x1);
  // bytes past the run as sentinels 0x100 + k (ascending, above every byte)
  uint32_t ob[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t b = ((k < 4 ? x0 : x1) >> (8 * (k & 3))) & 0xFFu;
    ob[k] = (uint32_t)k < old_len ? b : 0x100u + (uint32_t)k;
  }
  bool simple = len <= 8 && (n == 1 || r1 > r0);
#pragma unroll
  for (int k = 1; k < 8; k++) simple = simple && ob[k] > ob[k - 1];
  // r0 goes after the old bytes <= r0, r1 after those <= r1 and r0
  uint32_t a0 = 0, a1 = 1;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    a0 += ob[k] <= r0 ? 1u : 0u;
    a1 += ob[k] <= r1 ? 1u : 0u;
  }
  auto insert = [](uint64_t v, uint32_t at, uint32_t r) {  // byte r at `at`, the bytes from `at` up one
    at = at < 7u ? at : 7u;                                  // (only a run that is not simple passes 7)
    const uint64_t m = gqf_mask(8u * at);
    return (v & m) | ((uint64_t)r << (8u * at)) | ((v & ~m) << 8);
  };
  uint64_t x = ((uint64_t)x1 << 32 | x0) & gqf_mask(8u * old_len);
  x = insert(x, a0, r0);
  if (n == 2) x = insert(x, a1, r1);
  lo = (uint32_t)x;
  hi = (uint32_t)(x >> 32);
  return simple;
}

template <int W>
__device__ __forceinline__ void lean_write(Smem<W>& sm, uint32_t lo, uint32_t hi, uint32_t len, int32_t st_n) {
  const uint32_t a = sin_off(st_n);
  const uint32_t sh = (a & 3u) * 8u;
  const uint64_t u = (uint64_t)lo << sh, t = (uint64_t)hi << sh;
  const uint32_t y0 = (uint32_t)u, y1 = (uint32_t)(u >> 32) | (uint32_t)t, y2 = (uint32_t)(t >> 32);
  const uint32_t pad = 16u - ((uint32_t)(st_n & 63) >> 2);
  uint32_t* wp = sm.sout + (a >> 2);
  const uint32_t used = (a & 3u) + len;  // bytes of the three words the run takes
  atomicOr(&wp[0], y0);
  if (used > 4) atomicOr(&wp[pad == 1 ? 2 : 1], y1);
  if (used > 8) atomicOr(&wp[pad <= 2 ? 3 : 2], y2);
}

// The straight-line delete merge: one or two items, each removing one
// instance (an absent key's removal is a no-op), on an old run of at most 8
// strictly ascending bytes.  Such a run holds count-1 counters only (a
// count of 2 or more repeats its remainder), so each item's remainder, if
// present, loses its one byte and the rest stay in order.  Returns false
// (the general merge, merge_quotient) otherwise; lo / hi / len as
// lean_read's.  Reads only the old rows, so the measuring and the writing
// pass can each call it.
template <int W>
__device__ __forceinline__ bool lean_delete(const Smem<W>& sm, const Items& src, uint32_t it0, uint32_t it1,
                                            int32_t st_o, uint32_t old_len, uint32_t h_first, uint32_t& lo,
                                            uint32_t& hi, uint32_t& len) {
  const uint32_t n = it1 - it0;
  if (old_len == 0 || old_len > 8 || n > 2) return false;
  const uint32_t r0 = h_first & 0xFFu;
  const uint32_t r1 = n == 2 ? (src.hash(sm.items, it0 + 1) & 0xFFu) : r0;
  uint32_t x0, x1;
  read_slots8(sm, st_o, x0, x1);
  bool asc = true;
  uint32_t prev = 0, kept = 0;
  uint64_t y = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t b = ((k < 4 ? x0 : x1) >> (8 * (k & 3))) & 0xFFu;
    const bool in = (uint32_t)k < old_len;
    if (k > 0 && in && b <= prev) asc = false;
    prev = b;
    const bool keep = in && b != r0 && b != r1;
    y |= keep ? (uint64_t)b << (8u * kept) : 0ULL;
    kept += keep ? 1u : 0u;
  }
  lo = (uint32_t)y;
  hi = (uint32_t)(y >> 32);
  len = kept;
  return asc;
}

// ---------------------------------------------------------------------------
// Two-pointer merge of the old counters of one quotient (ascending
// remainders, old run at st_o .. st_o + old_len - 1 of sin) with its items
// [it0, it1) (ascending hashes; h_first is item it0's, which the caller
// has already read).  Returns the new run length; with WRITE the counters
// are written from st_n into raw.
//
// The items of one hash are applied in order, as a sequence of the
// reference's qf_insert (count + 1, src/gqf.cu:1523-1752) and qf_remove
// (count - 1, floored at 0: removing an absent key is a no-op,
// src/gqf.cu:1755-1933).  Each step maps the count x to max(m, x + s), and
// such maps compose: +d gives (m + d, s + d), -d gives (max(0, m - d), s - d),
// from the identity (0, 0).  m >= s and m >= 0 throughout, so the new count
// is old + s when old >= m - s, else m.

template <int W, bool WRITE>
__device__ __forceinline__ uint32_t merge_quotient(Smem<W>& sm, const Items& src, uint32_t it0, uint32_t it1,
                                                   int32_t st_o, uint32_t old_len, int32_t st_n, uint32_t h_first) {
  const uint8_t* in = (const uint8_t*)sm.sin;
  uint8_t* out = (uint8_t*)sm.sout;
  // Inserts (kAdd): one or two items, with distinct hashes, on a run of at
  // most 8 count-1 counters (strictly ascending bytes, see below) -- nearly
  // every quotient of a batch.  The new run is the old bytes and the new
  // ones merged, an old byte before an equal new one (r r is the count-2
  // counter of r).  Straight-line code, the same on every lane (lean_read,
  // lean_write), for merged runs of at most 8 slots.
  if (src.kind == kAdd && it1 - it0 <= 2 && old_len <= 8) {
    uint32_t lo, hi, len;
    if (lean_read(sm, src, it0, it1, st_o, old_len, h_first, lo, hi, len)) {
      if (WRITE) lean_write(sm, lo, hi, len, st_n);
      return len;
    }
  }
  // Common case: one item adding or removing 1 on a run (at most 8 slots) of
  // count-1 counters.  Such a run's bytes strictly ascend -- a counter of
  // count >= 2 repeats or drops a byte (r r, r 0 r, 0 0 0, r [0] d.. r,
  // 0 d.. 0 0) -- and each byte is a remainder.  Adding r puts a byte r before
  // the first larger byte (r r, if r was there, is its count-2 counter);
  // removing r drops its byte, or leaves the run as it is when r is absent.
  // The old bytes are read together, one LDS round trip.
  if (it1 == it0 + 1 && old_len <= 8) {
    const int64_t d = src.delta(it0);
    if (d == 1 || d == -1) {
      const uint32_t r = h_first & 0xFFu;
      uint32_t ob[8];
#pragma unroll
      for (int k = 0; k < 8; k++) ob[k] = (uint32_t)k < old_len ? in[sin_off(st_o + k)] : 0u;
      bool simple = true;
      uint32_t pos = old_len;  // first byte above r
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if ((uint32_t)k < old_len) {
          if (k && ob[k] <= ob[k - 1]) simple = false;
          if (pos == old_len && ob[k] > r) pos = (uint32_t)k;
        }
      }
      if (simple && d == 1) {
        if (WRITE) {
#pragma unroll
          for (int k = 0; k < 8; k++)
            if ((uint32_t)k < old_len) out[sin_off(st_n + k + ((uint32_t)k >= pos ? 1 : 0))] = (uint8_t)ob[k];
          out[sin_off(st_n + (int32_t)pos)] = (uint8_t)r;
        }
        return old_len + 1;
      }
      if (simple) {
        const bool present = pos > 0 && ob[pos - 1] == r;
        const uint32_t gone = present ? pos - 1 : 8u;  // the removed byte
        if (WRITE) {
#pragma unroll
          for (int k = 0; k < 8; k++)
            if ((uint32_t)k < old_len && (uint32_t)k != gone)
              out[sin_off(st_n + k - ((uint32_t)k > gone ? 1 : 0))] = (uint8_t)ob[k];
        }
        return present ? old_len - 1 : old_len;
      }
    }
  }
  uint32_t len = 0;
  auto old_slot = [&](uint32_t k) -> uint64_t { return in[sin_off(st_o + (int32_t)k)]; };
  auto put = [&](uint32_t k, uint64_t v) {
    if (WRITE) out[sin_off(st_n + (int32_t)(len + k))] = (uint8_t)v;
  };
  uint32_t oi = 0, o_end = 0;
  uint64_t o_rem = 0, o_cnt = 0;
  bool have_old = old_len > 0;
  if (have_old) o_end = gqf_decode_counter(old_slot, old_len, 0, &o_rem, &o_cnt);
  uint64_t n_rem = 0;
  int64_t n_m = 0, n_s = 0;
  bool have_new = false;
  uint32_t it = it0;
  uint32_t h_it = h_first;  // hash of item `it`, loaded once (items may be in global memory)
  auto next_new = [&]() {
    have_new = it < it1;
    if (!have_new) return;
    const uint32_t h0 = h_it;
    n_rem = h0 & 0xFF;
    n_m = 0;
    n_s = 0;
    do {
      const int64_t d = src.delta(it);
      n_s += d;
      n_m = (d > 0 || n_m + d > 0) ? n_m + d : 0;
      it++;
      h_it = it < it1 ? src.hash(sm.items, it) : 0u;
    } while (it < it1 && h_it == h0);
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
        for (uint32_t k = 0; k < clen; k++) out[sin_off(st_n + (int32_t)(len + k))] = in[sin_off(st_o + (int32_t)(oi + k))];
      len += clen;
      next_old();
    } else if (!have_old || n_rem < o_rem) {
      len += gqf_emit_counter(n_rem, (uint64_t)n_m, put);  // from count 0: max(m, s) = m
      next_new();
    } else {
      const uint64_t c = o_cnt >= (uint64_t)(n_m - n_s) ? o_cnt + (uint64_t)n_s : (uint64_t)n_m;
      len += gqf_emit_counter(n_rem, c, put);
      next_old();
      next_new();
    }
  }
  return len;
}

// ---------------------------------------------------------------------------
// One block's walk: the old runs of block j merged with its items.  Pass 1
// (WRITE = false) returns the block's max-plus aggregate and new occupieds
// word; pass 2 (WRITE = true) writes the runs into raw from xnew[j].

// Start and end bits of the runs a pass-2 walk writes from block j's
// quotients, inside block j (a run that leaves the block ends at bit 63).
// The runs are disjoint and ascend, so (e - s) | e sets exactly their
// slots; with the slots that earlier blocks' runs cover (the new offset)
// that is block j's new used-slot mask, without a walk over its slots.
struct RunBits {
  uint64_t s = 0, e = 0;
  __device__ inline void add(int32_t rs, int32_t re) {  // run [rs, re] relative to the block, rs >= 0
    if (rs < 64) {
      s |= 1ULL << rs;
      e |= 1ULL << (re < 63 ? re : 63);
    }
  }
  __device__ inline uint64_t used(uint32_t off) const {  // off: leading slots covered by earlier runs
    return off >= 64 ? ~0ULL : (gqf_mask(off) | (e - s) | e);
  }
};

template <int W, bool WRITE>
__device__ __forceinline__ bool walk_block(Smem<W>& sm, int j, int nwin, const Items& src, uint32_t it, uint32_t ie,
                           uint64_t occ, int32_t& agg_a, int32_t& agg_b, uint64_t& newocc, RunBits& rb) {
  const uint8_t* in = (const uint8_t*)sm.sin;
  uint8_t* out = (uint8_t*)sm.sout;
  const int32_t blk_base = 64 * j;
  int32_t xo = sm.xold[j];
  int32_t xn = WRITE ? sm.xnew[j] : 0;
  agg_a = kNegInf;
  agg_b = 0;
  newocc = 0;
  int cw_blk = -1;
  uint64_t cw = 0;

  uint32_t cur = it < ie ? src.hash(sm.items, it) : 0;
  for (;;) {
    const int qo = occ ? __builtin_ctzll(occ) : 64;
    const int qi = it < ie ? (int)((cur >> 8) & 63) : 64;
    const int qrel = qo < qi ? qo : qi;
    if (qrel == 64) break;
    const int32_t qa = blk_base + qrel;

    // old run of qa: from max(qa, xo + 1) to the next run end
    int32_t st_o = 0, en_o = -1;
    if (qo == qrel) {
      occ &= occ - 1;
      st_o = qa > xo + 1 ? qa : xo + 1;
      int bj = st_o >> 6;
      if (bj >= nwin) return false;
      if (bj != cw_blk) { cw_blk = bj; cw = sm.run_in[bj]; }
      uint64_t w = cw >> (st_o & 63);
      int32_t pos = st_o;
      while (!w) {
        bj++;
        if (bj >= nwin) return false;
        cw_blk = bj;
        cw = sm.run_in[bj];
        w = cw;
        pos = 64 * bj;
      }
      en_o = pos + __builtin_ctzll(w);
      xo = en_o;
    }
    const uint32_t old_len = (uint32_t)(en_o - st_o + 1);  // 0 if no old run

    int32_t st_n = 0;
    if (WRITE) st_n = qa > xn + 1 ? qa : xn + 1;
    uint32_t len = 0;

    if (qi != qrel) {
      // untouched run: same bytes at the new position
      len = old_len;
      if (WRITE)
        for (uint32_t k = 0; k < len; k++) out[sin_off(st_n + (int32_t)k)] = in[sin_off(st_o + (int32_t)k)];
    } else {
      // cur is item it's hash (its quotient is qa): each item is loaded once
      const uint32_t it0 = it;
      const uint32_t h0 = cur;
      do {
        it++;
        cur = it < ie ? src.hash(sm.items, it) : 0u;
      } while (it < ie && (int32_t)(cur >> 8) == qa);
      len = merge_quotient<W, WRITE>(sm, src, it0, it, st_o, old_len, st_n, h0);
    }

    if (len) {
      newocc |= 1ULL << qrel;
      compose(agg_a, agg_b, qa + (int32_t)len - 1, (int32_t)len);
      if (WRITE) {
        const int32_t en_n = st_n + (int32_t)len - 1;
        atomicOr((unsigned long long*)&sm.run_out[en_n >> 6], 1ULL << (en_n & 63));
        rb.add(st_n - blk_base, en_n - blk_base);
        xn = en_n;
      }
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// Fast path (all of a region's items cached in LDS).  The quotients the
// items touch are listed once (table tit0: each quotient's first item);
// each lane then locates the old run of one listed quotient straight from
// its block's header -- the rank
// of the quotient among the block's occupied quotients selects its run end
// among the run ends after the block's end-before -- and merges it with the
// quotient's items.  The per-block walk then only reads the table, so its
// loop body is the same for every lane.

template <int W>
__device__ __forceinline__ bool locate_old_run(const Smem<W>& sm, int nwin, int32_t q, int32_t& st_o,
                                               uint32_t& old_len) {
  const int j = q >> 6;
  const int bit = q & 63;
  const uint64_t occ = sm.occ_in[j];
  if (!((occ >> bit) & 1ULL)) {
    st_o = q;
    old_len = 0;
    return true;
  }
  const int rank = __popcll(occ & ((2ULL << bit) - 1ULL));
  int32_t prev = sm.xold[j], end = -1;
  const int32_t pos = sm.xold[j] + 1;
  int bj = pos >> 6;
  if (bj >= nwin) return false;
  uint64_t w = sm.run_in[bj] & (~0ULL << (pos & 63));
  int seen = 0;
  int32_t last = prev;  // the last run end seen in earlier words
  for (;;) {
    const int c = __popcll(w);
    if (rank - 1 < seen + c) {
      // q's run end; the one before it (the highest set bit below, or the
      // last one of an earlier word) ends q's predecessor
      const int e = gqf_select64(w, rank - 1 - seen);
      end = 64 * bj + e;
      const uint64_t below = w & gqf_mask((uint64_t)e);
      if (rank >= 2) prev = below ? 64 * bj + 63 - __builtin_clzll(below) : last;
      break;
    }
    if (c) last = 64 * bj + 63 - __builtin_clzll(w);
    seen += c;
    if (++bj >= nwin) return false;
    w = sm.run_in[bj];
  }
  st_o = q > prev + 1 ? q : prev + 1;
  old_len = (uint32_t)(end - st_o + 1);
  return true;
}

template <int W, bool WRITE>
__device__ __forceinline__ bool walk_block_table(Smem<W>& sm, int j, int nwin, uint32_t tbase, uint64_t occ,
                                                 uint64_t touched, int32_t& agg_a, int32_t& agg_b,
                                                 uint64_t& newocc, RunBits& rb) {
  const uint8_t* in = (const uint8_t*)sm.sin;
  uint8_t* out = (uint8_t*)sm.sout;
  const int32_t blk_base = 64 * j;
  int32_t xo = sm.xold[j];
  int32_t xn = WRITE ? sm.xnew[j] : 0;
  agg_a = kNegInf;
  agg_b = 0;
  newocc = 0;
  // run ends of this block's runs, as a bit stream from xo + 1 on: the k-th
  // occupied quotient owns the k-th set bit, so each one is ctz + clear
  int bj = (xo + 1) >> 6;
  uint64_t w = 0;
  if (occ) {
    if (bj >= nwin) return false;
    w = sm.run_in[bj] & (~0ULL << ((xo + 1) & 63));
  }
  uint64_t u = occ | touched;
  // the table entry of the block's next listed quotient: they are visited in
  // ascending order, so the k-th one met is entry tbase + k
  uint32_t tn = tbase;
  if constexpr (!WRITE) {
    // Pass 1 without branches apart from the step to the next run-end
    // word: each quotient's old run end, its table length when listed, and
    // the fold of its run into the aggregate are computed on every lane and
    // selected, so the lanes stay converged through the loop.
    while (u) {
      const int q = __builtin_ctzll(u);
      const uint64_t bit = 1ULL << q;
      u ^= bit;
      const int32_t qa = blk_base + q;
      const bool occupied = (occ & bit) != 0;
      if (occupied && !w) {
        do {
          if (++bj >= nwin) return false;
          w = sm.run_in[bj];
        } while (!w);
      }
      const int32_t en_o = 64 * bj + __builtin_ctzll(w | (1ULL << 63));  // w != 0 when occupied
      const int32_t st_o = qa > xo + 1 ? qa : xo + 1;
      const uint32_t tl = sm.tnl[tn < (uint32_t)kItemCap ? tn : (uint32_t)kItemCap - 1u];  // used when listed
      const bool listed = (touched & bit) != 0;
      const uint32_t len = listed ? tl : (occupied ? (uint32_t)(en_o - st_o + 1) : 0u);
      tn += listed ? 1u : 0u;
      w = occupied ? (w & (w - 1)) : w;
      xo = occupied ? en_o : xo;
      const int32_t na = agg_a + (int32_t)len, la = qa + (int32_t)len - 1;
      agg_a = len ? (na > la ? na : la) : agg_a;  // compose(agg, (qa + len - 1, len)) for a run
      agg_b += (int32_t)len;
      newocc |= len ? bit : 0ULL;
    }
    return true;
  }
  // Pass 2: the runs written at their new starts.  Only the listed
  // quotients' table update is a branch; an unlisted run's slots are copied
  // on every lane (a listed one copies 0 bytes: its counters are written
  // after the walk) and the bit updates are selects.
  while (u) {
    const int q = __builtin_ctzll(u);
    const uint64_t bit = 1ULL << q;
    u ^= bit;
    const int32_t qa = blk_base + q;
    const bool occupied = (occ & bit) != 0;
    if (occupied && !w) {
      do {
        if (++bj >= nwin) return false;
        w = sm.run_in[bj];
      } while (!w);
    }
    const int32_t en_o = 64 * bj + __builtin_ctzll(w | (1ULL << 63));  // w != 0 when occupied
    const int32_t st_o = qa > xo + 1 ? qa : xo + 1;
    w = occupied ? (w & (w - 1)) : w;
    xo = occupied ? en_o : xo;
    const int32_t st_n = qa > xn + 1 ? qa : xn + 1;
    uint32_t len, copy;
    if (touched & bit) {
      // lengths from the table; the counters are written after the walk
      len = sm.tnl[tn];
      sm.tsn[tn] = (int16_t)st_n;
      tn++;
      copy = 0;
    } else {
      len = copy = (uint32_t)(en_o - st_o + 1);  // u = occ | touched: an unlisted quotient is occupied
    }
    if (copy <= 8) {
      // as words: the old slots read aligned, ORed into the cleared rows (a
      // listed run ORs 0 into its block's first word)
      uint32_t x0, x1;
      read_slots8(sm, copy ? st_o : blk_base, x0, x1);
      const uint64_t x = ((uint64_t)x1 << 32 | x0) & gqf_mask(8u * copy);
      lean_write(sm, (uint32_t)x, (uint32_t)(x >> 32), copy, copy ? st_n : blk_base);
    } else {
      for (uint32_t k = 0; k < copy; k++) out[sin_off(st_n + (int32_t)k)] = in[sin_off(st_o + (int32_t)k)];
    }
    // the run's end bit, its start and end bits inside the block (RunBits),
    // the new end-before; nothing for a run deleted to length 0
    const int32_t en_n = st_n + (int32_t)len - 1;
    const int32_t ei = len ? en_n : blk_base;
    atomicOr((unsigned long long*)&sm.run_out[ei >> 6], len ? 1ULL << (ei & 63) : 0ULL);
    const int32_t rs = st_n - blk_base, re = en_n - blk_base;  // rs >= 0: st_n > xnew[j] >= 64 j - 1
    const bool in_blk = len && rs < 64;
    rb.s |= in_blk ? 1ULL << (rs & 63) : 0ULL;
    rb.e |= in_blk ? 1ULL << (re < 63 ? re : 63) : 0ULL;
    xn = len ? en_n : xn;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Shift path: insert batches whose items are all cached (one item adds one
// instance).  Inserts only grow runs, so the new layout is the old one with
// each byte moved right by a shift that
//   * grows by (new length - old length) at every listed quotient's event
//     position -- the start of its old run, or for a quotient without one
//     the slot a new run would take, max(q, end of the runs before it + 1);
//   * drops by one, down to 0, at every empty slot (an empty slot absorbs
//     one unit: the run after it is at home, max(q, ...) takes q);
//   * is unchanged across used slots.
// Each of these is a map x -> max(a, x + b), and maps compose, so lane j
// folds slot-block j's events and empty slots (from the block's stored
// used-slot mask) into one map and a wave scan gives the shift entering
// every block.  The untouched bytes then move as whole stretches between
// events and empty slots, instead of run by run; only the listed quotients'
// runs are re-encoded (merge_quotient).  Offsets come from the new run-end
// bits by rank: the runs of quotients < 64 j end at the k-th run end of the
// window, k = runs ending before it + occupied quotients of blocks < j.

// position of a listed quotient q's event: its old run [st, st + ol) if q
// is occupied, else st = the slot its new run would start at (ol = 0);
// false if the run ends past the window
template <int W>
__device__ __forceinline__ bool locate_run_or_slot(const Smem<W>& sm, int nwin, int32_t q, int32_t& st,
                                                   uint32_t& ol) {
  const int j = q >> 6;
  const int bit = q & 63;
  const uint64_t occ = sm.occ_in[j];
  const bool occupied = (occ >> bit) & 1ULL;
  const int before = __popcll(occ & ((1ULL << bit) - 1ULL));  // block j's occupied quotients below q
  const int need = occupied ? before + 1 : before;            // run ends to find after xold[j]
  int32_t prev = sm.xold[j], end = -1;
  if (need > 0) {
    // the need-th run end after xold[j]; for an occupied q that is q's own
    // run end, and the one before it (the highest set bit below, or the
    // last one of an earlier word) ends q's predecessor
    const int32_t pos = sm.xold[j] + 1;
    int bj = pos >> 6;
    if (bj >= nwin) return false;
    uint64_t w = sm.run_in[bj] & (~0ULL << (pos & 63));
    int seen = 0;
    int32_t last = prev;
    for (;;) {
      const int c = __popcll(w);
      if (need - 1 < seen + c) {
        const int e = gqf_select64(w, need - 1 - seen);
        if (occupied) {
          end = 64 * bj + e;
          const uint64_t below = w & gqf_mask((uint64_t)e);
          if (before >= 1) prev = below ? 64 * bj + 63 - __builtin_clzll(below) : last;
        } else {
          prev = 64 * bj + e;
        }
        break;
      }
      if (c) last = 64 * bj + 63 - __builtin_clzll(w);
      seen += c;
      if (++bj >= nwin) return false;
      w = sm.run_in[bj];
    }
  }
  st = q > prev + 1 ? q : prev + 1;
  ol = occupied ? (uint32_t)(end - st + 1) : 0u;
  return st < 64 * nwin;
}

__device__ inline uint64_t range_mask(int lo, int hi) {  // bits [lo, hi) of a word, 0 <= lo <= hi <= 64
  return gqf_mask((uint64_t)hi) & ~gqf_mask((uint64_t)lo);
}

// f(integral_constant<int, i>) for i = 0, 1, ..., in order, as straight-line code
template <typename F, int... I>
__device__ __forceinline__ void for_each_slot(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}

// Bit b of x as 0 or ~0u, bit b as 0 or 1, min(x, 1) and (x << sh) | acc:
// each one VALU instruction.  As asm the optimizer keeps them as values (it
// would otherwise rewrite the masks and bits applied with them into
// compare-and-select chains).
template <int b>
__device__ __forceinline__ uint32_t bit_mask(uint32_t x) {
  static_assert(b >= 0 && b < 32, "bit index");
  uint32_t r;
  asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(r) : "v"(x), "i"(b));
  return r;
}
template <int b>
__device__ __forceinline__ uint32_t bit_u(uint32_t x) {  // bit b of x, 0 or 1
  static_assert(b >= 0 && b < 32, "bit index");
  uint32_t r;
  asm("v_bfe_u32 %0, %1, %2, 1" : "=v"(r) : "v"(x), "i"(b));
  return r;
}
__device__ __forceinline__ uint32_t min1(uint32_t x) {
  uint32_t r;
  asm("v_min_u32 %0, %1, 1" : "=v"(r) : "v"(x));
  return r;
}
__device__ __forceinline__ uint32_t lshl_or(uint32_t x, uint32_t sh, uint32_t acc) {  // (x << sh) | acc
  uint32_t r;
  asm("v_lshl_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(sh), "v"(acc));
  return r;
}
template <int sh>
__device__ __forceinline__ uint32_t lshl_or_c(uint32_t x, uint32_t acc) {  // (x << sh) | acc
  uint32_t r;
  asm("v_lshl_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "i"(sh), "v"(acc));
  return r;
}

// ORs the low `len` bits of `bits` into the window bit array at position d
// (len <= 64; the bits may straddle two words)
__device__ inline void or_bits(uint64_t* arr, int32_t d, uint64_t bits, int len) {
  const int w = d >> 6, sh = d & 63;
  atomicOr((unsigned long long*)&arr[w], bits << sh);
  if (sh && sh + len > 64) atomicOr((unsigned long long*)&arr[w + 1], bits >> (64 - sh));
}

template <int W, bool kStash>
__device__ __forceinline__ int rebuild_insert_shift(const QF& qf, uint64_t B0, int R, int nwin, const Items& src,
                                                    uint32_t m, Smem<W>& sm, uint64_t occ, uint64_t used,
                                                    uint64_t touched, bool table_end_in_window) {
  const int lane = lane_id();
  const int j = lane;
  const uint32_t nq = sm.nq;
  const uint8_t* in = (const uint8_t*)sm.sin;
  uint8_t* out = (uint8_t*)sm.sout;

  // ---- event position, old length and new length of every listed quotient.
  // With kStash, lane takes quotients lane + 64 k in round k (at most
  // kItemCap = 4 x 64 of them), and a quotient that takes the straight-line
  // merge keeps its merged bytes in registers for the writing pass
  // (lean_lo / lean_hi; lean_meta 0: merge_quotient).  Without it (a caller
  // short of registers) both passes call merge_quotient, which redoes the
  // straight-line merge where it applies.
  static_assert(kItemCap <= 4 * 64, "four rounds of listed quotients");
  uint32_t lean_lo[4], lean_hi[4], lean_meta[4] = {0, 0, 0, 0};
  auto measure = [&](uint32_t t, int k) {
    const uint32_t it0 = sm.tit0[t];
    const uint32_t h0 = src.hash(sm.items, it0);
    const int32_t q = (int32_t)(h0 >> 8);
    int32_t st = q;
    uint32_t ol = 0, nl = 0;
    if (locate_run_or_slot(sm, nwin, q, st, ol)) {
      const uint32_t it1 = t + 1 < nq ? sm.tit0[t + 1] : m;
      uint32_t ln;
      if (kStash && it1 - it0 <= 2 && ol <= 8 && lean_read(sm, src, it0, it1, st, ol, h0, lean_lo[k], lean_hi[k], ln))
        nl = ln, lean_meta[k] = 1;
      else
        nl = merge_quotient<W, false>(sm, src, it0, it1, st, ol, 0, h0), lean_meta[k] = 0;
      atomicAdd((uint32_t*)&sm.xnew[st >> 6], 1u);  // events per slot-block (st < 64 nwin)
    } else {
      atomicMin(&sm.first_invalid, -1);
      st = 0;
    }
    if (ol > kRunLenMax<W> || nl > kRunLenMax<W>) {  // redone in a wider window
      atomicMin(&sm.first_invalid, -1);
      ol = nl = 0;
    }
    sm.tst[t] = (int16_t)st;
    sm.tol[t] = (RunLen<W>)ol;
    sm.tnl[t] = (RunLen<W>)nl;
  };
  sm.xnew[j] = 0;
  __syncthreads();
  if constexpr (kStash) {
#pragma unroll
    for (int k = 0; k < 4; k++)
      if ((uint32_t)lane + 64u * (uint32_t)k < nq) measure((uint32_t)lane + 64u * (uint32_t)k, k);
  } else {
    for (uint32_t t = lane; t < nq; t += 64) measure(t, 0);
  }
  __syncthreads();
  if (sm.first_invalid < 0) return kNeedWide;

  // ---- the events of slot-block j: listed quotients e0 .. e1 - 1 (their
  // positions ascend with the quotients, so a scan of the per-block counts
  // gives each block's range)
  uint32_t e0, e1;
  {
    const uint32_t c = (uint32_t)sm.xnew[j];
    const uint32_t incl = wave_scan_add(c);
    e0 = incl - c;
    e1 = incl;
  }

  // ---- lane j's shift map: events and (valid) empty slots of slot-block j
  uint32_t vlim = 0;
  if (j < nwin) {
    const uint64_t first = (B0 + (uint64_t)j) * 64;
    vlim = !table_end_in_window ? 64u : (first >= qf.xnslots ? 0u : (uint32_t)min((uint64_t)64, qf.xnslots - first));
  }
  const uint64_t zero_ok = ~used & gqf_mask(vlim);  // empty slots that absorb shift
  int32_t A = kNegInf, Bm = 0;
  {
    int pos = 0;
    for (uint32_t t = e0; t < e1; t++) {
      const int rel = sm.tst[t] - 64 * j;
      const int c = __popcll(zero_ok & range_mask(pos, rel));
      if (c) compose(A, Bm, 0, -c);
      compose(A, Bm, kNegInf, (int32_t)sm.tnl[t] - (int32_t)sm.tol[t]);
      pos = rel;
    }
    const int c = __popcll(zero_ok & range_mask(pos, 64));
    if (c) compose(A, Bm, 0, -c);
  }
  // inclusive scan of the maps over the lanes; the shift entering block j
  // is the exclusive prefix applied to 0
  int32_t ia = A, ib = Bm;
  wave_scan_compose(ia, ib);
  const int32_t ea = wave_shr1(ia, kNegInf), eb = wave_shr1(ib, 0);
  const int32_t delta_in = ea > eb ? ea : eb;  // max(a, 0 + b)
  const int32_t la = __builtin_amdgcn_readlane(ia, 63), lb = __builtin_amdgcn_readlane(ib, 63);
  const int32_t delta_end = la > lb ? la : lb;

  // ---- b*: the first block at or after R, past the last event's old run,
  // that no shift reaches; blocks [0, b*) are rewritten, the rest stay
  {
    int lo = R;
    if (nq) {
      const int32_t last = sm.tst[nq - 1] + (sm.tol[nq - 1] ? (int32_t)sm.tol[nq - 1] : 1) - 1;
      if ((last >> 6) + 1 > lo) lo = (last >> 6) + 1;
    }
    const uint64_t bm = __ballot(j >= lo && j < nwin && delta_in == 0);
    if (lane == 0) {
      int res = kOk, bstar = nwin;
      if (bm) {
        bstar = __builtin_ctzll(bm);
      } else if (!table_end_in_window) {
        res = kNeedWide;
      } else if (delta_end != 0) {
        res = kNoSpace;
      }
      if (sm.xold[0] + 1 > 64 * nwin) res = kNeedWide;  // spill-in longer than the window
      sm.bstar = bstar;
      sm.result = res;
    }
  }
  __syncthreads();
  const int res = sm.result;
  if (res != kOk) return res;
  const int bstar = sm.bstar;

  // ---- clear the rows and bit words to rewrite
  for (uint32_t c = lane; c < (uint32_t)bstar * 17; c += 64) sm.sout[c] = 0;
  if (j < bstar) sm.run_out[j] = 0;
  __syncthreads();

  // ---- lane j moves the untouched bytes of slot-block j and records the
  // new start of each of its events' runs; used_new is the block's new
  // used-slot mask (the new runs' slots and what block j - 1 pushes in
  // included), which only lane j writes and reads
  uint64_t used_new = 0;
  if (j < bstar) {
    const uint64_t run_j = sm.run_in[j];
    const int32_t base = 64 * j;
    int32_t cur = base;
    if (e0 > 0) {  // an earlier event's old run may reach into this block
      const int32_t pe = sm.tst[e0 - 1] + (int32_t)sm.tol[e0 - 1];
      if (pe > cur) cur = pe;
    }
    // The events, in order: each one's new start; the slots of the old runs
    // they replace (`skip`, with what an earlier block's last event covers);
    // and the shift each event position adds, in 2-bit fields (ev[0]:
    // positions 0..31, ev[1]: 32..63).  A position whose fields would pass
    // 3, or a shift above kShiftFast, sends the block to the general loop
    // below (at 2^24 inserts into 2^28 slots, one block in 60,000 passes
    // 24).
    constexpr int kShiftFast = 24;
    uint64_t skip = cur > base ? gqf_mask((uint64_t)min(64, cur - base)) : 0ULL;
    uint64_t ev[2] = {0ULL, 0ULL};
    bool general = false;
    {
      int32_t delta = delta_in, emax = delta_in;
      int pos = 0;
      for (uint32_t t = e0; t < e1; t++) {
        const int32_t st = sm.tst[t];
        const int rel = st - base;
        const int c = __popcll(zero_ok & range_mask(pos, rel));
        delta = delta > c ? delta - c : 0;
        sm.tsn[t] = (int16_t)(st + delta);
        const int32_t d = (int32_t)sm.tnl[t] - (int32_t)sm.tol[t];
        delta += d;
        if (delta > emax) emax = delta;
        pos = rel;
        if (d < 0) general = true;  // inserts never shorten a run; kept general all the same
        if (d > 0) {
          const int sh = 2 * (rel & 31);
          const uint64_t h = rel < 32 ? ev[0] : ev[1];
          if ((int)((h >> sh) & 3ULL) + d > 3) general = true;
          const uint64_t add = (uint64_t)(d & 3) << sh;
          if (rel < 32) ev[0] += add; else ev[1] += add;
        }
        const int tol = (int)sm.tol[t];
        if (tol) skip |= range_mask(rel, min(64, rel + tol));
        const int32_t pe = st + tol;
        if (pe > cur) cur = pe;
      }
      if (emax > kShiftFast) general = true;
    }
    if (!general) {
      // One straight pass over the block's 64 slots, the same code on every
      // lane.  e, the shift at slot i, grows by the slot's event fields; an
      // untouched byte at i goes to i + e; after the slot, e drops by one
      // (not below 0) if the slot is a valid empty.  Slot i of the block is
      // used afterwards iff it was used or e > 0 there (the shifted items
      // fill it), so the used mask needs no scatter.  The run ends of the
      // moved bytes collect in eight 32-bit windows: slot i's lands in
      // window i / 8 at bit e + i % 8 (<= kShiftFast + 7 < 32).
      uint32_t w[16];
      const uint32_t* row = sm.sin + 17 * j;
#pragma unroll
      for (int k = 0; k < 16; k++) w[k] = row[k];
      const uint64_t copy = used & ~skip;
      const uint64_t runc = run_j & copy;
      uint32_t e = (uint32_t)delta_in;
      uint32_t nz[2] = {0u, 0u};
      uint32_t rw[8];
#pragma unroll
      for (int k = 0; k < 8; k++) rw[k] = 0u;
      const uint32_t cw[2] = {(uint32_t)copy, (uint32_t)(copy >> 32)};
      const uint32_t rc[2] = {(uint32_t)runc, (uint32_t)(runc >> 32)};
      const uint32_t zw[2] = {(uint32_t)zero_ok, (uint32_t)(zero_ok >> 32)};
      const uint32_t ew[4] = {(uint32_t)ev[0], (uint32_t)(ev[0] >> 32), (uint32_t)ev[1], (uint32_t)(ev[1] >> 32)};
      uint8_t* dst = out + sin_off(base);  // row j; row j + 1 starts 68 bytes on
      auto slot = [&](auto ic) {
        constexpr int i = decltype(ic)::value;
        e += (ew[i >> 4] >> (2 * (i & 15))) & 3u;
        nz[i >> 5] = lshl_or_c<i & 31>(min1(e), nz[i >> 5]);
        const uint32_t sft = e + (uint32_t)(i & 7);      // < 32
        const uint32_t pdst = sft + (uint32_t)(i & ~7);  // i + e < 88
        // row j + 1 starts 68 bytes on; no shift reaches it from i < 40
        const uint32_t off = i + kShiftFast < 64 ? pdst : pdst + ((pdst >> 6) << 2);
        // slots that move nothing write to the row's pad byte (offset 64),
        // which no row read takes
        const uint32_t cm = bit_mask<i & 31>(cw[i >> 5]);
        const uint32_t woff = (off & cm) | (64u & ~cm);
        dst[woff] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
        rw[i >> 3] = lshl_or(bit_u<i & 31>(rc[i >> 5]), sft, rw[i >> 3]);
        const uint32_t z = (zw[i >> 5] >> (i & 31)) & 1u;
        e = e > z ? e - z : 0u;
      };
      for_each_slot(slot, std::make_integer_sequence<int, 64>{});
      // e is now the shift leaving the block: the bytes moved into block
      // j + 1 fill its first e slots
      uint64_t rlo = 0, rhi = 0;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        rlo |= (uint64_t)rw[k] << (8 * k);
        if (8 * k + 31 >= 64) rhi |= (uint64_t)rw[k] >> (64 - 8 * k);
      }
      used_new = (used | ((uint64_t)nz[1] << 32 | nz[0])) & (used | gqf_mask((uint64_t)vlim));
      atomicOr((unsigned long long*)&sm.run_out[j], rlo);
      if (e) {
        if (j + 1 < bstar) {  // always, by the choice of b*
          atomicOr((unsigned long long*)&sm.run_out[j + 1], rhi);
        } else {
          atomicOr(qf.status, GQF_STATUS_CLUSTER_OVERFLOW);
        }
      }
    } else {
      // the general loop: stretch by stretch, any shift
      int32_t delta = delta_in;
      int32_t gcur = base;
      if (e0 > 0) {
        const int32_t pe = sm.tst[e0 - 1] + (int32_t)sm.tol[e0 - 1];
        if (pe > gcur) gcur = pe;
      }
      // moves the used stretches of [a, b) (inside block j) right by `delta`;
      // each valid empty slot on the way lowers delta by one (not below 0)
      auto stretches = [&](int32_t a, int32_t b) {
        if (b > base + 64) b = base + 64;
        while (a < b) {
          const int ra = a - base, rb = b - base;
          const uint64_t nu = ~used & range_mask(ra, rb);
          const int re = nu ? __builtin_ctzll(nu) : rb;
          if (re > ra) {
            const int len = re - ra;
            const int32_t dd = a + delta;
            if (dd + len <= 64 * bstar) {  // always, by the choice of b*; checked so no write leaves the rows
              for (int k = 0; k < len; k++) out[sin_off(dd + k)] = in[sin_off(a + k)];
              const uint64_t lm = gqf_mask((uint64_t)len);
              or_bits(sm.run_out, dd, (run_j >> ra) & lm, len);
            } else {
              atomicOr(qf.status, GQF_STATUS_CLUSTER_OVERFLOW);
            }
          }
          if (re >= rb) break;
          // the gap of unused slots at re, up to rb
          const uint64_t rest = ~(~used >> re);
          int g = rest ? __builtin_ctzll(rest) : 64 - re;
          if (g > rb - re) g = rb - re;
          const int gv = re >= (int)vlim ? 0 : ((int)vlim - re < g ? (int)vlim - re : g);
          delta = delta > gv ? delta - gv : 0;
          a = base + re + g;
        }
      };
      for (uint32_t t = e0; t < e1; t++) {
        const int32_t st = sm.tst[t];
        stretches(gcur, st);
        delta += (int32_t)sm.tnl[t] - (int32_t)sm.tol[t];
        const int32_t pe = st + (int32_t)sm.tol[t];
        if (pe > gcur) gcur = pe;
      }
      stretches(gcur, base + 64);
      // used afterwards, as in the straight pass: slot i is used iff it was,
      // or the shift reaching it is above 0
      uint64_t mz = 0;
      int32_t ee = delta_in;
      uint32_t t = e0;
      for (int i = 0; i < 64; i++) {
        for (; t < e1 && sm.tst[t] - base == i; t++) ee += (int32_t)sm.tnl[t] - (int32_t)sm.tol[t];
        if (ee > 0) mz |= 1ULL << i;
        if ((zero_ok >> i) & 1ULL) ee = ee > 0 ? ee - 1 : 0;
      }
      used_new = (used | mz) & (used | gqf_mask((uint64_t)vlim));
    }
  }
  __syncthreads();

  // ---- the listed quotients' runs at their new starts
  auto write_run = [&](uint32_t t, int k) {
    const uint32_t nl = sm.tnl[t];
    const int32_t sn = sm.tsn[t];
    const int32_t en = sn + (int32_t)nl - 1;
    if (sn < 0 || en >= 64 * bstar) {  // cannot happen (see above); no write leaves the rows
      atomicOr(qf.status, GQF_STATUS_CLUSTER_OVERFLOW);
      return;
    }
    if (kStash && lean_meta[k]) {
      lean_write(sm, lean_lo[k], lean_hi[k], nl, sn);
    } else {
      const uint32_t it1 = t + 1 < nq ? sm.tit0[t + 1] : m;
      merge_quotient<W, true>(sm, src, sm.tit0[t], it1, sm.tst[t], sm.tol[t], sn, src.hash(sm.items, sm.tit0[t]));
    }
    atomicOr((unsigned long long*)&sm.run_out[en >> 6], 1ULL << (en & 63));
  };
  if constexpr (kStash) {
#pragma unroll
    for (int k = 0; k < 4; k++)
      if ((uint32_t)lane + 64u * (uint32_t)k < nq) write_run((uint32_t)lane + 64u * (uint32_t)k, k);
  } else {
    for (uint32_t t = lane; t < nq; t += 64) write_run(t, 0);
  }
  __syncthreads();

  // ---- offsets by rank over the new run ends: occupied quotients of the
  // window's blocks < j, plus the earlier regions' runs that end inside the
  // window (run ends at or before xold[0])
  const uint64_t occ_new = j < nwin ? (occ | touched) : 0ULL;
  const uint64_t re_w = j < bstar ? sm.run_out[j] : (j < nwin ? sm.run_in[j] : 0ULL);
  uint32_t c_occ = (uint32_t)__popcll(occ_new), c_re = (uint32_t)__popcll(re_w);
  uint32_t c_pre = 0;
  {
    const int32_t x0 = sm.xold[0];
    if (j < nwin && x0 >= 64 * j) {
      const int upto = x0 - 64 * j + 1;
      c_pre = (uint32_t)__popcll(sm.run_in[j] & gqf_mask((uint64_t)(upto < 64 ? upto : 64)));
    }
  }
  const uint32_t inc_occ = wave_scan_add(c_occ), inc_re = wave_scan_add(c_re);
  const uint32_t tot_pre = (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_add(c_pre), 63);
  sm.xnew[j] = (int32_t)(inc_re - c_re);  // run ends in blocks < j
  if (lane == 63) sm.xnew[64] = (int32_t)inc_re;
  __syncthreads();
  uint32_t off_new = 0;
  if (j < bstar) {
    const uint32_t k = tot_pre + inc_occ - c_occ;  // runs of quotients < 64 j ending in the window
    int32_t x = 64 * j - 1;
    if (k > 0) {
      if (k > (uint32_t)sm.xnew[nwin]) {
        x = sm.xold[j];  // ends past the window, where nothing moved
      } else {
        int lo = 0, hi = nwin - 1;  // the block holding the k-th run end: cum[b] < k <= cum[b + 1]
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if ((uint32_t)sm.xnew[mid] < k) lo = mid; else hi = mid - 1;
        }
        const uint64_t w = lo < bstar ? sm.run_out[lo] : sm.run_in[lo];
        const int32_t p = 64 * lo + gqf_select64(w, (int)(k - (uint32_t)sm.xnew[lo]) - 1);
        if (p > x) x = p;
      }
    }
    off_new = (uint32_t)(x - (64 * j - 1));
  }

  // ---- store blocks [0, b*): lane j writes block j's line
  if (j < bstar) {
    GQF_GLOBAL uint4* g = gqf_gptr_mut((uint4*)(qf.blocks + B0 * GQF_BLOCK_BYTES));
    const uint64_t run = sm.run_out[j];
    const uint64_t un = used_new;
    const uint32_t* row = sm.sout + 17 * j;
    GQF_GLOBAL uint4* b = g + (size_t)j * (GQF_BLOCK_BYTES / 16);
    b[0] = make_uint4((uint32_t)occ_new, (uint32_t)(occ_new >> 32), (uint32_t)run, (uint32_t)(run >> 32));
    b[1] = make_uint4(off_new, 0, (uint32_t)un, (uint32_t)(un >> 32));
    b[4] = make_uint4(row[0], row[1], row[2], row[3]);
    b[5] = make_uint4(row[4], row[5], row[6], row[7]);
    b[6] = make_uint4(row[8], row[9], row[10], row[11]);
    b[7] = make_uint4(row[12], row[13], row[14], row[15]);
  }
  return kOk;
}

// ---------------------------------------------------------------------------
// A region's window lines and cached items, in registers: lane j holds
// window block j's 128-byte line (round k: block j + 64 k) and items
// lane + 64 k (when the region's items fit the LDS cache).  Loading them
// is split from using them so that the persistent phase kernel can load the
// next region while it rebuilds the current one.
template <int kRounds>
struct Lines {
  uint4 hv[kRounds], h1[kRounds], s0[kRounds], s1[kRounds], s2[kRounds], s3[kRounds];
  uint64_t iv[4];
};

// blocks in region r's window: W, or 2 R when other waves rebuild the
// regions of the same parity at the same time (see rebuild_region), cut at
// the end of the table
__device__ inline int window_blocks(const QF& qf, uint64_t r, int W, bool concurrent) {
  const uint64_t R = qf.region_blocks;
  const uint64_t wmax = concurrent && 2 * R < (uint64_t)W ? 2 * R : (uint64_t)W;
  return (int)min(wmax, qf.nblocks - r * R);
}

template <int W>
__device__ __forceinline__ void load_lines(const QF& qf, uint64_t r, uint64_t s, uint64_t e, bool concurrent,
                                           const uint64_t* __restrict__ items, Lines<W / 64>& L) {
  constexpr int kRounds = W / 64;
  const int lane = lane_id();
  const int nwin = window_blocks(qf, r, W, concurrent);
  const GQF_GLOBAL uint4* g = gqf_block_ptr(qf, r * qf.region_blocks);
#pragma unroll
  for (int k = 0; k < kRounds; k++) {
    const int j = lane + 64 * k;
    const int jj = j < nwin ? j : nwin - 1;  // clamped: every lane loads, all in flight together
    const GQF_GLOBAL uint4* b = g + (size_t)jj * (GQF_BLOCK_BYTES / 16);
    L.hv[k] = b[0];
    L.h1[k] = b[1];  // offset, 0, used-slot mask
    L.s0[k] = b[4];
    L.s1[k] = b[5];
    L.s2[k] = b[6];
    L.s3[k] = b[7];
  }
  // at most kItemCap <= 256 items are cached: four per lane
  static_assert(kItemCap <= 256, "item cache: four items per lane");
  const uint64_t m = e - s;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t i = lane + 64 * k;
    L.iv[k] = m <= (uint64_t)kItemCap && i < m ? items[s + i] : 0;
  }
}

// ---------------------------------------------------------------------------
// Region rebuild.  All 64 lanes of the (single-wave) workgroup call it.
// Lane l handles window blocks l, l + 64, ...

// `tag` is added to every hash of `items` (a filter group sorts all its
// filters' hashes together, each tagged with its filter above the hash bits).
//
// `concurrent`: other waves rebuild the regions of the same parity at the
// same time (the phase and group kernels).  Regions of one parity are 2 R
// blocks apart, so the window is then cut to 2 R blocks: region r reads and
// writes only its own blocks and those of region r + 1, which is of the
// other parity and not running.  Without the cut, a region of R < 32 blocks
// would push into region r + 2 while that region's wave rebuilds it.  A push
// the cut window cannot absorb is redone by the serial wide pass.
//
// `L`: the region's lines, already loaded by load_lines when `loaded`,
// else loaded here.  With `has_next`, the lines of the region (rn, sn, en)
// are loaded into L as soon as this region's lines are in LDS, so they
// arrive while this one is rebuilt.
template <int W, bool kStash = true>
__device__ __forceinline__ int rebuild_region_lines(const QF& qf, uint64_t r, const uint64_t* __restrict__ items,
                              const uint64_t* __restrict__ counts, uint64_t s, uint64_t e, int mode,
                              Smem<W>& sm, uint32_t* lcnt, bool concurrent, uint64_t tag, Lines<W / 64>& L,
                              bool loaded, bool has_next, uint64_t rn, uint64_t sn, uint64_t en) {
  constexpr int kRounds = W / 64;
  const bool shift_ok = !(mode & GQF_REBUILD_NO_SHIFT);
  mode &= ~GQF_REBUILD_NO_SHIFT;
  const int lane = lane_id();
  const int R = (int)qf.region_blocks;
  const uint64_t B0 = r * (uint64_t)R;
  const int nwin = window_blocks(qf, r, W, concurrent);
  const int nreg = (int)min((uint64_t)R, qf.nblocks - B0);
  const int64_t win_base = (int64_t)(B0 * 64);
  const bool table_end_in_window = B0 + (uint64_t)nwin >= qf.nblocks;
  const uint32_t m = (uint32_t)(e - s);
  const uint64_t hmask = mode == kSeq ? GQF_SEQ_HASH_MASK : ~0ULL;
  const int kind = mode == kSeq ? kTagged : (counts ? kCounts : (mode == kDelete ? kSub : kAdd));
  const Items src{items, counts, lcnt, s, tag + (B0 << 14), hmask, kind, m <= (uint32_t)kItemCap};

  // ---- 1. stage the window and the items ----
  // lane j reads block j's line: header into registers, slots into LDS row j
  uint64_t occ_w[kRounds];
  uint32_t off_w[kRounds];
  uint64_t used_w[kRounds];
  {
    if (!loaded) load_lines<W>(qf, r, s, e, concurrent, items, L);
    uint32_t ov[kRounds];
#pragma unroll
    for (int k = 0; k < kRounds; k++) ov[k] = L.h1[k].x;
    if (src.cached) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint32_t i = lane + 64 * k;
        if (i < m) {
          sm.items[i] = (uint32_t)((L.iv[k] & hmask) - src.base);
          if (kind == kCounts || kind == kTagged) lcnt[i] = src.cache_delta(s + i);
        }
      }
    }
    if (!src.cached) sm.blk.cnt[lane] = 0;
    if (lane == 0) { sm.first_invalid = nwin; sm.bstar = -1; sm.result = kOk; }
#pragma unroll
    for (int k = 0; k < kRounds; k++) {
      const int j = lane + 64 * k;
      occ_w[k] = 0;
      off_w[k] = 0;
      used_w[k] = 0;
      if (j < nwin) {
        used_w[k] = (uint64_t)L.h1[k].z | ((uint64_t)L.h1[k].w << 32);
        const uint64_t occ = (uint64_t)L.hv[k].x | ((uint64_t)L.hv[k].y << 32);
        const uint64_t run = (uint64_t)L.hv[k].z | ((uint64_t)L.hv[k].w << 32);
        occ_w[k] = occ;
        off_w[k] = ov[k];
        sm.run_in[j] = run;
        sm.occ_in[j] = occ;
        sm.tmask[j] = 0;
        uint32_t* row = sm.sin + 17 * j;
        row[0] = L.s0[k].x; row[1] = L.s0[k].y; row[2] = L.s0[k].z; row[3] = L.s0[k].w;
        row[4] = L.s1[k].x; row[5] = L.s1[k].y; row[6] = L.s1[k].z; row[7] = L.s1[k].w;
        row[8] = L.s2[k].x; row[9] = L.s2[k].y; row[10] = L.s2[k].z; row[11] = L.s2[k].w;
        row[12] = L.s3[k].x; row[13] = L.s3[k].y; row[14] = L.s3[k].z; row[15] = L.s3[k].w;
        sm.xold[j] = 64 * j - 1 + (int32_t)ov[k];  // device offsets are exact (never saturated)
      }
    }
  }
  // the next region's lines start loading now (this region's are in LDS)
  if (has_next) load_lines<W>(qf, rn, sn, en, concurrent, items, L);
  __syncthreads();

  // ---- 2. per-block item ranges / quotient table ----
  const bool fast = src.cached;
  if (fast) {
    // list the distinct quotients of the items (sorted, so heads compact in order)
    uint32_t carry = 0;
    const uint64_t lt = lane ? (~0ULL >> (64 - lane)) : 0ULL;
    for (uint32_t base = 0; base < m; base += 64) {
      const uint32_t i = base + lane;
      const uint32_t h = i < m ? sm.items[i] : 0;
      const bool head = i < m && (i == 0 || (sm.items[i - 1] >> 8) != (h >> 8));
      const uint64_t hm = __ballot(head);
      if (head) {
        const uint32_t t = carry + (uint32_t)__popcll(hm & lt);
        sm.tit0[t] = (uint8_t)i;
        atomicOr((unsigned long long*)&sm.tmask[h >> 14], 1ULL << ((h >> 8) & 63));
      }
      carry += (uint32_t)__popcll(hm);
    }
    if (lane == 0) sm.nq = carry;
  } else {
    for (uint32_t i = lane; i < m; i += 64) atomicAdd(&sm.blk.cnt[src.hash(sm.items, i) >> 14], 1u);
  }
  __syncthreads();
  // block `lane` (a region block: there are at most 56) starts at table entry
  // tbase (fast path: its touched quotients, counted by popcount) or at item
  // blk.start[lane] (general path)
  uint32_t tbase;
  {
    const uint32_t c = fast ? (lane < nreg ? (uint32_t)__popcll(sm.tmask[lane]) : 0u) : sm.blk.cnt[lane];
    tbase = wave_scan_add(c) - c;
    if (!fast) sm.blk.start[lane] = tbase;
  }
  // the touched quotients of this lane's blocks, in registers from here on
  // (the tmask space becomes tsn in pass 2)
  uint64_t tm_w[kRounds];
#pragma unroll
  for (int k = 0; k < kRounds; k++) {
    const int j = lane + 64 * k;
    tm_w[k] = fast && j < nwin ? sm.tmask[j] : 0ULL;
  }
  if constexpr (W == kWinFast) {
    if (shift_ok && fast && kind == kAdd)
      return rebuild_insert_shift<W, kStash>(qf, B0, R, nwin, src, m, sm, occ_w[0], used_w[0], tm_w[0], table_end_in_window);
  }
  // With kStash, lane takes listed quotients lane + 64 k in round k (at
  // most kItemCap = 4 x 64 of them), and a delete that takes the
  // straight-line merge keeps its bytes in registers (del_lo / del_hi, bit k
  // of del_lean) for the writing pass; without it (a caller short of
  // registers) both passes call merge_quotient.
  uint32_t del_lo[4], del_hi[4], del_lean = 0;
  if (fast) {
    // old run and new length of every listed quotient, one quotient per lane
    const uint32_t nq = sm.nq;
    auto measure = [&](uint32_t t, int k) {
      const int32_t q = (int32_t)(sm.items[sm.tit0[t]] >> 8);
      int32_t st_o;
      uint32_t old_len;
      uint32_t nl = 0;
      if (locate_old_run(sm, nwin, q, st_o, old_len)) {
        const uint32_t it0 = sm.tit0[t];
        const uint32_t it1 = t + 1 < nq ? sm.tit0[t + 1] : m;
        const uint32_t h0 = src.hash(sm.items, it0);
        uint32_t ln;
        if (kStash && kind == kSub && lean_delete(sm, src, it0, it1, st_o, old_len, h0, del_lo[k], del_hi[k], ln)) {
          nl = ln;
          del_lean |= 1u << k;
        } else {
          nl = merge_quotient<W, false>(sm, src, it0, it1, st_o, old_len, 0, h0);
        }
      } else {
        atomicMin(&sm.first_invalid, q >> 6);
        st_o = q;
        old_len = 0;
      }
      if (old_len > kRunLenMax<W> || nl > kRunLenMax<W>) {  // redone in a wider window
        atomicMin(&sm.first_invalid, -1);
        old_len = nl = 0;
      }
      sm.tst[t] = (int16_t)st_o;
      sm.tol[t] = (RunLen<W>)old_len;
      sm.tnl[t] = (RunLen<W>)nl;
    };
    if constexpr (kStash) {
      static_assert(kItemCap <= 4 * 64, "four rounds of listed quotients");
#pragma unroll
      for (int k = 0; k < 4; k++)
        if ((uint32_t)lane + 64u * (uint32_t)k < nq) measure((uint32_t)lane + 64u * (uint32_t)k, k);
    } else {
      for (uint32_t t = lane; t < nq; t += 64) measure(t, 0);
    }
  }
  __syncthreads();

  // ---- 3. pass 1 + max-plus scan -> xnew ----
  uint64_t occ_new[kRounds];
  {
    const int32_t e_in = sm.xold[0];
    int32_t ca = kNegInf, cb = 0;  // carry over earlier rounds
#pragma unroll
    for (int k = 0; k < kRounds; k++) {
      const int j = lane + 64 * k;
      int32_t a = kNegInf, b = 0;
      occ_new[k] = 0;
      if (j < nwin) {
        uint32_t it = 0, ie = 0;
        if (j < nreg) {
          it = fast ? tbase : sm.blk.start[j];
          ie = fast ? 0u : it + sm.blk.cnt[j];
        }
        RunBits unused;
        const bool ok = fast ? walk_block_table<W, false>(sm, j, nwin, it, occ_w[k], tm_w[k], a, b, occ_new[k], unused)
                             : walk_block<W, false>(sm, j, nwin, src, it, ie, occ_w[k], a, b, occ_new[k], unused);
        if (!ok) atomicMin(&sm.first_invalid, j);
      }
      int32_t ia = a, ib = b;
      wave_scan_compose(ia, ib);
      const int32_t ea = wave_shr1(ia, kNegInf), eb = wave_shr1(ib, 0);
      int32_t pa = ca, pb = cb;
      compose(pa, pb, ea, eb);
      if (j < nwin) {
        const int32_t x = pa > e_in + pb ? pa : e_in + pb;
        sm.xnew[j] = x > 64 * j - 1 ? x : 64 * j - 1;
      }
      compose(ca, cb, __builtin_amdgcn_readlane(ia, 63), __builtin_amdgcn_readlane(ib, 63));
    }
    if (lane == 0) {
      const int32_t x = ca > e_in + cb ? ca : e_in + cb;
      sm.xnew[nwin] = x > 64 * nwin - 1 ? x : 64 * nwin - 1;
      sm.xend = x;  // true end of the last run in the window (not normalised)
    }
  }
  __syncthreads();

  // ---- 4. where does the new layout meet the old one? ----
  {
    const int first_invalid = sm.first_invalid;
    int cand = nwin;
    for (int base = R; base < nwin; base += 64) {
      const int j = base + lane;
      const bool ok = j < nwin && j <= first_invalid && sm.xnew[j] == sm.xold[j] &&
                      sm.xold[j] - (64 * j - 1) < 64;
      const uint64_t bm = __ballot(ok);
      if (bm) { cand = base + __builtin_ctzll(bm); break; }
    }
    if (lane == 0) {
      int res = kOk;
      const int64_t xn_end = win_base + (cand < nwin ? sm.xnew[cand] : sm.xend);
      if (cand == nwin && !(table_end_in_window && first_invalid == nwin)) res = kNeedWide;
      else if (xn_end >= (int64_t)qf.xnslots) res = kNoSpace;
      if (sm.xold[0] + 1 > 64 * nwin) res = kNeedWide;  // spill-in longer than the window
      sm.bstar = cand;
      sm.result = res;
    }
  }
  __syncthreads();
  const int res = sm.result;
  if (res != kOk) return res;
  const int bstar = sm.bstar;  // blocks [0, bstar) re-laid out, bstar keeps its own runs
  const int nwrite = bstar < nwin ? bstar + 1 : nwin;

  // ---- 5. pass 2: assemble the new image in LDS ----
  for (uint32_t c = lane; c < (uint32_t)nwrite * 17; c += 64) sm.sout[c] = 0;
  for (int j = lane; j < nwrite; j += 64) sm.run_out[j] = 0;
  __syncthreads();
  {
    const uint8_t* in = (const uint8_t*)sm.sin;
    uint8_t* out = (uint8_t*)sm.sout;
    // the previous region's runs spilling into this window are kept
    for (int32_t i = lane; i <= sm.xold[0]; i += 64) {
      out[sin_off(i)] = in[sin_off(i)];
      if ((sm.run_in[i >> 6] >> (i & 63)) & 1ULL)
        atomicOr((unsigned long long*)&sm.run_out[i >> 6], 1ULL << (i & 63));
    }
    // the meeting block keeps everything after the new end-before
    if (bstar < nwin) {
      for (int32_t i = sm.xnew[bstar] + 1 + lane; i < 64 * (bstar + 1); i += 64) {
        out[sin_off(i)] = in[sin_off(i)];
        if ((sm.run_in[i >> 6] >> (i & 63)) & 1ULL)
          atomicOr((unsigned long long*)&sm.run_out[i >> 6], 1ULL << (i & 63));
      }
    }
  }
  __syncthreads();
  const int nrelay = bstar < nwin ? bstar : nwin;
  RunBits rbits[kRounds];
#pragma unroll
  for (int k = 0; k < kRounds; k++) {
    const int j = lane + 64 * k;
    if (j < nrelay) {
      uint32_t it = 0, ie = 0;
      if (j < nreg) {
        it = fast ? tbase : sm.blk.start[j];
        ie = fast ? 0u : it + sm.blk.cnt[j];
      }
      int32_t a, b;
      uint64_t occn;
      if (fast) walk_block_table<W, true>(sm, j, nwin, it, occ_w[k], tm_w[k], a, b, occn, rbits[k]);
      else walk_block<W, true>(sm, j, nwin, src, it, ie, occ_w[k], a, b, occn, rbits[k]);
    }
  }
  __syncthreads();
  if (fast) {
    // counters of the listed quotients at their new starts
    const uint32_t nq = sm.nq;
    auto write = [&](uint32_t t, int k) {
      if (sm.tnl[t] == 0) return;
      if (kStash && ((del_lean >> k) & 1u)) {  // the measuring pass's straight-line delete
        lean_write(sm, del_lo[k], del_hi[k], sm.tnl[t], sm.tsn[t]);
        return;
      }
      const uint32_t it0 = sm.tit0[t];
      const uint32_t it1 = t + 1 < nq ? sm.tit0[t + 1] : m;
      merge_quotient<W, true>(sm, src, it0, it1, sm.tst[t], sm.tol[t], sm.tsn[t], src.hash(sm.items, it0));
    };
    if constexpr (kStash) {
#pragma unroll
      for (int k = 0; k < 4; k++)
        if ((uint32_t)lane + 64u * (uint32_t)k < nq) write((uint32_t)lane + 64u * (uint32_t)k, k);
    } else {
      for (uint32_t t = lane; t < nq; t += 64) write(t, 0);
    }
  }
  __syncthreads();
  // ---- 6. store blocks [0, nwrite): lane j writes block j's line ----
  {
    GQF_GLOBAL uint4* g = gqf_gptr_mut((uint4*)(qf.blocks + B0 * GQF_BLOCK_BYTES));
#pragma unroll
    for (int k = 0; k < kRounds; k++) {
      const int j = lane + 64 * k;
      if (j < nwrite) {
        uint32_t off;
        uint64_t occ;
        if (j == bstar) {
          off = off_w[k];  // unchanged end-before: unchanged header
          occ = occ_w[k];
        } else {
          const int32_t o = sm.xnew[j] - (64 * j - 1);
          off = (uint32_t)o;
          occ = occ_new[k];
        }
        const uint64_t run = sm.run_out[j];
        // the meeting block's slots are laid out as before (everything up to
        // its unchanged end-before is covered, the rest is copied)
        const uint64_t lim = qf.xnslots - (B0 + (uint64_t)j) * 64;
        const uint64_t used = j == bstar ? used_w[k] : rbits[k].used(off) & (lim < 64 ? gqf_mask(lim) : ~0ULL);
        const uint32_t* row = sm.sout + 17 * j;
        GQF_GLOBAL uint4* b = g + (size_t)j * (GQF_BLOCK_BYTES / 16);
        b[0] = make_uint4((uint32_t)occ, (uint32_t)(occ >> 32), (uint32_t)run, (uint32_t)(run >> 32));
        b[1] = make_uint4(off, 0, (uint32_t)used, (uint32_t)(used >> 32));
        b[4] = make_uint4(row[0], row[1], row[2], row[3]);
        b[5] = make_uint4(row[4], row[5], row[6], row[7]);
        b[6] = make_uint4(row[8], row[9], row[10], row[11]);
        b[7] = make_uint4(row[12], row[13], row[14], row[15]);
      }
    }
  }
  return kOk;
}

// rebuild_region_lines with the lines loaded here
template <int W, bool kStash = true>
__device__ __forceinline__ int rebuild_region(const QF& qf, uint64_t r, const uint64_t* __restrict__ items,
                                              const uint64_t* __restrict__ counts, uint64_t s, uint64_t e, int mode,
                                              Smem<W>& sm, uint32_t* lcnt, bool concurrent, uint64_t tag = 0) {
  Lines<W / 64> L;
  return rebuild_region_lines<W, kStash>(qf, r, items, counts, s, e, mode, sm, lcnt, concurrent, tag, L, false, false, 0, 0, 0);
}

}  // namespace

// ---------------------------------------------------------------------------
// kernels

// One workgroup per non-empty region of one parity: workgroup i takes
// region list[i] (the list gqf_region_bounds compacted; count[parity]
// entries), so a small batch launches only as many workgroups as it has
// regions with items, and the hardware hands regions to workgroups as
// earlier ones finish.  (A persistent grid that loads each region's window
// while rebuilding the previous one measured slower: 6.6 against 7.5 G
// inserts/s at 2^28 slots, with 168 registers against 85 and a static share
// of regions per workgroup.)  gqf_rebuild_phase serves insert and delete
// (one instance per item); gqf_rebuild_phase_counts takes per-item counts.
__device__ __forceinline__ void rebuild_phase_item(const QF* __restrict__ qfp, const uint64_t* __restrict__ items,
                                                   const uint64_t* __restrict__ counts,
                                                   const uint32_t* __restrict__ region_start,
                                                   const uint32_t* __restrict__ region_list,
                                                   const uint32_t* __restrict__ list_count, int parity, int mode,
                                                   uint8_t* __restrict__ wide_flags, uint32_t* __restrict__ wide_count,
                                                   Smem<kWinFast>& sm, uint32_t* lcnt) {
  const QF qf = *qfp;
  if (blockIdx.x >= list_count[parity]) return;
  const uint64_t r = region_list[blockIdx.x];
  if (r >= qf.nregions) return;
  const uint64_t s = region_start[r], e = region_start[r + 1];
  if (s == e) return;
  const int res = rebuild_region<kWinFast>(qf, r, items, counts, s, e, mode, sm, lcnt, true);
  if (lane_id() == 0) {
    if (res == kNeedWide) {
      wide_flags[r] = 1;
      const uint32_t k = atomicAdd(wide_count, 1u);
      if (k < GQF_WIDE_LIST) wide_count[1 + k] = (uint32_t)r;
    } else if (res == kNoSpace) {
      atomicOr(qf.status, GQF_STATUS_NO_SPACE);
    }
  }
}

extern "C" __global__ void __launch_bounds__(64)
gqf_rebuild_phase(const QF* __restrict__ qfp, const uint64_t* __restrict__ items,
                  const uint32_t* __restrict__ region_start, const uint32_t* __restrict__ region_list,
                  const uint32_t* __restrict__ list_count, int parity, int mode, uint8_t* __restrict__ wide_flags,
                  uint32_t* __restrict__ wide_count) {
  __shared__ Smem<kWinFast> sm;
  rebuild_phase_item(qfp, items, nullptr, region_start, region_list, list_count, parity, mode, wide_flags, wide_count,
                     sm, nullptr);
}

extern "C" __global__ void __launch_bounds__(64)
gqf_rebuild_phase_counts(const QF* __restrict__ qfp, const uint64_t* __restrict__ items,
                         const uint64_t* __restrict__ counts, const uint32_t* __restrict__ region_start,
                         const uint32_t* __restrict__ region_list, const uint32_t* __restrict__ list_count,
                         int parity, int mode, uint8_t* __restrict__ wide_flags, uint32_t* __restrict__ wide_count) {
  __shared__ Smem<kWinFast> sm;
  __shared__ uint32_t icnt[kItemCap];
  rebuild_phase_item(qfp, items, counts, region_start, region_list, list_count, parity, mode, wide_flags, wide_count,
                     sm, icnt);
}

// Redoes flagged regions of one parity with a wider window: 128 blocks,
// then 256 for a push that outruns that too (a region that needs a wider
// window has not written anything).  Regions whose windows could meet are
// redone in ascending order; regions kWinWide blocks or more apart cannot
// touch each other's blocks, so the flagged regions are cut into chains of
// closer ones, and each chain runs on its own workgroup (up to
// GQF_WIDE_LIST of them).  Up to GQF_WIDE_LIST flagged regions come from
// the phase's list, ordered by rank; more go to gqf_rebuild_wide_class.
// Every workgroup exits at once when nothing was flagged; otherwise the
// last one to finish clears the count for the next phase.
extern "C" __global__ void __launch_bounds__(64)
gqf_rebuild_wide(const QF* __restrict__ qfp, const uint64_t* __restrict__ items,
                 const uint64_t* __restrict__ counts, const uint32_t* __restrict__ region_start,
                 int parity, int mode, uint8_t* __restrict__ wide_flags,
                 uint32_t* __restrict__ wide_count) {
  __shared__ Smem<kWinMid> smm;
  __shared__ Smem<kWinWide> sm;
  __shared__ uint32_t icnt[kItemCap];
  const uint32_t nflag = *(volatile uint32_t*)wide_count;
  if (nflag == 0) return;
  const QF qf = *qfp;
  const int lane = lane_id();
  auto redo = [&](uint64_t rr) {
    const uint64_t s = region_start[rr], e = region_start[rr + 1];
    int res = rebuild_region<kWinMid>(qf, rr, items, counts, s, e, mode, smm, icnt, false);
    __syncthreads();
    if (res == kNeedWide) res = rebuild_region<kWinWide>(qf, rr, items, counts, s, e, mode, sm, icnt, false);
    __syncthreads();
    if (lane == 0) {
      wide_flags[rr] = 0;
      if (res == kNeedWide) atomicOr(qf.status, GQF_STATUS_CLUSTER_OVERFLOW);
      else if (res == kNoSpace) atomicOr(qf.status, GQF_STATUS_NO_SPACE);
    }
    __syncthreads();
  };
  if (nflag <= GQF_WIDE_LIST) {
    // distinct regions: each one's rank among them is its place in the order
    const uint32_t mine = (uint32_t)lane < nflag ? wide_count[1 + lane] : 0xFFFFFFFFu;
    uint32_t rank = 0;
    for (uint32_t j = 0; j < nflag; j++) rank += (uint32_t)__shfl((int)mine, (int)j, 64) < mine ? 1u : 0u;
    auto by_rank = [&](uint32_t k) -> uint64_t {
      const uint64_t bm = __ballot((uint32_t)lane < nflag && rank == k);
      return bm ? (uint64_t)(uint32_t)__shfl((int)mine, __builtin_ctzll(bm), 64) : ~0ULL;
    };
    // regions at least `apart` apart have disjoint windows
    const uint64_t apart = (kWinWide + qf.region_blocks - 1) / qf.region_blocks;
    const uint32_t k0 = blockIdx.x;
    if (k0 < nflag) {
      uint64_t cur = by_rank(k0);
      const uint64_t prev = k0 ? by_rank(k0 - 1) : ~0ULL;
      // workgroup k0 starts a chain when its region is far from the previous one
      if (cur != ~0ULL && (k0 == 0 || prev == ~0ULL || cur >= prev + apart)) {
        for (uint32_t k = k0;;) {
          redo(cur);
          if (++k >= nflag) break;
          const uint64_t nxt = by_rank(k);
          if (nxt == ~0ULL || nxt >= cur + apart) break;
          cur = nxt;
        }
      }
    }
  } else {
    return;  // more than GQF_WIDE_LIST flagged: the class launch takes them (gqf_rebuild_wide_class)
  }
  // the last workgroup to finish clears the count and the finish counter
  if (lane == 0) {
    __threadfence();
    if (atomicAdd(&wide_count[1 + GQF_WIDE_LIST], 1u) == gridDim.x - 1) {
      wide_count[1 + GQF_WIDE_LIST] = 0;
      *(volatile uint32_t*)wide_count = 0;
    }
  }
}

// The flagged regions of a phase when there are more than GQF_WIDE_LIST
// (loads near 95%): regions of one parity whose indices differ by at least
// `apart` (the widest window, 256 blocks, in regions) have disjoint
// windows, so region parity + 2 i is in class i mod ncls (ncls =
// ceil(apart / 2)) and the classes run one after another, in ONE launch of
// ncls x kWideClassGrid workgroups.  Each workgroup takes a ticket when it
// starts; tickets c * kWideClassGrid .. (c + 1) * kWideClassGrid - 1 run
// class c, after waiting for class c - 1's finished counter.  Every
// workgroup of class c - 1 took its ticket before any of class c did, so it
// is running or done, and it waits only on earlier classes: every wait
// ends.  (A wait past max_polls, 2^24 unless qf_set_wide_wait_bound says
// otherwise, which no correct run reaches, flags the class's regions as
// CLUSTER_OVERFLOW instead of rebuilding them.)  The last workgroup to
// finish clears every counter for the next phase.  Nothing runs when
// GQF_WIDE_LIST or fewer were flagged (gqf_rebuild_wide has taken them).
constexpr uint32_t kWideMaxPolls = 1u << 24;

__device__ inline bool wait_count(const uint32_t* ctr, uint32_t target, uint32_t max_polls) {
  for (uint32_t k = 0; k < max_polls; k++) {
    const uint32_t v = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)__builtin_amdgcn_readfirstlane(v) == target) return true;
    __builtin_amdgcn_s_sleep(4);
  }
  return false;
}

extern "C" __global__ void __launch_bounds__(64)
gqf_rebuild_wide_class(const QF* __restrict__ qfp, const uint64_t* __restrict__ items,
                       const uint64_t* __restrict__ counts, const uint32_t* __restrict__ region_start,
                       int parity, int mode, uint8_t* __restrict__ wide_flags, uint32_t* __restrict__ wide_count,
                       uint32_t ncls, uint32_t max_polls) {
  __shared__ Smem<kWinMid> smm;
  __shared__ Smem<kWinWide> sm;
  __shared__ uint32_t icnt[kItemCap];
  __shared__ uint32_t ticket;
  const uint32_t nflag = *(volatile uint32_t*)wide_count;
  if (nflag <= GQF_WIDE_LIST) return;
  const QF qf = *qfp;
  const int lane = lane_id();
  uint32_t* finished = &wide_count[1 + GQF_WIDE_LIST];
  uint32_t* tickets = &wide_count[2 + GQF_WIDE_LIST];
  uint32_t* class_done = &wide_count[3 + GQF_WIDE_LIST];
  if (lane == 0) ticket = atomicAdd(tickets, 1u);
  __syncthreads();
  const uint32_t t = ticket;
  const uint32_t cls = t / kWideClassGrid, w = t % kWideClassGrid;
  const bool ok = cls == 0 || wait_count(&class_done[cls - 1], kWideClassGrid, max_polls);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  const uint64_t nper = qf.nregions > (uint64_t)parity ? (qf.nregions - (uint64_t)parity + 1) / 2 : 0;
  // candidate j of this class is i = cls + ncls j, region parity + 2 i;
  // workgroups take 64 candidates at a time
  const uint64_t ncand = nper > cls ? (nper - cls + ncls - 1) / ncls : 0;
  for (uint64_t j0 = (uint64_t)w * 64; j0 < ncand; j0 += (uint64_t)kWideClassGrid * 64) {
    const uint64_t j = j0 + (uint64_t)lane;
    const uint64_t r = (uint64_t)parity + 2 * ((uint64_t)cls + (uint64_t)ncls * j);
    uint64_t m = __ballot(j < ncand && wide_flags[r]);
    while (m) {
      const int l = __builtin_ctzll(m);
      m &= m - 1;
      const uint64_t rr = (uint64_t)parity + 2 * ((uint64_t)cls + (uint64_t)ncls * (j0 + (uint64_t)l));
      int res = kNeedWide;
      if (ok) {
        const uint64_t s = region_start[rr], e = region_start[rr + 1];
        res = rebuild_region<kWinMid>(qf, rr, items, counts, s, e, mode, smm, icnt, false);
        __syncthreads();
        if (res == kNeedWide) res = rebuild_region<kWinWide>(qf, rr, items, counts, s, e, mode, sm, icnt, false);
        __syncthreads();
      }
      if (lane == 0) {
        wide_flags[rr] = 0;
        if (res == kNeedWide) atomicOr(qf.status, GQF_STATUS_CLUSTER_OVERFLOW);
        else if (res == kNoSpace) atomicOr(qf.status, GQF_STATUS_NO_SPACE);
      }
      __syncthreads();
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (lane == 0) {
    atomicAdd(&class_done[cls], 1u);
    if (atomicAdd(finished, 1u) == gridDim.x - 1) {  // everyone is done: clear for the next phase
      for (uint32_t c = 0; c < ncls; c++) class_done[c] = 0;
      *tickets = 0;
      *finished = 0;
      *(volatile uint32_t*)wide_count = 0;
    }
  }
}

// region_start[r] = first index i with (items[i] & hmask) >= r * (region_blocks * 64 * 256),
// for r in [0, nregions]; items at or above `range` fall past the last region.
// Also appends each non-empty region r < nregions to region_list (at
// parity r & 1, list stride `list_stride`), counting in list_count[parity]
// (zeroed by gqf_region_lists_reset just before).
//
// Each wave covers 63 regions plus the next wave's first: every lane runs
// one binary search, and a region is non-empty when the next lane's start
// is above its own (one chain of dependent loads per lane instead of two).
constexpr int kBoundsPerWave = 63;
constexpr int kBoundsThreads = 1024;

// Appends the non-empty regions of the workgroup to the list of their
// parity with ONE atomic per workgroup and parity: same-address atomics
// serialise, and one per wave (2,800 at 2^28 slots) took about 30 us.
__device__ __forceinline__ void append_region_lists(bool nonempty, uint64_t r, uint32_t* __restrict__ region_list,
                                                    uint32_t* __restrict__ list_count, uint64_t list_stride) {
  __shared__ uint32_t wcnt[2][kBoundsThreads / 64];
  __shared__ uint32_t wbase[2];
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = lane ? (~0ULL >> (64 - lane)) : 0ULL;
  const uint64_t m0 = __ballot(nonempty && (r & 1) == 0);
  const uint64_t m1 = __ballot(nonempty && (r & 1) == 1);
  if (lane == 0) {
    wcnt[0][w] = (uint32_t)__popcll(m0);
    wcnt[1][w] = (uint32_t)__popcll(m1);
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint32_t tot = 0;
    for (uint32_t k = 0; k < (blockDim.x >> 6); k++) tot += wcnt[threadIdx.x][k];
    wbase[threadIdx.x] = tot ? atomicAdd(&list_count[threadIdx.x], tot) : 0u;
  }
  __syncthreads();
  if (nonempty) {
    const uint32_t p = (uint32_t)(r & 1);
    uint32_t base = wbase[p];
    for (uint32_t k = 0; k < w; k++) base += wcnt[p][k];
    region_list[p * list_stride + base + (uint32_t)__popcll((p ? m1 : m0) & lt)] = (uint32_t)r;
  }
}

extern "C" __global__ void __launch_bounds__(kBoundsThreads)
gqf_region_bounds(const QF* __restrict__ qfp, const uint64_t* __restrict__ items, uint64_t n_host,
                  const uint64_t* __restrict__ n_dev, uint32_t* __restrict__ region_start,
                  uint32_t* __restrict__ region_list, uint32_t* __restrict__ list_count, uint64_t list_stride,
                  uint64_t hmask) {
  const QF qf = *qfp;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t r = wave * kBoundsPerWave + lane;
  const bool valid = r <= qf.nregions;
  const uint64_t n = n_dev ? *n_dev : n_host;
  auto lower = [&](uint64_t bound) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if ((items[mid] & hmask) < bound) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  uint64_t lo = 0;
  if (valid) {
    lo = lower(r == qf.nregions ? qf.range : (r * qf.region_blocks) << 14);
    region_start[r] = (uint32_t)lo;  // a wave's last lane repeats the next wave's first
    if (r == qf.nregions && lo < n) atomicOr(qf.status, GQF_STATUS_BAD_HASH);
  }
  const uint64_t next = __shfl_down(lo, 1, 64);
  const bool nonempty = valid && lane < (uint32_t)kBoundsPerWave && r < qf.nregions && next > lo;
  append_region_lists(nonempty, r, region_list, list_count, list_stride);
}

// The region table after a two-level sort that wrote start[r] for every
// region whose bound lies below range: the other entries (bounds at or past
// range, and r = nregions) are n, and the non-empty regions are listed as
// gqf_region_bounds lists them (same wave layout, no search).
extern "C" __global__ void __launch_bounds__(kBoundsThreads)
gqf_region_lists(const QF* __restrict__ qfp, uint64_t n, uint32_t* __restrict__ region_start,
                 uint32_t* __restrict__ region_list, uint32_t* __restrict__ list_count, uint64_t list_stride) {
  const QF qf = *qfp;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t r = wave * kBoundsPerWave + lane;
  const bool valid = r <= qf.nregions;
  uint64_t lo = 0;
  if (valid) {
    if (r == qf.nregions || ((r * qf.region_blocks) << 14) >= qf.range) {
      lo = n;
      region_start[r] = (uint32_t)n;
    } else {
      lo = region_start[r];
    }
  }
  const uint64_t next = __shfl_down(lo, 1, 64);
  const bool nonempty = valid && lane < (uint32_t)kBoundsPerWave && r < qf.nregions && next > lo;
  append_region_lists(nonempty, r, region_list, list_count, list_stride);
}

extern "C" __global__ void gqf_region_lists_reset(uint32_t* __restrict__ list_count) {
  if (threadIdx.x < 2) list_count[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// Filter groups: every filter's regions in one persistent launch.
//
// Work items are all even regions of all filters, then all odd regions.
// Workgroup w takes items w, w + G, w + 2G, ... (G = grid size, at most the
// number of workgroups the GPU holds at once, so all of them are resident):
// the item index is uniform by construction, with no atomics and no
// broadcast between lanes.  An odd region waits until its even neighbours
// (whose windows overlap its own) are done.  Every workgroup finishes its
// even items before it reaches any odd item, and even items wait for
// nothing, so each awaited item belongs to a resident workgroup that gets
// to it: the wait always ends.  It is bounded anyway (max_polls, which the
// host sets): a region whose wait runs out is flagged for the wide tail
// instead of being rebuilt.  Done flags carry the epoch of the call,
// so they never need clearing.  Writes are published with an agent-scope
// release before the flag and read after an acquire, since the two regions
// may run on different XCDs.

namespace {

__device__ inline uint64_t lower_bound_u64(const uint64_t* __restrict__ a, uint64_t lo, uint64_t hi, uint64_t v) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// filter f of work item k: the last f with base[f] <= k
__device__ inline uint32_t item_filter(const uint64_t* __restrict__ base, uint32_t nf, uint64_t k) {
  uint32_t lo = 0, hi = nf;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (base[mid] <= k) lo = mid; else hi = mid;
  }
  return lo;
}

// Bounded wait for a region's completion flag, executed by the whole wave
// with a wave-uniform exit (readfirstlane of the polled value), so the
// loop cannot be split between lanes; false if the bound ran out.
__device__ inline bool wait_done(const uint32_t* flag, uint32_t epoch, uint32_t max_polls) {
  for (uint32_t spin = 0; spin < max_polls; spin++) {
    const uint32_t v = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (__builtin_amdgcn_readfirstlane(v) == epoch) return true;
    __builtin_amdgcn_s_sleep(4);
  }
  return false;
}

__device__ inline void region_hash_bounds(const gqf_group_dev& g, const QF& qf, uint32_t f, uint64_t r,
                                          const uint64_t* __restrict__ items, uint64_t& s, uint64_t& e,
                                          uint64_t& tag) {
  // tag + hash, not tag | hash: the last bound, tag + range, reaches the
  // next filter's tag when range = 2^tag_shift
  tag = (uint64_t)f << g.tag_shift;
  const uint64_t lo_h = tag + ((r * qf.region_blocks) << 14);
  const uint64_t hi_h = r + 1 >= qf.nregions ? tag + qf.range : tag + (((r + 1) * qf.region_blocks) << 14);
  s = lower_bound_u64(items, g.seg[f], g.seg[f + 1], lo_h);
  e = lower_bound_u64(items, s, g.seg[f + 1], hi_h);
}

}  // namespace

// Every value that steers the item loop is wave-uniform: the item index
// comes from blockIdx, and the flag polls exit on readfirstlane of the
// polled value.  (A workgroup is one wave, so its __syncthreads() is only a
// wave barrier: it does not keep lanes in the same loop iteration, and a
// per-lane loop condition would let the compiler split the lanes.)
// Three waves per SIMD (168 VGPRs, no spills) keep 12 workgroups per CU
// resident, as many as the LDS allows.
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
gqf_group_rebuild(gqf_group_dev g, const uint64_t* __restrict__ items, int mode) {
  __shared__ Smem<kWinFast> sm;
  const int lane = lane_id();
  const uint32_t epoch = __builtin_amdgcn_readfirstlane(
      __hip_atomic_load(&g.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  const uint64_t n_even = g.even_base[g.nf];
  const uint64_t total = n_even + g.odd_base[g.nf];
  for (uint64_t idx = blockIdx.x; idx < total; idx += gridDim.x) {
    const bool odd = idx >= n_even;
    const uint64_t k = odd ? idx - n_even : idx;
    const uint64_t* base = odd ? g.odd_base : g.even_base;
    const uint32_t f = item_filter(base, g.nf, k);
    const uint64_t r = 2 * (k - base[f]) + (odd ? 1 : 0);
    const QF qf = g.descs[f];
    const uint64_t flag = g.reg_base[f] + r;
    // an odd region whose wait ran out is not rebuilt here (its neighbour
    // may still be writing its window): it goes to the wide tail, which
    // runs after this kernel has finished
    bool ok = true;
    if (odd) {
      ok = wait_done(&g.done[flag - 1], epoch, g.max_polls);
      if (ok && r + 1 < qf.nregions) ok = wait_done(&g.done[flag + 1], epoch, g.max_polls);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    uint64_t s, e, tag;
    region_hash_bounds(g, qf, f, r, items, s, e, tag);
    int res = kOk;
    if (!ok) res = s < e ? kNeedWide : kOk;
    else if (s < e) res = rebuild_region<kWinFast, false>(qf, r, items, nullptr, s, e, mode, sm, nullptr, true, tag);
    __syncthreads();
    if (lane == 0) {
      if (res == kNeedWide) {
        g.wide[flag] = 1;
        atomicAdd(&g.counters[2], 1u);
      } else if (res == kNoSpace) {
        atomicOr(qf.status, GQF_STATUS_NO_SPACE);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0) __hip_atomic_store(&g.done[flag], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Redoes flagged regions with the wide window, in ascending region order
// within each filter.  Filters share no blocks, so each filter has its own
// workgroup (one long serial tail took 7.5 ms per 256-filter insert).
// Every workgroup exits at once when nothing was flagged; otherwise the
// last one to finish clears the count (counters[2]) and its own finish
// counter (counters[3]).
extern "C" __global__ void __launch_bounds__(64)
gqf_group_wide(gqf_group_dev g, const uint64_t* __restrict__ items, int mode) {
  __shared__ Smem<kWinMid> smm;
  __shared__ Smem<kWinWide> sm;
  if (*(volatile uint32_t*)&g.counters[2] == 0) return;
  const int lane = lane_id();
  const uint32_t f = blockIdx.x;
  if (f < g.nf) {
    const QF qf = g.descs[f];
    const uint64_t f0 = g.reg_base[f];
    for (uint64_t base = 0; base < qf.nregions; base += 64) {
      const uint64_t r = base + (uint64_t)lane;
      uint64_t m = __ballot(r < qf.nregions && g.wide[f0 + r]);
      while (m) {
        const uint64_t rr = base + (uint64_t)__builtin_ctzll(m);
        m &= m - 1;
        uint64_t s, e, tag;
        region_hash_bounds(g, qf, f, rr, items, s, e, tag);
        int res = rebuild_region<kWinMid>(qf, rr, items, nullptr, s, e, mode, smm, nullptr, false, tag);
        __syncthreads();
        if (res == kNeedWide) res = rebuild_region<kWinWide>(qf, rr, items, nullptr, s, e, mode, sm, nullptr, false, tag);
        __syncthreads();
        if (lane == 0) {
          g.wide[f0 + rr] = 0;
          if (res == kNeedWide) atomicOr(qf.status, GQF_STATUS_CLUSTER_OVERFLOW);
          else if (res == kNoSpace) atomicOr(qf.status, GQF_STATUS_NO_SPACE);
        }
        __syncthreads();
      }
    }
  }
  if (lane == 0) {
    __threadfence();
    if (atomicAdd(&g.counters[3], 1u) == gridDim.x - 1) {
      g.counters[3] = 0;
      *(volatile uint32_t*)&g.counters[2] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// launchers

int gqf_rebuild_mode_flags() {
  static int flags = -1;
  if (flags < 0) {
    const char* v = getenv("GQF_SHIFT_PATH");
    flags = v && v[0] == '0' ? GQF_REBUILD_NO_SHIFT : 0;
  }
  return flags;
}

void gqf_launch_region_bounds(const QF* qf_dev, uint64_t nregions, const uint64_t* items, uint64_t n,
                              const uint64_t* n_dev, uint32_t* region_start, uint32_t* region_list,
                              uint32_t* list_count, uint64_t hmask, hipStream_t s) {
  // waves of kBoundsPerWave regions (+1 overlap), kBoundsThreads / 64 per workgroup
  const uint64_t waves = (nregions + 1 + kBoundsPerWave - 1) / kBoundsPerWave;
  const uint32_t g = (uint32_t)((waves + kBoundsThreads / 64 - 1) / (kBoundsThreads / 64));
  gqf_prof_begin("region_bounds", s);
  hipLaunchKernelGGL(gqf_region_lists_reset, dim3(1), dim3(64), 0, s, list_count);
  hipLaunchKernelGGL(gqf_region_bounds, dim3(g), dim3(kBoundsThreads), 0, s, qf_dev, items, n, n_dev, region_start,
                     region_list, list_count, (uint64_t)((nregions + 1) / 2 + 1), hmask);
  gqf_prof_end("region_bounds", s);
}

void gqf_launch_region_lists(const QF* qf_dev, uint64_t nregions, uint64_t n, uint32_t* region_start,
                             uint32_t* region_list, uint32_t* list_count, hipStream_t s) {
  const uint64_t waves = (nregions + 1 + kBoundsPerWave - 1) / kBoundsPerWave;
  const uint32_t g = (uint32_t)((waves + kBoundsThreads / 64 - 1) / (kBoundsThreads / 64));
  gqf_prof_begin("region_bounds", s);
  hipLaunchKernelGGL(gqf_region_lists_reset, dim3(1), dim3(64), 0, s, list_count);
  hipLaunchKernelGGL(gqf_region_lists, dim3(g), dim3(kBoundsThreads), 0, s, qf_dev, n, region_start, region_list, list_count,
                     (uint64_t)((nregions + 1) / 2 + 1));
  gqf_prof_end("region_bounds", s);
}

namespace {
std::atomic<uint32_t> g_wide_max_polls{kWideMaxPolls};
}  // namespace

uint32_t gqf_wide_wait_bound_set(uint32_t max_polls) { return g_wide_max_polls.exchange(max_polls); }

void gqf_launch_rebuild(const QF* qf_dev, uint64_t nregions, uint64_t region_blocks, const uint64_t* items,
                        const uint64_t* counts, uint64_t n_items, const uint32_t* region_start,
                        const uint32_t* region_list, const uint32_t* list_count, int mode, uint8_t* wide_flags,
                        uint32_t* wide_count, hipStream_t s) {
  // at most one workgroup per non-empty region: no more than the regions
  // of a parity, nor than the items
  const uint64_t per_parity = (nregions + 1) / 2;
  const uint32_t g = (uint32_t)(n_items < per_parity ? (n_items ? n_items : 1) : per_parity);
  const uint64_t stride = per_parity + 1;
  // classes of the wide pass (regions whose widest windows are disjoint)
  const uint32_t apart = (uint32_t)((kWinWide + region_blocks - 1) / region_blocks);
  const uint32_t ncls = (apart + 1) / 2;
  if (ncls < 1 || ncls > GQF_WIDE_MAX_CLASSES) {  // region_blocks is validated to 16..56 on the host
    fprintf(stderr, "gqf: %u wide classes for %llu-block regions\n", ncls, (unsigned long long)region_blocks);
    abort();
  }
  mode |= gqf_rebuild_mode_flags();
  for (int parity = 0; parity < 2; parity++) {
    gqf_prof_begin("rebuild_phase", s);
    if (counts || (mode & ~GQF_REBUILD_NO_SHIFT) == kSeq)
      hipLaunchKernelGGL(gqf_rebuild_phase_counts, dim3(g), dim3(64), 0, s, qf_dev, items, counts, region_start,
                         region_list + (size_t)parity * stride, list_count, parity, mode, wide_flags, wide_count);
    else
      hipLaunchKernelGGL(gqf_rebuild_phase, dim3(g), dim3(64), 0, s, qf_dev, items, region_start,
                         region_list + (size_t)parity * stride, list_count, parity, mode, wide_flags, wide_count);
    gqf_prof_end("rebuild_phase", s);
    gqf_prof_begin("rebuild_wide", s);
    hipLaunchKernelGGL(gqf_rebuild_wide, dim3(GQF_WIDE_LIST), dim3(64), 0, s, qf_dev, items, counts, region_start,
                       parity, mode, wide_flags, wide_count);
    hipLaunchKernelGGL(gqf_rebuild_wide_class, dim3(ncls * kWideClassGrid), dim3(64), 0, s, qf_dev, items, counts,
                       region_start, parity, mode, wide_flags, wide_count, ncls,
                       g_wide_max_polls.load(std::memory_order_relaxed));
    gqf_prof_end("rebuild_wide", s);
  }
}

// Workgroups per CU the persistent rebuild can keep resident (all of its
// grid must be: see gqf_group_rebuild), and the grid for `per_cu` of them
// (0: the default, clamped to what is resident).  Up to 12 are resident
// (3 waves per SIMD; 13,088 B of LDS each); the default is 8, and
// gqf_group_autotune times the candidates for a group's own shape (10 for
// the 256-filter k-mer workload).
uint32_t gqf_group_rebuild_max_per_cu() {
  static int per_cu = 0;
  if (!per_cu) {
    int p = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&p, (const void*)gqf_group_rebuild, 64, 0);
    per_cu = p < 1 ? 1 : p;
  }
  return (uint32_t)per_cu;
}

uint32_t gqf_group_rebuild_grid(uint32_t per_cu) {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus < 1) cus = 1;
  }
  if (per_cu == 0) per_cu = GQF_GROUP_PER_CU;
  const uint32_t mx = gqf_group_rebuild_max_per_cu();
  return (uint32_t)cus * (per_cu < mx ? per_cu : mx);
}

void gqf_launch_group_rebuild(const gqf_group_dev& g, const uint64_t* items, int mode, uint32_t grid,
                              hipStream_t s) {
  mode |= gqf_rebuild_mode_flags();
  gqf_prof_begin("group_rebuild", s);
  hipLaunchKernelGGL(gqf_group_rebuild, dim3(grid ? grid : 1), dim3(64), 0, s, g, items, mode);
  gqf_prof_end("group_rebuild", s);
  gqf_prof_begin("rebuild_wide", s);
  hipLaunchKernelGGL(gqf_group_wide, dim3(g.nf ? g.nf : 1), dim3(64), 0, s, g, items, mode);
  gqf_prof_end("rebuild_wide", s);
}

