/*
 * gqf_device.h -- device-side API of libgqf (HIP, gfx950): the table layout
 * and the point operations of PASSIONLab/cqf (include/gqf.cuh:94-103), for
 * kernels that insert or query single keys.  Header-only: device code has
 * no stable ABI across shared objects, so the reference's relocatable
 * device functions (built with -rdc) become inline functions here.
 *
 *   qf_returns point_insert(QF*, key, value, flags)                 src/gqf.cu:2970-3020
 *   qf_returns point_insert_not_exists(QF*, key, value, &ret, flags) src/gqf.cu:2911-2966
 *   uint64_t   point_query(QF*, key, value, &ret, flags)             src/gqf.cu:3024-3051
 *   uint64_t   point_query_concurrent(QF*, key, value, &ret, flags)  src/gqf.cu:3055-3108
 *
 * The QF* is the device pointer qf_malloc_device returns.  Mutating point
 * calls take the two region locks of the reference (8192-slot regions,
 * src/gqf.cu:58-59, 2993-3000), but wave64-safe: a lane takes both locks
 * or neither and runs its critical section inside the same loop iteration,
 * so no lane ever waits for another lane of its own wave to leave a loop.
 * Point calls must not run concurrently with bulk calls on the same filter
 * (the reference has the same rule).
 */
#ifndef GQF_DEVICE_H
#define GQF_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gqf.h"

/*
 * The reference stores the table as packed 81-byte blocks
 * (PASSIONLab/cqf include/gqf_int.cuh:38-57):
 *   byte 0       offset     u8   (saturates at 255 = "recompute")
 *   bytes 1..8   occupieds  u64 little endian, bit i = quotient 64b+i has a run
 *   bytes 9..16  runends    u64 little endian, bit i = slot 64b+i ends a run
 *   bytes 17..80 slots[64]  u8 remainders / counter digits
 * On the device each block occupies one 128-byte line, every field aligned:
 *   bytes 0..7   occupieds   bytes 8..15 runends   bytes 16..19 offset (u32)
 *   bytes 20..23 zero        bytes 24..31 used     bytes 32..63 zero
 *   bytes 64..127 slots[64]
 * `used` has bit i set when slot 64b+i lies inside some run (the slots
 * offset_lower_bound reports non-empty, src/gqf.cu:681-705).  It is
 * redundant -- gqf_used_mask derives it from the block's own header -- and
 * every writer keeps it, so the bulk rebuild finds a block's empty slots
 * without walking its runs.
 * A probe loads the whole line in one round of back-to-back vector loads:
 * header and, for ~94% of runs at 90% load, the run itself.  Random access
 * on MI355X is bound by the number of dependent line requests, not bytes
 * (independent 16-byte and 128-byte random reads run at the same rate), so
 * one line per probe beats the 81-byte packing (1.4 lines, two dependent
 * rounds) although the table is 58% larger.
 * The device offset is exact -- end of the runs with quotient < 64 b, minus
 * (64 b - 1), floored at 0 -- never saturated, so no reader has to walk
 * back to recompute it; the export saturates it to the reference's byte.
 * The reference's 81-byte image is produced by qf_export_image and read by
 * qf_import_image; it is the parity format, byte for byte.
 */

#define GQF_REF_BLOCK_BYTES 81   /* the reference's packed block (export format) */
#define GQF_BLOCK_BYTES 128      /* device stride of one 64-slot block */
#define GQF_SLOT_BASE 64         /* byte of slot 0 inside a device block */
#define GQF_SLOTS_PER_BLOCK 64
#define GQF_REF_LOCK_SLOTS 8192ULL  /* point-lock region (src/gqf.cu:58) */
#define GQF_REF_LOCK_DIST 64        /* lock word stride (src/gqf.cu:59) */
#define GQF_OFFSET_SAT 255          /* the reference offset byte saturates here */

// Device-resident filter descriptor.  The public QF* handle is a device
// pointer to this struct, as the reference's QF* is a device pointer
// (src/gqf.cu:2212-2217); callers never dereference it on the host.
struct quotient_filter {
  uint64_t nslots;              // 2^q
  uint64_t xnslots;             // nslots + 10 sqrt(nslots), truncated (src/gqf.cu:1950)
  uint64_t nblocks;             // ceil(xnslots / 64)
  uint64_t key_bits;            // hash bits: q + 8 - value_bits
  uint64_t key_remainder_bits;  // 8 - value_bits
  uint64_t bits_per_slot;       // 8
  uint64_t range;               // slot-hash range 2^(q+8) (the reference's range is 2^key_bits)
  uint64_t num_locks;           // xnslots / 8192 + 2 (src/gqf.cu:1996)
  uint64_t nregions;            // ceil(nblocks / region_blocks): even/odd rebuild regions
  uint64_t region_blocks;       // blocks per rebuild region (tunable, 16..56; default 48)
  uint8_t* blocks;              // nblocks * GQF_BLOCK_BYTES bytes (+ slack)
  uint32_t* locks;              // point-API region locks: num_locks words, 64 apart
  uint32_t* status;             // GQF_STATUS_* bits, sticky until cleared
  uint64_t value_bits;          // 0, or the low slot bits a resize freed (see gqf_key_hash)
};


// Pointers read out of the descriptor are generic; these casts let the
// compiler emit global_load / global_store (not flat) for table accesses.
#if defined(__HIP_DEVICE_COMPILE__)
#define GQF_GLOBAL __attribute__((address_space(1)))
#else
#define GQF_GLOBAL
#endif
template <typename T>
__device__ static inline const GQF_GLOBAL T* gqf_gptr(const T* p) { return (const GQF_GLOBAL T*)p; }
template <typename T>
__device__ static inline GQF_GLOBAL T* gqf_gptr_mut(T* p) { return (GQF_GLOBAL T*)p; }

// ---------------------------------------------------------------------------
// scalar helpers (host and device)

__host__ __device__ static inline uint64_t gqf_mask(uint64_t nbits) {
  return nbits >= 64 ? ~0ULL : ((1ULL << nbits) - 1ULL);
}

// Thomas Wang's invertible 64-bit mix under a power-of-two mask
// (reference src/hashutil.cu:132-142).  Every step is an add, a left shift
// or an xor with a right shift of the masked value, so the result depends
// only on key & mask.
__host__ __device__ static inline uint64_t gqf_hash64(uint64_t key, uint64_t mask) {
  key = (~key + (key << 21)) & mask;
  key = key ^ (key >> 24);
  key = ((key + (key << 3)) + (key << 8)) & mask;
  key = key ^ (key >> 14);
  key = ((key + (key << 2)) + (key << 4)) & mask;
  key = key ^ (key >> 28);
  key = (key + (key << 31)) & mask;
  return key;
}

// Inverse of an odd multiplier modulo 2^64 by Newton's iteration
// x <- x (2 - c x): each step doubles the number of correct low bits, and
// x = c is already right in the low 3 bits (c c = 1 mod 8 for odd c).
__host__ __device__ constexpr uint64_t gqf_inv_odd(uint64_t c) {
  uint64_t x = c;
  for (int i = 0; i < 5; i++) x *= 2 - c * x;
  return x;
}

// Undoes y = x ^ (x >> s) on 64-bit words: the top s bits of x equal those
// of y, and each further pass x = y ^ (x >> s) fixes s more bits.
__host__ __device__ static inline uint64_t gqf_unxorshift(uint64_t y, int s) {
  uint64_t x = y;
  for (int i = s; i < 64; i += s) x = y ^ (x >> s);
  return x;
}

// Inverse of gqf_hash64 under the same mask, for hashes < mask + 1
// (the key the reference's qfi_get_key recovers with hash_64i,
// src/hashutil.cu:144-181, src/gqf.cu:3791-3805).  The mix is a chain of
// multiplications by odd constants and xors with right shifts of a masked
// value; multiplication commutes with masking modulo a power of two, so
// each step is undone in reverse order: a multiplication through the
// constant's inverse modulo 2^64, an xor-shift by gqf_unxorshift.
//   forward: k (2^21 - 1) - 1,  ^ >>24,  * 265,  ^ >>14,  * 21,  ^ >>28,  * (2^31 + 1)
__host__ __device__ static inline uint64_t gqf_hash64_inv(uint64_t h, uint64_t mask) {
  constexpr uint64_t i31 = gqf_inv_odd((1ULL << 31) + 1);
  constexpr uint64_t i21 = gqf_inv_odd(21);
  constexpr uint64_t i265 = gqf_inv_odd(265);
  constexpr uint64_t ik = gqf_inv_odd((1ULL << 21) - 1);
  uint64_t k = (h * i31) & mask;
  k = gqf_unxorshift(k, 28);
  k = (k * i21) & mask;
  k = gqf_unxorshift(k, 14);
  k = (k * i265) & mask;
  k = gqf_unxorshift(k, 24);
  return ((k + 1) * ik) & mask;
}

// Slot hash of a key: the reference's (hash_64(key) << value_bits) | value
// with value 0 (qf_insert / qf_count_key_value, src/gqf.cu:2486,
// 3492); quotient = slot hash >> 8, remainder = its low 8 bits.
// value_bits is 0 for every filter qf_malloc_device makes; a filter grown by
// qf_resize_device keeps its hash width (key_bits) and gains value bits, so
// its remainders carry that many zero bits below the hash's.
__host__ __device__ static inline uint64_t gqf_key_hash(const quotient_filter& f, uint64_t key) {
  return gqf_hash64(key, gqf_mask(f.key_bits)) << f.value_bits;
}

// ---------------------------------------------------------------------------
// device bit helpers

__host__ __device__ static inline int gqf_popc(uint64_t v) { return __builtin_popcountll(v); }

// position of the k'th set bit (k = 0 is the lowest), 64 if fewer: a
// six-step binary descent on popcounts, no table.
__host__ __device__ static inline int gqf_select64(uint64_t v, int k) {
  if (k >= __builtin_popcountll(v)) return 64;
  int pos = 0;
  uint32_t lo = (uint32_t)v;
  int c = __builtin_popcount(lo);
  if (k >= c) { k -= c; v >>= 32; pos += 32; }
  uint32_t w = (uint32_t)v;
  c = __builtin_popcount(w & 0xFFFFu);
  if (k >= c) { k -= c; w >>= 16; pos += 16; }
  c = __builtin_popcount(w & 0xFFu);
  if (k >= c) { k -= c; w >>= 8; pos += 8; }
  c = __builtin_popcount(w & 0xFu);
  if (k >= c) { k -= c; w >>= 4; pos += 4; }
  c = __builtin_popcount(w & 0x3u);
  if (k >= c) { k -= c; w >>= 2; pos += 2; }
  c = (int)(w & 1u);
  if (k >= c) { pos += 1; }
  return pos;
}

// ---------------------------------------------------------------------------
// block header access in global memory

struct gqf_header {
  uint32_t offset;
  uint64_t occupieds;
  uint64_t runends;
};

__device__ static inline const GQF_GLOBAL uint4* gqf_block_ptr(const quotient_filter& qf, uint64_t b) {
  return gqf_gptr((const uint4*)(qf.blocks + b * GQF_BLOCK_BYTES));
}

__device__ static inline gqf_header gqf_load_header(const quotient_filter& qf, uint64_t b) {
  const GQF_GLOBAL uint4* p = gqf_block_ptr(qf, b);
  const uint4 v = p[0];
  gqf_header h;
  h.occupieds = (uint64_t)v.x | ((uint64_t)v.y << 32);
  h.runends = (uint64_t)v.z | ((uint64_t)v.w << 32);
  h.offset = ((const GQF_GLOBAL uint32_t*)p)[4];
  return h;
}

__device__ static inline uint64_t gqf_load_runends(const quotient_filter& qf, uint64_t b) {
  return ((const GQF_GLOBAL uint64_t*)gqf_block_ptr(qf, b))[1];
}

// The used-slot mask of a block from its header alone: the first `offset`
// slots hold runs of earlier blocks' quotients; past them, slot i is inside
// a run iff more of the block's quotients <= i have runs than runs have
// ended in [offset, i) (each run covers its start .. end without gaps).
// `lim` = slots of the block below xnslots (64 for every block but the last).
__host__ __device__ static inline uint64_t gqf_used_mask(uint64_t occ, uint64_t runends, uint32_t offset, uint32_t lim) {
  uint64_t used;
  if (offset >= 64) {
    used = ~0ULL;
  } else {
    used = gqf_mask(offset);
    int p = __builtin_popcountll(occ & gqf_mask(offset));
    int r = 0;
    for (uint32_t i = offset; i < 64; i++) {
      p += (int)((occ >> i) & 1);
      if (p > r) used |= 1ULL << i;
      r += (int)((runends >> i) & 1);
    }
  }
  return lim < 64 ? used & gqf_mask(lim) : used;
}

// byte address of slot s inside the block array
__host__ __device__ static inline uint64_t gqf_slot_addr(uint64_t s) {
  return (s >> 6) * GQF_BLOCK_BYTES + GQF_SLOT_BASE + (s & 63);
}

// device table bytes: the blocks plus slack for vector over-reads
__host__ __device__ static inline uint64_t gqf_table_bytes(uint64_t nblocks) {
  return nblocks * GQF_BLOCK_BYTES + 4096;
}

// ---------------------------------------------------------------------------
// Counter codec (the CQF counter format of src/gqf.cu:1207-1215):
//   count 1   : r
//   count 2   : r r
//   count 3   : r 0 r  (r > 0)      0 0 0  (r == 0)
//   count > 3 : r [0] d.. r  (r > 0, digits base 254 skipping 0 and r,
//                             a 0 is prepended when the top digit >= r)
//               0 d.. 0 0    (r == 0, digits base 255 skipping 0)
// gqf_counter_len / gqf_counter_emit produce exactly the slots that
// encode_counter (src/gqf.cu:1216-1272) writes, in slot order.

// A count > 3 as digits: their number n and whether a 0 precedes them
// (r > 0 and top digit >= r).  The digits themselves are recomputed where
// they are written, least significant first, so no per-lane buffer is
// needed (a buffer indexed by digit would live in scratch memory).
// c / base (returned) and c % base (*r) for a counter's digit base: 254
// for a nonzero remainder, 255 for remainder 0.  Two constant divisors let
// the compiler multiply by reciprocals instead of running a 64-bit division
// (the rebuild encodes every counter of count >= 3 this way).
__host__ __device__ static inline uint64_t gqf_digit_divmod(uint64_t rem, uint64_t c, uint64_t* r) {
  const uint64_t q = rem ? c / 254 : c / 255;
  *r = c - q * (rem ? 254 : 255);
  return q;
}

__host__ __device__ static inline void gqf_counter_shape(uint64_t rem, uint64_t count, int* n_out, int* pad0_out) {
  uint64_t c = rem ? count - 3 : count - 4;
  int n = 0;
  uint64_t digit = 0;
  do {
    uint64_t r;
    c = gqf_digit_divmod(rem, c, &r);
    digit = r + 1;
    if (rem && digit >= rem) digit++;
    n++;
  } while (c);
  *n_out = n;
  *pad0_out = (rem && digit >= rem) ? 1 : 0;
}

// put(k0 + n - 1 - i, digit i) for the n digits of a count > 3, digit 0 the
// least significant
template <typename Put>
__host__ __device__ static inline void gqf_put_counter_digits(uint64_t rem, uint64_t count, uint32_t k0, int n,
                                                              Put put) {
  uint64_t c = rem ? count - 3 : count - 4;
  for (int i = 0; i < n; i++) {
    uint64_t r;
    c = gqf_digit_divmod(rem, c, &r);
    uint64_t digit = r + 1;
    if (rem && digit >= rem) digit++;
    put(k0 + (uint32_t)(n - 1 - i), digit);
  }
}

__host__ __device__ static inline uint32_t gqf_counter_len(uint64_t rem, uint64_t count) {
  if (count == 0) return 0;
  if (count <= 3) return (uint32_t)count;
  int n, pad0;
  gqf_counter_shape(rem, count, &n, &pad0);
  return rem ? (uint32_t)(n + 2 + pad0) : (uint32_t)(n + 3);
}

// Decoder of one counter inside a run held as bytes run[0..len) whose last
// byte ends the run (decode_counter semantics, src/gqf.cu:1276-1345).
// Returns the index of the counter's last slot.
template <typename SlotFn>
__host__ __device__ static inline uint32_t gqf_decode_counter(SlotFn slot, uint32_t len, uint32_t index,
                                                             uint64_t* remainder, uint64_t* count) {
  uint64_t rem = slot(index);
  *remainder = rem;
  if (index + 1 == len) { *count = 1; return index; }
  uint64_t digit = slot(index + 1);
  if (index + 2 == len) {
    *count = digit == rem ? 2 : 1;
    return index + (digit == rem ? 1 : 0);
  }
  if (rem > 0 && digit >= rem) {
    *count = digit == rem ? 2 : 1;
    return index + (digit == rem ? 1 : 0);
  }
  if (rem > 0 && digit == 0 && slot(index + 2) == rem) { *count = 3; return index + 2; }
  if (rem == 0 && digit == 0) {
    if (slot(index + 2) == 0) { *count = 3; return index + 2; }
    *count = 2;
    return index + 1;
  }
  uint64_t cnt = 0;
  uint64_t base = rem ? 254 : 255;
  uint32_t end = index + 1;
  while (digit != rem && end + 1 != len) {
    if (digit > rem) digit--;
    if (digit && rem) digit--;
    cnt = cnt * base + digit;
    end++;
    digit = slot(end);
  }
  if (rem) { *count = cnt + 3; return end; }
  if (end + 1 == len || slot(end + 1) != 0) { *count = 1; return index; }
  *count = cnt + 4;
  return end + 1;
}

// Counter emission in slot order: put(k, byte) for k = 0.. (the bytes
// encode_counter writes, src/gqf.cu:1216-1272); returns the length.
template <typename Put>
__host__ __device__ static inline uint32_t gqf_emit_counter(uint64_t rem, uint64_t count, Put put) {
  if (count == 0) return 0;
  if (count == 1) { put(0, rem); return 1; }
  if (count == 2) { put(0, rem); put(1, rem); return 2; }
  if (count == 3) {  // r 0 r, or 0 0 0 for r == 0
    put(0, rem);
    put(1, 0);
    put(2, rem);
    return 3;
  }
  int n, pad0;
  gqf_counter_shape(rem, count, &n, &pad0);
  uint32_t k = 0;
  put(k++, rem);
  if (rem && pad0) put(k++, 0);
  gqf_put_counter_digits(rem, count, k, n, put);
  k += (uint32_t)n;
  if (rem == 0) put(k++, 0);
  put(k++, rem);
  return k;
}

// ---------------------------------------------------------------------------
// Point operations.  All positions are slot indices; the helpers read and
// write the table directly (no staging), one lane per key.  They are also
// compiled for the host, where the tests run them on a table in host memory
// against the CPU oracle.

namespace gqf_point {

__host__ __device__ static inline uint8_t* blk(const QF& f, uint64_t b) { return f.blocks + b * GQF_BLOCK_BYTES; }
__host__ __device__ static inline uint64_t occ_word(const QF& f, uint64_t b) { return ((const uint64_t*)blk(f, b))[0]; }
__host__ __device__ static inline uint64_t run_word(const QF& f, uint64_t b) { return ((const uint64_t*)blk(f, b))[1]; }
__host__ __device__ static inline uint32_t offset(const QF& f, uint64_t b) { return ((const uint32_t*)blk(f, b))[4]; }
__host__ __device__ static inline uint8_t slot(const QF& f, uint64_t s) { return f.blocks[gqf_slot_addr(s)]; }
__host__ __device__ static inline void set_slot(const QF& f, uint64_t s, uint8_t v) { f.blocks[gqf_slot_addr(s)] = v; }
__host__ __device__ static inline bool is_runend(const QF& f, uint64_t s) { return (run_word(f, s >> 6) >> (s & 63)) & 1ULL; }
__host__ __device__ static inline void set_runend(const QF& f, uint64_t s, bool on) {
  uint64_t* w = (uint64_t*)blk(f, s >> 6) + 1;
  const uint64_t bit = 1ULL << (s & 63);
  *w = on ? (*w | bit) : (*w & ~bit);
}
__host__ __device__ static inline void set_occupied(const QF& f, uint64_t q) {
  uint64_t* w = (uint64_t*)blk(f, q >> 6);
  *w |= 1ULL << (q & 63);
}
// marks slot s as inside a run (the used-slot mask at bytes 24..31)
__host__ __device__ static inline void set_used(const QF& f, uint64_t s) {
  uint64_t* w = (uint64_t*)blk(f, s >> 6) + 3;
  *w |= 1ULL << (s & 63);
}

// end of the runs of quotients < 64 b (at least 64 b - 1)
__host__ __device__ static inline int64_t end_before_block(const QF& f, uint64_t b) {
  return (int64_t)(b * 64) - 1 + (int64_t)offset(f, b);
}

// position of the k-th (k = 0 first) run end at or after `pos`, -1 if none
__host__ __device__ static inline int64_t runend_from(const QF& f, uint64_t pos, uint64_t k) {
  uint64_t b = pos >> 6;
  if (b >= f.nblocks) return -1;
  uint64_t w = run_word(f, b) & (~0ULL << (pos & 63));
  for (;;) {
    const uint64_t c = (uint64_t)__builtin_popcountll(w);
    if (k < c) return (int64_t)(b * 64) + gqf_select64(w, (int)k);
    k -= c;
    if (++b >= f.nblocks) return -1;
    w = run_word(f, b);
  }
}

// end of the runs of quotients <= q, at least q (run_end, src/gqf.cu:632-679)
__host__ __device__ static inline int64_t end_upto(const QF& f, uint64_t q) {
  const uint64_t b = q >> 6;
  const int64_t x = end_before_block(f, b);
  const uint64_t k = (uint64_t)__builtin_popcountll(occ_word(f, b) & gqf_mask((q & 63) + 1));
  int64_t e = k ? runend_from(f, (uint64_t)(x + 1), k - 1) : x;
  return e > (int64_t)q ? e : (int64_t)q;
}

// first empty slot at or after `from` (find_first_empty_slot,
// src/gqf.cu:728-771); nblocks * 64 if none.  Slot s = 64 b + i after the
// runs of earlier blocks is empty iff the runs of block b's quotients <= s
// have all ended before s.
__host__ __device__ static inline uint64_t first_empty(const QF& f, uint64_t from) {
  for (uint64_t b = from >> 6; b < f.nblocks; b++) {
    const int64_t x = end_before_block(f, b);
    const int64_t first = (int64_t)from > x + 1 ? (int64_t)from : x + 1;
    if (first >= (int64_t)(b * 64 + 64)) continue;
    const uint64_t o = occ_word(f, b);
    const int lo_x = (int)(x + 1 - (int64_t)(b * 64));  // run ends at block positions < lo_x belong to earlier blocks
    const uint64_t r = run_word(f, b) & ~(lo_x > 0 ? gqf_mask((uint64_t)lo_x) : 0ULL);
    for (int i = (int)(first - (int64_t)(b * 64)); i < 64; i++) {
      if (__builtin_popcountll(o & gqf_mask((uint64_t)i + 1)) == __builtin_popcountll(r & gqf_mask((uint64_t)i))) return b * 64 + (uint64_t)i;
    }
  }
  return f.nblocks * 64;
}

// Opens d free slots at p: the occupied stretch [p, e1) moves right by d,
// (e1, e2) by d - 1, ..., where e1 < e2 < ... are the first d empty slots at
// or after p; run-end bits move with their slots.  Slots [p, p + d) are then
// free (stale bytes and bits) for the caller to fill; the d empty slots
// become used, and the used-slot masks say so.  Returns the last empty slot
// used, or -1 if fewer than d empty slots exist below xnslots (nothing is
// moved then).
__host__ __device__ static inline int64_t open_slots(const QF& f, uint64_t p, int d) {
  uint64_t e[4];
  uint64_t from = p;
  for (int j = 0; j < d; j++) {
    e[j] = first_empty(f, from);
    if (e[j] >= f.xnslots) return -1;
    from = e[j] + 1;
  }
  for (int j = d - 1; j >= 0; j--) {
    const uint64_t lo = j == 0 ? p : e[j - 1] + 1;
    const uint64_t sh = (uint64_t)(d - j);
    for (uint64_t s = e[j]; s-- > lo;) {
      set_slot(f, s + sh, slot(f, s));
      set_runend(f, s + sh, is_runend(f, s));
    }
    set_used(f, e[j]);
  }
  return (int64_t)e[d - 1];
}

// Re-derives the exact offsets of blocks hb+1 .. last/64 from their
// predecessors after a shift that ended at slot `last`.
__host__ __device__ static inline void fix_offsets(const QF& f, uint64_t hb, uint64_t last) {
  for (uint64_t b = hb + 1; b <= (last >> 6) && b < f.nblocks; b++) {
    const int64_t x = end_before_block(f, b - 1);
    const uint64_t k = (uint64_t)__builtin_popcountll(occ_word(f, b - 1));
    int64_t e = k ? runend_from(f, (uint64_t)(x + 1), k - 1) : x;
    const int64_t lo = (int64_t)(b * 64) - 1;
    ((uint32_t*)blk(f, b))[4] = (uint32_t)(e > lo ? e - lo : 0);
  }
}

struct run_bounds {
  uint64_t start, end;
  bool ok;  // false only for an inconsistent table (no run end left)
};

// run of occupied quotient q: starts at max(q, end of the earlier runs + 1)
__host__ __device__ static inline run_bounds run_of(const QF& f, uint64_t q) {
  const int64_t prev = q ? end_upto(f, q - 1) : -1;
  run_bounds r;
  r.start = (uint64_t)(prev + 1 > (int64_t)q ? prev + 1 : (int64_t)q);
  const int64_t e = r.start < f.nblocks * 64 ? runend_from(f, r.start, 0) : -1;
  r.ok = e >= (int64_t)r.start;
  r.end = r.ok ? (uint64_t)e : r.start;
  return r;
}

// Count of `hash` (qf_query / qf_count_key_value, src/gqf.cu:3479-3557).
__host__ __device__ static inline uint64_t count_hash(const QF& f, uint64_t hash) {
  const uint64_t q = hash >> 8, rem = hash & 0xFF;
  if (!((occ_word(f, q >> 6) >> (q & 63)) & 1ULL)) return 0;
  const run_bounds rb = run_of(f, q);
  if (!rb.ok) return 0;
  const uint32_t len = (uint32_t)(rb.end - rb.start + 1);
  auto sl = [&](uint32_t k) -> uint64_t { return slot(f, rb.start + k); };
  uint32_t idx = 0;
  do {
    uint64_t r, c;
    const uint32_t e = gqf_decode_counter(sl, len, idx, &r, &c);
    if (r == rem) return c;
    idx = e + 1;
  } while (idx < len);
  return 0;
}

// Adds one instance of `hash` (insert1, src/gqf.cu:1523-1766): a new
// counter [r] at its sorted place in the run (a new run if q is empty), or
// the counter re-encoded with count + 1.  The layout stays the one
// encode_counter and the run rules define.  Returns 0, or QF_NO_SPACE.
__host__ __device__ static inline int insert1(const QF& f, uint64_t hash) {
  const uint64_t q = hash >> 8, rem = hash & 0xFF, hb = q >> 6;
  if (!((occ_word(f, hb) >> (q & 63)) & 1ULL)) {
    const int64_t prev = q ? end_upto(f, q - 1) : -1;
    const uint64_t p = (uint64_t)(prev + 1 > (int64_t)q ? prev + 1 : (int64_t)q);
    const int64_t last = open_slots(f, p, 1);
    if (last < 0) return QF_NO_SPACE;
    set_slot(f, p, (uint8_t)rem);
    set_runend(f, p, true);
    set_occupied(f, q);
    fix_offsets(f, hb, (uint64_t)last);
    return 0;
  }
  const run_bounds rb = run_of(f, q);
  if (!rb.ok) return QF_NO_SPACE;
  const uint32_t len = (uint32_t)(rb.end - rb.start + 1);
  auto sl = [&](uint32_t k) -> uint64_t { return slot(f, rb.start + k); };
  uint32_t idx = 0;
  uint64_t cnt = 0;
  uint32_t cend = 0;
  while (idx < len) {
    uint64_t r, c;
    const uint32_t e = gqf_decode_counter(sl, len, idx, &r, &c);
    if (r == rem) { cnt = c; cend = e; break; }
    if (r > rem) break;
    idx = e + 1;
  }
  if (cnt == 0) {  // new counter [rem] at slot start + idx
    const uint64_t p = rb.start + idx;
    const int64_t last = open_slots(f, p, 1);
    if (last < 0) return QF_NO_SPACE;
    set_slot(f, p, (uint8_t)rem);
    if (p == rb.end + 1) {
      set_runend(f, rb.end, false);
      set_runend(f, p, true);
    } else {
      set_runend(f, p, false);
    }
    fix_offsets(f, hb, (uint64_t)last);
    return 0;
  }
  const uint64_t cs = rb.start + idx, ce = rb.start + cend;
  const uint32_t old_len = cend - idx + 1;
  const uint32_t new_len = gqf_counter_len(rem, cnt + 1);
  const int d = (int)(new_len - old_len);
  int64_t last = -1;
  if (d > 0) {
    last = open_slots(f, ce + 1, d);
    if (last < 0) return QF_NO_SPACE;
    for (int k = 1; k <= d; k++) set_runend(f, ce + (uint64_t)k, false);
    if (ce == rb.end) {
      set_runend(f, ce, false);
      set_runend(f, ce + (uint64_t)d, true);
    }
  }
  gqf_emit_counter(rem, cnt + 1, [&](uint32_t k, uint64_t v) { set_slot(f, cs + k, (uint8_t)v); });
  if (d > 0) fix_offsets(f, hb, (uint64_t)last);
  return 0;
}

// insert1_if_not_exists (src/gqf.cu:1364-1518), including its slot-wise
// search: the run is scanned slot by slot (not counter by counter) for the
// first slot >= r; equal means found, otherwise a raw slot r is inserted
// there (or appended after the run).  *value is set to 0 when found
// (get_slot(...) && BITMASK(0) is always 0, src/gqf.cu:1454).
__host__ __device__ static inline qf_returns insert1_if_not_exists(const QF& f, uint64_t hash, uint8_t* value) {
  const uint64_t q = hash >> 8, rem = hash & 0xFF, hb = q >> 6;
  uint64_t p;
  int op;  // 0: new run, 1: append to the run, 2: insert before slot p
  run_bounds rb{0, 0, true};
  if (!((occ_word(f, hb) >> (q & 63)) & 1ULL)) {
    const int64_t prev = q ? end_upto(f, q - 1) : -1;
    p = (uint64_t)(prev + 1 > (int64_t)q ? prev + 1 : (int64_t)q);
    op = 0;
  } else {
    rb = run_of(f, q);
    if (!rb.ok) return QF_FULL;
    uint64_t i = rb.start;
    uint64_t cur = slot(f, i);
    while (cur < rem && i <= rb.end) {
      i++;
      cur = i < f.nblocks * 64 ? slot(f, i) : 0;
    }
    if (i > rb.end) {
      op = 1;
    } else if (cur != rem) {
      op = 2;
    } else {
      *value = 0;
      return QF_ITEM_FOUND;
    }
    p = i;
  }
  const int64_t last = open_slots(f, p, 1);
  if (last < 0) return QF_FULL;
  set_slot(f, p, (uint8_t)rem);
  if (op == 0) {
    set_runend(f, p, true);
  } else if (op == 1) {
    set_runend(f, p - 1, false);
    set_runend(f, p, true);
  } else {
    set_runend(f, p, false);
  }
  set_occupied(f, q);
  fix_offsets(f, hb, (uint64_t)last);
  return QF_ITEM_INSERTED;
}

// The reference reduces the (hashed or caller-given) key modulo its range
// 2^key_bits (src/gqf.cu:2973-2980) and ignores the point value.
__host__ __device__ static inline uint64_t key_to_hash(const QF& f, uint64_t key, uint8_t flags) {
  if (!(flags & QF_KEY_IS_HASH)) key = gqf_hash64(key, gqf_mask(f.key_bits));
  return (key & gqf_mask(f.key_bits)) << f.value_bits;
}

__device__ static inline bool try_lock(const QF& f, uint64_t idx) {
  uint32_t* w = &f.locks[idx * GQF_REF_LOCK_DIST];
  if (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
  uint32_t expected = 0;
  return __hip_atomic_compare_exchange_strong(w, &expected, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}

__device__ static inline void unlock(const QF& f, uint64_t idx) {
  __hip_atomic_store(&f.locks[idx * GQF_REF_LOCK_DIST], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Runs body() holding the region locks idx and idx + 1.  Both are tried
// without waiting; a lane that gets only the first releases it and tries
// again.  Of the lanes contending for a chain of regions the one with the
// highest idx always gets both, so every lane finishes.
//
// The reference waits on lock idx + 1 while holding idx
// (src/gqf.cu:2993-3000), which deadlocks lanes of a wave64 that run in
// lockstep.  The same deadlock returns if the compiler moves the critical
// section behind the loop exit (the natural shape of "loop until acquired,
// then work"): the lanes that got the locks would wait at the loop's
// reconvergence point for lanes spinning on those very locks.  So the loop
// exits only when no lane of the wave is left (a ballot, which the compiler
// cannot split per lane), and the body runs inside the iteration that took
// the locks.  Acquire / release fences are issued by lock holders only.
template <typename Body>
__device__ static inline void with_region_locks(const QF& f, uint64_t idx, Body body) {
  bool done = false;
  uint32_t backoff = 1;
  while (__any(!done)) {
    if (!done) {
      if (try_lock(f, idx)) {
        if (try_lock(f, idx + 1)) {
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          body();
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          unlock(f, idx + 1);
          unlock(f, idx);
          done = true;
        } else {
          unlock(f, idx);
        }
      }
      if (!done) {
        // exponential backoff keeps spinning lanes off the lock words
        for (uint32_t k = 0; k < backoff; k++) __builtin_amdgcn_s_sleep(8);
        backoff = backoff < 64 ? 2 * backoff : 64;
      }
    }
  }
}

}  // namespace gqf_point

// One instance of key (hashed unless flags has QF_KEY_IS_HASH); value is
// unused (0 value bits).  QF_ITEM_INSERTED, or QF_FULL when no empty slot
// is left (src/gqf.cu:2970-3020).
__device__ static inline qf_returns point_insert(QF* qf, uint64_t key, uint8_t value, uint8_t flags) {
  (void)value;
  const QF& f = *qf;
  const uint64_t hash = gqf_point::key_to_hash(f, key, flags);
  const uint64_t lock = (hash >> 8) / GQF_REF_LOCK_SLOTS;
  int ret = 0;
  gqf_point::with_region_locks(f, lock, [&]() { ret = gqf_point::insert1(f, hash); });
  if (ret == QF_NO_SPACE) {
    atomicOr(f.status, QF_STATUS_POINT_FULL);
    return QF_FULL;
  }
  return QF_ITEM_INSERTED;
}

// Inserts key unless its hash is present: QF_ITEM_INSERTED, QF_ITEM_FOUND
// (returnedVal = 0), or QF_FULL (src/gqf.cu:2911-2966).
__device__ static inline qf_returns point_insert_not_exists(QF* qf, uint64_t key, uint8_t value,
                                                            uint8_t& returnedVal, uint8_t flags) {
  (void)value;
  const QF& f = *qf;
  const uint64_t hash = gqf_point::key_to_hash(f, key, flags);
  const uint64_t lock = (hash >> 8) / GQF_REF_LOCK_SLOTS;
  qf_returns ret = QF_ITEM_INSERTED;
  uint8_t query = 0;
  gqf_point::with_region_locks(f, lock, [&]() { ret = gqf_point::insert1_if_not_exists(f, hash, &query); });
  if (ret == QF_ITEM_FOUND) returnedVal = query;
  if (ret == QF_FULL) atomicOr(f.status, QF_STATUS_POINT_FULL);
  return ret;
}

// Count of key, without locks (src/gqf.cu:3024-3051).  returnedVal is the
// value bits of the matching counter: always 0 here.  (The reference leaves
// it uninitialised when the quotient is empty; it is 0 here.)
__device__ static inline uint64_t point_query(QF* qf, uint64_t key, uint8_t value, uint8_t& returnedVal,
                                              uint8_t flags) {
  (void)value;
  const QF& f = *qf;
  returnedVal = 0;
  return gqf_point::count_hash(f, gqf_point::key_to_hash(f, key, flags));
}

// point_query under the region locks (src/gqf.cu:3055-3108).
__device__ static inline uint64_t point_query_concurrent(QF* qf, uint64_t key, uint8_t value,
                                                         uint8_t& returnedVal, uint8_t flags) {
  (void)value;
  const QF& f = *qf;
  const uint64_t hash = gqf_point::key_to_hash(f, key, flags);
  const uint64_t lock = (hash >> 8) / GQF_REF_LOCK_SLOTS;
  uint64_t c = 0;
  gqf_point::with_region_locks(f, lock, [&]() { c = gqf_point::count_hash(f, hash); });
  returnedVal = 0;
  return c;
}

// The reference's point_bulk_insert kernel (include/gqf.cuh:136,
// src/gqf.cu:3153-3162): thread i inserts one instance of hashes[i] with
// point_insert and flags 0, so the items are hashed like any key.  The
// reference's callers launch it as point_bulk_insert<<<(n - 1) / 32 + 1, 32>>>
// (include/point_wrapper.cuh:63); any launch shape works.  A full table
// sets QF_STATUS_POINT_FULL (the reference's assert compares a qf_returns
// with QF_NO_SPACE, which it never equals).  Defined here, in the header, so
// the caller's launch compiles unchanged: a kernel cannot be exported
// across a shared library.  GQF_NO_POINT_BULK_KERNEL leaves it out.
#ifndef GQF_NO_POINT_BULK_KERNEL
static __global__ void point_bulk_insert(QF* qf, uint64_t* hashes, uint64_t nitems) {
  const uint64_t tid = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  if (tid >= nitems) return;
  (void)point_insert(qf, hashes[tid], 0, 0);
}
#endif

#endif /* GQF_DEVICE_H */
