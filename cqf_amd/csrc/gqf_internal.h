// gqf_internal.h -- device-side layout, constants and helpers shared by the
// HIP kernels of libgqf (counting quotient filter for gfx950).
//
// The reference stores the table as packed 81-byte blocks
// (PASSIONLab/cqf include/gqf_int.cuh:38-57):
//   byte 0       offset     u8   (saturates at 255 = "recompute")
//   bytes 1..8   occupieds  u64 little endian, bit i = quotient 64b+i has a run
//   bytes 9..16  runends    u64 little endian, bit i = slot 64b+i ends a run
//   bytes 17..80 slots[64]  u8 remainders / counter digits
// On the device each block occupies one 128-byte line, every field aligned:
//   bytes 0..7   occupieds   bytes 8..15 runends   bytes 16..19 offset (u32)
//   bytes 20..63 zero        bytes 64..127 slots[64]
// A probe loads the whole line in one round of back-to-back vector loads:
// header and, for ~94% of runs at 90% load, the run itself.  Random access
// on MI355X is bound by the number of dependent line requests, not bytes
// (independent 16-byte and 128-byte random reads run at the same rate), so
// one line per probe beats the 81-byte packing (1.4 lines, two dependent
// rounds) although the table is 58% larger.
// The device offset is exact -- end of the runs with quotient < 64 b, minus
// (64 b - 1), floored at 0 -- never saturated, so no reader has to walk
// back to recompute it; the export saturates it to the reference's byte.
// The reference's 81-byte image is produced by qf_export_image and read by
// qf_import_image; it is the parity format, byte for byte.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GQF_REF_BLOCK_BYTES 81   // the reference's packed block (export format)
#define GQF_BLOCK_BYTES 128      // device stride of one 64-slot block
#define GQF_SLOT_BASE 64         // byte of slot 0 inside a device block
#define GQF_SLOTS_PER_BLOCK 64
#define GQF_REGION_BLOCKS 48                       // rebuild region; + 16 push blocks = one wave64
#define GQF_REGION_SLOTS (GQF_REGION_BLOCKS * GQF_SLOTS_PER_BLOCK)
#define GQF_REF_LOCK_SLOTS 8192ULL                 // reference point-lock region (src/gqf.cu:58)
#define GQF_REF_LOCK_DIST 64                       // reference lock stride (src/gqf.cu:59)
#define GQF_OFFSET_SAT 255                         // the reference offset byte saturates here

// status bits reported through qf_get_status()
#define GQF_STATUS_NO_SPACE 0x1u          // a run would end past xnslots; region left unchanged
#define GQF_STATUS_CLUSTER_OVERFLOW 0x2u  // push longer than the widest window; region left unchanged
#define GQF_STATUS_BAD_HASH 0x4u          // pre-hashed key >= range; key ignored
#define GQF_STATUS_POINT_FULL 0x8u        // a point insert found no empty slot

// Device-resident filter descriptor.  The public QF* handle is a device
// pointer to this struct, as the reference's QF* is a device pointer
// (src/gqf.cu:2212-2217); callers never dereference it on the host.
struct quotient_filter {
  uint64_t nslots;              // 2^q
  uint64_t xnslots;             // nslots + 10 sqrt(nslots), truncated (src/gqf.cu:1950)
  uint64_t nblocks;             // ceil(xnslots / 64)
  uint64_t key_bits;            // q + 8
  uint64_t key_remainder_bits;  // 8
  uint64_t bits_per_slot;       // 8
  uint64_t range;               // 2^(q+8)
  uint64_t num_locks;           // xnslots / 8192 + 2 (src/gqf.cu:1996)
  uint64_t nregions;            // ceil(nblocks / 64): even/odd rebuild regions
  uint8_t* blocks;              // nblocks * GQF_BLOCK_BYTES bytes (+ slack)
  uint32_t* locks;              // point-mode region locks (num_locks * 64 words), or null
  uint32_t* status;             // GQF_STATUS_* bits, sticky until cleared
};
typedef struct quotient_filter QF;

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

// ---------------------------------------------------------------------------
// device bit helpers

__device__ static inline int gqf_popc(uint64_t v) { return __popcll(v); }

// position of the k'th set bit (k = 0 is the lowest), 64 if fewer: a
// six-step binary descent on popcounts, no table.
__device__ static inline int gqf_select64(uint64_t v, int k) {
  if (k >= __popcll(v)) return 64;
  int pos = 0;
  uint32_t lo = (uint32_t)v;
  int c = __popc(lo);
  if (k >= c) { k -= c; v >>= 32; pos += 32; }
  uint32_t w = (uint32_t)v;
  c = __popc(w & 0xFFFFu);
  if (k >= c) { k -= c; w >>= 16; pos += 16; }
  c = __popc(w & 0xFFu);
  if (k >= c) { k -= c; w >>= 8; pos += 8; }
  c = __popc(w & 0xFu);
  if (k >= c) { k -= c; w >>= 4; pos += 4; }
  c = __popc(w & 0x3u);
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

struct gqf_digits {
  uint8_t d[12];  // most significant first
  int n;
  int pad0;       // leading 0 before the digits (r > 0 and top digit >= r)
};

__host__ __device__ static inline int gqf_counter_digits(uint64_t rem, uint64_t count, gqf_digits* out) {
  // count > 3 only
  uint64_t base = rem ? 254 : 255;
  uint64_t c = rem ? count - 3 : count - 4;
  uint8_t tmp[12];
  int n = 0;
  uint64_t digit = 0;
  do {
    digit = c % base + 1;
    if (rem && digit >= rem) digit++;
    tmp[n++] = (uint8_t)digit;
    c /= base;
  } while (c);
  for (int i = 0; i < n; i++) out->d[i] = tmp[n - 1 - i];
  out->n = n;
  out->pad0 = (rem && digit >= rem) ? 1 : 0;
  return n;
}

__host__ __device__ static inline uint32_t gqf_counter_len(uint64_t rem, uint64_t count) {
  if (count == 0) return 0;
  if (count <= 3) return (uint32_t)count;
  gqf_digits g;
  gqf_counter_digits(rem, count, &g);
  return rem ? (uint32_t)(g.n + 2 + g.pad0) : (uint32_t)(g.n + 3);
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
