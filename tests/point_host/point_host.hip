// Host build of the device point operations (include/gqf_device.h) for the
// CPU tests: the same insert1 / insert1_if_not_exists / count code the GPU
// runs, applied to a table in host memory so the tests can compare it with
// the CPU oracle without a GPU.  Test infrastructure, not part of libgqf.
#include <math.h>
#include <string.h>

#include "gqf_device.h"

namespace {

QF make_qf(uint8_t* blocks, int nbits, uint32_t* status) {
  QF f;
  memset(&f, 0, sizeof(f));
  f.nslots = 1ULL << nbits;
  f.xnslots = f.nslots + (uint64_t)(10 * sqrt((double)f.nslots));
  f.nblocks = (f.xnslots + 63) / 64;
  f.key_bits = (uint64_t)nbits + 8;
  f.key_remainder_bits = 8;
  f.bits_per_slot = 8;
  f.range = f.nslots << 8;
  f.num_locks = f.xnslots / GQF_REF_LOCK_SLOTS + 2;
  f.blocks = blocks;
  f.status = status;
  return f;
}

}  // namespace

extern "C" uint64_t ph_table_bytes(int nbits) {
  uint32_t st = 0;
  const QF f = make_qf(nullptr, nbits, &st);
  return gqf_table_bytes(f.nblocks);
}

extern "C" int ph_insert1(uint8_t* blocks, int nbits, uint64_t hash) {
  uint32_t st = 0;
  const QF f = make_qf(blocks, nbits, &st);
  return gqf_point::insert1(f, hash);
}

extern "C" int ph_insert_not_exists(uint8_t* blocks, int nbits, uint64_t hash, uint8_t* value) {
  uint32_t st = 0;
  const QF f = make_qf(blocks, nbits, &st);
  return (int)gqf_point::insert1_if_not_exists(f, hash, value);
}

extern "C" uint64_t ph_count(uint8_t* blocks, int nbits, uint64_t hash) {
  uint32_t st = 0;
  const QF f = make_qf(blocks, nbits, &st);
  return gqf_point::count_hash(f, hash);
}

// device layout -> the reference's 81-byte image (offset byte saturated)
extern "C" void ph_export(const uint8_t* blocks, int nbits, uint8_t* out) {
  uint32_t st = 0;
  const QF f = make_qf(nullptr, nbits, &st);
  for (uint64_t b = 0; b < f.nblocks; b++) {
    const uint8_t* src = blocks + b * GQF_BLOCK_BYTES;
    uint8_t* dst = out + b * GQF_REF_BLOCK_BYTES;
    uint32_t off;
    memcpy(&off, src + 16, 4);
    dst[0] = (uint8_t)(off < GQF_OFFSET_SAT ? off : GQF_OFFSET_SAT);
    memcpy(dst + 1, src, 16);
    memcpy(dst + 17, src + GQF_SLOT_BASE, 64);
  }
}

// blocks whose stored used-slot mask (bytes 24..31) disagrees with the one
// their header implies
extern "C" uint64_t ph_used_mismatches(const uint8_t* blocks, int nbits) {
  uint32_t st = 0;
  const QF f = make_qf(nullptr, nbits, &st);
  uint64_t bad = 0;
  for (uint64_t b = 0; b < f.nblocks; b++) {
    const uint8_t* src = blocks + b * GQF_BLOCK_BYTES;
    uint64_t occ, run, used;
    uint32_t off;
    memcpy(&occ, src, 8);
    memcpy(&run, src + 8, 8);
    memcpy(&off, src + 16, 4);
    memcpy(&used, src + 24, 8);
    const uint64_t lim = f.xnslots - b * 64;
    if (used != gqf_used_mask(occ, run, off, lim < 64 ? (uint32_t)lim : 64u)) bad++;
  }
  return bad;
}
