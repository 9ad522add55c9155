// gqf_internal.h -- constants and helpers private to the libgqf kernels
// (counting quotient filter for gfx950).  The table layout, the filter
// descriptor and the shared device helpers live in the public device
// header gqf_device.h.
#pragma once

#define GQF_NO_POINT_BULK_KERNEL  // the library launches its own point kernels (gqf_point.hip)
#include "gqf_device.h"

#define GQF_REGION_BLOCKS 48      // default rebuild region; + 16 push blocks = one wave64
#define GQF_REGION_BLOCKS_MIN 16  // tunable range (the push window is min(64, 2 x region) - region blocks)
#define GQF_REGION_BLOCKS_MAX 56

// Largest batch one pass of the bulk pipeline takes (its region starts,
// sort histograms and reduce counts are 32-bit); longer batches are cut
// into pieces of this size.
#define GQF_MAX_PIECE (1ULL << 30)

// Cuts n keys into consecutive pieces of at most max_piece keys (max_piece
// >= 1): piece k is [starts[k], starts[k + 1]).  Returns the number of
// pieces p (0 for n = 0) and writes starts[0..p] when p + 1 <= cap; p is
// returned either way, so a caller can size the array.
static inline uint32_t gqf_batch_pieces(uint64_t n, uint64_t max_piece, uint64_t* starts, uint32_t cap) {
  if (max_piece == 0) return 0;
  const uint64_t p = n / max_piece + (n % max_piece ? 1 : 0);
  if (p + 1 <= cap && starts) {
    for (uint64_t k = 0; k < p; k++) starts[k] = k * max_piece;
    starts[p] = n;
  }
  return (uint32_t)p;
}

// status bits reported through qf_get_status()
#define GQF_STATUS_NO_SPACE QF_STATUS_NO_SPACE                  // a run would end past xnslots; region left unchanged
#define GQF_STATUS_CLUSTER_OVERFLOW QF_STATUS_CLUSTER_OVERFLOW  // push longer than the widest window
#define GQF_STATUS_BAD_HASH QF_STATUS_BAD_HASH                  // pre-hashed key >= range; key ignored
#define GQF_STATUS_POINT_FULL QF_STATUS_POINT_FULL              // a point insert found no empty slot

// Inclusive sum over the 64 lanes of a wave on DPP row shifts (1, 2, 4, 8
// within each row of 16 lanes) and row broadcasts (lane 15 into rows 1 and
// 3, then lane 31 into rows 2 and 3): six VALU steps that read the
// neighbouring lanes' registers, no LDS round trip.  All lanes must be
// active.
__device__ __forceinline__ uint32_t wave_scan_add(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x;
}
