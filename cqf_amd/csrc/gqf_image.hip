// gqf_image.hip -- conversion from the device layout (128-byte blocks with an
// exact 32-bit offset) to the reference's packed 81-byte block image
// (include/gqf_int.cuh:38-57): offset u8 saturated at 255, occupieds u64,
// runends u64, 64 slot bytes.  This is the parity / serialisation format.
#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {

// one thread per 8-byte word of output, gathering its bytes from the arrays
__global__ void __launch_bounds__(256) gqf_export_kernel(const QF* __restrict__ qfp, uint8_t* __restrict__ out) {
  const QF& qf = *qfp;
  const uint64_t nblocks = qf.nblocks;
  const uint64_t total = nblocks * GQF_REF_BLOCK_BYTES;
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t first = w * 8;
  if (first >= total) return;
  uint8_t v[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint64_t i = first + k;
    uint8_t byte = 0;
    if (i < total) {
      const uint64_t b = i / GQF_REF_BLOCK_BYTES;
      const uint32_t o = (uint32_t)(i - b * GQF_REF_BLOCK_BYTES);
      const uint8_t* src = qf.blocks + b * GQF_BLOCK_BYTES;
      if (o == 0) {
        const uint32_t off = *(const uint32_t*)(src + 16);
        byte = (uint8_t)(off < GQF_OFFSET_SAT ? off : GQF_OFFSET_SAT);
      } else if (o < 17) {
        byte = src[o - 1];  // occupieds then runends, both little endian
      } else {
        byte = src[GQF_SLOT_BASE + (o - 17)];
      }
    }
    v[k] = byte;
  }
  if (first + 8 <= total && (((uintptr_t)out + first) & 7) == 0) {
    uint64_t word = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) word |= (uint64_t)v[k] << (8 * k);
    *(uint64_t*)(out + first) = word;
  } else {
    for (int k = 0; k < 8 && first + k < total; k++) out[first + k] = v[k];
  }
}

}  // namespace

void gqf_launch_export(const QF* qf_dev, uint64_t nblocks, uint8_t* out, hipStream_t s) {
  const uint64_t words = (nblocks * GQF_REF_BLOCK_BYTES + 7) / 8;
  if (words == 0) return;
  const uint64_t grid = (words + 255) / 256;
  hipLaunchKernelGGL(gqf_export_kernel, dim3((uint32_t)grid), dim3(256), 0, s, qf_dev, out);
}
