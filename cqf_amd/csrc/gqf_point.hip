// gqf_point.hip -- host-side drivers of the device point API (gqf_device.h):
// one lane per key, grid-stride.
//
// Reference: point_bulk_insert (src/gqf.cu:3153-3162, launched by
// include/point_wrapper.cuh:63), point_bulk_get / point_get_wrapper
// (src/gqf.cu:3112-3150).  The extra drivers for point_insert_not_exists
// and point_query(_concurrent) give host code (and the tests) the device
// functions' return values.
#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {

inline uint32_t point_grid(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  const uint64_t cap = 256ULL * 32;
  return (uint32_t)(g < cap ? (g ? g : 1) : cap);
}

}  // namespace

extern "C" __global__ void __launch_bounds__(256)
gqf_point_insert_kernel(QF* qf, const uint64_t* __restrict__ keys, uint64_t n, uint8_t flags, int32_t* rets) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const qf_returns r = point_insert(qf, keys[i], 0, flags);
    if (rets) rets[i] = (int32_t)r;
  }
}

extern "C" __global__ void __launch_bounds__(256)
gqf_point_insert_not_exists_kernel(QF* qf, const uint64_t* __restrict__ keys, uint64_t n, uint8_t flags,
                                   int32_t* rets) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint8_t val = 0;
    const qf_returns r = point_insert_not_exists(qf, keys[i], 0, val, flags);
    if (rets) rets[i] = (int32_t)r;
  }
}

extern "C" __global__ void __launch_bounds__(256)
gqf_point_query_kernel(QF* qf, const uint64_t* __restrict__ keys, uint64_t n, uint8_t flags, uint64_t* counts,
                       int concurrent, unsigned long long* misses) {
  uint32_t mine = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint8_t val = 0;
    const uint64_t c = concurrent ? point_query_concurrent(qf, keys[i], 0, val, flags)
                                  : point_query(qf, keys[i], 0, val, flags);
    if (counts) counts[i] = c;
    mine += c == 0;
  }
  if (misses) {
    for (int d = 32; d >= 1; d >>= 1) mine += __shfl_down(mine, d, 64);
    if ((threadIdx.x & 63) == 0 && mine) atomicAdd(misses, (unsigned long long)mine);
  }
}

// Sparse bulk insert: point_insert of each (already hashed) key; a key the
// table has no room for sets QF_STATUS_NO_SPACE as a bulk call would.
extern "C" __global__ void __launch_bounds__(256)
gqf_sparse_insert_kernel(QF* qf, const uint64_t* __restrict__ hashes, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (point_insert(qf, hashes[i], 0, QF_KEY_IS_HASH) == QF_FULL) atomicOr(qf->status, GQF_STATUS_NO_SPACE);
  }
}

void gqf_launch_sparse_insert(QF* qf_dev, const uint64_t* hashes, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  gqf_prof_begin("sparse_insert", s);
  hipLaunchKernelGGL(gqf_sparse_insert_kernel, dim3(point_grid(n)), dim3(256), 0, s, qf_dev, hashes, n);
  gqf_prof_end("sparse_insert", s);
}

void gqf_launch_point_insert(QF* qf_dev, const uint64_t* keys, uint64_t n, uint8_t flags, int32_t* rets,
                             int not_exists, hipStream_t s) {
  if (n == 0) return;
  if (not_exists)
    hipLaunchKernelGGL(gqf_point_insert_not_exists_kernel, dim3(point_grid(n)), dim3(256), 0, s, qf_dev, keys, n,
                       flags, rets);
  else
    hipLaunchKernelGGL(gqf_point_insert_kernel, dim3(point_grid(n)), dim3(256), 0, s, qf_dev, keys, n, flags, rets);
}

void gqf_launch_point_query(QF* qf_dev, const uint64_t* keys, uint64_t n, uint8_t flags, uint64_t* counts,
                            int concurrent, unsigned long long* misses, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(gqf_point_query_kernel, dim3(point_grid(n)), dim3(256), 0, s, qf_dev, keys, n, flags, counts,
                     concurrent, misses);
}
