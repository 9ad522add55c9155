// A reference-style caller of libgqf, written as the reference's own
// wrappers are (include/gqf_wrapper.cuh:36-86, include/point_wrapper.cuh:25-79)
// after the header swap of the HIP build: the wrappers keep their calls and
// the point_bulk_insert kernel launch compiles unchanged against
// include/gqf_device.h.  A user kernel calls point_insert directly, as an
// application with its own k-mer kernel would.  Test infrastructure: the
// extern "C" entry points only let the tests drive it through ctypes.
#include <hip/hip_runtime.h>

#include "gqf.h"         // was: gqf.cuh + gqf_int.cuh
#include "gqf_device.h"  // point_insert, point_bulk_insert

// ---- include/gqf_wrapper.cuh, HIP build -----------------------------------
static QF* g_quotient_filter;

extern inline int gqf_init(uint64_t nbits, uint64_t, uint64_t) {
  qf_malloc_device(&g_quotient_filter, (int)nbits, true);
  return 0;
}
extern inline int gqf_destroy() {
  qf_destroy_device(g_quotient_filter);
  return 0;
}
extern inline int gqf_bulk_insert(uint64_t* vals, uint64_t count) {
  bulk_insert(g_quotient_filter, count, vals, QF_NO_LOCK);
  (void)hipDeviceSynchronize();
  return 0;
}
extern inline uint64_t gqf_bulk_get(uint64_t* vals, uint64_t count) {
  return bulk_get_misses_wrapper(g_quotient_filter, vals, count);
}
extern inline void gqf_bulk_delete(uint64_t* vals, uint64_t count) {
  bulk_delete(g_quotient_filter, count, vals, QF_NO_LOCK);
  (void)hipDeviceSynchronize();
}

// ---- include/point_wrapper.cuh, HIP build ---------------------------------
static QF* point_quotient_filter;

extern inline int point_init(uint64_t nbits, uint64_t, uint64_t) {
  qf_malloc_device(&point_quotient_filter, (int)nbits, false);
  return 0;
}
extern inline int point_destroy() {
  qf_destroy_device(point_quotient_filter);
  return 0;
}
extern inline int point_bulk_insert_wrapper(uint64_t* vals, uint64_t count) {
  point_bulk_insert<<<(count - 1) / 32 + 1, 32>>>(point_quotient_filter, vals, count);
  (void)hipDeviceSynchronize();
  return 0;
}
extern inline uint64_t point_bulk_get(uint64_t* vals, uint64_t count) {
  return point_get_wrapper(point_quotient_filter, vals, count);
}

// ---- an application kernel calling the device API -------------------------
__global__ void count_kmers(QF* qf, const uint64_t* kmers, uint64_t n) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i < n) (void)point_insert(qf, kmers[i], 0, 0);
}

// ---- ctypes entry points ---------------------------------------------------
extern "C" {
QF* dropin_gqf_init(int nbits) {
  gqf_init((uint64_t)nbits, 0, 0);
  return g_quotient_filter;
}
void dropin_gqf_destroy() { gqf_destroy(); }
void dropin_gqf_bulk_insert(uint64_t* vals, uint64_t n) { gqf_bulk_insert(vals, n); }
uint64_t dropin_gqf_bulk_get(uint64_t* vals, uint64_t n) { return gqf_bulk_get(vals, n); }
void dropin_gqf_bulk_delete(uint64_t* vals, uint64_t n) { gqf_bulk_delete(vals, n); }
QF* dropin_point_init(int nbits) {
  point_init((uint64_t)nbits, 0, 0);
  return point_quotient_filter;
}
void dropin_point_destroy() { point_destroy(); }
void dropin_point_bulk_insert(uint64_t* vals, uint64_t n) { point_bulk_insert_wrapper(vals, n); }
uint64_t dropin_point_bulk_get(uint64_t* vals, uint64_t n) { return point_bulk_get(vals, n); }
void dropin_count_kmers(QF* qf, const uint64_t* kmers, uint64_t n) {
  if (n == 0) return;
  count_kmers<<<(unsigned)((n + 255) / 256), 256>>>(qf, kmers, n);
  (void)hipDeviceSynchronize();
}
}
