// gqf_query.hip -- bulk_get / bulk_get_misses probes.
//
// Reference: bulk_get_kernel -> qf_count_key_value (src/gqf.cu:3386-3401,
// 3479-3518): hash the key, return 0 if its quotient is unoccupied, else
// walk the counters of the quotient's run and return the count of the first
// counter whose remainder matches.
//
// Here the run bounds come straight from the home block's header: with
// x = end of the runs of earlier blocks (offset byte), the k-th occupied
// quotient of the block owns the k-th run end after x, so the run start and
// end are two selects over the runends words -- no run_end recursion unless
// the offset is saturated.
#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {

__device__ int64_t next_runend_g(const QF& qf, uint64_t pos) {
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

// end of the runs with quotient < 64 b (at least 64 b - 1), for a block
// whose stored offset is saturated (src/gqf.cu:619-630)
__device__ int64_t true_x_g(const QF& qf, uint64_t b) {
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
      int64_t en = next_runend_g(qf, (uint64_t)st);
      if (en < 0) return (int64_t)(qf.nblocks * 64);
      x = en;
    }
  }
  int64_t lo = (int64_t)(b * 64) - 1;
  return x > lo ? x : lo;
}

}  // namespace

// Count of `hash` (< range) in the table.
__device__ uint64_t gqf_count_hash(const QF& qf, uint64_t hash) {
  const uint64_t q = hash >> 8;
  const uint64_t rem = hash & 0xFF;
  const uint64_t b = q >> 6;
  const uint32_t bit = (uint32_t)(q & 63);
  const gqf_header hd = gqf_load_header(qf.blocks, b);
  if (!((hd.occupieds >> bit) & 1ULL)) return 0;

  const int64_t xb = hd.offset < GQF_OFFSET_SAT ? (int64_t)(b * 64) - 1 + (int64_t)hd.offset
                                                : true_x_g(qf, b);
  const int rank = __popcll(hd.occupieds & ((2ULL << bit) - 1ULL));  // >= 1

  // the (rank-2)'th and (rank-1)'th run ends at or after xb + 1
  int64_t prev_end = xb, end = -1;
  uint64_t pos = (uint64_t)(xb + 1);
  uint64_t bb = pos >> 6;
  uint64_t word = (bb == b ? hd.runends : gqf_load_header(qf.blocks, bb).runends) & (~0ULL << (pos & 63));
  int seen = 0;
  for (;;) {
    int c = __popcll(word);
    if (rank >= 2 && rank - 2 >= seen && rank - 2 < seen + c)
      prev_end = (int64_t)(bb * 64) + gqf_select64(word, rank - 2 - seen);
    if (rank - 1 < seen + c) {
      end = (int64_t)(bb * 64) + gqf_select64(word, rank - 1 - seen);
      break;
    }
    seen += c;
    bb++;
    if (bb >= qf.nblocks) return 0;
    word = gqf_load_header(qf.blocks, bb).runends;
  }
  int64_t start = prev_end + 1;
  if (start < (int64_t)q) start = (int64_t)q;
  const uint32_t len = (uint32_t)(end - start + 1);
  const GQF_GLOBAL uint8_t* blocks = gqf_gptr(qf.blocks);
  auto slot = [&](uint32_t k) -> uint64_t { return blocks[gqf_slot_addr((uint64_t)start + k)]; };
  uint32_t idx = 0;
  do {
    uint64_t r, c;
    uint32_t e = gqf_decode_counter(slot, len, idx, &r, &c);
    if (r == rem) return c;
    idx = e + 1;
  } while (idx < len);
  return 0;
}

extern "C" __global__ void __launch_bounds__(256)
gqf_bulk_get_kernel(const QF* __restrict__ qfp, const uint64_t* __restrict__ vals, uint64_t n,
                    uint64_t* __restrict__ returns) {
  const QF qf = *qfp;
  const uint64_t mask = qf.range - 1;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    returns[i] = gqf_count_hash(qf, gqf_hash64(vals[i], mask));
  }
}

// number of keys whose count is < 1 (src/gqf.cu:3370-3384, 3404-3419)
extern "C" __global__ void __launch_bounds__(256)
gqf_bulk_misses_kernel(const QF* __restrict__ qfp, const uint64_t* __restrict__ vals, uint64_t n,
                       unsigned long long* __restrict__ misses) {
  const QF qf = *qfp;
  const uint64_t mask = qf.range - 1;
  uint32_t mine = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    mine += gqf_count_hash(qf, gqf_hash64(vals[i], mask)) < 1;
  }
  // wave reduction, one atomic per wave
  for (int d = 32; d >= 1; d >>= 1) mine += __shfl_down(mine, d, 64);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(misses, (unsigned long long)mine);
}

// ---------------------------------------------------------------------------
// launchers

static inline uint32_t probe_grid(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  const uint64_t cap = 256ULL * 64;  // grid-stride beyond ~16 waves per CU
  return (uint32_t)(g < cap ? (g ? g : 1) : cap);
}

void gqf_launch_bulk_get(const QF* qf_dev, const uint64_t* vals, uint64_t n, uint64_t* returns, hipStream_t s) {
  if (n == 0) return;
  gqf_prof_begin("bulk_get", s);
  hipLaunchKernelGGL(gqf_bulk_get_kernel, dim3(probe_grid(n)), dim3(256), 0, s, qf_dev, vals, n, returns);
  gqf_prof_end("bulk_get", s);
}

void gqf_launch_misses(const QF* qf_dev, const uint64_t* vals, uint64_t n, unsigned long long* misses,
                       hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(gqf_bulk_misses_kernel, dim3(probe_grid(n)), dim3(256), 0, s, qf_dev, vals, n, misses);
}

// ---------------------------------------------------------------------------
// synthetic key streams for benchmarks and tests: out[i] = splitmix64 output
// number (skip + i) of the stream seeded with `seed` (Steele, Lea & Flood
// 2014; the state advances by the golden gamma, then two xor-shift-multiply
// rounds)

extern "C" __global__ void __launch_bounds__(256)
gqf_splitmix64_kernel(uint64_t* __restrict__ out, uint64_t n, uint64_t seed, uint64_t skip) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + (skip + i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    out[i] = z ^ (z >> 31);
  }
}

void gqf_launch_splitmix64(uint64_t* out, uint64_t n, uint64_t seed, uint64_t skip, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(gqf_splitmix64_kernel, dim3(probe_grid(n)), dim3(256), 0, s, out, n, seed, skip);
}
