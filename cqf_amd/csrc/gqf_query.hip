// gqf_query.hip -- bulk_get / bulk_get_misses probes.
//
// Reference: bulk_get_kernel -> qf_count_key_value (src/gqf.cu:3386-3401,
// 3479-3518): hash the key, return 0 if its quotient is unoccupied, else
// walk the counters of the quotient's run and return the count of the first
// counter whose remainder matches.
//
// Here the run bounds come straight from the home block's header: with
// x = end of the runs of earlier blocks (the exact 32-bit offset the device
// layout keeps), the k-th occupied quotient of the block owns the k-th run
// end after x, so the run start and end are two selects over the runends
// words -- no run_end recursion.
#include "gqf_internal.h"
#include "gqf_launch.h"

// Count of `hash` (< range) in the table (count_from_line, given the home
// block's line from load_home).  `row` is the calling thread's 17-word LDS
// scratch row.
//
// One round of loads fetches the home block's whole line (header and the
// 64 slots, six 16-byte loads to one 128-byte line).  When the quotient's
// run lies inside the home block -- ~94% of occupied probes at 90% load --
// the run is decoded from that copy (through the LDS row, since lanes index
// it at different bytes).  Otherwise the run ends of later blocks are read,
// then the slots of the block the run starts in (four 16-byte loads, into
// the row); only bytes of a run that also leaves that block are read one by
// one.  A probe is latency bound (one dependent round after the key), so
// bulk_get keeps two keys' lines in flight per thread: at 90% load, 2^28
// slots, present keys 6.1 -> 5.3 ms per 2.4e8 (the row path for runs pushed
// out of the home block is part of that), absent keys unchanged.
struct HomeLine {
  uint4 h0, h1, s0, s1, s2, s3;
};

__device__ __forceinline__ HomeLine load_home(const QF& qf, uint64_t hash) {
  const GQF_GLOBAL uint4* p = gqf_block_ptr(qf, hash >> 14);
  HomeLine l;
  l.h0 = p[0];
  l.h1 = p[1];
  l.s0 = p[4]; l.s1 = p[5]; l.s2 = p[6]; l.s3 = p[7];
  return l;
}

// the line's registers taken as read here: every load of the line is issued
// before the decode branches (the compiler would otherwise sink the slot
// loads behind the header test, a second round), and both keys' lines of a
// step are in flight before either is decoded
__device__ __forceinline__ void keep4(const uint4& v) { asm volatile("" ::"v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w)); }
__device__ __forceinline__ void keep_line(const HomeLine& l) {
  keep4(l.h0); keep4(l.h1); keep4(l.s0); keep4(l.s1); keep4(l.s2); keep4(l.s3);
}

__device__ uint64_t count_from_line(const QF& qf, uint64_t hash, const HomeLine& l, uint32_t* row) {
  const uint64_t q = hash >> 8;
  const uint64_t rem = hash & 0xFF;
  const uint64_t b = q >> 6;
  const uint32_t bit = (uint32_t)(q & 63);
  const uint4 h0 = l.h0, h1 = l.h1, s0 = l.s0, s1 = l.s1, s2 = l.s2, s3 = l.s3;
  const uint64_t occ = (uint64_t)h0.x | ((uint64_t)h0.y << 32);
  const uint64_t runends = (uint64_t)h0.z | ((uint64_t)h0.w << 32);
  if (!((occ >> bit) & 1ULL)) return 0;

  // device offsets are exact: the end of the runs with quotient < 64 b
  const int64_t xb = (int64_t)(b * 64) - 1 + (int64_t)h1.x;
  const int rank = __popcll(occ & ((2ULL << bit) - 1ULL));  // >= 1

  // the (rank-2)'th and (rank-1)'th run ends at or after xb + 1
  int64_t prev_end = xb, end = -1;
  uint64_t pos = (uint64_t)(xb + 1);
  uint64_t bb = pos >> 6;
  uint64_t word = (bb == b ? runends : gqf_load_runends(qf, bb)) & (~0ULL << (pos & 63));
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
    word = gqf_load_runends(qf, bb);
  }
  int64_t start = prev_end + 1;
  if (start < (int64_t)q) start = (int64_t)q;
  const uint32_t len = (uint32_t)(end - start + 1);
  // the slots of the block the run starts in go through the LDS row (lanes
  // index it at different bytes): the home block's from the loaded line, a
  // later block's (the run pushed out of its home block) by one more round
  const uint64_t sb = (uint64_t)start >> 6;
  uint4 t0 = s0, t1 = s1, t2 = s2, t3 = s3;
  if (sb != b) {
    const GQF_GLOBAL uint4* pn = gqf_block_ptr(qf, sb);
    t0 = pn[4]; t1 = pn[5]; t2 = pn[6]; t3 = pn[7];
  }
  row[0] = t0.x; row[1] = t0.y; row[2] = t0.z; row[3] = t0.w;
  row[4] = t1.x; row[5] = t1.y; row[6] = t1.z; row[7] = t1.w;
  row[8] = t2.x; row[9] = t2.y; row[10] = t2.z; row[11] = t2.w;
  row[12] = t3.x; row[13] = t3.y; row[14] = t3.z; row[15] = t3.w;
  const uint8_t* bytes = (const uint8_t*)row + (start & 63);
  uint32_t idx = 0;
  if ((uint64_t)end >> 6 == sb) {
    auto slot = [&](uint32_t k) -> uint64_t { return bytes[k]; };
    do {
      uint64_t r, c;
      uint32_t e = gqf_decode_counter(slot, len, idx, &r, &c);
      if (r == rem) return c;
      idx = e + 1;
    } while (idx < len);
    return 0;
  }
  // a run that leaves its start block: the bytes past that block one by one
  const uint32_t in_row = 64u - (uint32_t)(start & 63);
  const GQF_GLOBAL uint8_t* blocks = gqf_gptr(qf.blocks);
  auto slot = [&](uint32_t k) -> uint64_t {
    return k < in_row ? (uint64_t)bytes[k] : (uint64_t)blocks[gqf_slot_addr((uint64_t)start + k)];
  };
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
  __shared__ uint32_t rows[256 * 17];
  const QF qf = *qfp;
  uint32_t* row = rows + 17 * threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // keys i and i + stride per step; keys and results stream through once:
  // non-temporal, so they do not displace table lines from the caches
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += 2 * stride) {
    const uint64_t i1 = i + stride;
    const bool two = i1 < n;
    const uint64_t h0 = gqf_key_hash(qf, __builtin_nontemporal_load(&vals[i]));
    const uint64_t h1 = gqf_key_hash(qf, two ? __builtin_nontemporal_load(&vals[i1]) : 0);
    const HomeLine l0 = load_home(qf, h0);
    const HomeLine l1 = load_home(qf, two ? h1 : h0);
    keep_line(l0);
    keep_line(l1);
    __builtin_nontemporal_store(count_from_line(qf, h0, l0, row), &returns[i]);
    if (two) __builtin_nontemporal_store(count_from_line(qf, h1, l1, row), &returns[i1]);
  }
}

// number of keys whose count is < 1 (src/gqf.cu:3370-3384, 3404-3419)
extern "C" __global__ void __launch_bounds__(256)
gqf_bulk_misses_kernel(const QF* __restrict__ qfp, const uint64_t* __restrict__ vals, uint64_t n,
                       unsigned long long* __restrict__ misses) {
  __shared__ uint32_t rows[256 * 17];
  const QF qf = *qfp;
  uint32_t* row = rows + 17 * threadIdx.x;
  uint32_t mine = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // keys i and i + stride per step, both lines in flight (as bulk_get)
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += 2 * stride) {
    const uint64_t i1 = i + stride;
    const bool two = i1 < n;
    const uint64_t h0 = gqf_key_hash(qf, __builtin_nontemporal_load(&vals[i]));
    const uint64_t h1 = two ? gqf_key_hash(qf, __builtin_nontemporal_load(&vals[i1])) : h0;
    const HomeLine l0 = load_home(qf, h0);
    const HomeLine l1 = load_home(qf, h1);
    keep_line(l0);
    keep_line(l1);
    mine += count_from_line(qf, h0, l0, row) < 1;
    if (two) mine += count_from_line(qf, h1, l1, row) < 1;
  }
  // wave reduction, one atomic per wave
  for (int d = 32; d >= 1; d >>= 1) mine += __shfl_down(mine, d, 64);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(misses, (unsigned long long)mine);
}

// One launch over the keys of a filter group: key i belongs to the filter
// whose segment [seg[f], seg[f+1]) holds i (the last such f, so empty
// segments are passed over).  Workgroup w takes the contiguous keys
// [w chunk, (w + 1) chunk), chunk a multiple of 512, two per thread and
// step (i and i + 256, both lines in flight as in bulk_get).  The filter
// index only moves forward inside a chunk: one test of the next segment
// start per key, a binary search only when it is passed.

// the filter holding key i, given a filter lo with seg[lo] <= i
__device__ inline uint32_t group_filter_from(const gqf_group_dev& g, uint64_t i, uint32_t lo) {
  if (lo + 1 >= g.nf || g.seg[lo + 1] > i) return lo;
  uint32_t hi = g.nf;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (g.seg[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

extern "C" __global__ void __launch_bounds__(256)
gqf_group_get_kernel(gqf_group_dev g, const uint64_t* __restrict__ vals, uint64_t n, uint64_t chunk,
                     uint64_t* __restrict__ returns) {
  __shared__ uint32_t rows[256 * 17];
  uint32_t* row = rows + 17 * threadIdx.x;
  const uint64_t c0 = (uint64_t)blockIdx.x * chunk;
  if (c0 >= n) return;
  const uint64_t c1 = n - c0 > chunk ? c0 + chunk : n;
  uint32_t lo = group_filter_from(g, c0, 0);  // seg[0] = 0
  for (uint64_t i = c0 + threadIdx.x; i < c1; i += 512) {
    const uint64_t i1 = i + 256;
    const bool two = i1 < c1;
    const uint32_t f0 = group_filter_from(g, i, lo);
    const uint32_t f1 = two ? group_filter_from(g, i1, f0) : f0;
    const QF& d0 = g.descs[f0];
    const QF& d1 = g.descs[f1];
    QF q0, q1;
    q0.blocks = d0.blocks;
    q0.nblocks = d0.nblocks;
    q1.blocks = d1.blocks;
    q1.nblocks = d1.nblocks;
    const uint64_t h0 = gqf_key_hash(d0, vals[i]);
    const uint64_t h1 = two ? gqf_key_hash(d1, vals[i1]) : h0;
    const HomeLine l0 = load_home(q0, h0);
    const HomeLine l1 = load_home(q1, h1);
    keep_line(l0);
    keep_line(l1);
    returns[i] = count_from_line(q0, h0, l0, row);
    if (two) returns[i1] = count_from_line(q1, h1, l1, row);
    lo = f1;
  }
}

// ---------------------------------------------------------------------------
// launchers

static inline uint32_t probe_grid(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  // grid-stride beyond 65536 workgroups (bulk_get of 2.1e8 keys, G keys/s
  // by workgroup cap: 8K 41.3, 16K 41.9, 32K 42.2, 64K 42.4)
  const uint64_t cap = 256ULL * 256;
  return (uint32_t)(g < cap ? (g ? g : 1) : cap);
}

void gqf_launch_bulk_get(const QF* qf_dev, const uint64_t* vals, uint64_t n, uint64_t* returns, hipStream_t s) {
  if (n == 0) return;
  gqf_prof_begin("bulk_get", s);
  // two keys per thread and step: half the threads
  hipLaunchKernelGGL(gqf_bulk_get_kernel, dim3(probe_grid((n + 1) / 2)), dim3(256), 0, s, qf_dev, vals, n, returns);
  gqf_prof_end("bulk_get", s);
}

void gqf_launch_misses(const QF* qf_dev, const uint64_t* vals, uint64_t n, unsigned long long* misses,
                       hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(gqf_bulk_misses_kernel, dim3(probe_grid((n + 1) / 2)), dim3(256), 0, s, qf_dev, vals, n, misses);
}

// ---------------------------------------------------------------------------
// Sequenced batches: one filter's requests in arrival order.  Mutation keys
// become tagged slot hashes (gqf_seq_tag), are sorted by hash alone (stable,
// so a hash's items stay in request order) and rebuilt in one pass; the
// batch's queries are counted before the mutation and then corrected by
// gqf_seq_adjust with the mutation items of the same hash that precede them.

namespace {

// the request boundaries into LDS, then request of key i = first r with end[r] > i
struct SeqReqsLds {
  uint32_t end[QF_SEQ_MAX_REQUESTS];
  uint32_t tag[QF_SEQ_MAX_REQUESTS];
};

__device__ inline void seq_reqs_to_lds(const gqf_seq_reqs& reqs, SeqReqsLds& l) {
  for (uint32_t r = threadIdx.x; r < reqs.nreq; r += blockDim.x) {
    l.end[r] = reqs.end[r];
    l.tag[r] = reqs.tag[r];
  }
  __syncthreads();
}

__device__ inline uint32_t seq_request_of(const SeqReqsLds& l, uint32_t nreq, uint64_t i) {
  uint32_t lo = 0, hi = nreq;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)l.end[mid] > i) hi = mid; else lo = mid + 1;
  }
  return lo < nreq ? lo : nreq - 1;
}

}  // namespace

extern "C" __global__ void __launch_bounds__(256)
gqf_seq_tag_kernel(const QF* __restrict__ qfp, uint64_t* __restrict__ keys, uint64_t n, gqf_seq_reqs reqs) {
  __shared__ SeqReqsLds l;
  seq_reqs_to_lds(reqs, l);
  const QF qf = *qfp;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t t = l.tag[seq_request_of(l, reqs.nreq, i)];
    keys[i] = gqf_key_hash(qf, keys[i]) | ((uint64_t)(t >> 1) << GQF_SEQ_SHIFT) | ((uint64_t)(t & 1u) << 63);
  }
}

// Query i (of request with sequence number qs) replays, on its count before
// the mutation, the items of its hash from requests before qs, in order:
// +1 for an insert, -1 floored at 0 for a delete (the reference's qf_insert
// and qf_remove on one hash, src/gqf.cu:1523-1933).
extern "C" __global__ void __launch_bounds__(256)
gqf_seq_adjust_kernel(const QF* __restrict__ qfp, const uint64_t* __restrict__ keys, uint64_t n, gqf_seq_reqs reqs,
                      uint64_t* __restrict__ counts, const uint64_t* __restrict__ mut, uint64_t n_mut) {
  __shared__ SeqReqsLds l;
  seq_reqs_to_lds(reqs, l);
  const QF qf = *qfp;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t qs = l.tag[seq_request_of(l, reqs.nreq, i)];
    const uint64_t h = gqf_key_hash(qf, keys[i]);
    uint64_t lo = 0, hi = n_mut;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if ((mut[mid] & GQF_SEQ_HASH_MASK) < h) lo = mid + 1; else hi = mid;
    }
    if (lo == n_mut || (mut[lo] & GQF_SEQ_HASH_MASK) != h) continue;
    uint64_t c = counts[i];
    for (uint64_t j = lo; j < n_mut; j++) {
      const uint64_t t = mut[j];
      if ((t & GQF_SEQ_HASH_MASK) != h) break;
      if (((t >> GQF_SEQ_SHIFT) & 0x7FFFu) >= qs) break;  // later requests of this hash follow in order
      c = (t >> 63) ? (c ? c - 1 : 0) : c + 1;
    }
    counts[i] = c;
  }
}

void gqf_launch_seq_tag(const QF* qf_dev, uint64_t* keys, uint64_t n, const gqf_seq_reqs& reqs, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(gqf_seq_tag_kernel, dim3(probe_grid(n)), dim3(256), 0, s, qf_dev, keys, n, reqs);
}

void gqf_launch_seq_adjust(const QF* qf_dev, const uint64_t* keys, uint64_t n, const gqf_seq_reqs& reqs,
                           uint64_t* counts, const uint64_t* mut, uint64_t n_mut, hipStream_t s) {
  if (n == 0 || n_mut == 0) return;
  hipLaunchKernelGGL(gqf_seq_adjust_kernel, dim3(probe_grid(n)), dim3(256), 0, s, qf_dev, keys, n, reqs, counts, mut,
                     n_mut);
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

void gqf_launch_group_get(const gqf_group_dev& g, const uint64_t* vals, uint64_t n, uint64_t* returns,
                          hipStream_t s) {
  if (n == 0) return;
  gqf_prof_begin("group_get", s);
  // contiguous chunks of whole 512-key steps, one per workgroup
  const uint32_t grid = probe_grid((n + 1) / 2);
  const uint64_t chunk = ((n + grid - 1) / grid + 511) / 512 * 512;
  hipLaunchKernelGGL(gqf_group_get_kernel, dim3(grid), dim3(256), 0, s, g, vals, n, chunk, returns);
  gqf_prof_end("group_get", s);
}

