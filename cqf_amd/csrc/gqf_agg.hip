// gqf_agg.hip -- bulk_insert_reduce of duplicate-heavy batches without a
// sort of every key.
//
// The reference's bulk_insert_reduce sorts all keys (leaving `keys` holding
// the ascending hashes), reduces equal hashes to (hash, count) with
// thrust::reduce_by_key and inserts the counts (src/gqf.cu:3238-3314).  The
// table it builds depends only on the (hash, count) multiset, so a batch
// whose keys repeat heavily can count its hashes first and sort only the
// distinct ones:
//   1. gqf_agg_sample reads 4096 keys spread over the batch, counts the
//      distinct ones (the host skips this path when more than half are
//      distinct) and lists the most frequent (seen at least twice, up to
//      kHot) as hot keys;
//   2. gqf_agg_tiles: each workgroup counts 131072 keys.  A key equal to a
//      hot key adds to a per-thread register count (at Zipf(1.5) the eight
//      most frequent ids are three in four draws); the other keys of a wave
//      are queued in LDS and taken 64 at a time (so the wave's lanes are
//      full), hashed and counted in an 8192-entry LDS hash table.  The
//      workgroup's (hash, count) entries then go to staging buckets chosen
//      by hash, with one reserving atomic per bucket per workgroup;
//   3. gqf_agg_merge: one workgroup per bucket sums its entries by hash in
//      LDS and writes the distinct ones out as (index << width) | hash, plus
//      the count by index; an LSD sort of the low `width` bits (the bits the
//      sort orders for the hash width, see gqf_sort_width) orders them;
//   4. gqf_agg_extract / _scan_tiles / _positions split them into sorted
//      hashes and counts (the reduce_by_key outputs) and the exclusive
//      prefix sums of the counts;
//   5. gqf_agg_tile_starts / gqf_agg_expand write every hash count times
//      into `keys`: the ascending hashes the sort would have left there.
// Any table or bucket that runs out of room sets ctrl[2] and the batch
// takes the sort instead; `keys` is only read until step 5.
#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {

constexpr uint64_t kEmpty = ~0ULL - 1;  // no entry: keys and hashes here are below 2^40
constexpr int kThreads = 256;
constexpr int kSample = GQF_AGG_SAMPLE;
constexpr int kSampleLog = 13;
constexpr int kSampleCap = 1 << kSampleLog;
constexpr int kHot = GQF_AGG_HOT;
constexpr int kBuckets = GQF_AGG_BUCKETS;
constexpr uint32_t kBucketCap = GQF_AGG_BUCKET_CAP;
constexpr int kTileThreads = 1024;       // threads per workgroup of gqf_agg_tiles
constexpr int kTileWaves = kTileThreads / 64;
constexpr int kTileKeys = 131072;        // keys per workgroup of gqf_agg_tiles
constexpr int kLoadsPerRound = 8;        // keys each thread loads before counting them
constexpr int kLocalLog = 13;
constexpr int kLocalCap = 1 << kLocalLog;  // LDS table entries
constexpr int kCountBits = 24;             // count bits of an LDS entry and a staging entry
constexpr unsigned long long kLocalEmpty = ~0ULL;
static_assert(kTileKeys < (1 << kCountBits), "a workgroup's counts fit an entry");
constexpr int kLocalProbe = 32;
constexpr int kMergeThreads = 1024;
constexpr int kMergeLog = 13;
constexpr int kMergeCap = 1 << kMergeLog;  // distinct hashes a bucket may hold (3/4 of it used)
constexpr int kMergeProbe = 64;
constexpr int kScanItems = 16;
constexpr int kScanTile = kThreads * kScanItems;  // 4096 entries per scan tile
constexpr int kExpandTile = kThreads * 16;        // 4096 output keys per expand workgroup

__device__ inline uint32_t slot_of(uint64_t h, int log_cap) {
  return (uint32_t)((h * 0x9E3779B97F4A7C15ULL) >> (64 - log_cap));
}

__device__ inline uint32_t bucket_of(uint64_t h) {  // staging bucket of a hash
  return (uint32_t)(((h ^ (h >> 29)) * 0xBF58476D1CE4E5B9ULL) >> 55) & (kBuckets - 1);
}

// exclusive scan of one value per thread over a workgroup of T threads;
// returns the prefix, *total the sum
template <int T>
__device__ inline uint64_t block_scan(uint64_t v, uint64_t* total) {
  __shared__ uint64_t wsum[T / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t t = __shfl_up(incl, d, 64);
    if (lane >= d) incl += t;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  uint64_t before = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < T / 64; k++) {
    if (k < w) before += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *total = tot;
  return before + incl - v;
}

__device__ inline uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += (uint32_t)__shfl_xor((int)v, d, 64);
  return v;
}

}  // namespace

// ctrl[0] = distinct keys among kSample keys spread over the batch;
// ctrl[1..3] = 0 (distinct hashes, overflow flag, output cursor);
// fill[0 .. kBuckets) = 0; hot[0] = number of hot keys, hot[1 ..] = the
// masked keys seen at least twice in the sample, most frequent first
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_agg_sample(const uint64_t* __restrict__ keys, uint64_t n, uint64_t mask, uint32_t* __restrict__ ctrl,
               uint32_t* __restrict__ fill, uint64_t* __restrict__ hot) {
  __shared__ unsigned long long tab[kSampleCap];
  __shared__ uint32_t cnt[kSampleCap];
  __shared__ uint32_t distinct;
  __shared__ uint32_t best[kThreads / 64];
  for (int i = threadIdx.x; i < kSampleCap; i += kThreads) {
    tab[i] = kEmpty;
    cnt[i] = 0;
  }
  if (threadIdx.x == 0) distinct = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < kSample; i += kThreads) {
    const uint64_t k = keys[(uint64_t)i * (n / kSample)] & mask;
    uint32_t s = slot_of(k, kSampleLog);
    // at most kSample keys in kSampleCap slots: a free slot is always found
    for (int p = 0; p < kSampleCap; p++) {
      const unsigned long long prev = atomicCAS(&tab[s], kEmpty, k);
      if (prev == kEmpty) atomicAdd(&distinct, 1u);
      if (prev == kEmpty || prev == k) {
        atomicAdd(&cnt[s], 1u);
        break;
      }
      s = (s + 1) & (kSampleCap - 1);
    }
  }
  __syncthreads();
  // the kHot most frequent keys: each round takes the largest (count, slot)
  uint32_t nh = 0;
  for (int r = 0; r < kHot; r++) {
    uint32_t mx = 0;  // (count << kSampleLog) | slot; counts <= kSample < 2^13
    for (int i = threadIdx.x; i < kSampleCap; i += kThreads) {
      const uint32_t v = (cnt[i] << kSampleLog) | (uint32_t)i;
      mx = v > mx ? v : mx;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      const uint32_t o = (uint32_t)__shfl_xor((int)mx, d, 64);
      mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) best[threadIdx.x >> 6] = mx;
    __syncthreads();
    uint32_t b = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; w++) b = best[w] > b ? best[w] : b;
    __syncthreads();
    if ((b >> kSampleLog) < 2) break;  // uniform: every thread read the same best
    const uint32_t slot = b & (kSampleCap - 1);
    if (threadIdx.x == 0) {
      hot[1 + r] = tab[slot];
      cnt[slot] = 0;
    }
    nh++;
    __syncthreads();
  }
  for (int i = threadIdx.x; i < kBuckets; i += kThreads) fill[i] = 0;
  if (threadIdx.x == 0) {
    hot[0] = nh;
    ctrl[0] = distinct;
    ctrl[1] = 0;
    ctrl[2] = 0;
    ctrl[3] = 0;
  }
}

extern "C" __global__ void __launch_bounds__(kTileThreads)
gqf_agg_tiles(const uint64_t* __restrict__ keys, uint64_t n, uint64_t mask, int vshift,
              const uint64_t* __restrict__ hot, uint64_t* __restrict__ stage, uint32_t* __restrict__ fill,
              uint32_t* __restrict__ ctrl) {
  // LDS entry: (hash << kCountBits) | count, kLocalEmpty when free (hashes
  // are below 2^40 here and a workgroup counts fewer than 2^24 keys), so an
  // 8192-entry table is 64 KB; with the wave queues two workgroups fit a CU
  // the wave queues (a wave holds at most 63 + 64 keys) are dead once the
  // keys are counted; the bucket counts take their space then.  64 KB +
  // 16 KB: two workgroups per CU
  __shared__ unsigned long long tk[kLocalCap];
  __shared__ union {
    uint64_t queue[kTileWaves][127];
    struct {
      uint32_t bcount[kBuckets];
      uint32_t bbase[kBuckets];
    } b;
  } u;
  __shared__ uint32_t lfill;
  __shared__ uint32_t hotsum[kHot];
  static_assert(sizeof(u.b) <= sizeof(u.queue), "bucket counts fit the queues");
  for (int i = threadIdx.x; i < kLocalCap; i += kTileThreads) tk[i] = kLocalEmpty;
  if (threadIdx.x < kHot) hotsum[threadIdx.x] = 0;
  if (threadIdx.x == 0) lfill = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = lane ? (~0ULL >> (64 - lane)) : 0ULL;
  const uint32_t nhot = (uint32_t)hot[0];
  uint64_t hk[kHot];
  uint32_t hc[kHot];
#pragma unroll
  for (int j = 0; j < kHot; j++) {
    hk[j] = (uint32_t)j < nhot ? hot[1 + j] : kEmpty + 1;  // kEmpty + 1 = ~0: no masked key
    hc[j] = 0;
  }
  // a cold key: hashed and counted in the LDS table; once the table is half
  // full a hash not yet in it goes to its staging bucket alone
  auto count_cold = [&](uint64_t m) {
    const uint64_t h = gqf_hash64(m, mask) << vshift;
    uint32_t s = slot_of(h, kLocalLog);
    const unsigned long long tag = (unsigned long long)h << kCountBits;
    for (int p = 0; p < kLocalProbe; p++) {
      unsigned long long cur = *(volatile unsigned long long*)&tk[s];
      if (cur == kLocalEmpty) {
        if (*(volatile uint32_t*)&lfill >= kLocalCap / 2) break;
        cur = atomicCAS(&tk[s], kLocalEmpty, tag);
        if (cur == kLocalEmpty) {
          atomicAdd(&lfill, 1u);
          cur = tag;
        }
      }
      if ((cur >> kCountBits) == h) {
        atomicAdd(&tk[s], 1ULL);
        return;
      }
      s = (s + 1) & (kLocalCap - 1);
    }
    const uint32_t b = bucket_of(h);
    const uint32_t at = atomicAdd(&fill[b], 1u);
    if (at < kBucketCap) stage[(uint64_t)b * kBucketCap + at] = tag | 1ULL;
    else atomicOr(&ctrl[2], 1u);
  };
  uint32_t qlen = 0;  // the wave's queued cold keys (the same in every lane)
  for (uint64_t base = (uint64_t)blockIdx.x * kTileKeys; base < n; base += (uint64_t)gridDim.x * kTileKeys) {
    for (int r = 0; r < kTileKeys; r += kTileThreads * kLoadsPerRound) {
      uint64_t m[kLoadsPerRound];
#pragma unroll
      for (int k = 0; k < kLoadsPerRound; k++) {
        const uint64_t i = base + (uint64_t)(r + k * kTileThreads + (int)threadIdx.x);
        m[k] = i < n ? keys[i] & mask : kEmpty;
      }
#pragma unroll
      for (int k = 0; k < kLoadsPerRound; k++) {
        bool hit = false;
#pragma unroll
        for (int j = 0; j < kHot; j++) {
          const bool e = m[k] == hk[j];
          hc[j] += e ? 1u : 0u;
          hit = hit || e;
        }
        const bool cold = m[k] != kEmpty && !hit;
        const uint64_t cm = __ballot(cold);
        if (cold) u.queue[w][qlen + (uint32_t)__popcll(cm & lt)] = m[k];
        qlen += (uint32_t)__popcll(cm);
        if (qlen >= 64) {
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          const uint64_t x = u.queue[w][qlen - 64 + (uint32_t)lane];
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          qlen -= 64;
          count_cold(x);
        }
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if ((uint32_t)lane < qlen) count_cold(u.queue[w][lane]);
#pragma unroll
  for (int j = 0; j < kHot; j++) {
    if ((uint32_t)j < nhot) {
      const uint32_t t = wave_sum(hc[j]);
      if (lane == 0 && t) atomicAdd(&hotsum[j], t);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kBuckets; i += kTileThreads) u.b.bcount[i] = 0;
  __syncthreads();
  // this workgroup's entries into the staging buckets: count per bucket,
  // reserve each bucket's run with one atomic, write
  constexpr int kPer = kLocalCap / kTileThreads;
  uint64_t e[kPer + 1];
  uint32_t off[kPer + 1];
#pragma unroll
  for (int k = 0; k <= kPer; k++) {
    if (k < kPer) {
      e[k] = tk[threadIdx.x + k * kTileThreads];
    } else {
      const uint32_t j = threadIdx.x;
      e[k] = j < nhot && hotsum[j] ? (((gqf_hash64(hot[1 + j], mask) << vshift) << kCountBits) | hotsum[j]) : kLocalEmpty;
    }
    off[k] = e[k] != kLocalEmpty ? atomicAdd(&u.b.bcount[bucket_of(e[k] >> kCountBits)], 1u) : 0u;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kBuckets; b += kTileThreads)
    u.b.bbase[b] = u.b.bcount[b] ? atomicAdd(&fill[b], u.b.bcount[b]) : 0u;
  __syncthreads();
#pragma unroll
  for (int k = 0; k <= kPer; k++) {
    if (e[k] == kLocalEmpty) continue;
    const uint32_t b = bucket_of(e[k] >> kCountBits);
    const uint32_t at = u.b.bbase[b] + off[k];
    if (at < kBucketCap) stage[(uint64_t)b * kBucketCap + at] = e[k];
    else atomicOr(&ctrl[2], 1u);
  }
}

// one workgroup per staging bucket: sums the bucket's entries by hash in
// LDS, then writes its distinct hashes as comp[idx] = (idx << width) | hash,
// cnt[idx] = count, idx from the cursor ctrl[3] (one atomic per
// workgroup); ctrl[1] += distinct hashes; ctrl[2] set when the LDS table
// or cnt (cnt_cap entries) has no room
extern "C" __global__ void __launch_bounds__(kMergeThreads)
gqf_agg_merge(const uint64_t* __restrict__ stage, const uint32_t* __restrict__ fill, int width,
              uint64_t* __restrict__ comp, uint64_t* __restrict__ cnt, uint64_t cnt_cap, uint32_t* __restrict__ ctrl) {
  __shared__ unsigned long long mk[kMergeCap];
  __shared__ unsigned long long mc[kMergeCap];
  __shared__ uint32_t mfill;
  __shared__ uint32_t wg_base;
  for (int i = threadIdx.x; i < kMergeCap; i += kMergeThreads) {
    mk[i] = kEmpty;
    mc[i] = 0;
  }
  if (threadIdx.x == 0) mfill = 0;
  __syncthreads();
  const uint32_t b = blockIdx.x;
  const uint32_t nb = min(fill[b], kBucketCap);  // a fuller bucket has already set ctrl[2]
  const uint64_t cmask = (1ULL << kCountBits) - 1;
  for (uint32_t i = threadIdx.x; i < nb; i += kMergeThreads) {
    const uint64_t en = stage[(uint64_t)b * kBucketCap + i];
    const uint64_t h = en >> kCountBits;
    uint32_t s = slot_of(h, kMergeLog);
    bool done = false;
    for (int p = 0; p < kMergeProbe; p++) {
      unsigned long long cur = *(volatile unsigned long long*)&mk[s];
      if (cur == kEmpty) {
        if (*(volatile uint32_t*)&mfill >= kMergeCap / 4 * 3) break;
        cur = atomicCAS(&mk[s], kEmpty, h);
        if (cur == kEmpty) {
          atomicAdd(&mfill, 1u);
          cur = h;
        }
      }
      if (cur == h) {
        atomicAdd(&mc[s], en & cmask);
        done = true;
        break;
      }
      s = (s + 1) & (kMergeCap - 1);
    }
    if (!done) atomicOr(&ctrl[2], 1u);
  }
  __syncthreads();
  constexpr int kPer = kMergeCap / kMergeThreads;
  uint32_t mine = 0;
#pragma unroll
  for (int k = 0; k < kPer; k++) mine += mk[threadIdx.x * kPer + k] != kEmpty ? 1u : 0u;
  uint64_t total;
  const uint32_t before = (uint32_t)block_scan<kMergeThreads>(mine, &total);
  if (threadIdx.x == 0) {
    wg_base = total ? atomicAdd(&ctrl[3], (uint32_t)total) : 0u;
    if (total) atomicAdd(&ctrl[1], (uint32_t)total);
  }
  __syncthreads();
  uint64_t idx = (uint64_t)wg_base + before;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const uint32_t s = threadIdx.x * kPer + k;
    const uint64_t h = mk[s];
    if (h == kEmpty) continue;
    if (idx < cnt_cap) {
      comp[idx] = (idx << width) | h;
      cnt[idx] = mc[s];
    } else {
      atomicOr(&ctrl[2], 1u);
    }
    idx++;
  }
}

// sorted comp -> ukeys (hashes) and ucounts; tile_sums[t] = sum of the counts of scan tile t
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_agg_extract(const uint64_t* __restrict__ comp, uint64_t m, int width, const uint64_t* __restrict__ cnt,
                uint64_t* __restrict__ ukeys, uint64_t* __restrict__ ucounts, uint64_t* __restrict__ tile_sums) {
  const uint64_t hmask = (1ULL << width) - 1;
  uint64_t sum = 0;
  for (int k = 0; k < kScanItems; k++) {
    const uint64_t i = (uint64_t)blockIdx.x * kScanTile + (uint64_t)k * kThreads + threadIdx.x;
    if (i < m) {
      const uint64_t c = comp[i];
      const uint64_t v = cnt[c >> width];
      ukeys[i] = c & hmask;
      ucounts[i] = v;
      sum += v;
    }
  }
  uint64_t total;
  (void)block_scan<kThreads>(sum, &total);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// exclusive scan in place of ntiles (<= kScanTile) tile sums, one workgroup
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_agg_scan_tiles(uint64_t* __restrict__ tile_sums, uint32_t ntiles) {
  uint64_t v[kScanItems];
  uint64_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const uint32_t i = threadIdx.x * kScanItems + k;
    v[k] = i < ntiles ? tile_sums[i] : 0;
    sum += v[k];
  }
  uint64_t total;
  uint64_t run = block_scan<kThreads>(sum, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const uint32_t i = threadIdx.x * kScanItems + k;
    if (i < ntiles) tile_sums[i] = run;
    run += v[k];
  }
}

// upos[i] = sum of ucounts[0 .. i) for i < m, upos[m] = n
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_agg_positions(const uint64_t* __restrict__ ucounts, uint64_t m, uint64_t n, const uint64_t* __restrict__ tile_sums,
                  uint64_t* __restrict__ upos) {
  // thread t takes kScanItems consecutive entries of the tile
  const uint64_t first = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint64_t v[kScanItems];
  uint64_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    v[k] = first + k < m ? ucounts[first + k] : 0;
    sum += v[k];
  }
  uint64_t total;
  uint64_t run = tile_sums[blockIdx.x] + block_scan<kThreads>(sum, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    if (first + k < m) upos[first + k] = run;
    run += v[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) upos[m] = n;
}

// start[t] = the distinct index whose keys cover output position
// min(t * kExpandTile, n - 1), for t in [0, ntiles]
extern "C" __global__ void gqf_agg_tile_starts(const uint64_t* __restrict__ upos, uint64_t m, uint64_t n,
                                               uint32_t ntiles, uint32_t* __restrict__ start) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  uint64_t j = (uint64_t)t * kExpandTile;
  if (j > n - 1) j = n - 1;
  uint64_t lo = 0, hi = m;  // upos[lo] <= j < upos[hi]
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (upos[mid] <= j) lo = mid;
    else hi = mid;
  }
  start[t] = (uint32_t)lo;
}

// keys[j] = ukeys[i] for upos[i] <= j < upos[i + 1]: one workgroup per
// kExpandTile outputs, the tile's boundaries staged in LDS
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_agg_expand(const uint64_t* __restrict__ ukeys, const uint64_t* __restrict__ upos, uint64_t n,
               const uint32_t* __restrict__ start, uint64_t* __restrict__ keys) {
  __shared__ uint64_t bnd[kExpandTile + 2];
  const uint64_t b = (uint64_t)blockIdx.x * kExpandTile;
  const uint32_t i_lo = start[blockIdx.x];
  // the tile's last output lies in a distinct index <= start[t + 1], so
  // boundaries i_lo .. start[t + 1] + 1 bracket every output of the tile
  const uint32_t d = start[blockIdx.x + 1] - i_lo + 1;  // candidate indices, <= kExpandTile + 1
  for (uint32_t k = threadIdx.x; k <= d; k += kThreads) bnd[k] = upos[i_lo + k];
  __syncthreads();
  uint32_t lo = 0;
#pragma unroll 4
  for (int k = 0; k < 16; k++) {
    const uint64_t j = b + (uint64_t)k * kThreads + threadIdx.x;
    if (j >= n) break;
    // largest idx in [lo, d) with bnd[idx] <= j (outputs ascend per thread)
    uint32_t hi = d;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (bnd[mid] <= j) lo = mid;
      else hi = mid;
    }
    keys[j] = ukeys[i_lo + lo];
  }
}

// ---------------------------------------------------------------------------
// launchers

void gqf_launch_agg_sample(const uint64_t* keys, uint64_t n, uint64_t mask, uint32_t* ctrl, uint32_t* fill,
                           uint64_t* hot, hipStream_t s) {
  hipLaunchKernelGGL(gqf_agg_sample, dim3(1), dim3(kThreads), 0, s, keys, n, mask, ctrl, fill, hot);
}

void gqf_launch_agg_tiles(const uint64_t* keys, uint64_t n, uint64_t mask, int vshift, const uint64_t* hot,
                          uint64_t* stage, uint32_t* fill, uint32_t* ctrl, hipStream_t s) {
  // one workgroup per tile (131072 keys), two resident per CU
  const uint64_t g = (n + kTileKeys - 1) / kTileKeys;
  gqf_prof_begin("reduce_count", s);
  hipLaunchKernelGGL(gqf_agg_tiles, dim3((uint32_t)g), dim3(kTileThreads), 0, s, keys, n, mask, vshift, hot, stage,
                     fill, ctrl);
  gqf_prof_end("reduce_count", s);
}

void gqf_launch_agg_merge(const uint64_t* stage, const uint32_t* fill, int width, uint64_t* comp, uint64_t* cnt,
                          uint64_t cnt_cap, uint32_t* ctrl, hipStream_t s) {
  gqf_prof_begin("reduce_merge", s);
  hipLaunchKernelGGL(gqf_agg_merge, dim3(kBuckets), dim3(kMergeThreads), 0, s, stage, fill, width, comp, cnt, cnt_cap,
                     ctrl);
  gqf_prof_end("reduce_merge", s);
}

void gqf_launch_agg_finish(uint64_t* keys, uint64_t n, uint64_t m, int width, uint64_t* comp, uint64_t* cnt,
                           uint64_t* tmp, uint32_t* hist, uint32_t* totals, uint64_t* tile_sums, uint32_t* starts,
                           uint64_t* ukeys, uint64_t* ucounts, uint64_t* upos, hipStream_t s) {
  gqf_prof_begin("reduce_finish", s);
  // distinct hashes: the sort orders exactly the low `width` bits (its
  // shape for `width` is the one that defined it) and carries the index along
  gqf_launch_sort(comp, tmp, m, 0, 0, 0, width, 0, hist, totals, s);
  const uint32_t stiles = (uint32_t)((m + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(gqf_agg_extract, dim3(stiles), dim3(kThreads), 0, s, (const uint64_t*)comp, m, width,
                     (const uint64_t*)cnt, ukeys, ucounts, tile_sums);
  hipLaunchKernelGGL(gqf_agg_scan_tiles, dim3(1), dim3(kThreads), 0, s, tile_sums, stiles);
  hipLaunchKernelGGL(gqf_agg_positions, dim3(stiles), dim3(kThreads), 0, s, (const uint64_t*)ucounts, m, n,
                     (const uint64_t*)tile_sums, upos);
  const uint32_t etiles = (uint32_t)((n + kExpandTile - 1) / kExpandTile);
  hipLaunchKernelGGL(gqf_agg_tile_starts, dim3((etiles + 1 + 255) / 256), dim3(256), 0, s, (const uint64_t*)upos, m,
                     n, etiles, starts);
  hipLaunchKernelGGL(gqf_agg_expand, dim3(etiles), dim3(kThreads), 0, s, (const uint64_t*)ukeys,
                     (const uint64_t*)upos, n, (const uint32_t*)starts, keys);
  gqf_prof_end("reduce_finish", s);
}
