// gqf_agg.hip -- bulk_insert_reduce of duplicate-heavy batches without a
// sort of every key.
//
// The reference's bulk_insert_reduce sorts all keys (leaving `keys` holding
// the ascending hashes), reduces equal hashes to (hash, count) with
// thrust::reduce_by_key and inserts the counts (src/gqf.cu:3238-3314).  The
// table it builds depends only on the (hash, count) multiset, so a batch
// whose keys repeat heavily can count its hashes first and sort only the
// distinct ones:
//   1. gqf_agg_sample hashes 4096 keys spread over the batch and counts
//      the distinct hashes among them (the host skips this path when more
//      than half are distinct);
//   2. gqf_agg_tiles: each workgroup counts the hashes of 131072 keys in an
//      8192-entry LDS hash table, then adds its (hash, count) entries to one global
//      open-addressing table (a key its LDS table has no room for goes
//      straight to the global table).  Before its LDS inserts, a wave
//      merges the lanes holding its most frequent hashes by ballot, so a
//      hash that fills many lanes costs one atomic, not one per lane;
//   3. gqf_agg_compact moves the m distinct entries out (emptying the table
//      for the next call) as (index << width) | hash, plus the count by
//      index; an LSD sort of the low `width` bits (the bits the sort orders
//      for the hash width, see gqf_sort_width) orders them by hash;
//   4. gqf_agg_extract / _scan_tiles / _positions split them into sorted
//      hashes and counts (the reduce_by_key outputs) and the exclusive
//      prefix sums of the counts;
//   5. gqf_agg_tile_starts / gqf_agg_expand write every hash count times
//      into `keys`: the ascending hashes the sort would have left there.
#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {

constexpr uint64_t kEmpty = ~0ULL - 1;  // no entry in the global table: hashes here are below 2^40
constexpr int kThreads = 256;
constexpr int kSample = GQF_AGG_SAMPLE;
constexpr int kSampleCap = 8192;
constexpr int kTileThreads = 1024;       // threads per workgroup of gqf_agg_tiles
constexpr int kTileKeys = 131072;        // keys per workgroup of gqf_agg_tiles
// (keys per tile / LDS entries / threads, ms for 2^28 Zipf(1.5) draws:
// 16K / 2048 / 256 4.19; 64K / 8192 / 1024 4.13; 128K / 8192 / 1024 3.25;
// 256K / 8192 / 1024 4.60; 128K / 16384 / 1024 4.35, which fits one
// workgroup per CU)
constexpr int kWaveRounds = 2;           // hashes a wave merges by ballot before its LDS inserts
                                         // (1: 3.28 ms, 2: 3.25 with 128K tiles; 1: 4.18, 2: 4.20,
                                         // 4: 4.49, 6: 4.78 with 16K tiles)
constexpr int kLoadsPerRound = 8;        // keys each thread loads before inserting them
constexpr int kLocalLog = 13;
constexpr int kLocalCap = 1 << kLocalLog;  // LDS table entries
constexpr int kCountBits = 24;             // count bits of an LDS entry
constexpr unsigned long long kLocalEmpty = ~0ULL;
static_assert(kTileKeys < (1 << kCountBits), "a workgroup's counts fit an LDS entry");
constexpr int kLocalProbe = 32;
constexpr int kGlobalProbe = 128;
constexpr int kScanItems = 16;
constexpr int kScanTile = kThreads * kScanItems;  // 4096 entries per scan tile
constexpr int kExpandTile = kThreads * 16;        // 4096 output keys per expand workgroup

__device__ inline uint32_t slot_of(uint64_t h, int log_cap) {
  return (uint32_t)((h * 0x9E3779B97F4A7C15ULL) >> (64 - log_cap));
}

// adds c to hash h in the global table; sets ctrl[2] when no slot is free
// within kGlobalProbe (the caller then falls back to the sort)
__device__ inline void global_add(unsigned long long* __restrict__ gkeys, unsigned long long* __restrict__ gcnt,
                                  int log_cap, uint64_t h, uint64_t c, uint32_t* __restrict__ ctrl) {
  const uint32_t mask = (1u << log_cap) - 1;
  uint32_t s = slot_of(h, log_cap);
  for (int p = 0; p < kGlobalProbe; p++) {
    // a slot changes only from empty to its hash, so a plain read that
    // shows the hash (or another hash) is final
    unsigned long long cur = __atomic_load_n(&gkeys[s], __ATOMIC_RELAXED);
    if (cur == kEmpty) {
      cur = atomicCAS(&gkeys[s], kEmpty, h);
      if (cur == kEmpty) atomicAdd(&ctrl[1], 1u);
    }
    if (cur == kEmpty || cur == h) {
      atomicAdd(&gcnt[s], (unsigned long long)c);
      return;
    }
    s = (s + 1) & mask;
  }
  atomicOr(&ctrl[2], 1u);
}

// exclusive scan of one value per thread over the workgroup; returns the
// prefix, *total the sum
__device__ inline uint64_t block_scan_u64(uint64_t v, uint64_t* total) {
  __shared__ uint64_t wsum[kThreads / 64];
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
  for (int k = 0; k < kThreads / 64; k++) {
    if (k < w) before += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *total = tot;
  return before + incl - v;
}

}  // namespace

// ctrl[0] = distinct hashes among kSample keys spread over the batch;
// ctrl[1..3] = 0 (distinct count, overflow flag, compaction cursor)
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_agg_sample(const uint64_t* __restrict__ keys, uint64_t n, uint64_t mask, int vshift, uint32_t* __restrict__ ctrl) {
  __shared__ unsigned long long tab[kSampleCap];
  __shared__ uint32_t distinct;
  for (int i = threadIdx.x; i < kSampleCap; i += kThreads) tab[i] = kEmpty;
  if (threadIdx.x == 0) distinct = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < kSample; i += kThreads) {
    const uint64_t h = gqf_hash64(keys[(uint64_t)i * (n / kSample)], mask) << vshift;
    uint32_t s = slot_of(h, 13);
    // at most kSample hashes in kSampleCap slots: a free slot is always found
    for (int p = 0; p < kSampleCap; p++) {
      const unsigned long long prev = atomicCAS(&tab[s], kEmpty, h);
      if (prev == kEmpty) atomicAdd(&distinct, 1u);
      if (prev == kEmpty || prev == h) break;
      s = (s + 1) & (kSampleCap - 1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ctrl[0] = distinct;
    ctrl[1] = 0;
    ctrl[2] = 0;
    ctrl[3] = 0;
  }
}

extern "C" __global__ void __launch_bounds__(kTileThreads)
gqf_agg_tiles(const uint64_t* __restrict__ keys, uint64_t n, uint64_t mask, int vshift,
              unsigned long long* __restrict__ gkeys, unsigned long long* __restrict__ gcnt, int log_cap,
              uint32_t* __restrict__ ctrl) {
  // Each distinct hash of a workgroup's tile costs one add to the global
  // table, and at Zipf(1.5) a tile of T draws holds about T^(2/3) distinct
  // hashes, so bigger tiles make fewer adds (T^(-1/3) per draw).  With
  // 16K-draw tiles (about 16M adds) 1.7 ms of the kernel's 4.2 ms were the
  // loads and the LDS counting, the rest the adds.  Eight copies of the
  // counts, one per XCD's workgroups, left the time as it was: the adds are
  // bound by device-scope atomic throughput, not by same-address chains.
  // LDS entry: (hash << kCountBits) | count, kLocalEmpty when free (hashes
  // are below 2^40 here and a workgroup counts fewer than 2^24 keys), so an
  // 8192-entry table is 64 KB and two workgroups fit a CU
  __shared__ unsigned long long tk[kLocalCap];
  __shared__ uint32_t fill;
  for (int i = threadIdx.x; i < kLocalCap; i += kTileThreads) tk[i] = kLocalEmpty;
  if (threadIdx.x == 0) fill = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  // once the table is half full a hash not yet in it goes to the global
  // table directly (short probes whatever the batch)
  for (uint64_t base = (uint64_t)blockIdx.x * kTileKeys; base < n; base += (uint64_t)gridDim.x * kTileKeys) {
    for (int r = 0; r < kTileKeys; r += kTileThreads * kLoadsPerRound) {
      uint64_t h[kLoadsPerRound];
#pragma unroll
      for (int k = 0; k < kLoadsPerRound; k++) {
        const uint64_t i = base + (uint64_t)(r + k * kTileThreads + (int)threadIdx.x);
        h[k] = i < n ? gqf_hash64(keys[i], mask) << vshift : kEmpty;
      }
#pragma unroll
      for (int k = 0; k < kLoadsPerRound; k++) {
        const uint64_t hk = h[k];
        // the lowest still-open lane's hash is matched across the wave: that
        // lane carries the number of lanes holding it, the others drop out
        bool open = hk != kEmpty, pend = open;
        uint32_t c = 1;
        for (int round = 0; round < kWaveRounds; round++) {
          const uint64_t om = __ballot(open);
          if (om == 0) break;
          const int leader = __builtin_ctzll(om);
          const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)hk, leader);
          const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(hk >> 32), leader);
          const bool same = open && hk == (((uint64_t)hi << 32) | lo);
          const uint64_t sm = __ballot(same);
          if (same) {
            open = false;
            if (lane != leader) pend = false;
            else c = (uint32_t)__popcll(sm);
          }
        }
        if (!pend) continue;
        uint32_t s = slot_of(hk, kLocalLog);
        bool done = false;
        const unsigned long long tag = (unsigned long long)hk << kCountBits;
        for (int p = 0; p < kLocalProbe; p++) {
          unsigned long long cur = *(volatile unsigned long long*)&tk[s];
          if (cur == kLocalEmpty) {
            if (*(volatile uint32_t*)&fill >= kLocalCap / 2) break;
            cur = atomicCAS(&tk[s], kLocalEmpty, tag);
            if (cur == kLocalEmpty) {
              atomicAdd(&fill, 1u);
              cur = tag;
            }
          }
          if ((cur >> kCountBits) == hk) {
            atomicAdd(&tk[s], (unsigned long long)c);
            done = true;
            break;
          }
          s = (s + 1) & (kLocalCap - 1);
        }
        if (!done) global_add(gkeys, gcnt, log_cap, hk, c, ctrl);
      }
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < kLocalCap; s += kTileThreads) {
    const uint64_t e = tk[s];
    if (e != kLocalEmpty) global_add(gkeys, gcnt, log_cap, e >> kCountBits, e & ((1ULL << kCountBits) - 1), ctrl);
  }
}

// empties the whole table (after a call that fell back to the sort)
extern "C" __global__ void gqf_agg_clear(unsigned long long* __restrict__ gkeys, unsigned long long* __restrict__ gcnt,
                                         uint64_t cap) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < cap; s += (uint64_t)gridDim.x * blockDim.x) {
    gkeys[s] = kEmpty;
    gcnt[s] = 0;
  }
}

// entry -> comp[idx] = (idx << width) | hash, cnt[idx] = count; empties
// the table.  Workgroup b takes the kCompactChunk slots from b *
// kCompactChunk and reserves its indices with ONE atomic on the cursor
// ctrl[3] (one per wave measured 0.75 ms at 2^22 slots: same-address
// atomics serialise).
constexpr int kCompactItems = 16;
constexpr uint64_t kCompactChunk = (uint64_t)kThreads * kCompactItems;

extern "C" __global__ void __launch_bounds__(kThreads)
gqf_agg_compact(unsigned long long* __restrict__ gkeys, unsigned long long* __restrict__ gcnt, uint64_t cap, int width,
                uint64_t* __restrict__ comp, uint64_t* __restrict__ cnt, uint32_t* __restrict__ ctrl) {
  __shared__ uint32_t wg_base;
  const uint64_t first = (uint64_t)blockIdx.x * kCompactChunk;
  uint64_t h[kCompactItems];
  uint32_t mine = 0;
#pragma unroll
  for (int k = 0; k < kCompactItems; k++) {
    const uint64_t s = first + (uint64_t)k * kThreads + threadIdx.x;
    h[k] = s < cap ? gkeys[s] : kEmpty;
    mine += h[k] != kEmpty ? 1u : 0u;
  }
  uint64_t total;
  const uint32_t before = (uint32_t)block_scan_u64(mine, &total);
  if (threadIdx.x == 0) wg_base = total ? atomicAdd(&ctrl[3], (uint32_t)total) : 0u;
  __syncthreads();
  uint64_t idx = (uint64_t)wg_base + before;
#pragma unroll
  for (int k = 0; k < kCompactItems; k++) {
    if (h[k] == kEmpty) continue;
    const uint64_t s = first + (uint64_t)k * kThreads + threadIdx.x;
    comp[idx] = (idx << width) | h[k];
    cnt[idx] = gcnt[s];
    gkeys[s] = kEmpty;
    gcnt[s] = 0;
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
  (void)block_scan_u64(sum, &total);
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
  uint64_t run = block_scan_u64(sum, &total);
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
  uint64_t run = tile_sums[blockIdx.x] + block_scan_u64(sum, &total);
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

void gqf_launch_agg_sample(const uint64_t* keys, uint64_t n, uint64_t mask, int vshift, uint32_t* ctrl,
                           hipStream_t s) {
  hipLaunchKernelGGL(gqf_agg_sample, dim3(1), dim3(kThreads), 0, s, keys, n, mask, vshift, ctrl);
}

void gqf_launch_agg_tiles(const uint64_t* keys, uint64_t n, uint64_t mask, int vshift, uint64_t* gkeys,
                          uint64_t* gcnt, int log_cap, uint32_t* ctrl, hipStream_t s) {
  // one workgroup per tile, two resident per CU (64 KB of LDS each).  With
  // 16K-draw tiles, 1024 workgroups striding over the tiles took 6.7 ms
  // against 4.2 ms for one per tile (the hashes beyond their tables' first
  // half cost one global add each)
  const uint64_t g = (n + kTileKeys - 1) / kTileKeys;
  gqf_prof_begin("reduce_count", s);
  hipLaunchKernelGGL(gqf_agg_tiles, dim3((uint32_t)g), dim3(kTileThreads), 0, s, keys, n, mask, vshift,
                     (unsigned long long*)gkeys, (unsigned long long*)gcnt, log_cap, ctrl);
  gqf_prof_end("reduce_count", s);
}

void gqf_launch_agg_clear(uint64_t* gkeys, uint64_t* gcnt, uint64_t cap, hipStream_t s) {
  hipLaunchKernelGGL(gqf_agg_clear, dim3(1024), dim3(kThreads), 0, s, (unsigned long long*)gkeys,
                     (unsigned long long*)gcnt, cap);
}

void gqf_launch_agg_finish(uint64_t* keys, uint64_t n, uint64_t m, int width, uint64_t* gkeys, uint64_t* gcnt,
                           uint64_t cap, uint32_t* ctrl, uint64_t* comp, uint64_t* cnt, uint64_t* tmp, uint32_t* hist,
                           uint32_t* totals, uint64_t* tile_sums, uint32_t* starts, uint64_t* ukeys,
                           uint64_t* ucounts, uint64_t* upos, hipStream_t s) {
  gqf_prof_begin("reduce_finish", s);
  hipLaunchKernelGGL(gqf_agg_compact, dim3((uint32_t)((cap + kCompactChunk - 1) / kCompactChunk)), dim3(kThreads), 0, s,
                     (unsigned long long*)gkeys, (unsigned long long*)gcnt, cap, width, comp, cnt, ctrl);
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
