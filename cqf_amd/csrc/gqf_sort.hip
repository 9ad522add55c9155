// gqf_sort.hip -- hash + LSD radix sort + run-length reduce for the bulk path.
//
// Replaces hash_all + thrust::sort + thrust::fill/reduce_by_key of the
// reference bulk API (src/gqf.cu:2615-2637, 3206-3209, 3258-3284).
//
// Sort: least-significant-digit radix sort over only the bits a hash can
// have (q + 8 of them; all 64 for pre-hashed keys), always an even number of
// passes so the result lands back in the caller's buffer, as the reference
// leaves `keys` holding the ascending hashes (src/gqf.cu:3180-3181).
// Each pass is upsweep (per-tile digit histogram in LDS) -> scan of the
// [digit][tile] matrix -> downsweep (stable rank with wave ballots, staged in
// LDS so each digit's keys leave the tile as one contiguous run).
// The first pass hashes on the fly.
#include "gqf_internal.h"
#include "gqf_launch.h"

namespace {
constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;  // 4096 keys
constexpr int kItemsSmall = 4;            // tiles of 1024 keys for inputs up to kSmallN
constexpr uint64_t kSmallN = 1ULL << 16;
constexpr uint32_t kScanSmallTiles = 64;  // one-workgroup scan up to this many tiles
constexpr int kWaves = kThreads / 64;
constexpr int kMaxRadix = 1024;           // digits of up to 10 bits

// hash_it: 0 keys as they are; 1 hash under `mask`, then shift left by
// `vshift` (the filter's value bits); 2 pre-hashed keys shifted by vshift,
// a key above `mask` becoming ~0 (past every region, flagged there).
__device__ inline uint64_t load_key(const uint64_t* in, uint64_t i, uint64_t n, int hash_it, uint64_t mask,
                                    int vshift) {
  if (i >= n) return ~0ULL;
  uint64_t k = in[i];
  if (hash_it == 1) return gqf_hash64(k, mask) << vshift;
  if (hash_it == 2) return (k & ~mask) ? ~0ULL : (k << vshift);
  return k;
}
}  // namespace

// Where tile `tile`'s count of digit d lives: [digit][tile] rows for the
// many-tile scan (gqf_sort_scan_rows, one workgroup per row), [tile][digit]
// up to kScanSmallTiles tiles, where ONE workgroup scans everything and
// strided rows made every wave access touch 64 lines (15 us per scan).
__device__ inline uint64_t hist_at(uint32_t d, uint32_t tile, uint32_t ntiles, int radix) {
  return ntiles <= kScanSmallTiles ? (uint64_t)tile * (uint32_t)radix + d : (uint64_t)d * ntiles + tile;
}

// per-tile digit histogram -> hist[hist_at(d, tile)]
template <int ITEMS>
__device__ __forceinline__ void upsweep_body(const uint64_t* __restrict__ in, uint64_t n, int shift, int bits,
                                             int hash_it, uint64_t mask, int vshift, uint32_t* __restrict__ hist,
                                             uint32_t ntiles) {
  __shared__ uint32_t h[kMaxRadix];
  const int radix = 1 << bits;
  for (int d = threadIdx.x; d < radix; d += kThreads) h[d] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * (kThreads * ITEMS);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll 4
  for (int k = 0; k < ITEMS; k++) {
    uint64_t i = base + (uint64_t)w * (64 * ITEMS) + (uint64_t)k * 64 + lane;
    if (i < n) {
      uint64_t key = load_key(in, i, n, hash_it, mask, vshift);
      atomicAdd(&h[(uint32_t)(key >> shift) & (radix - 1)], 1u);
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < radix; d += kThreads) hist[hist_at(d, blockIdx.x, ntiles, radix)] = h[d];
}

extern "C" __global__ void __launch_bounds__(kThreads)
gqf_sort_upsweep(const uint64_t* __restrict__ in, uint64_t n, int shift, int bits, int hash_it,
                 uint64_t mask, int vshift, uint32_t* __restrict__ hist, uint32_t ntiles) {
  upsweep_body<kItems>(in, n, shift, bits, hash_it, mask, vshift, hist, ntiles);
}

extern "C" __global__ void __launch_bounds__(kThreads)
gqf_sort_upsweep_small(const uint64_t* __restrict__ in, uint64_t n, int shift, int bits, int hash_it,
                       uint64_t mask, int vshift, uint32_t* __restrict__ hist, uint32_t ntiles) {
  upsweep_body<kItemsSmall>(in, n, shift, bits, hash_it, mask, vshift, hist, ntiles);
}

// Exclusive scan of each digit's row of tile counts; row totals -> totals[d].
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_sort_scan_rows(uint32_t* __restrict__ hist, uint32_t ntiles, uint32_t* __restrict__ totals) {
  __shared__ uint32_t wsum[kWaves];
  uint32_t* row = hist + (uint64_t)blockIdx.x * ntiles;
  uint32_t carry = 0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint32_t base = 0; base < ntiles; base += kThreads) {
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < ntiles ? row[i] : 0;
    uint32_t incl = wave_scan_add(v);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int k = 0; k < kWaves; k++) {
      if (k < w) before += wsum[k];
      total += wsum[k];
    }
    if (i < ntiles) row[i] = carry + before + incl - v;
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Exclusive scan in place of count (<= 4 * kThreads) values in LDS by one
// workgroup: each thread takes four consecutive entries, the threads' sums
// are scanned by wave shuffles and across the four waves.
__device__ __forceinline__ void block_exclusive_scan(uint32_t* s, int count) {
  __shared__ uint32_t wsum[kWaves];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t v[4];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int d = 4 * t + k;
    v[k] = d < count ? s[d] : 0;
    sum += v[k];
  }
  uint32_t incl = wave_scan_add(sum);
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  uint32_t run = incl - sum;
  for (int k = 0; k < w; k++) run += wsum[k];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int d = 4 * t + k;
    if (d < count) s[d] = run;
    run += v[k];
  }
  __syncthreads();
}

// Exclusive scan of the digit totals (one workgroup, radix <= 1024).
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_sort_scan_totals(uint32_t* __restrict__ totals, int radix) {
  __shared__ uint32_t s[kMaxRadix];
  for (int d = threadIdx.x; d < radix; d += kThreads) s[d] = totals[d];
  __syncthreads();
  block_exclusive_scan(s, radix);
  for (int d = threadIdx.x; d < radix; d += kThreads) totals[d] = s[d];
}

// Both scans for up to kScanSmallTiles tiles in one workgroup, over the
// [tile][digit] histogram: each digit's column of tile counts becomes its
// exclusive prefix (loads and stores coalesced across the digits' threads),
// and the digit totals their exclusive scan (as gqf_sort_scan_totals).
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_sort_scan_small(uint32_t* __restrict__ hist, uint32_t ntiles, int radix, uint32_t* __restrict__ totals) {
  __shared__ uint32_t s[kMaxRadix];
  for (int d = threadIdx.x; d < radix; d += kThreads) {
    uint32_t v[kScanSmallTiles];
#pragma unroll
    for (uint32_t i = 0; i < kScanSmallTiles; i++) v[i] = i < ntiles ? hist[(uint64_t)i * radix + d] : 0u;
    uint32_t run = 0;
#pragma unroll
    for (uint32_t i = 0; i < kScanSmallTiles; i++) {
      if (i < ntiles) hist[(uint64_t)i * radix + d] = run;
      run += v[i];
    }
    s[d] = run;
  }
  __syncthreads();
  block_exclusive_scan(s, radix);
  for (int d = threadIdx.x; d < radix; d += kThreads) totals[d] = s[d];
}

// Stable scatter of one tile.
template <int ITEMS>
__device__ __forceinline__ void downsweep_body(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                               uint64_t n, int shift, int bits, int hash_it, uint64_t mask,
                                               int vshift, const uint32_t* __restrict__ hist,
                                               const uint32_t* __restrict__ bin_base, uint32_t ntiles) {
  constexpr int TILE = kThreads * ITEMS;
  static_assert(TILE < 65536, "per-wave digit counts are 16-bit");
  // 16-bit per-wave counts keep the kernel at 48 KB of LDS: three
  // workgroups per CU instead of two
  __shared__ uint16_t wcnt[kWaves][kMaxRadix];
  __shared__ uint32_t tile_start[kMaxRadix];
  __shared__ uint32_t glob[kMaxRadix];
  __shared__ uint64_t stage[TILE];
  const int radix = 1 << bits;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int d = threadIdx.x; d < radix; d += kThreads) {
#pragma unroll
    for (int k = 0; k < kWaves; k++) wcnt[k][d] = 0;
    glob[d] = bin_base[d] + hist[hist_at(d, blockIdx.x, ntiles, radix)];
  }
  __syncthreads();

  const uint64_t base = (uint64_t)blockIdx.x * TILE;
  const uint64_t lt_mask = lane ? (~0ULL >> (64 - lane)) : 0ULL;
  uint64_t keys[ITEMS];
  uint32_t rank[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    uint64_t i = base + (uint64_t)w * (64 * ITEMS) + (uint64_t)k * 64 + lane;
    bool valid = i < n;
    uint64_t key = load_key(in, i, n, hash_it, mask, vshift);
    uint32_t d = (uint32_t)(key >> shift) & (radix - 1);
    // lanes holding the same digit (and a valid key)
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; b++) {
      uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    uint32_t before = wcnt[w][d];
    rank[k] = before + (uint32_t)__popcll(peers & lt_mask);
    keys[k] = key;
    // lowest peer bumps the wave's digit counter
    if (valid && (peers & lt_mask) == 0) wcnt[w][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
  }
  __syncthreads();
  // prefix over waves per digit, tile totals -> tile_start (exclusive over digits)
  for (int d = threadIdx.x; d < radix; d += kThreads) {
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < kWaves; k++) {
      uint32_t t = wcnt[k][d];
      wcnt[k][d] = (uint16_t)run;
      run += t;
    }
    tile_start[d] = run;  // tile count, scanned below
  }
  __syncthreads();
  block_exclusive_scan(tile_start, radix);  // all waves, 4 digits per thread
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    uint64_t i = base + (uint64_t)w * (64 * ITEMS) + (uint64_t)k * 64 + lane;
    if (i < n) {
      uint32_t d = (uint32_t)(keys[k] >> shift) & (radix - 1);
      stage[tile_start[d] + wcnt[w][d] + rank[k]] = keys[k];
    }
  }
  __syncthreads();
  const uint32_t cnt = (uint32_t)min((uint64_t)TILE, n - base);
  for (uint32_t i = threadIdx.x; i < cnt; i += kThreads) {
    uint64_t key = stage[i];
    uint32_t d = (uint32_t)(key >> shift) & (radix - 1);
    out[(uint64_t)glob[d] + (i - tile_start[d])] = key;
  }
}

extern "C" __global__ void __launch_bounds__(kThreads)
gqf_sort_downsweep(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n, int shift,
                   int bits, int hash_it, uint64_t mask, int vshift, const uint32_t* __restrict__ hist,
                   const uint32_t* __restrict__ bin_base, uint32_t ntiles) {
  downsweep_body<kItems>(in, out, n, shift, bits, hash_it, mask, vshift, hist, bin_base, ntiles);
}

extern "C" __global__ void __launch_bounds__(kThreads)
gqf_sort_downsweep_small(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n, int shift,
                         int bits, int hash_it, uint64_t mask, int vshift, const uint32_t* __restrict__ hist,
                         const uint32_t* __restrict__ bin_base, uint32_t ntiles) {
  downsweep_body<kItemsSmall>(in, out, n, shift, bits, hash_it, mask, vshift, hist, bin_base, ntiles);
}

// ---------------------------------------------------------------------------
// Top-digit split of hashed keys for the two-level sort.  Only the bucket
// of each key matters here: the bucket sort then orders every bucket in
// full, and keys with equal hashes are equal values, so the split need not
// be stable.  That lets it use big tiles (kMRound keys per round, up to 8
// rounds per workgroup: a 2 MB [digit][tile] histogram at 2^24 keys instead
// of 16.8 MB of scattered counts) and rank keys within a round by LDS
// atomics; each round's keys leave the workgroup grouped by digit, about
// 64 bytes per digit at 1024 digits.

namespace {
constexpr int kMThreads = 1024;
constexpr int kMWaves = kMThreads / 64;
constexpr int kMItems = 8;
constexpr uint32_t kMRound = kMThreads * kMItems;  // 8192 keys, 64 KB staged
constexpr int kMRankShift = 48;                    // ranks < kMRound ride above the hash bits
static_assert(kMRound <= (1u << (64 - kMRankShift)), "rank fits above the hash");
static_assert(kMaxRadix <= kMThreads, "one digit per thread");
}  // namespace

// per-tile digit histogram of the hashed keys -> hist[d * ntiles + tile]
extern "C" __global__ void __launch_bounds__(kMThreads)
gqf_msd_count(const uint64_t* __restrict__ in, uint64_t n, uint64_t tile_keys, int shift, int bits, uint64_t mask,
              int vshift, uint32_t* __restrict__ hist, uint32_t ntiles) {
  __shared__ uint32_t h[kMaxRadix];
  const int radix = 1 << bits;
  const int t = threadIdx.x;
  if (t < radix) h[t] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * tile_keys;
  const uint64_t end = min(n, base + tile_keys);
  for (uint64_t r0 = base; r0 < end; r0 += kMRound) {
    uint64_t key[kMItems];
#pragma unroll
    for (int k = 0; k < kMItems; k++) {
      const uint64_t i = r0 + (uint64_t)(k * kMThreads + t);
      key[k] = i < end ? in[i] : 0;
    }
#pragma unroll
    for (int k = 0; k < kMItems; k++) {
      const uint64_t i = r0 + (uint64_t)(k * kMThreads + t);
      if (i < end) atomicAdd(&h[(uint32_t)((gqf_hash64(key[k], mask) << vshift) >> shift) & (radix - 1)], 1u);
    }
  }
  __syncthreads();
  if (t < radix) hist[(uint64_t)t * ntiles + blockIdx.x] = h[t];
}

// scatter of one tile into the digit buckets (hist scanned per digit row,
// bin_base the exclusive scan of the digit totals)
extern "C" __global__ void __launch_bounds__(kMThreads, 8)
gqf_msd_scatter(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n, uint64_t tile_keys,
                int shift, int bits, uint64_t mask, int vshift, const uint32_t* __restrict__ hist,
                const uint32_t* __restrict__ bin_base, uint32_t ntiles) {
  __shared__ uint64_t stage[kMRound];
  __shared__ uint32_t cnt[kMaxRadix];     // keys of the round per digit
  __shared__ uint32_t lstart[kMaxRadix];  // their first index in stage
  __shared__ uint32_t glob[kMaxRadix];    // next output index of each digit
  __shared__ uint32_t wsum[kMWaves];
  const int radix = 1 << bits;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  if (t < radix) glob[t] = bin_base[t] + hist[(uint64_t)t * ntiles + blockIdx.x];
  const uint64_t base = (uint64_t)blockIdx.x * tile_keys;
  const uint64_t end = min(n, base + tile_keys);
  for (uint64_t r0 = base; r0 < end; r0 += kMRound) {
    // 32-bit offsets from the round's (uniform) base: one VGPR per address
    const uint64_t* __restrict__ src = in + r0;
    const uint32_t m = (uint32_t)min((uint64_t)kMRound, end - r0);
    if (t < radix) cnt[t] = 0;
    __syncthreads();
    // each key's rank within its digit rides in bits 48.. of the key (the
    // hashes here are below 2^42), which keeps the kernel within 64 VGPRs
    uint64_t key[kMItems];
#pragma unroll
    for (int k = 0; k < kMItems; k++) {
      const uint32_t i = (uint32_t)(k * kMThreads + t);
      key[k] = i < m ? src[i] : 0;
    }
#pragma unroll
    for (int k = 0; k < kMItems; k++) {
      const uint32_t i = (uint32_t)(k * kMThreads + t);
      const uint64_t h = gqf_hash64(key[k], mask) << vshift;
      const uint32_t rank = i < m ? atomicAdd(&cnt[(uint32_t)(h >> shift) & (radix - 1)], 1u) : 0u;
      key[k] = h | ((uint64_t)rank << kMRankShift);
    }
    __syncthreads();
    // exclusive scan of the digit counts, one digit per thread
    const uint32_t c = t < radix ? cnt[t] : 0u;
    uint32_t incl = wave_scan_add(c);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int k = 0; k < w; k++) before += wsum[k];
    if (t < radix) lstart[t] = before + incl - c;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kMItems; k++) {
      const uint32_t i = (uint32_t)(k * kMThreads + t);
      const uint64_t h = key[k] & ((1ULL << kMRankShift) - 1);
      if (i < m) stage[lstart[(uint32_t)(h >> shift) & (radix - 1)] + (uint32_t)(key[k] >> kMRankShift)] = h;
    }
    __syncthreads();
    for (uint32_t i = t; i < m; i += kMThreads) {
      const uint64_t v = stage[i];
      const uint32_t d = (uint32_t)(v >> shift) & (radix - 1);
      out[(uint64_t)glob[d] + (i - lstart[d])] = v;
    }
    __syncthreads();
    if (t < radix) glob[t] += c;
  }
}

// ---------------------------------------------------------------------------
// Two-level sort of hashed keys (batches of 2^20..2^24): the top-digit
// split above (gqf_msd_count / gqf_msd_scatter) puts the keys into 2^msd
// buckets by their top `msd` hash bits; then one workgroup per bucket sorts the bucket's remaining
// low bits (at most 32) in LDS -- up to kBucketCap keys as 32-bit values in
// two 68 KB buffers of gfx950's 160 KB -- with stable 8-bit LSD passes whose
// ranks come from wave ballots over peers of equal digit, as in the
// downsweep.  The bucket's top bits are its index, so the low bits order it.
// A bucket larger than kBucketCap (heavily repeated keys) runs the same
// passes over its 64-bit keys in global memory.

namespace {
constexpr int kBThreads = 1024;
constexpr int kBWaves = kBThreads / 64;
constexpr int kBItems = 8;                           // keys per lane per round
constexpr uint32_t kBRound = kBThreads * kBItems;    // 8192
constexpr uint32_t kBucketCap = 17408;               // 2^24 keys in 1024 buckets: mean 16384, sd 128
constexpr int kBMaxPasses = 5;

struct BucketSmem {
  uint32_t buf[2][kBucketCap];
  uint32_t wcnt[kBWaves][256];        // per wave digit counts, then positions
  uint32_t tot[kBMaxPasses][256];     // digit totals of every pass
  uint32_t run[256];                  // next free position of each digit
};

// Stable scatter of src[0, m) into dst by digit (v >> sh) & 255, given the
// digit totals; every thread of the workgroup calls it.
template <typename V>
__device__ __forceinline__ void bucket_pass(const V* src, V* dst, uint32_t m, int sh, const uint32_t* tot,
                                            BucketSmem& sm) {
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  if (w == 0) {  // exclusive scan of the totals, 4 digits per lane
    uint32_t v[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      v[k] = tot[4 * lane + k];
      sum += v[k];
    }
    uint32_t incl = wave_scan_add(sum);
    uint32_t r = incl - sum;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      sm.run[4 * lane + k] = r;
      r += v[k];
    }
  }
  __syncthreads();
  const uint64_t lt = lane ? (~0ULL >> (64 - lane)) : 0ULL;
  for (uint32_t base = 0; base < m; base += kBRound) {
    for (int d = lane; d < 256; d += 64) sm.wcnt[w][d] = 0;
    V keys[kBItems];
    uint32_t rank[kBItems];
#pragma unroll
    for (int k = 0; k < kBItems; k++) {
      const uint32_t i = base + (uint32_t)(w * 64 * kBItems + k * 64 + lane);
      const bool valid = i < m;
      const V key = valid ? src[i] : (V)0;
      const uint32_t d = (uint32_t)(key >> sh) & 255u;
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const uint64_t mb = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? mb : ~mb;
      }
      const uint32_t before = sm.wcnt[w][d];
      rank[k] = before + (uint32_t)__popcll(peers & lt);
      keys[k] = key;
      if (valid && (peers & lt) == 0) sm.wcnt[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    if (t < 256) {  // wave counts -> positions, in wave order
      uint32_t r = sm.run[t];
#pragma unroll
      for (int k = 0; k < kBWaves; k++) {
        const uint32_t x = sm.wcnt[k][t];
        sm.wcnt[k][t] = r;
        r += x;
      }
      sm.run[t] = r;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kBItems; k++) {
      const uint32_t i = base + (uint32_t)(w * 64 * kBItems + k * 64 + lane);
      if (i < m) dst[sm.wcnt[w][(uint32_t)(keys[k] >> sh) & 255u] + rank[k]] = keys[k];
    }
  }
}

}  // namespace

// start[r] of the regions whose bounds lie in bucket b's hash range
// [b << low_bits, (b + 1) << low_bits): the bucket's first index s plus
// below(bound - bucket base), the number of its keys whose low bits are
// below that (a bound at or past range is left to gqf_region_lists)
template <typename F>
__device__ __forceinline__ void bucket_region_starts(const gqf_region_args& ra, uint32_t b, int low_bits, uint64_t s,
                                                     F below) {
  if (!ra.start) return;
  const uint64_t lo = (uint64_t)b << low_bits;
  const uint64_t hi = lo + (1ULL << low_bits);
  const uint64_t r0 = (lo + ra.step - 1) / ra.step;
  uint64_t r1 = (hi + ra.step - 1) / ra.step;
  if (r1 > ra.nregions) r1 = ra.nregions;
  for (uint64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    const uint64_t bound = r * ra.step;
    if (bound < ra.range) ra.start[r] = (uint32_t)(s + below((uint32_t)(bound - lo)));
  }
}

// Common case, two workgroups per CU: the bucket's keys stay in registers
// (17 per thread) while they are counted by sub-bucket -- the top (up to) 11
// of the low bits, ~8 keys each for uniform hashes -- then go to one LDS
// buffer in sub-bucket order (any order within), and each key's place in
// its sub-bucket is the number of keys there below it (ties by buffer
// index), so every key is written straight to its sorted position.  A
// bucket above kBucketCap keys, or with a sub-bucket above kSubCap
// (repeated keys), is flagged for gqf_sort_buckets_slow.
constexpr int kBLoad = (int)((kBucketCap + kBThreads - 1) / kBThreads);  // 17
constexpr int kSubBitsMax = 11;
constexpr uint32_t kSubMax = 1u << kSubBitsMax;
constexpr uint32_t kSubCap = 64;
static_assert(kSubMax == 2 * kBThreads, "two sub-buckets per thread in the scan");

extern "C" __global__ void __launch_bounds__(kBThreads)
gqf_sort_buckets(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n, int low_bits,
                 const uint32_t* __restrict__ starts, uint32_t nbuckets, uint32_t* __restrict__ slow,
                 gqf_region_args ra, uint64_t tag) {
  __shared__ uint32_t buf[kBucketCap];
  __shared__ uint32_t sub_end[kSubMax];  // counts, then cursors: after the scatter, each sub-bucket's end
  __shared__ uint32_t wsum[kBWaves];
  __shared__ uint32_t oversize;
  const uint32_t b = blockIdx.x;
  const uint64_t s = starts[b];
  const uint64_t e = b + 1 < nbuckets ? (uint64_t)starts[b + 1] : n;
  if (e <= s) {
    bucket_region_starts(ra, b, low_bits, s, [](uint32_t) { return 0u; });
    return;
  }
  const uint32_t m = (uint32_t)(e - s);
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  if (m > kBucketCap) {
    if (t == 0) slow[b] = 1;
    return;
  }
  const uint64_t hi = (uint64_t)b << low_bits;  // the bucket's top bits
  const int db = low_bits < kSubBitsMax ? low_bits : kSubBitsMax;
  const int sb = low_bits - db;
  const uint32_t nsub = 1u << db;
  for (uint32_t d = t; d < nsub; d += kBThreads) sub_end[d] = 0;
  if (t == 0) oversize = 0;
  uint32_t v[kBLoad];
#pragma unroll
  for (int k = 0; k < kBLoad; k++) {
    const uint32_t i = (uint32_t)(k * kBThreads + t);
    v[k] = i < m ? (uint32_t)(in[s + i] - hi) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kBLoad; k++)
    if ((uint32_t)(k * kBThreads + t) < m) atomicAdd(&sub_end[v[k] >> sb], 1u);
  __syncthreads();
  {  // exclusive scan of the counts, two sub-buckets per thread
    const uint32_t d0 = 2 * (uint32_t)t;
    const uint32_t c0 = d0 < nsub ? sub_end[d0] : 0u;
    const uint32_t c1 = d0 + 1 < nsub ? sub_end[d0 + 1] : 0u;
    const uint32_t sum = c0 + c1;
    uint32_t incl = wave_scan_add(sum);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    for (int k = 0; k < w; k++) run += wsum[k];
    if (d0 < nsub) sub_end[d0] = run;
    if (d0 + 1 < nsub) sub_end[d0 + 1] = run + c0;
    if (c0 > kSubCap || c1 > kSubCap) oversize = 1;
  }
  __syncthreads();
  if (oversize) {
    if (t == 0) slow[b] = 1;
    return;
  }
#pragma unroll
  for (int k = 0; k < kBLoad; k++)
    if ((uint32_t)(k * kBThreads + t) < m) buf[atomicAdd(&sub_end[v[k] >> sb], 1u)] = v[k];
  __syncthreads();
  for (uint32_t i = t; i < m; i += kBThreads) {
    const uint32_t x = buf[i];
    const uint32_t d = x >> sb;
    const uint32_t st = d ? sub_end[d - 1] : 0u;
    const uint32_t en = sub_end[d];
    uint32_t r = 0;
    for (uint32_t j = st; j < en; j++) {
      const uint32_t y = buf[j];
      r += (y < x || (y == x && j < i)) ? 1u : 0u;
    }
    out[s + st + r] = tag | hi | x;
  }
  bucket_region_starts(ra, b, low_bits, s, [&](uint32_t x) {
    const uint32_t d = x >> sb;
    const uint32_t st = d ? sub_end[d - 1] : 0u, en = sub_end[d];
    uint32_t c = st;
    for (uint32_t j = st; j < en; j++) c += buf[j] < x ? 1u : 0u;
    return c;
  });
}

__device__ __forceinline__ void sort_flagged_bucket(uint64_t* in, uint64_t* out, uint64_t n, int low_bits,
                                                 const uint32_t* __restrict__ starts, uint32_t nbuckets,
                                                 uint32_t* __restrict__ slow, uint32_t b, BucketSmem& sm,
                                                 const gqf_region_args& ra, uint64_t tag) {
  const uint64_t s = starts[b];
  const uint64_t e = b + 1 < nbuckets ? (uint64_t)starts[b + 1] : n;
  const uint32_t m = (uint32_t)(e - s);
  const int t = threadIdx.x;
  const int passes = (low_bits + 7) / 8;
  for (int k = t; k < kBMaxPasses * 256; k += kBThreads) (&sm.tot[0][0])[k] = 0;
  __syncthreads();
  if (t == 0) slow[b] = 0;
  if (m <= kBucketCap) {
    const uint64_t hi = (uint64_t)b << low_bits;
    for (uint32_t i = t; i < m; i += kBThreads) {
      const uint32_t v = (uint32_t)(in[s + i] - hi);
      sm.buf[0][i] = v;
      for (int p = 0; p < passes; p++) atomicAdd(&sm.tot[p][(v >> (8 * p)) & 255u], 1u);
    }
    __syncthreads();
    int src = 0;
    for (int p = 0; p < passes; p++) {
      bucket_pass<uint32_t>(sm.buf[src], sm.buf[src ^ 1], m, 8 * p, sm.tot[p], sm);
      __syncthreads();
      src ^= 1;
    }
    for (uint32_t i = t; i < m; i += kBThreads) out[s + i] = tag | hi | sm.buf[src][i];
    const uint32_t* sorted = sm.buf[src];
    bucket_region_starts(ra, b, low_bits, s, [&](uint32_t x) {
      uint32_t lo = 0, hi2 = m;
      while (lo < hi2) {
        const uint32_t mid = (lo + hi2) >> 1;
        if (sorted[mid] < x) lo = mid + 1; else hi2 = mid;
      }
      return lo;
    });
    return;
  }
  // too large for LDS: the same passes between the bucket's ranges of in and
  // out, plus a copy pass on the (shared) top bits when the count is even, so
  // the result lands in out; other waves' stores are published between
  // passes with agent-scope fences around the barrier
  const int gpasses = passes | 1;
  uint64_t* a = in + s;
  uint64_t* o = out + s;
  for (uint32_t i = t; i < m; i += kBThreads) {
    const uint64_t v = a[i];
    for (int p = 0; p < gpasses; p++) atomicAdd(&sm.tot[p][(uint32_t)(v >> (8 * p)) & 255u], 1u);
  }
  __syncthreads();
  for (int p = 0; p < gpasses; p++) {
    bucket_pass<uint64_t>(p & 1 ? o : a, p & 1 ? a : o, m, 8 * p, sm.tot[p], sm);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  // the bucket's keys now ascend in o (their top bits are the bucket's)
  const uint64_t base = (uint64_t)b << low_bits;
  bucket_region_starts(ra, b, low_bits, s, [&](uint32_t x) {
    uint32_t lo = 0, hi2 = m;
    while (lo < hi2) {
      const uint32_t mid = (lo + hi2) >> 1;
      if (o[mid] - base < (uint64_t)x) lo = mid + 1; else hi2 = mid;
    }
    return lo;
  });
  if (tag) {
    __syncthreads();  // the searches above read o untagged
    for (uint32_t i = t; i < m; i += kBThreads) o[i] |= tag;
  }
}

// The flagged buckets (skips the others, and clears the flags it takes):
// stable 8-bit passes over the bucket's low bits in two LDS buffers, or, for
// a bucket above kBucketCap, the same passes over its 64-bit keys between
// its ranges of in and out in global memory.  A grid of at most one
// workgroup per CU strides over the buckets, so an all-clear call costs
// one short launch.
extern "C" __global__ void __launch_bounds__(kBThreads)
gqf_sort_buckets_slow(uint64_t* in, uint64_t* out, uint64_t n, int low_bits, const uint32_t* __restrict__ starts,
                      uint32_t nbuckets, uint32_t* __restrict__ slow, gqf_region_args ra, uint64_t tag) {
  __shared__ BucketSmem sm;
  for (uint32_t b = blockIdx.x; b < nbuckets; b += gridDim.x) {
    if (slow[b]) sort_flagged_bucket(in, out, n, low_bits, starts, nbuckets, slow, b, sm, ra, tag);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// run-length reduce of sorted keys: unique keys + their multiplicities

extern "C" __global__ void __launch_bounds__(kThreads)
gqf_rle_count(const uint64_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ tile_heads) {
  __shared__ uint32_t c;
  if (threadIdx.x == 0) c = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  uint32_t mine = 0;
  for (int k = 0; k < kItems; k++) {
    uint64_t i = base + (uint64_t)k * kThreads + threadIdx.x;
    if (i < n && (i == 0 || keys[i] != keys[i - 1])) mine++;
  }
  atomicAdd(&c, mine);
  __syncthreads();
  if (threadIdx.x == 0) tile_heads[blockIdx.x] = c;
}

// exclusive scan of tile head counts (one workgroup); total -> *m
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_rle_scan(uint32_t* __restrict__ tile_heads, uint32_t ntiles, uint64_t* __restrict__ m) {
  __shared__ uint32_t wsum[kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < ntiles; base += kThreads) {
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < ntiles ? tile_heads[i] : 0;
    uint32_t incl = wave_scan_add(v);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int k = 0; k < kWaves; k++) {
      if (k < w) before += wsum[k];
      total += wsum[k];
    }
    if (i < ntiles) tile_heads[i] = carry + before + incl - v;
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) *m = carry;
}

// write unique keys and the index where each run starts
extern "C" __global__ void __launch_bounds__(kThreads)
gqf_rle_emit(const uint64_t* __restrict__ keys, uint64_t n, const uint32_t* __restrict__ tile_base,
             uint64_t* __restrict__ ukeys, uint64_t* __restrict__ upos) {
  __shared__ uint32_t wsum[kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  uint32_t carry = tile_base[blockIdx.x];
  const uint64_t lt_mask = lane ? (~0ULL >> (64 - lane)) : 0ULL;
  for (int k = 0; k < kItems; k++) {
    uint64_t i = base + (uint64_t)k * kThreads + threadIdx.x;
    bool head = i < n && (i == 0 || keys[i] != keys[i - 1]);
    uint64_t b = __ballot(head);
    if (lane == 0) wsum[w] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int q = 0; q < kWaves; q++) {
      if (q < w) before += wsum[q];
      total += wsum[q];
    }
    if (head) {
      uint32_t u = carry + before + (uint32_t)__popcll(b & lt_mask);
      ukeys[u] = keys[i];
      upos[u] = i;
    }
    carry += total;
    __syncthreads();
  }
}

// counts[u] = next run start - this run start
extern "C" __global__ void gqf_rle_counts(const uint64_t* __restrict__ upos, const uint64_t* __restrict__ m,
                                          uint64_t n, uint64_t* __restrict__ counts) {
  uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t mm = *m;
  if (u >= mm) return;
  uint64_t next = u + 1 < mm ? upos[u + 1] : n;
  counts[u] = next - upos[u];
}

// ---------------------------------------------------------------------------
// launchers

static inline uint32_t sort_ntiles(uint64_t n) { return (uint32_t)((n + kTile - 1) / kTile); }

// sized for the larger of the 4096-key tiling and the small-input tiling
uint64_t gqf_sort_hist_words(uint64_t n) {
  const uint64_t t = sort_ntiles(n);
  return (uint64_t)kMaxRadix * (t > kScanSmallTiles ? t : kScanSmallTiles);
}

namespace {
// radix digit and number of passes for `bits` key bits; an even number of
// passes leaves the result in `keys`
void sort_shape(int bits, int digit_bits, int* passes, int* digit) {
  if (digit_bits >= 6 && digit_bits <= 10) {
    *digit = digit_bits;
    *passes = (bits + digit_bits - 1) / digit_bits;
    *passes += *passes & 1;
  } else if (bits <= 40) {
    *passes = 4;
    *digit = (bits + 3) / 4;
    if (*digit < 6) *digit = 6;
  } else {
    *passes = 8;
    *digit = 8;
  }
}
}  // namespace

int gqf_sort_width(int bits, int digit_bits) {
  int passes, digit;
  sort_shape(bits, digit_bits, &passes, &digit);
  return passes * digit;
}

void gqf_launch_sort(uint64_t* keys, uint64_t* tmp, uint64_t n, int hash_it, uint64_t mask, int vshift, int bits,
                     int digit_bits, uint32_t* hist, uint32_t* totals, hipStream_t s) {
  if (n == 0) return;
  int passes, digit;
  sort_shape(bits, digit_bits, &passes, &digit);
  // small inputs: 1024-key tiles (more workgroups, shorter tiles)
  const bool small = n <= kSmallN;
  const uint32_t tile = small ? kThreads * kItemsSmall : kTile;
  const uint32_t ntiles = (uint32_t)((n + tile - 1) / tile);
  const int radix = 1 << digit;
  uint64_t* in = keys;
  uint64_t* out = tmp;
  for (int p = 0; p < passes; p++) {
    const int shift = p * digit;
    const int h = p == 0 ? hash_it : 0;
    gqf_prof_begin("sort_pass", s);
    hipLaunchKernelGGL(small ? gqf_sort_upsweep_small : gqf_sort_upsweep, dim3(ntiles), dim3(kThreads), 0, s, in, n,
                       shift, digit, h, mask, vshift, hist, ntiles);
    if (ntiles <= kScanSmallTiles) {
      hipLaunchKernelGGL(gqf_sort_scan_small, dim3(1), dim3(kThreads), 0, s, hist, ntiles, radix, totals);
    } else {
      hipLaunchKernelGGL(gqf_sort_scan_rows, dim3(radix), dim3(kThreads), 0, s, hist, ntiles, totals);
      hipLaunchKernelGGL(gqf_sort_scan_totals, dim3(1), dim3(kThreads), 0, s, totals, radix);
    }
    hipLaunchKernelGGL(small ? gqf_sort_downsweep_small : gqf_sort_downsweep, dim3(ntiles), dim3(kThreads), 0, s,
                       (const uint64_t*)in, out, n, shift, digit, h, mask, vshift, (const uint32_t*)hist,
                       (const uint32_t*)totals, ntiles);
    gqf_prof_end("sort_pass", s);
    uint64_t* t = in;
    in = out;
    out = t;
  }
}

bool gqf_sort_two_level_fits(uint64_t n, int bits) {
  return n >= (1ULL << 20) && n <= (1ULL << 24) && bits >= 18 && bits <= 42;
}

void gqf_launch_sort_two_level(uint64_t* keys, uint64_t* tmp, uint64_t n, uint64_t mask, int vshift, int bits,
                               uint32_t* hist, uint32_t* totals, uint32_t* slow, hipStream_t s,
                               const gqf_region_args* regions, uint64_t tag) {
  const gqf_region_args ra = regions ? *regions : gqf_region_args{};
  // 2^msd buckets of about 12K keys (at most 1024), low bits <= 32
  int msd = 4;
  while (msd < 10 && (n >> msd) > 12288) msd++;
  if (bits - msd > 32) msd = bits - 32;
  const int radix = 1 << msd;
  // about 512 tiles of 1..8 rounds: 32K keys per tile at 2^24 keys.  The
  // histogram (radix x ntiles <= 1024 x n / 8192 words) fits the LSD
  // sort's (gqf_sort_hist_words)
  uint64_t rounds = n / ((uint64_t)kMRound * 512);
  rounds = rounds < 1 ? 1 : (rounds > 8 ? 8 : rounds);
  const uint64_t tile_keys = rounds * kMRound;
  const uint32_t ntiles = (uint32_t)((n + tile_keys - 1) / tile_keys);
  gqf_prof_begin("sort_pass", s);
  hipLaunchKernelGGL(gqf_msd_count, dim3(ntiles), dim3(kMThreads), 0, s, (const uint64_t*)keys, n, tile_keys,
                     bits - msd, msd, mask, vshift, hist, ntiles);
  hipLaunchKernelGGL(gqf_sort_scan_rows, dim3(radix), dim3(kThreads), 0, s, hist, ntiles, totals);
  hipLaunchKernelGGL(gqf_sort_scan_totals, dim3(1), dim3(kThreads), 0, s, totals, radix);
  hipLaunchKernelGGL(gqf_msd_scatter, dim3(ntiles), dim3(kMThreads), 0, s, (const uint64_t*)keys, tmp, n, tile_keys,
                     bits - msd, msd, mask, vshift, (const uint32_t*)hist, (const uint32_t*)totals, ntiles);
  gqf_prof_end("sort_pass", s);
  gqf_prof_begin("sort_buckets", s);
  hipLaunchKernelGGL(gqf_sort_buckets, dim3(radix), dim3(kBThreads), 0, s, (const uint64_t*)tmp, keys, n, bits - msd,
                     (const uint32_t*)totals, (uint32_t)radix, slow, ra, tag);
  hipLaunchKernelGGL(gqf_sort_buckets_slow, dim3(radix < 256 ? radix : 256), dim3(kBThreads), 0, s, tmp, keys, n, bits - msd,
                     (const uint32_t*)totals, (uint32_t)radix, slow, ra, tag);
  gqf_prof_end("sort_buckets", s);
}

void gqf_launch_rle(const uint64_t* keys, uint64_t n, uint32_t* tile_heads, uint64_t* ukeys, uint64_t* upos,
                    uint64_t* ucounts, uint64_t* m_dev, hipStream_t s) {
  if (n == 0) return;
  const uint32_t ntiles = sort_ntiles(n);
  hipLaunchKernelGGL(gqf_rle_count, dim3(ntiles), dim3(kThreads), 0, s, keys, n, tile_heads);
  hipLaunchKernelGGL(gqf_rle_scan, dim3(1), dim3(kThreads), 0, s, tile_heads, ntiles, m_dev);
  hipLaunchKernelGGL(gqf_rle_emit, dim3(ntiles), dim3(kThreads), 0, s, keys, n, (const uint32_t*)tile_heads, ukeys,
                     upos);
  const uint32_t g = (uint32_t)((n + 255) / 256);
  hipLaunchKernelGGL(gqf_rle_counts, dim3(g), dim3(256), 0, s, (const uint64_t*)upos, (const uint64_t*)m_dev, n,
                     ucounts);
}
