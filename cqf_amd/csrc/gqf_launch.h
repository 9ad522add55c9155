// gqf_launch.h -- host-side launchers of the libgqf kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gqf_internal.h"

// radix sort of n keys in place (tmp holds n keys); the first pass turns
// keys into slot hashes: hash_it 1 hashes under `mask` and shifts left by
// vshift, 2 shifts pre-hashed keys (a key above mask becomes ~0), 0 keeps
// them.  hist must hold gqf_sort_hist_words(n).
uint64_t gqf_sort_hist_words(uint64_t n);
// digit_bits: 6..10 fixes the radix digit (passes rounded up to even), 0 picks 4 passes
void gqf_launch_sort(uint64_t* keys, uint64_t* tmp, uint64_t n, int hash_it, uint64_t mask, int vshift, int bits,
                     int digit_bits, uint32_t* hist, uint32_t* totals, hipStream_t s);
// the same result for keys to be hashed (hash_it 1), for batches where
// gqf_sort_two_level_fits: one stable pass on the top hash bits into
// buckets of about 12K keys, then an LDS sort per bucket; not stable on
// equal hashes (plain keys only: equal hashes are indistinguishable)
// the low bits gqf_launch_sort orders for these arguments (passes x digit,
// at least `bits`): bits above them ride along in their stable order
int gqf_sort_width(int bits, int digit_bits);
bool gqf_sort_two_level_fits(uint64_t n, int bits);
// slow: GQF_SORT_MAX_BUCKETS flags, zero between calls (the kernels leave them so)
#define GQF_SORT_MAX_BUCKETS 1024
// Region starts from the bucket sort (optional): for every region r <
// nregions whose bound r * step lies below range, start[r] = the first index
// of the sorted keys at or above the bound (what gqf_region_bounds would
// write); gqf_launch_region_lists then completes the table and the lists.
struct gqf_region_args {
  uint32_t* start = nullptr;
  uint64_t step = 0;      // hash values per region: region_blocks << 14
  uint64_t nregions = 0;
  uint64_t range = 0;
};
// tag: ORed into every sorted key on its way out (a filter group's filter
// index above the hash bits; the region starts are of the untagged hashes)
void gqf_launch_sort_two_level(uint64_t* keys, uint64_t* tmp, uint64_t n, uint64_t mask, int vshift, int bits,
                               uint32_t* hist, uint32_t* totals, uint32_t* slow, hipStream_t s,
                               const gqf_region_args* regions = nullptr, uint64_t tag = 0);

// Counting reduce of duplicate-heavy batches (gqf_agg.hip).  ctrl: 4 words
// (distinct keys in the sample, distinct hashes, overflow flag, output
// cursor); fill: GQF_AGG_BUCKETS words; hot: GQF_AGG_HOT + 1 words; stage:
// GQF_AGG_BUCKETS x GQF_AGG_BUCKET_CAP words.
#define GQF_AGG_SAMPLE 4096
#define GQF_AGG_HOT 8                // keys counted in registers (the sample's most frequent)
#define GQF_AGG_BUCKETS 512          // staging buckets, by hash
#define GQF_AGG_BUCKET_CAP 32768     // staging entries per bucket (2^24 in all)
#define GQF_AGG_LOG_CAP 22           // distinct hashes a counted batch may hold
// sample: ctrl[0] = distinct keys among 4096 spread over the batch, the hot
// keys, ctrl[1..3] and fill zeroed
void gqf_launch_agg_sample(const uint64_t* keys, uint64_t n, uint64_t mask, uint32_t* ctrl, uint32_t* fill,
                           uint64_t* hot, hipStream_t s);
// counts every key's slot hash per 131072-key tile and stages the tiles'
// (hash, count) entries by bucket; ctrl[2] != 0 if a bucket ran out of room
void gqf_launch_agg_tiles(const uint64_t* keys, uint64_t n, uint64_t mask, int vshift, const uint64_t* hot,
                          uint64_t* stage, uint32_t* fill, uint32_t* ctrl, hipStream_t s);
// sums the staged entries by hash: comp[idx] = (idx << width) | hash,
// cnt[idx] = count for the ctrl[1] distinct hashes (ctrl[2] != 0 if they
// did not fit: LDS per bucket or cnt_cap)
void gqf_launch_agg_merge(const uint64_t* stage, const uint32_t* fill, int width, uint64_t* comp, uint64_t* cnt,
                          uint64_t cnt_cap, uint32_t* ctrl, hipStream_t s);
// the m merged entries -> ascending ukeys / ucounts (m each), upos (m + 1
// prefix sums) and keys = each hash count times in ascending order.
// width = gqf_sort_width(hash bits, 0); comp: m words (may be upos);
// tmp / hist / totals: a sort of m keys; tile_sums: 4096 words; starts:
// n / 4096 + 2 words.
void gqf_launch_agg_finish(uint64_t* keys, uint64_t n, uint64_t m, int width, uint64_t* comp, uint64_t* cnt,
                           uint64_t* tmp, uint32_t* hist, uint32_t* totals, uint64_t* tile_sums, uint32_t* starts,
                           uint64_t* ukeys, uint64_t* ucounts, uint64_t* upos, hipStream_t s);

// run-length reduce of sorted keys -> ukeys/ucounts, count -> *m_dev.
// tile_heads: ceil(n/4096) words; upos: n words.
void gqf_launch_rle(const uint64_t* keys, uint64_t n, uint32_t* tile_heads, uint64_t* ukeys,
                    uint64_t* upos, uint64_t* ucounts, uint64_t* m_dev, hipStream_t s);

// region bounds of sorted items (count n, or *n_dev when n_dev is set),
// plus the non-empty regions of each parity compacted into region_list
// (2 lists of (nregions + 1) / 2 + 1 entries) with their counts
void gqf_launch_region_bounds(const QF* qf_dev, uint64_t nregions, const uint64_t* items, uint64_t n,
                              const uint64_t* n_dev, uint32_t* region_start, uint32_t* region_list,
                              uint32_t* list_count, uint64_t hmask, hipStream_t s);
// region_start entries that gqf_launch_sort_two_level leaves (bounds at or
// past range, and r = nregions: n) and the non-empty region lists, as
// gqf_launch_region_bounds makes them
void gqf_launch_region_lists(const QF* qf_dev, uint64_t nregions, uint64_t n, uint32_t* region_start,
                             uint32_t* region_list, uint32_t* list_count, hipStream_t s);

// Sequenced batches: a sorted item is its slot hash with the request tag
// above GQF_SEQ_SHIFT -- the request's sequence number in bits 48..62 and
// bit 63 set for a delete.  Even radix passes of any digit the sort takes
// (6..10 bits) over at most 40 hash bits stay below bit 48.
#define GQF_SEQ_SHIFT 48
#define GQF_SEQ_HASH_MASK ((1ULL << GQF_SEQ_SHIFT) - 1)
#define GQF_REBUILD_INSERT 0
#define GQF_REBUILD_DELETE 1
#define GQF_REBUILD_SEQ 2    // items tagged as above, applied in tag order per hash
// ORed into the mode: plain insert batches take the run-by-run walks
// instead of the shift path (both give the same table)
#define GQF_REBUILD_NO_SHIFT 0x100
// the mode flag the launchers add: GQF_REBUILD_NO_SHIFT when the
// environment sets GQF_SHIFT_PATH=0 (read once), else 0
int gqf_rebuild_mode_flags();

// request boundaries of a sequenced call, passed by value to its kernels:
// request r owns keys [end[r - 1], end[r]) and has tag[r] = seq << 1 | delete
struct gqf_seq_reqs {
  uint32_t nreq;
  uint32_t end[QF_SEQ_MAX_REQUESTS];
  uint32_t tag[QF_SEQ_MAX_REQUESTS];
};
// keys[i] = slot hash of keys[i] | its request's tag
void gqf_launch_seq_tag(const QF* qf_dev, uint64_t* keys, uint64_t n, const gqf_seq_reqs& reqs, hipStream_t s);
// counts[i] (a count before the sequenced mutation `mut`, n_mut items sorted
// by hash) += the effect of the mutation items of its hash whose request
// precedes query i's request (reqs.tag holds the queries' sequence numbers)
void gqf_launch_seq_adjust(const QF* qf_dev, const uint64_t* keys, uint64_t n, const gqf_seq_reqs& reqs,
                           uint64_t* counts, const uint64_t* mut, uint64_t n_mut, hipStream_t s);

// both even/odd phases (+ wide fallback passes) of a merge-rebuild over
// the listed regions; n_items bounds the number of non-empty regions.
// wide_flags: one byte per region; wide_count: GQF_WIDE_WORDS words (the
// number of regions flagged in a phase, the first GQF_WIDE_LIST of them, a
// counter of finished wide workgroups, a ticket counter and one finished
// counter per class of the wide pass); all zero between calls (the kernels
// leave them so)
// the class waits' poll bound (qf_set_wide_wait_bound); returns the previous one
uint32_t gqf_wide_wait_bound_set(uint32_t max_polls);
#define GQF_WIDE_LIST 64
#define GQF_WIDE_MAX_CLASSES 8  // classes of the wide pass: ceil(ceil(256 / R) / 2), R >= 16
#define GQF_WIDE_WORDS (3 + GQF_WIDE_LIST + GQF_WIDE_MAX_CLASSES)
void gqf_launch_rebuild(const QF* qf_dev, uint64_t nregions, uint64_t region_blocks, const uint64_t* items,
                        const uint64_t* counts, uint64_t n_items, const uint32_t* region_start,
                        const uint32_t* region_list, const uint32_t* list_count, int mode, uint8_t* wide_flags,
                        uint32_t* wide_count, hipStream_t s);

void gqf_launch_bulk_get(const QF* qf_dev, const uint64_t* vals, uint64_t n, uint64_t* returns, hipStream_t s);
void gqf_launch_misses(const QF* qf_dev, const uint64_t* vals, uint64_t n, unsigned long long* misses,
                       hipStream_t s);

// device layout -> the reference's packed 81-byte image (nblocks * 81 bytes)
void gqf_launch_export(const QF* qf_dev, uint64_t nblocks, uint8_t* out, hipStream_t s);

// point API drivers: one lane per key (rets / counts / misses may be null)
void gqf_launch_point_insert(QF* qf_dev, const uint64_t* keys, uint64_t n, uint8_t flags, int32_t* rets,
                             int not_exists, hipStream_t s);
// bulk insert of hashed keys one lane per key through point_insert (the
// region locks); full table -> QF_STATUS_NO_SPACE
void gqf_launch_sparse_insert(QF* qf_dev, const uint64_t* hashes, uint64_t n, hipStream_t s);
void gqf_launch_point_query(QF* qf_dev, const uint64_t* keys, uint64_t n, uint8_t flags, uint64_t* counts,
                            int concurrent, unsigned long long* misses, hipStream_t s);

void gqf_launch_splitmix64(uint64_t* out, uint64_t n, uint64_t seed, uint64_t skip, hipStream_t s);

// ---------------------------------------------------------------------------
// Whole-table passes (gqf_enum.hip).  One thread per table block, 256
// blocks per workgroup.

#define GQF_ENUM_WG_BLOCKS 256  // table blocks per workgroup of these passes
// output forms of the enumeration write pass
enum { GQF_ENUM_OUT_HASH = 0, GQF_ENUM_OUT_KEY = 1, GQF_ENUM_OUT_MOVE = 2 };
uint64_t gqf_enum_wg_count(uint64_t nblocks);
// count pass over blocks [b_lo, b_hi), items of quotients [q_lo, q_hi):
// items per workgroup -> wg_items, their exclusive prefix sums -> wg_base
// (nwg + 1 entries, the last the total); stats[0..2] += items, sum of
// counts, slots used by the runs.  With `starts` (zeroed, one word per
// block from b_lo to the end of the table) it also sets the bit of every
// item's first slot and raises *max_word to the last word it marked.
void gqf_launch_enum_count(const QF* qf_dev, uint64_t b_lo, uint64_t b_hi, uint64_t q_lo, uint64_t q_hi,
                           uint32_t* wg_items, uint64_t* wg_base, unsigned long long* stats,
                           unsigned long long* starts, unsigned long long* max_word, hipStream_t s);
// write pass over those marks (words w_lo .. w_lo + nwords): the item whose
// counter starts at the r-th mark -> hashes[r] / counts[r] / values[r] when
// r < cap, in the form `mode` (MOVE: the slot hash the item has in a filter
// with dst_vb value bits); clears the marks.  Scratch: local_prefix nwords
// words, group_sums and group_base (+1) one per 256 words.
void gqf_launch_enum_emit(const QF* qf_dev, uint64_t w_lo, uint64_t nwords, unsigned long long* starts,
                          const unsigned long long* max_word, uint32_t* local_prefix, uint32_t* group_sums,
                          uint64_t* group_base, int mode, int vb, int dst_vb, uint64_t key_mask, uint64_t* hashes,
                          uint64_t* counts, uint64_t* values, uint64_t cap, hipStream_t s);
// cluster census: used-slot masks (nblocks words), cluster starts per
// workgroup and their bases (nwg + 1, the last = number of clusters);
// lengths[rank] (rank < cap), hist[min(len, nbins - 1)] += 1 (nbins <= 4096,
// 0 for none), *max_len = max(*max_len, longest)
// *bad += number of blocks whose stored used-slot mask disagrees with
// their header
void gqf_launch_check_used(const QF* qf_dev, uint64_t nblocks, unsigned long long* bad, hipStream_t s);
void gqf_launch_census(const QF* qf_dev, uint64_t nblocks, uint64_t* used, uint32_t* wg_starts, uint64_t* wg_base,
                       uint64_t* lengths, uint64_t cap, unsigned long long* hist, uint32_t nbins,
                       unsigned long long* max_len, hipStream_t s);

// Per-kernel event timing (off by default).  When enabled, the launchers
// bracket each kernel with hipEvents on its stream; gqf_profile_read sums
// the completed intervals of one kernel name.
void gqf_prof_begin(const char* name, hipStream_t s);
void gqf_prof_end(const char* name, hipStream_t s);
// while on (this thread), gqf_prof_begin / _end do nothing
void gqf_prof_suppress(bool on);

// ---------------------------------------------------------------------------
// Filter groups (gqf_group.hip): several filters served by shared launches.

// Device-side tables of a group, passed by value to its kernels.
struct gqf_group_dev {
  const QF* descs;            // nf filter descriptors (copies of each filter's)
  const uint64_t* seg;        // nf + 1 offsets of each filter's keys
  const uint64_t* even_base;  // nf + 1 prefix sums of the filters' even-region counts
  const uint64_t* odd_base;   // nf + 1 prefix sums of the odd-region counts
  const uint64_t* reg_base;   // nf + 1 prefix sums of the region counts (flag index)
  uint32_t nf;
  int tag_shift;              // hash bits below the filter tag
  uint32_t* counters;         // [1] epoch of the done flags, [2] wide-redo count, [3] finished wide workgroups
  uint32_t* done;             // per region: epoch in which it finished
  uint8_t* wide;              // per region: redo with the wide window
  uint32_t max_polls;         // bound of an odd region's wait for an even neighbour's done flag
};
#define GQF_GROUP_MAX_POLLS (1u << 22)  // default bound: about 1 s of polls, far above any region's rebuild

// the persistent rebuild over every filter's regions (even items, then odd
// items that wait for their even neighbours), then the wide redo tail
void gqf_launch_group_rebuild(const gqf_group_dev& g, const uint64_t* items, int mode, uint32_t grid,
                              hipStream_t s);
#define GQF_GROUP_PER_CU 8  // workgroups per CU of the persistent rebuild unless tuned
uint32_t gqf_group_rebuild_max_per_cu();
uint32_t gqf_group_rebuild_grid(uint32_t per_cu);
// returns[i] = count of vals[i] in the filter whose segment holds i
void gqf_launch_group_get(const gqf_group_dev& g, const uint64_t* vals, uint64_t n, uint64_t* returns,
                          hipStream_t s);
// the host copy of a filter's descriptor (false if the handle is unknown)
bool gqf_filter_host(QF* handle, QF* out);

