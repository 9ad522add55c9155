// gqf_launch.h -- host-side launchers of the libgqf kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gqf_internal.h"

// radix sort of n keys in place (tmp holds n keys); the first pass hashes
// with `mask` when hash_it is set.  hist must hold gqf_sort_hist_words(n).
uint64_t gqf_sort_hist_words(uint64_t n);
void gqf_launch_sort(uint64_t* keys, uint64_t* tmp, uint64_t n, int hash_it, uint64_t mask, int bits,
                     uint32_t* hist, uint32_t* totals, hipStream_t s);

// run-length reduce of sorted keys -> ukeys/ucounts, count -> *m_dev.
// tile_heads: ceil(n/4096) words; upos: n words.
void gqf_launch_rle(const uint64_t* keys, uint64_t n, uint32_t* tile_heads, uint64_t* ukeys,
                    uint64_t* upos, uint64_t* ucounts, uint64_t* m_dev, hipStream_t s);

// region bounds of sorted items (count n, or *n_dev when n_dev is set)
void gqf_launch_region_bounds(const QF* qf_dev, uint64_t nregions, const uint64_t* items, uint64_t n,
                              const uint64_t* n_dev, uint32_t* region_start, hipStream_t s);

// both even/odd phases (+ wide fallback passes) of a merge-rebuild
void gqf_launch_rebuild(const QF* qf_dev, uint64_t nregions, const uint64_t* items, const uint64_t* counts,
                        const uint32_t* region_start, int mode, uint8_t* wide_flags, uint32_t* wide_count,
                        hipStream_t s);

void gqf_launch_bulk_get(const QF* qf_dev, const uint64_t* vals, uint64_t n, uint64_t* returns, hipStream_t s);
void gqf_launch_misses(const QF* qf_dev, const uint64_t* vals, uint64_t n, unsigned long long* misses,
                       hipStream_t s);

// device layout -> the reference's packed 81-byte image (nblocks * 81 bytes)
void gqf_launch_export(const QF* qf_dev, uint64_t nblocks, uint8_t* out, hipStream_t s);

// point API drivers: one lane per key (rets / counts / misses may be null)
void gqf_launch_point_insert(QF* qf_dev, const uint64_t* keys, uint64_t n, uint8_t flags, int32_t* rets,
                             int not_exists, hipStream_t s);
void gqf_launch_point_query(QF* qf_dev, const uint64_t* keys, uint64_t n, uint8_t flags, uint64_t* counts,
                            int concurrent, unsigned long long* misses, hipStream_t s);

void gqf_launch_splitmix64(uint64_t* out, uint64_t n, uint64_t seed, uint64_t skip, hipStream_t s);

// Per-kernel event timing (off by default).  When enabled, the launchers
// bracket each kernel with hipEvents on its stream; gqf_profile_read sums
// the completed intervals of one kernel name.
void gqf_prof_begin(const char* name, hipStream_t s);
void gqf_prof_end(const char* name, hipStream_t s);
