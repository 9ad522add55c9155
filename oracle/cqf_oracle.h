/*
 * cqf_oracle.h -- CPU restatement of the reference GQF host path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker / the timed CPU baseline.
 *
 * Every routine restates an algorithm of PASSIONLab/cqf (file:line cited
 * at each definition in cqf_oracle.c) on the reference's own 81-byte
 * packed block image:
 *   byte 0        offset      (u8)
 *   bytes 1..8    occupieds   (u64, little endian)
 *   bytes 9..16   runends     (u64, little endian)
 *   bytes 17..80  slots[64]   (u8 remainders, QF_BITS_PER_SLOT = 8)
 * (include/gqf_int.cuh:38-57).
 *
 * How each part is pinned before it is trusted (tests/test_oracle.py; the
 * reference ships no golden vectors and its host path needs cuda.h and
 * Thrust, so it is not built here):
 *   hash_64 / hash_64i   multiplicative restatement, bijectivity, inverse
 *   counter codec        encoder written from the format comment; decode of
 *                        every encoding; insert1 increments reach it
 *   insert / counted insert / remove images
 *                        a layout builder written from the format's
 *                        invariants alone (tests/spec_model.py); full delete
 *                        gives the zero image
 *   counts               brute-force masked multisets
 *   sizing               closed-form table
 *   enumerate            the multiset the image was built from (canonical
 *                        images of spec_model.py and of counted inserts)
 *   cluster census       maximal used-slot stretches of layouts built by
 *                        spec_model.py from the format's invariants
 *   insert1_if_not_exists  return codes and returnedVal on fixed cases;
 *                        PARITY UNPINNED for its slot-wise search on runs
 *                        holding multi-slot counters (restated literally from
 *                        src/gqf.cu:1405-1440, no independent check exists)
 */
#ifndef CQF_ORACLE_H
#define CQF_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_BLOCK_BYTES 81
#define ORC_SLOTS_PER_BLOCK 64

/* Return codes of the per-key operations (include/gqf.cuh:52-56,213-215). */
#define ORC_NO_SPACE (-1)
#define ORC_ITEM_INSERTED 0
#define ORC_ITEM_FOUND 1
#define ORC_FULL 2

/* Flags (include/gqf.cuh:74-80). */
#define ORC_NO_LOCK 0x01
#define ORC_KEY_IS_HASH 0x08

typedef struct orc_qf {
    uint64_t nslots;             /* 2^q */
    uint64_t xnslots;            /* nslots + 10*sqrt(nslots), truncated */
    uint64_t nblocks;            /* ceil(xnslots / 64) */
    uint64_t key_bits;           /* q + 8 */
    uint64_t key_remainder_bits; /* 8 */
    uint64_t bits_per_slot;      /* 8 */
    uint64_t range;              /* 2^(q+8) */
    uint64_t num_locks;          /* xnslots / 8192 + 2 */
    uint64_t total_size_in_bytes;/* nblocks * 81 */
    uint8_t *image;              /* nblocks * 81 bytes, owned */
    uint64_t overflow_events;    /* find_first_empty_slot crossed > 1 region */
} orc_qf;

/* ---- sizing / lifecycle (src/gqf.cu:1939-2105, 2123-2140) ---- */
int      orc_init(orc_qf *qf, int nbits);
void     orc_free(orc_qf *qf);
void     orc_reset(orc_qf *qf);
uint64_t orc_sizes(int nbits, uint64_t out[8]);

/* ---- hashing (src/hashutil.cu:132-181) ---- */
uint64_t orc_hash_64(uint64_t key, uint64_t mask);
uint64_t orc_hash_64i(uint64_t key, uint64_t mask);
void     orc_hash_many(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t mask);

/* ---- counter codec (src/gqf.cu:1207-1345) ---- */
/* Writes the encoding of (remainder, count) in slot order to out[0..len),
 * returns len (0 for count 0).  out must hold 67 entries. */
int      orc_encode_counter(uint64_t remainder, uint64_t count, uint8_t *out);
/* decode_counter over a stand-alone run held in slots[0..n) whose last slot
 * is the run end; returns the index of the counter's last slot. */
uint64_t orc_decode_counter_run(const uint8_t *slots, uint64_t n, uint64_t index,
                                uint64_t *remainder, uint64_t *count);

/* ---- per-key operations, reference semantics ---- */
int      orc_qf_insert(orc_qf *qf, uint64_t key, uint64_t count, uint8_t flags);
int      orc_qf_remove(orc_qf *qf, uint64_t key, uint64_t count, uint8_t flags);
uint64_t orc_qf_count_key_value(const orc_qf *qf, uint64_t key, uint8_t flags);
uint64_t orc_qf_query(const orc_qf *qf, uint64_t key, uint64_t *value, uint8_t flags);
int      orc_qf_insert_not_exists(orc_qf *qf, uint64_t key, uint8_t flags, uint8_t *retval);
uint64_t orc_run_end(const orc_qf *qf, uint64_t hash_bucket_index);

/* ---- batch drivers (semantics of the bulk API, run sequentially) ---- */
/* bulk_insert: hash (unless KEY_IS_HASH), sort, insert1 each; keys are
 * overwritten with the ascending hashes (src/gqf.cu:3184-3232).
 * Returns the number of QF_NO_SPACE results. */
uint64_t orc_bulk_insert(orc_qf *qf, uint64_t *keys, uint64_t n, uint8_t flags);
/* bulk_insert_reduce: hash, sort, run-length reduce, counted insert
 * (src/gqf.cu:3238-3314). */
uint64_t orc_bulk_insert_reduce(orc_qf *qf, uint64_t *keys, uint64_t n, uint8_t flags);
/* bulk_delete: hash, sort, remove one instance each, reverse order within
 * each 8192-slot region (src/gqf.cu:3316-3363, 2860-2906). */
uint64_t orc_bulk_delete(orc_qf *qf, uint64_t *keys, uint64_t n, uint8_t flags);
/* bulk_get: returns[i] = qf_count_key_value(vals[i]) (src/gqf.cu:3386-3401). */
void     orc_bulk_get(const orc_qf *qf, const uint64_t *vals, uint64_t n, uint64_t *returns);
uint64_t orc_bulk_get_misses(const orc_qf *qf, const uint64_t *vals, uint64_t n);

/* ---- canonical image from a multiset (the layout the invariants fix) ---- */
/* items: ascending unique hashes (< range) with counts > 0.  Writes the
 * image into qf (which must be freshly reset).  Returns 0, or -1 when the
 * layout runs past xnslots. */
int      orc_build_canonical(orc_qf *qf, const uint64_t *hashes, const uint64_t *counts,
                             uint64_t n);
/* Enumerates (hash, count) items of the image in ascending hash order using
 * the reference decoder; returns the number written (<= cap). */
uint64_t orc_enumerate(const orc_qf *qf, uint64_t *hashes, uint64_t *counts, uint64_t cap);

/* Cluster lengths in slot order (the reference's census loop); returns the
 * number of clusters, writes the first cap lengths. */
uint64_t orc_cluster_census(orc_qf *qf, uint64_t *lengths, uint64_t cap);

/* ---- key generators ---- */
void     orc_splitmix64_fill(uint64_t seed, uint64_t skip, uint64_t n, uint64_t *out);

/* ---- timing helpers for the CPU baseline ---- */
double   orc_time_inserts(orc_qf *qf, const uint64_t *keys, uint64_t n);
double   orc_time_queries(const orc_qf *qf, const uint64_t *keys, uint64_t n, uint64_t *hits);
double   orc_time_insert_query(int nbits, const uint64_t *keys, uint64_t n_ins,
                               const uint64_t *qkeys, uint64_t n_q, uint64_t *hits,
                               double *t_insert, double *t_query);

#ifdef __cplusplus
}
#endif
#endif
