/*
 * gqf.h -- C ABI of libgqf, a counting quotient filter for one AMD MI355X
 * (gfx950).  Drop-in for the GQF API of PASSIONLab/cqf (include/gqf.cuh).
 *
 * Signatures, names and semantics of the reference entry points are kept:
 *   - QF* is an opaque DEVICE pointer owned by the library;
 *   - keys / vals / returns are DEVICE arrays owned by the caller;
 *   - bulk_insert, bulk_insert_reduce and bulk_delete overwrite `keys` with
 *     the ascending hashes (src/gqf.cu:3180-3181, 3206-3209);
 *   - bulk calls are asynchronous on the null stream (src/gqf.cu:3184-3419);
 *     the *_wrapper calls synchronise and return a host value.
 * Differences, all additive: nvals == 0 is a no-op (the reference underflows
 * its grid size); a mutation of more than 2^30 keys runs as pieces of 2^30
 * (see qf_batch_pieces); table-full and cluster-overflow conditions are recorded
 * in a status word (qf_get_status) instead of being printed and ignored;
 * stream-taking variants and image export/import exist for tests and the
 * serving loop.  No signature carries a HIP or torch type: streams are
 * passed as void* (a hipStream_t, or NULL for the null stream).
 */
#ifndef GQF_H
#define GQF_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct quotient_filter quotient_filter;
typedef quotient_filter QF;

/* include/gqf.cuh:45-49 */
enum qf_hashmode { QF_HASH_DEFAULT, QF_HASH_INVERTIBLE, QF_HASH_NONE };
/* include/gqf.cuh:52-56 */
enum qf_returns { QF_ITEM_INSERTED, QF_ITEM_FOUND, QF_FULL };

/* include/gqf.cuh:74-80 */
#define QF_NO_LOCK (0x01)
#define QF_TRY_ONCE_LOCK (0x02)
#define QF_WAIT_FOR_LOCK (0x04)
#define QF_KEY_IS_HASH (0x08)

/* include/gqf.cuh:213-215 */
#define QF_NO_SPACE (-1)
#define QF_COULDNT_LOCK (-2)
#define QF_DOESNT_EXIST (-3)

/* status bits (qf_get_status) */
#define QF_STATUS_NO_SPACE 0x1u
#define QF_STATUS_CLUSTER_OVERFLOW 0x2u
#define QF_STATUS_BAD_HASH 0x4u
#define QF_STATUS_POINT_FULL 0x8u

/* ---- lifecycle ----------------------------------------------------- */

/* replaces include/gqf.cuh:87 (src/gqf.cu:2123-2221): 2^nbits slots, 8-bit
 * remainders, invertible hash, seed 0.  The point-API region locks are
 * allocated whatever bulk_config says. */
void qf_malloc_device(QF **qf, int nbits, bool bulk_config);
/* replaces include/gqf.cuh:88 (src/gqf.cu:2225-2269) */
void qf_destroy_device(QF *qf);

/* ---- bulk API ------------------------------------------------------ */

/* replaces include/gqf.cuh:113 (src/gqf.cu:3184-3232) */
void bulk_insert(QF *qf, uint64_t nvals, uint64_t *keys, uint8_t flags);
/* replaces include/gqf.cuh:116 (src/gqf.cu:3238-3314) */
void bulk_insert_reduce(QF *qf, uint64_t nvals, uint64_t *keys, uint8_t flags);
/* replaces include/gqf.cuh:119 (src/gqf.cu:3316-3363) */
void bulk_delete(QF *qf, uint64_t nvals, uint64_t *keys, uint8_t flags);
/* replaces include/gqf.cuh:122 (src/gqf.cu:3397-3401); vals unchanged */
void bulk_get(QF *qf, uint64_t nvals, uint64_t *vals, uint64_t *returns);
/* replaces include/gqf.cuh:140 (src/gqf.cu:3404-3419): keys with count < 1 */
uint64_t bulk_get_misses_wrapper(QF *qf, uint64_t *vals, uint64_t nvals);

/* ---- point API, host side ------------------------------------------ */

/* replaces include/gqf.cuh:134 (src/gqf.cu:3112-3150): point_query of
 * every key, returns the number not found */
uint64_t point_get_wrapper(QF *qf, uint64_t *hashes, uint64_t nitems);
/* replaces the point_bulk_insert kernel launch of include/gqf.cuh:136
 * (src/gqf.cu:3153-3162; launched by include/point_wrapper.cuh:63): one
 * locked point_insert per key; synchronises. */
void point_bulk_insert_wrapper(QF *qf, uint64_t *hashes, uint64_t nitems);
/* the device point functions (gqf_device.h) over a device array of keys,
 * one lane per key, stream-ordered; rets[i] / counts[i] receive each
 * call's return value (either may be NULL).  point_insert_async is the
 * point_bulk_insert kernel (src/gqf.cu:3153-3162) with return codes. */
void point_insert_async(QF *qf, uint64_t nvals, uint64_t *keys, uint8_t flags, int32_t *rets, void *stream);
void point_insert_not_exists_async(QF *qf, uint64_t nvals, uint64_t *keys, uint8_t flags, int32_t *rets,
                                   void *stream);
void point_query_async(QF *qf, uint64_t nvals, uint64_t *keys, uint8_t flags, uint64_t *counts, int concurrent,
                       void *stream);

/* ---- additions (no reference counterpart) -------------------------- */

typedef struct gqf_info {
  uint64_t nslots, xnslots, nblocks, key_bits, key_remainder_bits, bits_per_slot;
  uint64_t range, num_locks, nregions;
  uint64_t total_size_in_bytes;  /* bytes of the 81-byte-block image (as the reference's) */
  uint64_t device_bytes;         /* bytes of the device table: 128 per block + 4096 slack */
  uint64_t region_blocks;        /* blocks per rebuild region (tuning) */
  uint64_t sort_digit_bits;      /* radix digit of the bulk sort, 0 = automatic (tuning) */
  uint64_t value_bits;           /* 0, or the slot bits below the hash a resize added */
} gqf_info;

int qf_get_info(QF *qf, gqf_info *out);
/* OR of QF_STATUS_* bits since the last clear (synchronises) */
uint32_t qf_get_status(QF *qf);
void qf_clear_status(QF *qf);
/* copy the 81-byte block image (total_size_in_bytes) to / from host memory */
int qf_export_image(QF *qf, void *host_dst, uint64_t nbytes);
int qf_import_image(QF *qf, const void *host_src, uint64_t nbytes);
/* the same image in device memory: export is stream-ordered, import
 * synchronises the stream (offsets are rebuilt on the host) */
int qf_export_image_device(QF *qf, void *dev_dst, uint64_t nbytes, void *stream);
int qf_import_image_device(QF *qf, const void *dev_src, uint64_t nbytes, void *stream);
/* empty the table (stream-ordered); mirrors qf_reset, src/gqf.cu:2283-2299 */
int qf_reset_device(QF *qf, void *stream);
/* size the per-filter workspace for batches up to max_batch keys so that
 * later bulk calls allocate nothing */
int qf_reserve(QF *qf, uint64_t max_batch);
/* tuning (0 keeps a value): blocks per rebuild region, 16..56 (the push
 * window is min(64, 2 x region) minus that), and radix digit bits of the
 * bulk sort, 6..10; the results do not depend on them, only the speed does */
int qf_set_tuning(QF *qf, uint32_t region_blocks, uint32_t sort_digit_bits);
/* No counterpart in the reference.  Bound on the polls of each wait between
 * the classes of the wide pass (more than 64 regions of a rebuild phase
 * whose push outran their window; default 2^24, which no correct run
 * reaches).  A wait that runs out leaves that class's flagged regions
 * unchanged and sets QF_STATUS_CLUSTER_OVERFLOW on the filter, so every
 * kernel still exits.  Process-wide; returns the previous bound.  For
 * tests of the failure path. */
uint32_t qf_set_wide_wait_bound(uint32_t max_polls);
/* stream-ordered variants of the bulk calls */
void bulk_insert_async(QF *qf, uint64_t nvals, uint64_t *keys, uint8_t flags, void *stream);
void bulk_insert_reduce_async(QF *qf, uint64_t nvals, uint64_t *keys, uint8_t flags, void *stream);
void bulk_delete_async(QF *qf, uint64_t nvals, uint64_t *keys, uint8_t flags, void *stream);
void bulk_get_async(QF *qf, uint64_t nvals, uint64_t *vals, uint64_t *returns, void *stream);

/* bulk_insert_async without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same result and key contract as bulk_insert.  Its workspace is a fixed
 * device array of the library (20.5 MB per GPU, held until the process
 * ends), shared by every call on that GPU: calls that overlap on one GPU
 * race, so issue them on one stream or synchronise between them.  Takes at
 * most 2^20 keys per call and filters of at most 1,200,000 rebuild regions
 * (2^30 slots at the smallest region size); returns 0, -1 for an unknown
 * handle, -2 for a batch or filter beyond those limits (nothing launched),
 * -3 for a HIP error. */
int bulk_insert_noalloc(QF *qf, uint64_t nvals, uint64_t *keys, uint8_t flags, void *stream);

/* ---- sequenced batches (no reference counterpart) -------------------
 * One filter's requests, in arrival order, as a few launches instead of
 * one bulk call each, with the results of the calls one by one.  For
 * serving loops that coalesce requests of all kinds.
 *
 * A call covers nreq <= QF_SEQ_MAX_REQUESTS requests of one device key
 * array: request r owns keys [req_end[r - 1], req_end[r]) (req_end[-1] = 0,
 * req_end[nreq - 1] = n; host array) and has sequence number seq_r, which
 * must ascend with r and stay below 2^15.
 *
 * qf_bulk_mutate_seq applies inserts and deletes (req_tag[r] = seq_r << 1,
 * | 1 for a delete; host array) as bulk_insert / bulk_delete called one
 * request at a time: on each hash, inserts add 1 and deletes remove 1 down
 * to 0, in request order.  One sort and one rebuild for all requests.
 * keys is left holding the tagged hashes sorted by hash: slot hash | seq_r
 * << 48 | delete << 63 (the input of qf_bulk_get_seq_adjust).  Filters of
 * at most 2^32 slots (key bits + value bits <= 40).
 *
 * qf_bulk_get_seq_adjust completes queries of the same batch: counts
 * holds bulk_get of the query keys taken BEFORE the mutation (stream
 * order), req_seq[r] the sequence number of query request r; each count
 * becomes the count after exactly the mutation requests whose sequence
 * numbers are below its request's.  mut / n_mut: the sorted keys the
 * mutation call left.
 *
 * Both return 0, or a negative value for bad arguments (nothing launched).
 * Stream-ordered; calls on one filter must use one stream. */
#define QF_SEQ_MAX_REQUESTS 256
int qf_bulk_mutate_seq(QF *qf, uint64_t n, uint64_t *keys, uint32_t nreq, const uint32_t *req_end,
                       const uint32_t *req_tag, void *stream);
int qf_bulk_get_seq_adjust(QF *qf, uint64_t n, const uint64_t *keys, uint32_t nreq, const uint32_t *req_end,
                           const uint32_t *req_seq, uint64_t *counts, uint64_t n_mut, const uint64_t *mut,
                           void *stream);

/* ---- filter groups and tuning (no reference counterpart) ---------- */

/* A group of existing filters served by shared launches.  keys / vals are
 * ONE device array cut into one segment per filter: segment f is
 * [offsets[f], offsets[f+1]) (offsets: host array of nfilters + 1, offsets[0]
 * = 0).  gqf_group_bulk_insert gives every segment bulk_insert semantics
 * (the segment is left holding its filter's ascending hashes) with one
 * radix sort and ONE persistent rebuild launch for all filters, captured
 * into a hipGraph per shape and replayed; gqf_group_bulk_get fills
 * returns[i] with the count of vals[i] in its segment's filter, in one
 * launch.  Stream-ordered; a call whose segments or filter tuning differ
 * from the previous one synchronises first.  A group insert takes at most
 * 2^30 keys in all (-2 above that). */
typedef struct gqf_group gqf_group;
int gqf_group_create(gqf_group **g, QF *const *filters, uint32_t nfilters);
void gqf_group_destroy(gqf_group *g);
int gqf_group_bulk_insert(gqf_group *g, uint64_t *keys, const uint64_t *offsets, uint8_t flags, void *stream);
int gqf_group_bulk_get(gqf_group *g, const uint64_t *vals, const uint64_t *offsets, uint64_t *returns, void *stream);
/* graphs on (default) or off (every call launches eagerly) */
int gqf_group_set_graphs(gqf_group *g, int on);
/* autotuner: times bulk_insert of a 2^nbits-slot table loaded to `load`
 * in one batch for each candidate region size and sort digit, returns the
 * fastest (cached per nbits and load rounded to 0.05); synchronises */
int gqf_autotune(int nbits, double load, uint32_t *region_blocks, uint32_t *sort_digit_bits);
/* records a choice in the autotuner's cache (as a measurement would): a
 * table of choices measured once and committed (cqf_amd/tuning_gfx950.json)
 * is loaded this way, so later gqf_autotune calls for those sizes time
 * nothing.  -2 for nbits outside 8..32, load outside (0, 1], a region size
 * outside 16..56 or a digit width other than 0 (default) or 6..10 */
int gqf_autotune_set(int nbits, double load, uint32_t region_blocks, uint32_t sort_digit_bits);
/* gqf_autotune for every filter of the group at the load its segment
 * brings, applied with qf_set_tuning; then the workgroups per CU of the
 * group's persistent rebuild, timed for a batch cut like `offsets` on
 * scratch filters of the same sizes (candidates 4..16 up to what the GPU
 * keeps resident); synchronises */
int gqf_group_autotune(gqf_group *g, const uint64_t *offsets);
/* bound on the polls of an odd region's wait for its even neighbours in
 * the persistent rebuild (default 2^22).  A region whose wait runs out is
 * not rebuilt in that launch but redone by the serial wide pass after it,
 * so results never depend on the bound; 0 sends every odd region there. */
int gqf_group_set_wait_bound(gqf_group *g, uint32_t max_polls);
/* workgroups per CU of the persistent rebuild (0: the default, 8); above
 * what the GPU keeps resident -> -2.  The getter reports the value used. */
int gqf_group_set_workgroups_per_cu(gqf_group *g, uint32_t per_cu);
int gqf_group_get_workgroups_per_cu(gqf_group *g, uint32_t *per_cu);

/* ---- serving loop (no reference counterpart) -----------------------
 * Many resident filters, requests of any size and kind arriving over time,
 * served by continuous batching on a native thread: each round, every
 * filter with queued requests and fewer than max_inflight batches on the
 * device takes the head of its queue -- gets, inserts and deletes in
 * arrival order, up to max_batch_keys keys and QF_SEQ_MAX_REQUESTS requests
 * per kind -- as one sequenced batch (bulk_get_async, qf_bulk_mutate_seq,
 * qf_bulk_get_seq_adjust) on the filter's stream, filters spread over
 * n_streams streams.  Each get returns the counts after exactly the
 * requests that arrived before it, as if the requests ran one at a time.
 * A filter is created on its first request if its device memory fits
 * budget_bytes; otherwise least recently used idle filters are evicted to
 * host images (qf_export_image) until it fits, and re-admitted by import.
 * plain != 0 makes the baseline instead: every request its own bulk call,
 * in arrival order, on one stream, every filter resident.
 *
 * Use: create, add the filters, submit every request with its arrival time
 * (ns after qf_server_run; times must not decrease), run, wait.  keys is a
 * device array of n <= max_batch_keys keys that stays valid and unchanged
 * until the request is done; a get's counts is a device array of n counts.
 * Returns 0 / a request id, or < 0 for bad arguments. */
#define QF_SERVER_GET 0
#define QF_SERVER_INSERT 1
#define QF_SERVER_DELETE 2
typedef struct qf_server qf_server;
typedef struct qf_server_stats {
  uint64_t batches, evictions, admissions;
} qf_server_stats;
int qf_server_create(qf_server **s, uint64_t budget_bytes, uint32_t n_streams, uint64_t max_batch_keys,
                     uint32_t max_inflight, int plain);
int qf_server_add_filter(qf_server *s, uint32_t fid, int nbits);
int64_t qf_server_submit(qf_server *s, uint64_t t_arrival_ns, uint32_t fid, int op, const uint64_t *keys,
                         uint64_t n, uint64_t *counts);
/* starts the loop thread; the arrival clock starts here */
int qf_server_run(qf_server *s);
/* blocks until every submitted request is done; *wall_ns = completion time
 * of the last one (ns after qf_server_run) */
int qf_server_wait(qf_server *s, uint64_t *wall_ns);
/* times of one request (ns after qf_server_run); returns 1 if it is done */
int qf_server_request_times(qf_server *s, int64_t rid, uint64_t *t_arrival_ns, uint64_t *t_launch_ns,
                            uint64_t *t_done_ns);
int qf_server_get_stats(qf_server *s, qf_server_stats *out);
/* the resident handle of a filter (NULL if evicted or unknown; not while running) */
QF *qf_server_filter(qf_server *s, uint32_t fid);
void qf_server_destroy(qf_server *s);

/* ---- whole-table operations ------------------------------------- */

/* qf_malloc (src/gqf.cu:2073-2105) with value bits: 2^nbits slots of 8
 * bits, key_bits = nbits + 8 - value_bits, value_bits 0..6 (the reference
 * asserts key_remainder_bits >= 2, src/gqf.cu:1957).  Keys are stored as
 * the slot hash (hash_64(key) << value_bits); values are always 0.  This
 * is the layout qf_resize_device produces. */
void qf_malloc_device_kv(QF **qf, int nbits, int value_bits);

/* Output forms of qf_enumerate */
#define QF_ENUM_HASH 0 /* the hash, (quotient << key_remainder_bits) | remainder (qfi_get_hash) */
#define QF_ENUM_KEY 1  /* the key, hash_64i of the hash (qfi_get_key, src/gqf.cu:3791-3805) */

/* Items of the runs of quotients in [q_lo, q_hi), in ascending hash order:
 * the sequence the reference's iterator visits from
 * qf_iterator_from_position(q_lo) (src/gqf.cu:3655-3700, qfi_next
 * :3814-3840), one (hash or key, count, value) per counter.  hashes, counts
 * and values are device arrays of cap entries, each may be NULL; *n_items
 * (host) receives the number of items in the range, of which the first
 * min(n_items, cap) are written.  q_hi is clamped to nslots.  Synchronises
 * `stream`.  Returns 0, or < 0 for a bad handle or form. */
int qf_enumerate(QF *qf, uint64_t q_lo, uint64_t q_hi, int form, uint64_t *hashes, uint64_t *counts,
                 uint64_t *values, uint64_t cap, uint64_t *n_items, void *stream);

/* The metadata counters the reference declares but whose GPU updates are
 * commented out (src/gqf.cu:1094, 1384-1401, 1545-1547; read by
 * qf_get_sum_of_counts, qf_get_num_distinct_key_value_pairs and
 * qf_get_num_occupied_slots, src/gqf.cu:3619-3643), computed from the
 * table: sum of counts, items, and slots the runs use.  Synchronises. */
int qf_get_stats(QF *qf, uint64_t *nelts, uint64_t *ndistinct_elts, uint64_t *noccupied_slots);

/* Cluster census (src/cluster_length_test.cu:37-120): the lengths of the
 * maximal stretches of used slots, in slot order.  lengths: device array
 * of cap entries or NULL; hist: device array of nbins (<= 4096) u64
 * counters or NULL, ADDED to (bin i counts clusters of length i, the last
 * bin those of length >= nbins - 1); *num_clusters and *max_len are host
 * outputs, either may be NULL.  Allocates 8 bytes per 64-slot block for
 * the call.  Synchronises. */
int qf_cluster_census(QF *qf, uint64_t *lengths, uint64_t cap, uint64_t *hist, uint32_t nbins,
                      uint64_t *num_clusters, uint64_t *max_len);

/* File format: the reference's qfmetadata (120 bytes,
 * include/gqf_int.cuh:96-113, magic 1018874902021329732) followed by the
 * packed 81-byte blocks -- the buffer qf_use maps (src/gqf.cu:2025-2053),
 * for the qf_serialize / qf_deserialize the reference declares
 * (include/gqf_file.cuh:45-49).  nelts, ndistinct_elts and noccupied_slots
 * are written from qf_get_stats.  Returns the bytes written, 0 on error. */
uint64_t qf_serialize(QF *qf, const char *filename);
/* A new device filter from such a file; the metadata (magic, invertible
 * hash mode, 8-bit slots, sizes) and the block image are validated.
 * Returns the bytes read, 0 on error (*qf = NULL then). */
uint64_t qf_deserialize_device(QF **qf, const char *filename);

/* Grows the table to 2^new_nbits slots keeping the hash (key_bits): every
 * (hash, count) moves to the larger table, as qf_resize_malloc re-inserts
 * them (src/gqf.cu:2301-2330).  The freed remainder bits become value
 * bits, so new_nbits - nbits + value_bits must stay <= 6.  The handle stays
 * valid; the table and workspace are reallocated.  Returns the number of
 * items moved, or < 0 on error (the filter is unchanged then). */
int64_t qf_resize_device(QF *qf, int new_nbits);

/* qfc += every item of qfa and qfb, counts added (qf_merge,
 * src/gqf.cu:3896-3938).  All filters must have the same key_bits, and qfc
 * at least the value bits of each input.  Returns the number of items
 * inserted, or < 0 on error. */
int64_t qf_merge_device(QF *qfa, QF *qfb, QF *qfc);
/* qfr += every item of the nqf filters (qf_multi_merge, src/gqf.cu:3943-4008) */
int64_t qf_multi_merge_device(QF *const *qfs, int nqf, QF *qfr);

/* How a bulk mutation of n keys is cut: pieces of at most max_piece keys
 * (the library uses 2^30), piece k = [starts[k], starts[k + 1]).  Returns
 * the number of pieces p and fills starts[0..p] if cap >= p + 1.  A batch of
 * more than 2^30 keys is applied piece by piece; each piece of `keys` is
 * then left holding its own ascending hashes (the reference sorts the whole
 * array, src/gqf.cu:3209).  Pure host arithmetic. */
uint32_t qf_batch_pieces(uint64_t n, uint64_t max_piece, uint64_t *starts, uint32_t cap);

/* Consistency check of the device table (no reference counterpart; for
 * tests): *bad_blocks = number of blocks whose stored used-slot mask (an
 * internal field of the device layout) disagrees with their runs.
 * Synchronises; returns 0 or < 0 on error. */
int qf_check_layout(QF *qf, uint64_t *bad_blocks);

/* last HIP error seen by the library (0 = none); clears it */
int qf_last_error(void);

/* ---- measurement support (no reference counterpart) ---------------- */

/* per-kernel hipEvent timing on the launch stream: enable (clears), then
 * read the summed milliseconds and launch count of one kernel name:
 * "sort_pass", "region_bounds", "rebuild_phase", "rebuild_wide", "bulk_get",
 * "enum_count", "enum_emit", "census_mask", "census_write" */
void gqf_profile_enable(int on);
int gqf_profile_read(const char *name, double *total_ms, uint64_t *launches);
/* dev_out[i] = splitmix64 output number skip+i of the stream seeded with seed */
void gqf_fill_splitmix64(uint64_t *dev_out, uint64_t n, uint64_t seed, uint64_t skip, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GQF_H */
