/*
 * cqf_oracle.c -- CPU restatement of the reference GQF host path.
 *
 * TEST INFRASTRUCTURE ONLY (see cqf_oracle.h).  It is the checker the GPU
 * library is compared against and the timed CPU baseline; the shipped
 * library never links or calls it.
 *
 * Each function names the reference routine it restates, as
 * src/gqf.cu:<lines> or src/hashutil.cu:<lines> of PASSIONLab/cqf.  The
 * control flow follows the reference step by step, including its quirks
 * (offset saturation at 255, the remove path's unsaturated offset store,
 * the reverse per-region order of bulk_delete), because the byte image is
 * what the GPU library must reproduce.  The only deliberate differences:
 *   - bounds guards where the reference would read or write past the table
 *     (it is undefined behaviour there; the oracle clamps and counts it);
 *   - the host select uses pdep/tzcnt when BMI2 is available and a portable
 *     byte-wise select otherwise (the reference aborts on the host without
 *     -D__SSE4_2_, src/gqf.cu:486,496-510).
 */
#include "cqf_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#if defined(__BMI2__)
#include <immintrin.h>
#endif

#define SPB 64ULL                       /* QF_SLOTS_PER_BLOCK, include/gqf_int.cuh:33-35 */
#define REGION_SLOTS 8192ULL            /* NUM_SLOTS_TO_LOCK, src/gqf.cu:58 */
#define BMASK(n) ((n) >= 64 ? ~0ULL : ((1ULL << (n)) - 1ULL)) /* src/gqf.cu:55-57 */
#define OFFSET_MAX 255ULL

/* ------------------------------------------------------------------ */
/* image accessors (include/gqf_int.cuh:38-57; src/gqf.cu:62-63,113-116,517-544) */

static inline uint64_t lim_slots(const orc_qf *qf) { return qf->nblocks * SPB; }

static inline uint8_t *blk(const orc_qf *qf, uint64_t b) { return qf->image + b * ORC_BLOCK_BYTES; }

static inline uint64_t ld64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }
static inline void st64(uint8_t *p, uint64_t v) { memcpy(p, &v, 8); }

static inline uint64_t occ_w(const orc_qf *qf, uint64_t b) { return b < qf->nblocks ? ld64(blk(qf, b) + 1) : 0; }
static inline uint64_t run_w(const orc_qf *qf, uint64_t b) { return b < qf->nblocks ? ld64(blk(qf, b) + 9) : 0; }
static inline void set_occ_w(orc_qf *qf, uint64_t b, uint64_t v) { if (b < qf->nblocks) st64(blk(qf, b) + 1, v); }
static inline void set_run_w(orc_qf *qf, uint64_t b, uint64_t v) { if (b < qf->nblocks) st64(blk(qf, b) + 9, v); }
static inline uint64_t off_b(const orc_qf *qf, uint64_t b) { return b < qf->nblocks ? blk(qf, b)[0] : 0; }
static inline void set_off_b(orc_qf *qf, uint64_t b, uint64_t v) { if (b < qf->nblocks) blk(qf, b)[0] = (uint8_t)v; }

static inline int is_runend(const orc_qf *qf, uint64_t i) { return (int)((run_w(qf, i / SPB) >> (i % SPB)) & 1ULL); }
static inline int is_occupied(const orc_qf *qf, uint64_t i) { return (int)((occ_w(qf, i / SPB) >> (i % SPB)) & 1ULL); }
static inline void runend_or(orc_qf *qf, uint64_t i) { set_run_w(qf, i / SPB, run_w(qf, i / SPB) | (1ULL << (i % SPB))); }
static inline void runend_clear(orc_qf *qf, uint64_t i) { set_run_w(qf, i / SPB, run_w(qf, i / SPB) & ~(1ULL << (i % SPB))); }
static inline void runend_xor(orc_qf *qf, uint64_t i) { set_run_w(qf, i / SPB, run_w(qf, i / SPB) ^ (1ULL << (i % SPB))); }
static inline void occ_or(orc_qf *qf, uint64_t i) { set_occ_w(qf, i / SPB, occ_w(qf, i / SPB) | (1ULL << (i % SPB))); }
static inline void occ_clear(orc_qf *qf, uint64_t i) { set_occ_w(qf, i / SPB, occ_w(qf, i / SPB) & ~(1ULL << (i % SPB))); }

static inline uint64_t get_slot(const orc_qf *qf, uint64_t i)
{
    if (i >= lim_slots(qf)) return 0;
    return blk(qf, i / SPB)[17 + i % SPB];
}
static inline void set_slot(orc_qf *qf, uint64_t i, uint64_t v)
{
    if (i >= lim_slots(qf)) return;
    blk(qf, i / SPB)[17 + i % SPB] = (uint8_t)(v & 0xFFULL);
}

/* ------------------------------------------------------------------ */
/* bit primitives (src/gqf.cu:143-215, 464-515) */

static inline int popcnt(uint64_t v) { return __builtin_popcountll(v); }

static inline int popcntv(uint64_t v, int ignore)
{
    if (ignore % 64) return popcnt(v & ~BMASK(ignore % 64));
    return popcnt(v);
}

/* number of ones in bits [0, pos] */
static inline int bitrank(uint64_t v, int pos) { return popcnt(v & ((2ULL << pos) - 1ULL)); }

/* position of the rank'th one (rank 0 = lowest), 64 when there are fewer */
static inline uint64_t bitselect(uint64_t v, uint64_t rank)
{
    if (rank >= (uint64_t)popcnt(v)) return 64;
#if defined(__BMI2__)
    return (uint64_t)__builtin_ctzll(_pdep_u64(1ULL << rank, v));
#else
    for (int byte = 0; byte < 8; byte++) {
        int c = popcnt((v >> (8 * byte)) & 0xFFULL);
        if (rank < (uint64_t)c) {
            uint64_t w = (v >> (8 * byte)) & 0xFFULL;
            while (rank--) w &= w - 1;
            return (uint64_t)(8 * byte + __builtin_ctzll(w));
        }
        rank -= (uint64_t)c;
    }
    return 64;
#endif
}

static inline uint64_t bitselectv(uint64_t v, int ignore, uint64_t rank)
{
    return bitselect(v & ~BMASK(ignore % 64), rank);
}

/* ------------------------------------------------------------------ */
/* run navigation (src/gqf.cu:617-771) */

static uint64_t run_end_impl(const orc_qf *qf, uint64_t h);

/* src/gqf.cu:619-630: an offset of 255 means "recompute from the run ends". */
static uint64_t block_offset(const orc_qf *qf, uint64_t b)
{
    uint64_t o = off_b(qf, b);
    if (o < OFFSET_MAX) return o;
    return run_end_impl(qf, SPB * b - 1) - SPB * b + 1;
}

/* src/gqf.cu:632-679 */
static uint64_t run_end_impl(const orc_qf *qf, uint64_t h)
{
    uint64_t bb = h / SPB, bio = h % SPB;
    uint64_t boff = block_offset(qf, bb);
    uint64_t rank = (uint64_t)bitrank(occ_w(qf, bb), (int)bio);
    if (rank == 0) {
        if (boff <= bio) return h;
        return SPB * bb + boff - 1;
    }
    uint64_t rb = bb + boff / SPB;
    int ignore = (int)(boff % SPB);
    uint64_t rrank = rank - 1;
    uint64_t ro = bitselectv(run_w(qf, rb), ignore, rrank);
    if (ro == SPB) {
        do {
            rrank -= (uint64_t)popcntv(run_w(qf, rb), ignore);
            rb++;
            ignore = 0;
            if (rb >= qf->nblocks) return lim_slots(qf); /* guard: corrupt / full */
            ro = bitselectv(run_w(qf, rb), ignore, rrank);
        } while (ro == SPB);
    }
    uint64_t ri = SPB * rb + ro;
    return ri < h ? h : ri;
}

uint64_t orc_run_end(const orc_qf *qf, uint64_t h) { return run_end_impl(qf, h); }

/* src/gqf.cu:681-704 (uses the raw stored offset byte, not block_offset) */
static int offset_lower_bound(const orc_qf *qf, uint64_t slot)
{
    uint64_t b = slot / SPB, so = slot % SPB;
    uint64_t boff = off_b(qf, b);
    uint64_t occ = occ_w(qf, b) & BMASK(so + 1);
    if (boff <= so) {
        uint64_t runends = (run_w(qf, b) & BMASK(so)) >> boff;
        return popcnt(occ) - popcnt(runends);
    }
    return (int)(boff - so) + popcnt(occ);
}

static inline int is_empty(const orc_qf *qf, uint64_t slot) { return offset_lower_bound(qf, slot) == 0; }

/* src/gqf.cu:712-716 */
static inline int might_be_empty(const orc_qf *qf, uint64_t slot)
{
    return !is_occupied(qf, slot) && !is_runend(qf, slot);
}

/* src/gqf.cu:728-771 */
static uint64_t find_first_empty_slot(orc_qf *qf, uint64_t from)
{
    uint64_t start = from;
    for (;;) {
        if (from >= lim_slots(qf)) break; /* guard: past the last block */
        int t = offset_lower_bound(qf, from);
        if (t <= 0) break;                 /* reference asserts t >= 0 */
        from += (uint64_t)t;
    }
    if (from / REGION_SLOTS > start / REGION_SLOTS + 1) qf->overflow_events++;
    return from;
}

/* ------------------------------------------------------------------ */
/* shifting (src/gqf.cu:828-1011) */

/* src/gqf.cu:828-837 */
static inline uint64_t shift_into_b(uint64_t a, uint64_t b, int bstart, int bend, int amount)
{
    uint64_t a_comp = bstart == 0 ? (amount ? (a >> (64 - amount)) : 0) : 0;
    uint64_t mask = BMASK((uint64_t)(bend - bstart)) << bstart;
    uint64_t b_shifted = ((mask & b) << amount) & mask;
    return a_comp | b_shifted | (b & ~mask);
}

/* src/gqf.cu:897-933 (memmove per block, then the carried slot) */
static void shift_remainders(orc_qf *qf, uint64_t start_index, uint64_t empty_index)
{
    uint64_t sb = start_index / SPB, so = start_index % SPB;
    uint64_t eb = empty_index / SPB, eo = empty_index % SPB;
    if (empty_index >= lim_slots(qf) || start_index > empty_index) return;
    while (sb < eb) {
        uint8_t *s = blk(qf, eb) + 17;
        memmove(s + 1, s, eo);
        s[0] = blk(qf, eb - 1)[17 + SPB - 1];
        eb--;
        eo = SPB - 1;
    }
    uint8_t *s = blk(qf, eb) + 17;
    memmove(s + so + 1, s + so, eo - so);
}

/* src/gqf.cu:964-971 */
static void find_next_n_empty_slots(orc_qf *qf, uint64_t from, uint64_t n, uint64_t *idx)
{
    while (n) {
        idx[--n] = find_first_empty_slot(qf, from);
        from = idx[n] + 1;
    }
}

/* src/gqf.cu:973-982 */
static void shift_slots(orc_qf *qf, int64_t first, uint64_t last, uint64_t distance)
{
    if (distance == 1) {
        shift_remainders(qf, (uint64_t)first, last + 1);
    } else {
        for (int64_t i = (int64_t)last; i >= first; i--)
            set_slot(qf, (uint64_t)i + distance, get_slot(qf, (uint64_t)i));
    }
}

/* src/gqf.cu:984-1011 */
static void shift_runends(orc_qf *qf, int64_t first, uint64_t last, uint64_t distance)
{
    uint64_t fw = (uint64_t)first / 64;
    int bstart = (int)((uint64_t)first % 64);
    uint64_t lw = (last + distance + 1) / 64;
    int bend = (int)((last + distance + 1) % 64);
    if (lw != fw) {
        set_run_w(qf, lw, shift_into_b(run_w(qf, lw - 1), run_w(qf, lw), 0, bend, (int)distance));
        bend = 64;
        lw--;
        while (lw != fw) {
            set_run_w(qf, lw, shift_into_b(run_w(qf, lw - 1), run_w(qf, lw), 0, bend, (int)distance));
            lw--;
        }
    }
    set_run_w(qf, lw, shift_into_b(0, run_w(qf, lw), bstart, bend, (int)distance));
}

/* src/gqf.cu:1013-1097 */
static int insert_replace_slots(orc_qf *qf, int operation, uint64_t bucket_index,
                                uint64_t overwrite_index, const uint64_t *rems,
                                uint64_t total, uint64_t noverwrites)
{
    uint64_t empties[67];
    int64_t ninserts = (int64_t)total - (int64_t)noverwrites;
    uint64_t insert_index = overwrite_index + noverwrites;

    if (ninserts > 0) {
        find_next_n_empty_slots(qf, insert_index, (uint64_t)ninserts, empties);
        if (empties[0] >= qf->xnslots) return 0;
        for (int64_t j = 0; j < ninserts - 1; j++)
            shift_slots(qf, (int64_t)empties[j + 1] + 1, empties[j] - 1, (uint64_t)j + 1);
        shift_slots(qf, (int64_t)insert_index, empties[ninserts - 1] - 1, (uint64_t)ninserts);
        for (int64_t j = 0; j < ninserts - 1; j++)
            shift_runends(qf, (int64_t)empties[j + 1] + 1, empties[j] - 1, (uint64_t)j + 1);
        shift_runends(qf, (int64_t)insert_index, empties[ninserts - 1] - 1, (uint64_t)ninserts);

        for (uint64_t i = noverwrites; i + 1 < total; i++)
            runend_clear(qf, overwrite_index + i);

        switch (operation) {
        case 0: /* insert into empty bucket */
            runend_or(qf, overwrite_index + total - 1);
            break;
        case 1: /* append to bucket */
            runend_clear(qf, overwrite_index + noverwrites - 1);
            runend_or(qf, overwrite_index + total - 1);
            break;
        case 2: /* insert into bucket */
            runend_clear(qf, overwrite_index + total - 1);
            break;
        default:
            abort();
        }

        uint64_t npre = 0;
        for (uint64_t i = bucket_index / SPB + 1; i <= empties[0] / SPB; i++) {
            while ((int64_t)npre < ninserts && empties[ninserts - 1 - (int64_t)npre] / SPB < i)
                npre++;
            uint64_t add = (uint64_t)ninserts - npre;
            if (off_b(qf, i) + add < OFFSET_MAX)
                set_off_b(qf, i, off_b(qf, i) + add);
            else
                set_off_b(qf, i, OFFSET_MAX);
        }
    }
    for (uint64_t i = 0; i < total; i++)
        set_slot(qf, overwrite_index + i, rems[i]);
    return 1;
}

/* src/gqf.cu:1099-1201 */
static int remove_replace_slots(orc_qf *qf, int operation, uint64_t bucket_index,
                                uint64_t overwrite_index, const uint64_t *rems,
                                uint64_t total, uint64_t old_length)
{
    for (uint64_t i = 0; i < total; i++)
        set_slot(qf, overwrite_index + i, rems[i]);

    if (is_runend(qf, overwrite_index + old_length - 1)) {
        if (total > 0) {
            runend_or(qf, overwrite_index + total - 1);
        } else if (overwrite_index > bucket_index && !is_runend(qf, overwrite_index - 1)) {
            runend_or(qf, overwrite_index - 1);
        }
    }

    uint64_t original_bucket = bucket_index;
    uint64_t current_bucket = bucket_index;
    uint64_t current_slot = overwrite_index + total;
    uint64_t current_distance = old_length - total;
    int ret = (int)current_distance;

    while (current_distance > 0) {
        if (current_slot + current_distance > lim_slots(qf)) break; /* guard */
        if (is_runend(qf, current_slot + current_distance - 1)) {
            do {
                current_bucket++;
            } while (current_bucket < current_slot + current_distance &&
                     !is_occupied(qf, current_bucket));
        }
        if (current_bucket <= current_slot) {
            set_slot(qf, current_slot, get_slot(qf, current_slot + current_distance));
            if (is_runend(qf, current_slot) != is_runend(qf, current_slot + current_distance))
                runend_xor(qf, current_slot);
            current_slot++;
        } else if (current_bucket <= current_slot + current_distance) {
            for (uint64_t i = current_slot; i < current_slot + current_distance; i++) {
                set_slot(qf, i, 0);
                runend_clear(qf, i);
            }
            current_distance = current_slot + current_distance - current_bucket;
            current_slot = current_bucket;
        } else {
            current_distance = 0;
        }
    }

    if (operation && !total)
        occ_clear(qf, bucket_index);

    uint64_t ob = original_bucket / SPB;
    if (old_length > total) {
        while (ob + 1 < qf->nblocks) {
            uint64_t last_h = SPB * ob + (SPB - 1);
            uint64_t re = run_end_impl(qf, last_h);
            if (re / SPB == ob) {
                if (off_b(qf, ob + 1) == 0) break;
                set_off_b(qf, ob + 1, 0);
            } else {
                if (off_b(qf, ob + 1) == re - last_h) break;
                set_off_b(qf, ob + 1, re - last_h); /* truncated to u8, no saturation */
            }
            ob++;
        }
    }
    return ret;
}

/* ------------------------------------------------------------------ */
/* counter codec (src/gqf.cu:1207-1345) */

/* src/gqf.cu:1216-1272: fills backwards from `end`, returns the first slot. */
static uint64_t *encode_counter(uint64_t remainder, uint64_t counter, uint64_t *end)
{
    uint64_t digit = remainder;
    uint64_t base = (1ULL << 8) - 1;
    uint64_t *p = end;
    if (counter == 0) return p;
    *--p = remainder;
    if (counter == 1) return p;
    if (counter == 2) { *--p = remainder; return p; }
    if (counter == 3 && remainder == 0) { *--p = remainder; *--p = remainder; return p; }
    if (counter == 3 && remainder > 0) { *--p = 0; *--p = remainder; return p; }
    if (remainder == 0) *--p = remainder;
    else base--;
    if (remainder) counter -= 3;
    else counter -= 4;
    do {
        digit = counter % base;
        digit++;
        if (remainder && digit >= remainder) digit++;
        *--p = digit;
        counter /= base;
    } while (counter);
    if (remainder && digit >= remainder) *--p = 0;
    *--p = remainder;
    return p;
}

int orc_encode_counter(uint64_t remainder, uint64_t count, uint8_t *out)
{
    uint64_t buf[67];
    uint64_t *p = encode_counter(remainder & 0xFF, count, &buf[67]);
    int n = (int)(&buf[67] - p);
    for (int i = 0; i < n; i++) out[i] = (uint8_t)p[i];
    return n;
}

/* The decoder, written against two accessors so that it can read either a
 * table image or a stand-alone run (src/gqf.cu:1276-1345). */
typedef struct {
    const orc_qf *qf;
    const uint8_t *slots;
    uint64_t n;
} slot_src;

static inline uint64_t src_slot(const slot_src *s, uint64_t i)
{
    if (s->qf) return get_slot(s->qf, i);
    return i < s->n ? s->slots[i] : 0;
}
static inline int src_runend(const slot_src *s, uint64_t i)
{
    if (s->qf) return is_runend(s->qf, i);
    return i + 1 == s->n;
}

static uint64_t decode_counter_src(const slot_src *s, uint64_t index, uint64_t *remainder,
                                   uint64_t *count)
{
    uint64_t base, rem, cnt, digit, end;
    *remainder = rem = src_slot(s, index);
    if (src_runend(s, index)) { *count = 1; return index; }
    digit = src_slot(s, index + 1);
    if (src_runend(s, index + 1)) {
        *count = digit == rem ? 2 : 1;
        return index + (digit == rem ? 1 : 0);
    }
    if (rem > 0 && digit >= rem) {
        *count = digit == rem ? 2 : 1;
        return index + (digit == rem ? 1 : 0);
    }
    if (rem > 0 && digit == 0 && src_slot(s, index + 2) == rem) {
        *count = 3;
        return index + 2;
    }
    if (rem == 0 && digit == 0) {
        if (src_slot(s, index + 2) == 0) { *count = 3; return index + 2; }
        *count = 2;
        return index + 1;
    }
    cnt = 0;
    base = (1ULL << 8) - (rem ? 2 : 1);
    end = index + 1;
    while (digit != rem && !src_runend(s, end)) {
        if (digit > rem) digit--;
        if (digit && rem) digit--;
        cnt = cnt * base + digit;
        end++;
        digit = src_slot(s, end);
    }
    if (rem) { *count = cnt + 3; return end; }
    if (src_runend(s, end) || src_slot(s, end + 1) != 0) { *count = 1; return index; }
    *count = cnt + 4;
    return end + 1;
}

static uint64_t decode_counter(const orc_qf *qf, uint64_t index, uint64_t *rem, uint64_t *count)
{
    slot_src s = {qf, NULL, 0};
    return decode_counter_src(&s, index, rem, count);
}

uint64_t orc_decode_counter_run(const uint8_t *slots, uint64_t n, uint64_t index,
                                uint64_t *remainder, uint64_t *count)
{
    slot_src s = {NULL, slots, n};
    return decode_counter_src(&s, index, remainder, count);
}

/* ------------------------------------------------------------------ */
/* per-key operations */

/* src/gqf.cu:1364-1518 (point_insert_not_exists path) */
static int insert1_if_not_exists(orc_qf *qf, uint64_t hash, uint8_t *value)
{
    uint64_t rem = hash & 0xFFULL;
    uint64_t hb = hash >> 8;
    uint64_t hbo = hb % SPB;
    uint64_t compare_rem = rem; /* value_bits = 0 */

    if (is_empty(qf, hb)) {
        set_run_w(qf, hb / SPB, run_w(qf, hb / SPB) | (1ULL << hbo));
        set_slot(qf, hb, rem);
        set_occ_w(qf, hb / SPB, occ_w(qf, hb / SPB) | (1ULL << hbo));
        return ORC_ITEM_INSERTED;
    }
    uint64_t runend_index = run_end_impl(qf, hb);
    int operation = 0;
    uint64_t insert_index = runend_index + 1;
    uint64_t new_value = rem;
    uint64_t runstart_index = hb == 0 ? 0 : run_end_impl(qf, hb - 1) + 1;

    if (is_occupied(qf, hb)) {
        uint64_t cur = get_slot(qf, runstart_index);
        while (cur < compare_rem && runstart_index <= runend_index) {
            runstart_index++;
            cur = get_slot(qf, runstart_index);
        }
        if (runstart_index > runend_index) {
            operation = 1;
            insert_index = runstart_index;
            new_value = rem;
        } else if (cur != compare_rem) {
            operation = 2;
            insert_index = runstart_index;
            new_value = rem;
        } else {
            /* `get_slot(...) && BITMASK(value_bits)` with value_bits 0 is 0 */
            *value = 0;
            return ORC_ITEM_FOUND;
        }
    }
    uint64_t empty = find_first_empty_slot(qf, runend_index + 1);
    if (empty >= qf->xnslots) return ORC_FULL;
    shift_remainders(qf, insert_index, empty);
    set_slot(qf, insert_index, new_value);
    shift_runends(qf, (int64_t)insert_index, empty - 1, 1);
    switch (operation) {
    case 0: runend_or(qf, insert_index); break;
    case 1: runend_clear(qf, insert_index - 1); runend_or(qf, insert_index); break;
    case 2: runend_clear(qf, insert_index); break;
    default: abort();
    }
    for (uint64_t i = hb / SPB + 1; i <= empty / SPB; i++)
        if (off_b(qf, i) < OFFSET_MAX) set_off_b(qf, i, off_b(qf, i) + 1);
    occ_or(qf, hb);
    return ORC_ITEM_INSERTED;
}

/* src/gqf.cu:1523-1766 (count == 1) */
static int insert1(orc_qf *qf, uint64_t hash)
{
    int ret_distance = 0;
    uint64_t rem = hash & 0xFFULL;
    uint64_t hb = hash >> 8;
    uint64_t hbo = hb % SPB;

    if (is_empty(qf, hb)) {
        set_run_w(qf, hb / SPB, run_w(qf, hb / SPB) | (1ULL << hbo));
        set_slot(qf, hb, rem);
        set_occ_w(qf, hb / SPB, occ_w(qf, hb / SPB) | (1ULL << hbo));
        return 0;
    }
    uint64_t runend_index = run_end_impl(qf, hb);
    int operation = 0;
    uint64_t insert_index = runend_index + 1;
    uint64_t new_value = rem;
    uint64_t runstart_index = hb == 0 ? 0 : run_end_impl(qf, hb - 1) + 1;

    if (is_occupied(qf, hb)) {
        uint64_t cur = get_slot(qf, runstart_index);
        uint64_t zero_terminator = runstart_index;

        if (cur == 0) { /* the counter for 0 is special */
            uint64_t t = runstart_index + 1;
            while (t < runend_index && get_slot(qf, t) != 0) t++;
            if (t < runend_index && get_slot(qf, t + 1) == 0)
                zero_terminator = t + 1;
            else if (runstart_index < runend_index && get_slot(qf, runstart_index + 1) == 0)
                zero_terminator = runstart_index + 1;
            if (rem != 0) {
                runstart_index = zero_terminator + 1;
                cur = get_slot(qf, runstart_index);
            }
        }
        while (cur < rem && runstart_index <= runend_index) {
            if (runstart_index < runend_index && get_slot(qf, runstart_index + 1) < cur) {
                runstart_index = runstart_index + 2;
                while (runstart_index < runend_index && get_slot(qf, runstart_index) != cur)
                    runstart_index++;
                runstart_index++;
            } else {
                runstart_index++;
            }
            cur = get_slot(qf, runstart_index);
        }

        if (runstart_index > runend_index) {
            operation = 1;
            insert_index = runstart_index;
            new_value = rem;
        } else if (cur != rem) {
            operation = 2;
            insert_index = runstart_index;
            new_value = rem;
        } else if (runstart_index == runend_index ||
                   (rem > 0 && get_slot(qf, runstart_index + 1) > rem) ||
                   (rem == 0 && zero_terminator == runstart_index)) {
            operation = 2;
            insert_index = runstart_index;
            new_value = rem;
        } else if ((rem > 0 && get_slot(qf, runstart_index + 1) == rem) ||
                   (rem == 0 && zero_terminator == runstart_index + 1)) {
            operation = 2;
            insert_index = runstart_index + 1;
            new_value = 0;
        } else if (rem == 0 && zero_terminator == runstart_index + 2) {
            operation = 2;
            insert_index = runstart_index + 1;
            new_value = 1;
        } else {
            /* extended counter: increment in place (src/gqf.cu:1650-1706) */
            insert_index = runstart_index + 1;
            while (get_slot(qf, insert_index + 1) != rem) insert_index++;
            uint64_t digit, carry;
            do {
                carry = 0;
                digit = get_slot(qf, insert_index);
                if (digit == 0) {
                    digit++;
                    if (digit == cur) digit++;
                }
                digit = (digit + 1) & 0xFFULL;
                if (digit == 0) { digit++; carry = 1; }
                if (digit == cur) digit = (digit + 1) & 0xFFULL;
                if (digit == 0) { digit++; carry = 1; }
                set_slot(qf, insert_index, digit);
                insert_index--;
            } while (insert_index > runstart_index && carry);

            if (insert_index == runstart_index && (carry > 0 || (cur != 0 && digit >= cur))) {
                operation = 2;
                insert_index = runstart_index + 1;
                if (!carry) new_value = 0;
                else new_value = 2;
            } else {
                operation = -1;
            }
        }
    }

    if (operation >= 0) {
        uint64_t empty = find_first_empty_slot(qf, runend_index + 1);
        if (empty >= qf->xnslots) return ORC_NO_SPACE;
        shift_remainders(qf, insert_index, empty);
        set_slot(qf, insert_index, new_value);
        ret_distance = (int)(insert_index - hb);
        shift_runends(qf, (int64_t)insert_index, empty - 1, 1);
        switch (operation) {
        case 0: runend_or(qf, insert_index); break;
        case 1: runend_clear(qf, insert_index - 1); runend_or(qf, insert_index); break;
        case 2: runend_clear(qf, insert_index); break;
        default: abort();
        }
        for (uint64_t i = hb / SPB + 1; i <= empty / SPB; i++)
            if (off_b(qf, i) < OFFSET_MAX) set_off_b(qf, i, off_b(qf, i) + 1);
    }
    occ_or(qf, hb);
    return ret_distance;
}

/* src/gqf.cu:1768-1874 (count > 1) */
static int insert_counted(orc_qf *qf, uint64_t hash, uint64_t count)
{
    int ret_distance = 0;
    uint64_t rem = hash & 0xFFULL;
    uint64_t hb = hash >> 8;
    uint64_t hbo = hb % SPB;
    uint64_t runend_index = run_end_impl(qf, hb);

    if (might_be_empty(qf, hb) && runend_index == hb) {
        set_run_w(qf, hb / SPB, run_w(qf, hb / SPB) | (1ULL << hbo));
        set_slot(qf, hb, rem);
        set_occ_w(qf, hb / SPB, occ_w(qf, hb / SPB) | (1ULL << hbo));
        if (count > 1) return insert_counted(qf, hash, count - 1);
        return 0;
    }
    uint64_t nv[67];
    int64_t runstart_index = hb == 0 ? 0 : (int64_t)run_end_impl(qf, hb - 1) + 1;
    int ok;
    if (!is_occupied(qf, hb)) {
        uint64_t *p = encode_counter(rem, count, &nv[67]);
        ok = insert_replace_slots(qf, 0, hb, (uint64_t)runstart_index, p, (uint64_t)(&nv[67] - p), 0);
        if (!ok) return ORC_NO_SPACE;
        ret_distance = (int)((uint64_t)runstart_index - hb);
    } else {
        uint64_t cur_rem, cur_cnt, cur_end;
        cur_end = decode_counter(qf, (uint64_t)runstart_index, &cur_rem, &cur_cnt);
        while (cur_rem < rem && !is_runend(qf, cur_end)) {
            runstart_index = (int64_t)cur_end + 1;
            cur_end = decode_counter(qf, (uint64_t)runstart_index, &cur_rem, &cur_cnt);
        }
        if (cur_rem < rem) {
            uint64_t *p = encode_counter(rem, count, &nv[67]);
            ok = insert_replace_slots(qf, 1, hb, cur_end + 1, p, (uint64_t)(&nv[67] - p), 0);
            if (!ok) return ORC_NO_SPACE;
            ret_distance = (int)((cur_end + 1) - hb);
        } else if (cur_rem == rem) {
            uint64_t *p = encode_counter(rem, cur_cnt + count, &nv[67]);
            ok = insert_replace_slots(qf, is_runend(qf, cur_end) ? 1 : 2, hb,
                                      (uint64_t)runstart_index, p, (uint64_t)(&nv[67] - p),
                                      cur_end - (uint64_t)runstart_index + 1);
            if (!ok) return ORC_NO_SPACE;
            ret_distance = (int)((uint64_t)runstart_index - hb);
        } else {
            uint64_t *p = encode_counter(rem, count, &nv[67]);
            ok = insert_replace_slots(qf, 2, hb, (uint64_t)runstart_index, p,
                                      (uint64_t)(&nv[67] - p), 0);
            if (!ok) return ORC_NO_SPACE;
            ret_distance = (int)((uint64_t)runstart_index - hb);
        }
    }
    occ_or(qf, hb);
    return ret_distance;
}

/* src/gqf.cu:1876-1933 */
static int remove_impl(orc_qf *qf, uint64_t hash, uint64_t count)
{
    uint64_t rem = hash & 0xFFULL;
    uint64_t hb = hash >> 8;
    uint64_t cur_rem, cur_cnt, cur_end;
    uint64_t nv[67];

    if (!is_occupied(qf, hb)) return -1;
    uint64_t runstart_index = hb == 0 ? 0 : run_end_impl(qf, hb - 1) + 1;
    uint64_t original_runstart_index = runstart_index;
    int only_item = 0;

    cur_end = decode_counter(qf, runstart_index, &cur_rem, &cur_cnt);
    while (cur_rem < rem && !is_runend(qf, cur_end)) {
        runstart_index = cur_end + 1;
        cur_end = decode_counter(qf, runstart_index, &cur_rem, &cur_cnt);
    }
    if (cur_rem != rem) return -1;
    if (original_runstart_index == runstart_index && is_runend(qf, cur_end)) only_item = 1;

    uint64_t *p = encode_counter(rem, count > cur_cnt ? 0 : cur_cnt - count, &nv[67]);
    return remove_replace_slots(qf, only_item, hb, runstart_index, p, (uint64_t)(&nv[67] - p),
                                cur_end - runstart_index + 1);
}

static inline uint64_t key_to_hash(const orc_qf *qf, uint64_t key, uint8_t flags)
{
    if (!(flags & ORC_KEY_IS_HASH)) key = orc_hash_64(key, BMASK(qf->key_bits));
    return key;
}

/* src/gqf.cu:2456-2525 */
int orc_qf_insert(orc_qf *qf, uint64_t key, uint64_t count, uint8_t flags)
{
    if (count == 0) return 0;
    uint64_t hash = key_to_hash(qf, key, flags);
    if ((hash >> 8) >= qf->nslots) return ORC_NO_SPACE; /* guard for unmasked pre-hashed keys */
    if (count == 1) return insert1(qf, hash);
    return insert_counted(qf, hash, count);
}

/* src/gqf.cu:3443-3459 */
int orc_qf_remove(orc_qf *qf, uint64_t key, uint64_t count, uint8_t flags)
{
    if (count == 0) return 1;
    uint64_t hash = key_to_hash(qf, key, flags);
    if ((hash >> 8) >= qf->nslots) return -1;
    return remove_impl(qf, hash, count);
}

/* src/gqf.cu:3479-3518 */
uint64_t orc_qf_count_key_value(const orc_qf *qf, uint64_t key, uint8_t flags)
{
    uint64_t hash = key_to_hash(qf, key, flags);
    uint64_t rem = hash & 0xFFULL;
    uint64_t hb = hash >> 8;
    if (hb >= qf->nslots) return 0;
    if (!is_occupied(qf, hb)) return 0;
    uint64_t rs = hb == 0 ? 0 : run_end_impl(qf, hb - 1) + 1;
    if (rs < hb) rs = hb;
    uint64_t cr, cc, ce;
    do {
        ce = decode_counter(qf, rs, &cr, &cc);
        if (cr == rem) return cc;
        rs = ce + 1;
    } while (!is_runend(qf, ce) && ce < lim_slots(qf));
    return 0;
}

/* src/gqf.cu:3520-3557 (value_bits = 0, so *value is always 0) */
uint64_t orc_qf_query(const orc_qf *qf, uint64_t key, uint64_t *value, uint8_t flags)
{
    *value = 0;
    return orc_qf_count_key_value(qf, key, flags);
}

/* src/gqf.cu:2383-2452 (count is always 1 on this path) */
int orc_qf_insert_not_exists(orc_qf *qf, uint64_t key, uint8_t flags, uint8_t *retval)
{
    uint64_t hash = key_to_hash(qf, key, flags);
    if ((hash >> 8) >= qf->nslots) return ORC_FULL;
    return insert1_if_not_exists(qf, hash, retval);
}

/* ------------------------------------------------------------------ */
/* hashing (src/hashutil.cu:132-181) */

uint64_t orc_hash_64(uint64_t key, uint64_t mask)
{
    key = (~key + (key << 21)) & mask;
    key = key ^ key >> 24;
    key = ((key + (key << 3)) + (key << 8)) & mask;
    key = key ^ key >> 14;
    key = ((key + (key << 2)) + (key << 4)) & mask;
    key = key ^ key >> 28;
    key = (key + (key << 31)) & mask;
    return key;
}

uint64_t orc_hash_64i(uint64_t key, uint64_t mask)
{
    uint64_t tmp;
    tmp = key - (key << 31);
    key = (key - (tmp << 31)) & mask;
    tmp = key ^ key >> 28;
    key = key ^ tmp >> 28;
    key = (key * 14933078535860113213ULL) & mask;
    tmp = key ^ key >> 14;
    tmp = key ^ tmp >> 14;
    tmp = key ^ tmp >> 14;
    key = key ^ tmp >> 14;
    key = (key * 15244667743933553977ULL) & mask;
    tmp = key ^ key >> 24;
    key = key ^ tmp >> 24;
    tmp = ~key;
    tmp = ~(key - (tmp << 21));
    tmp = ~(key - (tmp << 21));
    key = ~(key - (tmp << 21)) & mask;
    return key;
}

void orc_hash_many(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t mask)
{
    for (uint64_t i = 0; i < n; i++) out[i] = orc_hash_64(in[i], mask);
}

/* ------------------------------------------------------------------ */
/* lifecycle (src/gqf.cu:1939-2022, 2073-2105, 2123-2140) */

uint64_t orc_sizes(int nbits, uint64_t out[8])
{
    uint64_t nslots = 1ULL << nbits;
    uint64_t xnslots = nslots + (uint64_t)(10 * sqrt((double)nslots));
    uint64_t nblocks = (xnslots + SPB - 1) / SPB;
    uint64_t key_bits = (uint64_t)nbits + 8;
    uint64_t krb = key_bits;
    uint64_t t = nslots;
    while (t > 1 && krb > 0) { krb--; t >>= 1; }
    out[0] = nslots;
    out[1] = xnslots;
    out[2] = nblocks;
    out[3] = key_bits;
    out[4] = krb;
    out[5] = nslots << krb;               /* range */
    out[6] = xnslots / REGION_SLOTS + 2;  /* num_locks */
    out[7] = nblocks * ORC_BLOCK_BYTES;   /* total_size_in_bytes */
    return out[7];
}

int orc_init(orc_qf *qf, int nbits)
{
    uint64_t s[8];
    if (nbits < 1 || nbits > 40) return -1;
    orc_sizes(nbits, s);
    memset(qf, 0, sizeof(*qf));
    qf->nslots = s[0];
    qf->xnslots = s[1];
    qf->nblocks = s[2];
    qf->key_bits = s[3];
    qf->key_remainder_bits = s[4];
    qf->bits_per_slot = s[4];
    qf->range = s[5];
    qf->num_locks = s[6];
    qf->total_size_in_bytes = s[7];
    qf->image = (uint8_t *)calloc(s[7], 1);
    return qf->image ? 0 : -1;
}

void orc_free(orc_qf *qf)
{
    free(qf->image);
    qf->image = NULL;
}

void orc_reset(orc_qf *qf)
{
    memset(qf->image, 0, qf->total_size_in_bytes);
    qf->overflow_events = 0;
}

/* ------------------------------------------------------------------ */
/* batch drivers */

/* LSD radix sort of u64 keys (the ordering thrust::sort gives). */
static void sort_u64(uint64_t *a, uint64_t n)
{
    if (n < 2) return;
    uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *src = a, *dst = tmp;
    uint64_t all_or = 0;
    for (uint64_t i = 0; i < n; i++) all_or |= a[i];
    int bits = 64 - (all_or ? __builtin_clzll(all_or) : 64);
    for (int shift = 0; shift < bits; shift += 11) {
        uint64_t cnt[2048] = {0};
        for (uint64_t i = 0; i < n; i++) cnt[(src[i] >> shift) & 2047]++;
        uint64_t sum = 0;
        for (int d = 0; d < 2048; d++) { uint64_t c = cnt[d]; cnt[d] = sum; sum += c; }
        for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> shift) & 2047]++] = src[i];
        uint64_t *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof(uint64_t));
    free(tmp);
}

static void hash_and_sort(const orc_qf *qf, uint64_t *keys, uint64_t n, uint8_t flags)
{
    if (!(flags & ORC_KEY_IS_HASH))
        for (uint64_t i = 0; i < n; i++) keys[i] = orc_hash_64(keys[i], BMASK(qf->key_bits));
    sort_u64(keys, n);
}

/* Region bounds of the sorted hashes (src/gqf.cu:2646-2753): region r holds
 * quotients [8192 r, 8192 (r+1)). */
static uint64_t *region_starts(const orc_qf *qf, const uint64_t *sorted, uint64_t n)
{
    uint64_t nr = qf->num_locks;
    uint64_t *st = (uint64_t *)malloc((nr + 1) * sizeof(uint64_t));
    uint64_t k = 0;
    for (uint64_t r = 0; r < nr; r++) {
        while (k < n && (sorted[k] >> 8) < REGION_SLOTS * r) k++;
        st[r] = k;
    }
    st[nr] = n;
    return st;
}

uint64_t orc_bulk_insert(orc_qf *qf, uint64_t *keys, uint64_t n, uint8_t flags)
{
    uint64_t nospace = 0;
    hash_and_sort(qf, keys, n, flags);
    uint64_t *st = region_starts(qf, keys, n);
    for (int e = 0; e < 2; e++)
        for (uint64_t r = (uint64_t)e; r < qf->num_locks; r += 2)
            for (uint64_t i = st[r]; i < st[r + 1]; i++)
                if (orc_qf_insert(qf, keys[i], 1, ORC_NO_LOCK | ORC_KEY_IS_HASH) == ORC_NO_SPACE)
                    nospace++;
    free(st);
    return nospace;
}

uint64_t orc_bulk_insert_reduce(orc_qf *qf, uint64_t *keys, uint64_t n, uint8_t flags)
{
    uint64_t nospace = 0;
    hash_and_sort(qf, keys, n, flags);
    uint64_t *uk = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t *uc = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (m && uk[m - 1] == keys[i]) uc[m - 1]++;
        else { uk[m] = keys[i]; uc[m] = 1; m++; }
    }
    uint64_t *st = region_starts(qf, uk, m);
    for (int e = 0; e < 2; e++)
        for (uint64_t r = (uint64_t)e; r < qf->num_locks; r += 2)
            for (uint64_t i = st[r]; i < st[r + 1]; i++)
                if (orc_qf_insert(qf, uk[i], uc[i], ORC_NO_LOCK | ORC_KEY_IS_HASH) == ORC_NO_SPACE)
                    nospace++;
    free(st);
    free(uk);
    free(uc);
    return nospace;
}

uint64_t orc_bulk_delete(orc_qf *qf, uint64_t *keys, uint64_t n, uint8_t flags)
{
    uint64_t absent = 0;
    hash_and_sort(qf, keys, n, flags);
    uint64_t *st = region_starts(qf, keys, n);
    for (int e = 0; e < 2; e++)
        for (uint64_t r = (uint64_t)e; r < qf->num_locks; r += 2)
            for (uint64_t i = st[r + 1]; i > st[r]; i--)
                if (orc_qf_remove(qf, keys[i - 1], 1, ORC_NO_LOCK | ORC_KEY_IS_HASH) < 0)
                    absent++;
    free(st);
    return absent;
}

void orc_bulk_get(const orc_qf *qf, const uint64_t *vals, uint64_t n, uint64_t *returns)
{
    for (uint64_t i = 0; i < n; i++) returns[i] = orc_qf_count_key_value(qf, vals[i], ORC_NO_LOCK);
}

uint64_t orc_bulk_get_misses(const orc_qf *qf, const uint64_t *vals, uint64_t n)
{
    uint64_t miss = 0;
    for (uint64_t i = 0; i < n; i++) miss += orc_qf_count_key_value(qf, vals[i], ORC_NO_LOCK) < 1;
    return miss;
}

/* ------------------------------------------------------------------ */
/* canonical layout from a multiset.
 *
 * The invariants of the reference image fix it completely:
 *  - the run of quotient q starts at max(q, end of the previous run + 1)
 *    (the position run_end/offset arithmetic assumes, src/gqf.cu:632-704);
 *  - a run is the concatenation of the counters of its remainders in
 *    ascending order, each the encode_counter output (src/gqf.cu:1216-1272);
 *  - runends mark the last slot of each run, occupieds mark q;
 *  - block i's offset is min(255, run_end(64 i - 1) - 64 i + 1)
 *    (block_offset's recompute formula, src/gqf.cu:628-629, saturated as the
 *    insert path saturates it, src/gqf.cu:1748-1753);
 *  - every other slot is 0 (the memset image, src/gqf.cu:2081). */
int orc_build_canonical(orc_qf *qf, const uint64_t *hashes, const uint64_t *counts, uint64_t n)
{
    int64_t prev_end = -1;
    uint64_t next_boundary_block = 1;
    uint64_t i = 0;
    uint64_t enc[67];
    while (i < n) {
        uint64_t q = hashes[i] >> 8;
        /* offsets of blocks whose boundary lies before this quotient */
        while (next_boundary_block < qf->nblocks && SPB * next_boundary_block <= q) {
            int64_t bnd = (int64_t)(SPB * next_boundary_block) - 1;
            int64_t o = prev_end > bnd ? prev_end - bnd : 0;
            set_off_b(qf, next_boundary_block, o > 255 ? 255 : (uint64_t)o);
            next_boundary_block++;
        }
        int64_t start = (int64_t)q > prev_end + 1 ? (int64_t)q : prev_end + 1;
        int64_t pos = start;
        while (i < n && (hashes[i] >> 8) == q) {
            uint64_t *p = encode_counter(hashes[i] & 0xFF, counts[i], &enc[67]);
            for (; p < &enc[67]; p++) {
                if ((uint64_t)pos >= qf->xnslots) return -1;
                set_slot(qf, (uint64_t)pos, *p);
                pos++;
            }
            i++;
        }
        if (pos == start) continue; /* all counts zero */
        runend_or(qf, (uint64_t)(pos - 1));
        occ_or(qf, q);
        prev_end = pos - 1;
    }
    while (next_boundary_block < qf->nblocks) {
        int64_t bnd = (int64_t)(SPB * next_boundary_block) - 1;
        int64_t o = prev_end > bnd ? prev_end - bnd : 0;
        set_off_b(qf, next_boundary_block, o > 255 ? 255 : (uint64_t)o);
        next_boundary_block++;
    }
    return 0;
}

uint64_t orc_enumerate(const orc_qf *qf, uint64_t *hashes, uint64_t *counts, uint64_t cap)
{
    uint64_t m = 0;
    for (uint64_t b = 0; b < qf->nblocks; b++) {
        uint64_t occ = occ_w(qf, b);
        while (occ) {
            uint64_t q = b * SPB + (uint64_t)__builtin_ctzll(occ);
            occ &= occ - 1;
            if (q >= qf->nslots) continue;
            uint64_t rs = q == 0 ? 0 : run_end_impl(qf, q - 1) + 1;
            if (rs < q) rs = q;
            uint64_t cr, cc, ce;
            do {
                ce = decode_counter(qf, rs, &cr, &cc);
                if (m < cap) { hashes[m] = (q << 8) | cr; counts[m] = cc; }
                m++;
                rs = ce + 1;
            } while (!is_runend(qf, ce) && ce < lim_slots(qf));
        }
    }
    return m;
}

/* Cluster census: one_thread_cluster_write (src/cluster_length_test.cu:82-120)
 * over first_empty_slot_wrapper -> find_first_empty_slot (src/gqf.cu:728-776):
 * from each slot, the first empty slot at or after it; a used slot starts a
 * cluster that ends there. */
uint64_t orc_cluster_census(orc_qf *qf, uint64_t *lengths, uint64_t cap)
{
    uint64_t i = 0, start = 0;
    while (start < qf->xnslots) {
        uint64_t next = find_first_empty_slot(qf, start);
        if (start == next) {
            start++;
        } else {
            if (i < cap) lengths[i] = next - start;
            start = next;
            i++;
        }
    }
    return i;
}

/* ------------------------------------------------------------------ */
/* generators */

/* splitmix64 (Steele, Lea & Flood 2014; Vigna's public-domain form):
 * state += golden gamma, then two xor-shift-multiply rounds. */
void orc_splitmix64_fill(uint64_t seed, uint64_t skip, uint64_t n, uint64_t *out)
{
    uint64_t state = seed + skip * 0x9E3779B97F4A7C15ULL;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        out[i] = z ^ (z >> 31);
    }
}

/* ------------------------------------------------------------------ */
/* CPU baseline: the reference host path (qf_insert / qf_count_key_value),
 * single thread, keys pre-generated, steady clock around each loop. */

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Seconds of n sequential qf_insert calls (count 1, keys hashed) on an
 * existing table: the reference host path's insert loop. */
double orc_time_inserts(orc_qf *qf, const uint64_t *keys, uint64_t n)
{
    double t0 = now_s();
    for (uint64_t i = 0; i < n; i++) orc_qf_insert(qf, keys[i], 1, ORC_NO_LOCK);
    return now_s() - t0;
}

/* Seconds of n qf_count_key_value calls; *hits = keys found. */
double orc_time_queries(const orc_qf *qf, const uint64_t *keys, uint64_t n, uint64_t *hits)
{
    double t0 = now_s();
    uint64_t h = 0;
    for (uint64_t i = 0; i < n; i++) h += orc_qf_count_key_value(qf, keys[i], ORC_NO_LOCK) > 0;
    *hits = h;
    return now_s() - t0;
}

double orc_time_insert_query(int nbits, const uint64_t *keys, uint64_t n_ins,
                             const uint64_t *qkeys, uint64_t n_q, uint64_t *hits,
                             double *t_insert, double *t_query)
{
    orc_qf qf;
    if (orc_init(&qf, nbits) != 0) return -1.0;
    double t0 = now_s();
    for (uint64_t i = 0; i < n_ins; i++) orc_qf_insert(&qf, keys[i], 1, ORC_NO_LOCK);
    double t1 = now_s();
    uint64_t h = 0;
    for (uint64_t i = 0; i < n_q; i++) h += orc_qf_count_key_value(&qf, qkeys[i], ORC_NO_LOCK) > 0;
    double t2 = now_s();
    *hits = h;
    *t_insert = t1 - t0;
    *t_query = t2 - t1;
    orc_free(&qf);
    return (t2 - t0);
}
