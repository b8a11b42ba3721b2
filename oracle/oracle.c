/*
 * CPU oracle for the GENIE match-count top-k path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker / CPU baseline.  See oracle.h for the contract.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ====================================================================== */
/* Index build                                                            */
/* ====================================================================== */

typedef struct {
    int64_t n, cap;
    int32_t *v;
} ivec;

static void ivec_push(ivec *a, int32_t x)
{
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 16;
        a->v = (int32_t *)realloc(a->v, (size_t)a->cap * sizeof(int32_t));
    }
    a->v[a->n++] = x;
}

/* One inverted "list" object: bounds and one id vector per value
 * (inv_list::invert, inv_list.cc:295-320; invert_bijectMap 51-86). */
typedef struct {
    int32_t lo, hi;
    ivec *lists; /* hi - lo + 1 */
} orc_invlist;

/* Flatten a sequence of inverted lists (inv_table::build, inv_table.cu:268-350). */
static int flatten(orc_invlist *L, int32_t n_dims, int32_t n_objects,
                   int64_t max_length, int use_load_balance, orc_index *out)
{
    memset(out, 0, sizeof(*out));
    if (!use_load_balance || max_length <= 0)
        max_length = INT64_MAX;
    out->n_objects = n_objects;
    out->n_dims = n_dims;
    out->lo = (int32_t *)malloc(sizeof(int32_t) * (n_dims ? n_dims : 1));
    out->hi = (int32_t *)malloc(sizeof(int32_t) * (n_dims ? n_dims : 1));
    out->keyseg_off = (int64_t *)malloc(sizeof(int64_t) * (n_dims + 1));
    int64_t n_post = 0, n_seg = 0, n_ks = 0;
    for (int32_t d = 0; d < n_dims; ++d) {
        if (d + 1 < n_dims && L[d].hi - L[d].lo + 1 >= 65536)
            return -1; /* keys of neighbouring dims would collide (shifter 16) */
        for (int32_t v = 0; v <= L[d].hi - L[d].lo; ++v) {
            int64_t len = L[d].lists[v].n;
            n_post += len;
            if (len > 0)
                n_seg += 1 + (len - 1) / max_length;
        }
        n_ks += (int64_t)(L[d].hi - L[d].lo) + 2;
    }
    out->n_post = n_post;
    out->n_seg = n_seg;
    out->inv = (int32_t *)malloc(sizeof(int32_t) * (n_post ? n_post : 1));
    out->inv_pos = (int32_t *)malloc(sizeof(int32_t) * (n_seg + 1));
    out->keyseg = (int32_t *)malloc(sizeof(int32_t) * (n_ks ? n_ks : 1));
    int64_t p = 0, s = 0, ks = 0;
    for (int32_t d = 0; d < n_dims; ++d) {
        out->lo[d] = L[d].lo;
        out->hi[d] = L[d].hi;
        out->keyseg_off[d] = ks;
        for (int32_t v = 0; v <= L[d].hi - L[d].lo; ++v) {
            /* _inv_index[key] = _inv_pos.size() before the key's postings */
            out->keyseg[ks++] = (int32_t)s;
            ivec *lst = &L[d].lists[v];
            for (int64_t j = 0; j < lst->n; ++j) {
                if (j % max_length == 0)
                    out->inv_pos[s++] = (int32_t)p;
                out->inv[p++] = lst->v[j];
            }
        }
        /* key hi+1: the first segment of whatever follows (next dim or end) */
        out->keyseg[ks++] = (int32_t)s;
    }
    out->keyseg_off[n_dims] = ks;
    out->inv_pos[s] = (int32_t)p;
    return 0;
}

static void free_invlists(orc_invlist *L, int32_t n)
{
    for (int32_t d = 0; d < n; ++d) {
        if (!L[d].lists)
            continue;
        for (int32_t v = 0; v <= L[d].hi - L[d].lo; ++v)
            free(L[d].lists[v].v);
        free(L[d].lists);
    }
    free(L);
}

int orc_build_single(const int32_t *rows, int32_t n_rows, int32_t n_dims,
                     int64_t max_length, int use_load_balance, orc_index *out)
{
    if (n_rows <= 0 || n_dims <= 0)
        return -1;
    orc_invlist *L = (orc_invlist *)calloc((size_t)n_dims, sizeof(orc_invlist));
    for (int32_t d = 0; d < n_dims; ++d) {
        int32_t lo = rows[d], hi = rows[d];
        for (int32_t i = 0; i < n_rows; ++i) {
            int32_t x = rows[(int64_t)i * n_dims + d];
            if (x < lo) lo = x;
            if (x > hi) hi = x;
        }
        L[d].lo = lo;
        L[d].hi = hi;
        L[d].lists = (ivec *)calloc((size_t)(hi - lo + 1), sizeof(ivec));
        for (int32_t i = 0; i < n_rows; ++i)
            ivec_push(&L[d].lists[rows[(int64_t)i * n_dims + d] - lo], i);
    }
    int rc = flatten(L, n_dims, n_rows, max_length, use_load_balance, out);
    free_invlists(L, n_dims);
    return rc;
}

int orc_build_biject(const int64_t *row_ptr, const int32_t *vals, int32_t n_rows,
                     int64_t max_length, int use_load_balance, orc_index *out)
{
    if (n_rows <= 0 || row_ptr[n_rows] <= 0)
        return -1;
    orc_invlist *L = (orc_invlist *)calloc(1, sizeof(orc_invlist));
    int32_t lo = vals[0], hi = vals[0];
    for (int64_t j = 0; j < row_ptr[n_rows]; ++j) {
        if (vals[j] < lo) lo = vals[j];
        if (vals[j] > hi) hi = vals[j];
    }
    L[0].lo = lo;
    L[0].hi = hi;
    L[0].lists = (ivec *)calloc((size_t)(hi - lo + 1), sizeof(ivec));
    for (int32_t i = 0; i < n_rows; ++i)
        for (int64_t j = row_ptr[i]; j < row_ptr[i + 1]; ++j)
            ivec_push(&L[0].lists[vals[j] - lo], i);
    int rc = flatten(L, 1, n_rows, max_length, use_load_balance, out);
    free_invlists(L, 1);
    return rc;
}

void orc_index_free(orc_index *ix)
{
    free(ix->lo); free(ix->hi); free(ix->inv); free(ix->inv_pos);
    free(ix->keyseg_off); free(ix->keyseg);
    memset(ix, 0, sizeof(*ix));
}

/* ====================================================================== */
/* Codecs                                                                 */
/* ====================================================================== */

/* Bit width of the largest value: gccbits (DeviceBitPackingHelpers.h:922-924). */
static uint32_t bitwidth(uint32_t v) { return v ? 32u - (uint32_t)__builtin_clz(v) : 0u; }

/* Pack 32 values of width b LSB-first into b words (value j occupies bits
 * [j*b, (j+1)*b) of the little-endian word stream), as fastpackwithoutmask
 * lays them out (DeviceBitPackingHelpers.cu). */
static void pack32(const uint32_t *in, uint32_t *out, uint32_t b)
{
    if (b == 0)
        return;
    memset(out, 0, b * sizeof(uint32_t));
    for (uint32_t j = 0; j < 32; ++j) {
        uint64_t bit = (uint64_t)j * b;
        uint32_t w = (uint32_t)(bit >> 5), sh = (uint32_t)(bit & 31);
        uint64_t v = (uint64_t)in[j] << sh;
        out[w] |= (uint32_t)v;
        if (sh + b > 32)
            out[w + 1] |= (uint32_t)(v >> 32);
    }
}

static void unpack32(const uint32_t *in, uint32_t *out, uint32_t b)
{
    for (uint32_t j = 0; j < 32; ++j) {
        if (b == 0) { out[j] = 0; continue; }
        uint64_t bit = (uint64_t)j * b;
        uint32_t w = (uint32_t)(bit >> 5), sh = (uint32_t)(bit & 31);
        uint64_t lo = in[w];
        uint64_t hi = (sh + b > 32) ? in[w + 1] : 0;
        uint64_t v = ((hi << 32) | lo) >> sh;
        out[j] = (uint32_t)(b == 32 ? v : (v & ((1ull << b) - 1)));
    }
}

/* BP32 (DeviceBitPackingCodec.cu:16-58 encode, 60-99 decode): w[0] = length;
 * per 128-value block a header b0<<24|b1<<16|b2<<8|b3 then the packed
 * miniblocks; a tail uses ceil(rem/32) zero-padded miniblocks. */
static int64_t bp32_encode(const uint32_t *in, int64_t n, uint32_t *out)
{
    uint32_t *o = out;
    *o++ = (uint32_t)n;
    int64_t i = 0;
    while (i < n) {
        int64_t rem = n - i;
        uint32_t howmany = rem >= 128 ? 4u : (uint32_t)((rem + 31) / 32);
        uint32_t blk[128];
        memset(blk, 0, sizeof(blk));
        int64_t take = rem >= 128 ? 128 : rem;
        memcpy(blk, in + i, (size_t)take * sizeof(uint32_t));
        uint32_t B[4] = {0, 0, 0, 0};
        for (uint32_t m = 0; m < howmany; ++m) {
            uint32_t acc = 0;
            for (uint32_t j = 0; j < 32; ++j)
                acc |= blk[m * 32 + j];
            B[m] = bitwidth(acc);
        }
        *o++ = (B[0] << 24) | (B[1] << 16) | (B[2] << 8) | B[3];
        for (uint32_t m = 0; m < howmany; ++m) {
            pack32(blk + m * 32, o, B[m]);
            o += B[m];
        }
        i += take;
    }
    return o - out;
}

static int64_t bp32_decode(const uint32_t *in, int64_t n_words, uint32_t *out, int64_t cap,
                           int64_t *used_words)
{
    if (n_words < 1)
        return -1;
    const uint32_t *p = in, *end = in + n_words;
    int64_t n = *p++;
    if (n > cap)
        return -1;
    int64_t i = 0;
    while (i < n) {
        if (p >= end)
            return -1;
        int64_t rem = n - i;
        uint32_t howmany = rem >= 128 ? 4u : (uint32_t)((rem + 31) / 32);
        uint32_t h = *p++;
        uint32_t B[4] = {h >> 24, (h >> 16) & 0xFF, (h >> 8) & 0xFF, h & 0xFF};
        for (uint32_t m = 0; m < howmany; ++m) {
            if (B[m] > 32 || p + B[m] > end)
                return -1;
            uint32_t tmp[32];
            unpack32(p, tmp, B[m]);
            p += B[m];
            for (uint32_t j = 0; j < 32 && i + m * 32 + j < n; ++j)
                out[i + m * 32 + j] = tmp[j];
        }
        i += rem >= 128 ? 128 : rem;
    }
    if (used_words)
        *used_words = p - in;
    return n;
}

/* Varint (DeviceVarintCodec.cu:14-76): 7-bit little-endian groups, the high
 * bit marks the LAST byte of a value; zero bytes pad to a word boundary. */
static int64_t varint_encode(const uint32_t *in, int64_t n, uint32_t *out)
{
    uint8_t *b = (uint8_t *)out, *b0 = b;
    for (int64_t k = 0; k < n; ++k) {
        uint32_t v = in[k];
        int nb = v < (1u << 7) ? 1 : v < (1u << 14) ? 2 : v < (1u << 21) ? 3 : v < (1u << 28) ? 4 : 5;
        for (int j = 0; j < nb; ++j) {
            uint8_t c = (uint8_t)((v >> (7 * j)) & (j == nb - 1 ? 0xFFu : 0x7Fu));
            if (j == nb - 1)
                c |= 0x80;
            *b++ = c;
        }
    }
    while ((b - b0) & 3)
        *b++ = 0;
    return (b - b0) / 4;
}

static int64_t varint_decode(const uint32_t *in, int64_t n_words, uint32_t *out, int64_t cap)
{
    const uint8_t *b = (const uint8_t *)in, *e = b + n_words * 4;
    int64_t n = 0;
    uint32_t v = 0, shift = 0;
    for (; b < e; ++b) {
        uint8_t c = *b;
        v += (uint32_t)(c & 127) << shift;
        if (c & 128) {
            if (n >= cap)
                return -1;
            out[n++] = v;
            v = 0;
            shift = 0;
        } else {
            shift += 7;
        }
    }
    return n;
}

static int64_t delta_encode(const uint32_t *in, int64_t n, uint32_t *out)
{
    uint32_t prev = 0;
    for (int64_t i = 0; i < n; ++i) {
        out[i] = in[i] - prev; /* DeviceDeltaHelper::delta, DeviceDeltaHelper.h:18-26 */
        prev = in[i];
    }
    return n;
}

static void delta_decode_inplace(uint32_t *x, int64_t n)
{
    for (int64_t i = 1; i < n; ++i)
        x[i] += x[i - 1]; /* inverseDelta, DeviceDeltaHelper.h:28-50 */
}

enum { LEAF_COPY, LEAF_DELTA, LEAF_BP32, LEAF_VARINT };

static int64_t leaf_encode(int leaf, const uint32_t *in, int64_t n, uint32_t *out)
{
    switch (leaf) {
    case LEAF_COPY: memcpy(out, in, (size_t)n * 4); return n;
    case LEAF_DELTA: return delta_encode(in, n, out);
    case LEAF_BP32: return bp32_encode(in, n, out);
    case LEAF_VARINT: return varint_encode(in, n, out);
    }
    return -1;
}

static int64_t leaf_decode(int leaf, const uint32_t *in, int64_t nw, uint32_t *out, int64_t cap)
{
    switch (leaf) {
    case LEAF_COPY:
        if (nw > cap) return -1;
        memcpy(out, in, (size_t)nw * 4);
        return nw;
    case LEAF_DELTA:
        if (nw > cap) return -1;
        memcpy(out, in, (size_t)nw * 4);
        delta_decode_inplace(out, nw);
        return nw;
    case LEAF_BP32: return bp32_decode(in, nw, out, cap, NULL);
    case LEAF_VARINT: return varint_decode(in, nw, out, cap);
    }
    return -1;
}

/* Composite(BP32, X) (DeviceCompositeCodec.cu:28-97): out[0] = BP32 word
 * count for the first floor(n/8)*8 values (minEffectiveLength 8,
 * DeviceBitPackingCodec.h:117), then BP32 words, then X for the rest. */
static int64_t composite_encode(int leaf2, const uint32_t *in, int64_t n, uint32_t *out)
{
    int64_t n1 = (n / 8) * 8, w1 = 0, w2 = 0;
    if (n1)
        w1 = bp32_encode(in, n1, out + 1);
    if (n - n1)
        w2 = leaf_encode(leaf2, in + n1, n - n1, out + 1 + w1);
    out[0] = (uint32_t)w1;
    return 1 + w1 + w2;
}

static int64_t composite_decode(int leaf2, const uint32_t *in, int64_t nw, uint32_t *out, int64_t cap)
{
    if (nw < 1)
        return -1;
    int64_t w1 = in[0], n1 = 0;
    if (w1 + 1 > nw)
        return -1;
    if (w1) {
        int64_t used = 0;
        n1 = bp32_decode(in + 1, w1, out, cap, &used);
        if (n1 < 0 || used != w1)
            return -1;
    }
    int64_t w2 = nw - 1 - w1, n2 = 0;
    if (w2) {
        n2 = leaf_decode(leaf2, in + 1 + w1, w2, out + n1, cap - n1);
        if (n2 < 0)
            return -1;
    }
    return n1 + n2;
}

static int inner_kind(int codec, int *leaf2)
{
    /* returns 0: plain leaf, 1: composite(BP32, leaf2) */
    switch (codec) {
    case ORC_COPY: case ORC_SERIAL_COPY_COPY: case ORC_SERIAL_DELTA_COPY: *leaf2 = LEAF_COPY; return 0;
    case ORC_DELTA: case ORC_SERIAL_DELTA_DELTA: *leaf2 = LEAF_DELTA; return 0;
    case ORC_BP32: case ORC_SERIAL_DELTA_BP32: *leaf2 = LEAF_BP32; return 0;
    case ORC_VARINT: case ORC_SERIAL_DELTA_VARINT: *leaf2 = LEAF_VARINT; return 0;
    case ORC_COMP_BP32_COPY: case ORC_SERIAL_DELTA_COMP_BP32_COPY: *leaf2 = LEAF_COPY; return 1;
    case ORC_COMP_BP32_VARINT: case ORC_SERIAL_DELTA_COMP_BP32_VARINT: *leaf2 = LEAF_VARINT; return 1;
    }
    return -1;
}

/* Serial(Delta, X) runs delta first, X on its output (DeviceSerialCodec.cu:
 * 47-91); Serial(Copy, Copy) is Copy twice. */
static int serial_delta(int codec)
{
    return codec >= ORC_SERIAL_DELTA_COPY && codec <= ORC_SERIAL_DELTA_COMP_BP32_VARINT;
}

int64_t orc_codec_bound(int64_t n) { return 2 * n + 128; }

int64_t orc_encode(int codec, const uint32_t *in, int64_t n, uint32_t *out)
{
    int leaf2;
    int kind = inner_kind(codec, &leaf2);
    if (kind < 0 || n < 0)
        return -1;
    const uint32_t *src = in;
    uint32_t *tmp = NULL;
    if (serial_delta(codec)) {
        tmp = (uint32_t *)malloc((size_t)(n ? n : 1) * 4);
        delta_encode(in, n, tmp);
        src = tmp;
    }
    int64_t w = kind == 1 ? composite_encode(leaf2, src, n, out) : leaf_encode(leaf2, src, n, out);
    free(tmp);
    return w;
}

int64_t orc_decode(int codec, const uint32_t *in, int64_t n_words, uint32_t *out, int64_t cap)
{
    int leaf2;
    int kind = inner_kind(codec, &leaf2);
    if (kind < 0 || n_words < 0)
        return -1;
    int64_t n = kind == 1 ? composite_decode(leaf2, in, n_words, out, cap)
                          : leaf_decode(leaf2, in, n_words, out, cap);
    if (n >= 0 && serial_delta(codec))
        delta_decode_inplace(out, n);
    return n;
}

int64_t orc_compress_index(const orc_index *ix, int codec, int64_t max_length,
                           uint32_t *out_words, int32_t *out_pos)
{
    int64_t w = 0;
    out_pos[0] = 0;
    for (int64_t s = 0; s < ix->n_seg; ++s) {
        int64_t a = ix->inv_pos[s], b = ix->inv_pos[s + 1];
        if (b - a <= 0 || b - a > max_length)
            return -1;
        int64_t nw = orc_encode(codec, (const uint32_t *)ix->inv + a, b - a, out_words + w);
        if (nw <= 0 || nw > max_length)
            return -1;
        w += nw;
        out_pos[s + 1] = (int32_t)w;
    }
    return w;
}

/* ====================================================================== */
/* Query compilation                                                      */
/* ====================================================================== */

static int cmp_range_dim(const void *a, const void *b)
{
    const orc_range *x = *(const orc_range *const *)a, *y = *(const orc_range *const *)b;
    if (x->dim != y->dim)
        return x->dim < y->dim ? -1 : 1;
    return x < y ? -1 : (x > y); /* stable: insertion order */
}

int64_t orc_compile_query(const orc_index *ix, const orc_range *ranges, int64_t n_ranges,
                          int32_t max_load, const int32_t *compr_pos,
                          orc_item *out, int64_t out_cap)
{
    const orc_range **ord = (const orc_range **)malloc(sizeof(void *) * (n_ranges ? n_ranges : 1));
    int64_t m = 0;
    for (int64_t i = 0; i < n_ranges; ++i)
        if (ranges[i].dim >= 0 && ranges[i].dim < ix->n_dims)
            ord[m++] = &ranges[i];
    qsort(ord, (size_t)m, sizeof(void *), cmp_range_dim);
    int64_t n = 0;
    for (int64_t i = 0; i < m; ++i) {
        const orc_range *r = ord[i];
        int32_t d = r->dim, low = r->low, up = r->up;
        if (low > up || low > ix->hi[d] || up < ix->lo[d])
            continue;
        if (low < ix->lo[d]) low = ix->lo[d];
        if (up > ix->hi[d]) up = ix->hi[d];
        const int32_t *ks = ix->keyseg + ix->keyseg_off[d];
        int32_t s0 = ks[low - ix->lo[d]], s1 = ks[up - ix->lo[d] + 1];
        if (compr_pos) {
            for (int32_t s = s0; s < s1; ++s) {
                if (n >= out_cap) { free(ord); return -1; }
                orc_item it = {r->query, r->order, compr_pos[s], compr_pos[s + 1], r->weight};
                out[n++] = it;
            }
            continue;
        }
        int32_t a = ix->inv_pos[s0], b = ix->inv_pos[s1];
        if (max_load > 0 && (int64_t)(b - a) >= max_load) {
            int32_t len = b - a, j = 1;
            for (; (int64_t)max_load * j <= len; ++j) {
                if (n >= out_cap) { free(ord); return -1; }
                int32_t st = max_load * (j - 1) + a;
                orc_item it = {r->query, r->order, st, st + max_load, r->weight};
                out[n++] = it;
            }
            if (max_load * (j - 1) != len) {
                if (n >= out_cap) { free(ord); return -1; }
                orc_item it = {r->query, r->order, max_load * (j - 1) + a, b, r->weight};
                out[n++] = it;
            }
            continue;
        }
        if (n >= out_cap) { free(ord); return -1; }
        orc_item it = {r->query, r->order, a, b, r->weight};
        out[n++] = it;
    }
    free(ord);
    return n;
}

/* ====================================================================== */
/* Counting and top-k                                                     */
/* ====================================================================== */

void orc_count_exhaustive_single(const int32_t *rows, int32_t n_rows, int32_t n_dims,
                                 const orc_range *ranges, int64_t n_ranges, uint32_t *counts)
{
    memset(counts, 0, sizeof(uint32_t) * (size_t)n_rows);
    for (int64_t r = 0; r < n_ranges; ++r) {
        int32_t d = ranges[r].dim;
        if (d < 0 || d >= n_dims || ranges[r].low > ranges[r].up)
            continue;
        uint32_t w = (uint32_t)ranges[r].weight;
        for (int32_t o = 0; o < n_rows; ++o) {
            int32_t x = rows[(int64_t)o * n_dims + d];
            if (x >= ranges[r].low && x <= ranges[r].up)
                counts[o] += w;
        }
    }
}

void orc_count_exhaustive_biject(const int64_t *row_ptr, const int32_t *vals, int32_t n_rows,
                                 const orc_range *ranges, int64_t n_ranges, uint32_t *counts)
{
    memset(counts, 0, sizeof(uint32_t) * (size_t)n_rows);
    for (int64_t r = 0; r < n_ranges; ++r) {
        if (ranges[r].dim != 0 || ranges[r].low > ranges[r].up)
            continue;
        uint32_t w = (uint32_t)ranges[r].weight;
        for (int32_t o = 0; o < n_rows; ++o)
            for (int64_t j = row_ptr[o]; j < row_ptr[o + 1]; ++j)
                if (vals[j] >= ranges[r].low && vals[j] <= ranges[r].up)
                    counts[o] += w;
    }
}

static int count_items(const uint32_t *words, int codec, const orc_item *items, int64_t n_items,
                       uint32_t *counts, int32_t n_objects, uint32_t *scratch)
{
    for (int64_t i = 0; i < n_items; ++i) {
        const orc_item *it = &items[i];
        uint32_t w = (uint32_t)it->weight;
        if (it->start_pos >= it->end_pos)
            continue;
        if (codec == ORC_NO_COMPRESSION) {
            for (int32_t p = it->start_pos; p < it->end_pos; ++p) {
                uint32_t id = words[p];
                if ((int64_t)id >= n_objects)
                    return -1;
                counts[id] += w;
            }
        } else {
            int64_t n = orc_decode(codec, words + it->start_pos, it->end_pos - it->start_pos,
                                   scratch, 4096);
            if (n < 0)
                return -1;
            for (int64_t j = 0; j < n; ++j) {
                if ((int64_t)scratch[j] >= n_objects)
                    return -1;
                counts[scratch[j]] += w;
            }
        }
    }
    return 0;
}

int orc_count_index(const uint32_t *words, int codec, const orc_item *items, int64_t n_items,
                    uint32_t *counts, int32_t n_objects)
{
    uint32_t scratch[4096];
    memset(counts, 0, sizeof(uint32_t) * (size_t)n_objects);
    return count_items(words, codec, items, n_items, counts, n_objects, scratch);
}

typedef struct {
    uint32_t c;
    int32_t id;
} cand;

static int cmp_cand(const void *a, const void *b)
{
    const cand *x = (const cand *)a, *y = (const cand *)b;
    if (x->c != y->c)
        return x->c > y->c ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id);
}

/* Exact top-k: threshold c_k from a pass over the counts, then every id with
 * count > c_k and the smallest ids with count == c_k, finally sorted. */
static void topk_core(const uint32_t *counts, int32_t n, int32_t k, int32_t *ids, int32_t *cnt,
                      cand *buf)
{
    int64_t m = 0;
    if (k <= 0)
        return;
    /* find c_k by counting: first the max, then a histogram when counts are small */
    uint32_t mx = 0;
    for (int32_t o = 0; o < n; ++o)
        if (counts[o] > mx) mx = counts[o];
    uint32_t ck = 1;
    if (mx > 0) {
        if (mx <= 65536) {
            int64_t *h = (int64_t *)calloc(mx + 1, sizeof(int64_t));
            for (int32_t o = 0; o < n; ++o) h[counts[o]]++;
            int64_t acc = 0;
            ck = 1;
            for (uint32_t c = mx; c >= 1; --c) {
                acc += h[c];
                if (acc >= k) { ck = c; break; }
            }
            free(h);
        } else {
            ck = 1; /* rare: fall back to collecting all nonzero */
        }
        for (int32_t o = 0; o < n; ++o)
            if (counts[o] >= ck)
                buf[m++] = (cand){counts[o], o};
        qsort(buf, (size_t)m, sizeof(cand), cmp_cand);
    }
    for (int32_t j = 0; j < k; ++j) {
        if (j < m) { ids[j] = buf[j].id; cnt[j] = (int32_t)buf[j].c; }
        else { ids[j] = 0; cnt[j] = 0; }
    }
}

void orc_topk(const uint32_t *counts, int32_t n_objects, int32_t k,
              int32_t *out_ids, int32_t *out_counts)
{
    cand *buf = (cand *)malloc(sizeof(cand) * (size_t)(n_objects ? n_objects : 1));
    topk_core(counts, n_objects, k, out_ids, out_counts, buf);
    free(buf);
}

typedef struct {
    const uint32_t *words;
    int codec;
    int32_t n_objects;
    const orc_item *items;
    const int64_t *q_ptr;
    int32_t n_queries;
    const int32_t *kq;
    const int64_t *off;
    int32_t *ids, *cnt;
    int32_t next; /* shared work counter */
    int failed;
} mt_job;

static void *mt_worker(void *arg)
{
    mt_job *J = (mt_job *)arg;
    uint32_t *counts = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)J->n_objects);
    cand *buf = (cand *)malloc(sizeof(cand) * (size_t)J->n_objects);
    uint32_t *scratch = (uint32_t *)malloc(sizeof(uint32_t) * 4096);
    for (;;) {
        int32_t q = __atomic_fetch_add(&J->next, 1, __ATOMIC_RELAXED);
        if (q >= J->n_queries)
            break;
        memset(counts, 0, sizeof(uint32_t) * (size_t)J->n_objects);
        if (count_items(J->words, J->codec, J->items + J->q_ptr[q], J->q_ptr[q + 1] - J->q_ptr[q],
                        counts, J->n_objects, scratch))
            J->failed = 1;
        topk_core(counts, J->n_objects, J->kq[q], J->ids + J->off[q], J->cnt + J->off[q], buf);
    }
    free(counts);
    free(buf);
    free(scratch);
    return NULL;
}

int orc_search_mt(const uint32_t *words, int codec, int32_t n_objects,
                  const orc_item *items, const int64_t *q_ptr, int32_t n_queries,
                  const int32_t *k_per_query, const int64_t *out_off,
                  int32_t *out_ids, int32_t *out_counts, int n_threads)
{
    if (n_threads <= 0)
        n_threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (n_threads < 1)
        n_threads = 1;
    mt_job J = {words, codec, n_objects, items, q_ptr, n_queries, k_per_query, out_off,
                out_ids, out_counts, 0, 0};
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
    for (int t = 0; t < n_threads; ++t)
        pthread_create(&th[t], NULL, mt_worker, &J);
    for (int t = 0; t < n_threads; ++t)
        pthread_join(th[t], NULL);
    free(th);
    return J.failed ? -1 : n_threads;
}
