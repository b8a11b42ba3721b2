/*
 * CPU oracle for the GENIE match-count top-k path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg, and by nothing else.  The product (genie_amd/, include/) never calls it.
 *
 * Every function restates the reference algorithm it is named after, with the
 * reference file:line it follows (paths relative to the reference root).
 * Pinning of each part (golden vectors, closed forms, brute force) lives in
 * tests/test_oracle_*.py; no function here is "parity unpinned".
 */
#ifndef GENIE_ORACLE_H_
#define GENIE_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flattened inverted index with the reference's layout
 * (src/genie/table/inv_table.cu:268-350, inv_table.h:110-174).
 * key(d, v) = (d << 16) + v - lo[d]; keyseg holds, for every dim d, the
 * entries of _inv_index_map for v = lo[d] .. hi[d] + 1 (so hi-lo+2 entries),
 * starting at keyseg_off[d]. */
typedef struct orc_index {
    int32_t  n_objects;     /* inv_table::i_size() */
    int32_t  n_dims;        /* inv_table::m_size() */
    int32_t *lo;            /* [n_dims] inv_list_lowerbound */
    int32_t *hi;            /* [n_dims] inv_list_upperbound */
    int64_t  n_post;        /* _inv.size() */
    int32_t *inv;           /* [n_post] */
    int64_t  n_seg;         /* _inv_pos.size() - 1 */
    int32_t *inv_pos;       /* [n_seg + 1] */
    int64_t *keyseg_off;    /* [n_dims + 1] */
    int32_t *keyseg;        /* [keyseg_off[n_dims]] */
} orc_index;

/* Work item with the reference's Query::dim layout (src/genie/query/query.h:59-68). */
typedef struct orc_item {
    int32_t query;
    int32_t order;
    int32_t start_pos;
    int32_t end_pos;
    float   weight;
} orc_item;

/* Query range (Query::range subset, query.h:43-54). */
typedef struct orc_range {
    int32_t query;
    int32_t dim;
    int32_t low;
    int32_t up;
    float   weight;
    int32_t order;
} orc_range;

/* ---- index build ---------------------------------------------------- */
/* Single-value mode: inv_list::invert per column (inv_list.cc:295-320) then
 * inv_table::append / build (inv_table.cu:107-128, 268-350).  rows is
 * row-major [n_rows][n_dims].  Returns 0 on success. */
int orc_build_single(const int32_t *rows, int32_t n_rows, int32_t n_dims,
                     int64_t max_length, int use_load_balance, orc_index *out);
/* bijectMap mode: inv_list::invert_bijectMap (inv_list.cc:51-86): one dim,
 * row ids repeated once per occurrence.  CSR rows (row_ptr [n_rows+1]). */
int orc_build_biject(const int64_t *row_ptr, const int32_t *vals, int32_t n_rows,
                     int64_t max_length, int use_load_balance, orc_index *out);
void orc_index_free(orc_index *ix);

/* ---- codecs (src/genie/compression/) --------------------------------- */
/* Codec ids follow COMPRESSION_TYPE (DeviceCodecFactory.h:38-53). */
enum {
    ORC_NO_COMPRESSION = 0, ORC_COPY, ORC_DELTA, ORC_BP32, ORC_VARINT,
    ORC_COMP_BP32_COPY, ORC_COMP_BP32_VARINT, ORC_SERIAL_COPY_COPY,
    ORC_SERIAL_DELTA_COPY, ORC_SERIAL_DELTA_DELTA, ORC_SERIAL_DELTA_VARINT,
    ORC_SERIAL_DELTA_BP32, ORC_SERIAL_DELTA_COMP_BP32_COPY,
    ORC_SERIAL_DELTA_COMP_BP32_VARINT, ORC_N_CODECS
};
/* Encode n values; out must hold at least orc_codec_bound(n) words.
 * Returns the number of words written, or -1 for an unknown codec. */
int64_t orc_encode(int codec, const uint32_t *in, int64_t n, uint32_t *out);
/* Decode a compressed array of n_words words; out must hold cap values.
 * Returns the number of values decoded, or -1 on error. */
int64_t orc_decode(int codec, const uint32_t *in, int64_t n_words, uint32_t *out, int64_t cap);
int64_t orc_codec_bound(int64_t n);

/* inv_compr_table::build (inv_compr_table.cu:31-119): encode every
 * [inv_pos[s], inv_pos[s+1]) segment independently.  out_words must hold
 * orc_codec_bound per segment summed; out_pos has n_seg+1 entries.
 * Returns total words, or -1 if a segment is longer than max_length or its
 * compressed form exceeds max_length words. */
int64_t orc_compress_index(const orc_index *ix, int codec, int64_t max_length,
                           uint32_t *out_words, int32_t *out_pos);

/* ---- query compilation (src/genie/query/query.cc) --------------------- */
/* Query::build (query.cc:273-333) for the ranges of ONE query, given in the
 * order they were attr()-ed; ranges on dims outside [0, n_dims) are dropped as
 * Query::attr does (query.cc:94-95); output ordered by dim then insertion
 * (std::map order, query.cc:279, 536-560).  With max_load > 0 each item is
 * split as build_and_apply_load_balance does (query.cc:141-191).  With
 * compr_pos != NULL items are per segment and carry compressed positions, as
 * build_compressed does (query.cc:464-533).  Returns #items written to out
 * (out must hold at most orc_compile_bound items). */
int64_t orc_compile_query(const orc_index *ix, const orc_range *ranges, int64_t n_ranges,
                          int32_t max_load, const int32_t *compr_pos,
                          orc_item *out, int64_t out_cap);

/* ---- counting -------------------------------------------------------- */
/* Closed-form count straight from the rows, independent of the index:
 * count(o) = sum over ranges (d, low, up, w) of w * #{values of o in dim d
 * inside [low, up]}.  Single-value rows: [n_rows][n_dims]; bijectMap rows are
 * CSR with every range on dim 0.  counts has n_rows entries. */
void orc_count_exhaustive_single(const int32_t *rows, int32_t n_rows, int32_t n_dims,
                                 const orc_range *ranges, int64_t n_ranges, uint32_t *counts);
void orc_count_exhaustive_biject(const int64_t *row_ptr, const int32_t *vals, int32_t n_rows,
                                 const orc_range *ranges, int64_t n_ranges, uint32_t *counts);
/* Index-based count (the CPU-Idx restatement, test/test_compression_query.cu:
 * 256-257): count[inv[p]] += w for p in every item span.  With codec != 0 the
 * spans are compressed word ranges decoded first.  Returns 0, or -1 on error. */
int orc_count_index(const uint32_t *words, int codec, const orc_item *items, int64_t n_items,
                    uint32_t *counts, int32_t n_objects);

/* Top-k with the build's order: count descending then id ascending; only
 * counts >= 1; padding (id 0, count 0) (heap_count.cu:254-255). */
void orc_topk(const uint32_t *counts, int32_t n_objects, int32_t k,
              int32_t *out_ids, int32_t *out_counts);

/* Multi-threaded CPU match-count search over a batch: per query, a dense
 * counter array, += w over every item span (decoding compressed segments
 * when codec != 0), then the exact top-k above.  items are grouped by
 * q_ptr (CSR over queries, [n_queries+1]); k_per_query may differ.
 * out rows start at out_off[q] (k_q slots each). n_threads <= 0 means all
 * cores.  Returns the thread count used. */
int orc_search_mt(const uint32_t *words, int codec, int32_t n_objects,
                  const orc_item *items, const int64_t *q_ptr, int32_t n_queries,
                  const int32_t *k_per_query, const int64_t *out_off,
                  int32_t *out_ids, int32_t *out_counts, int n_threads);

#ifdef __cplusplus
}
#endif

#endif
