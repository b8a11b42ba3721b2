/*
 * genie_table.h -- C entry points over the C++ host library (table build,
 * query compilation, legacy knn_search), for callers without C++ (the Python
 * tests and benchmark use it through ctypes).  They wrap exactly the C++ API
 * of include/genie/ and report its exceptions as status codes:
 * 0 ok, GENIE_E_* of genie_hip.h otherwise, message in genie_table_last_error().
 */
#ifndef GENIE_TABLE_C_H_
#define GENIE_TABLE_C_H_

#include <stdint.h>

#include "genie_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct genie_table genie_table;

/* Query range (Query::attr arguments, src/genie/query/query.h:43-54). */
typedef struct genie_range_t {
    int32_t query;
    int32_t dim;
    int32_t low;
    int32_t up;
    float weight;
    int32_t order;
} genie_range_t;

/* Single-value table from row-major rows [n][d].  via_lists = 1 builds it
 * through inv_list::invert + inv_table::append + build (the GENIE path);
 * 0 uses the direct row builder (same layout).  codec 0 = uncompressed,
 * otherwise an inv_compr_table with that COMPRESSION_TYPE. */
int genie_table_build_rows(const int32_t* rows, uint32_t n, uint32_t d, uint32_t max_length, int use_load_balance,
                           int codec, int via_lists, genie_table** out);
/* bijectMap table from CSR set rows. */
int genie_table_build_sets(const int64_t* row_ptr, const int32_t* vals, uint32_t n, uint32_t max_length,
                           int use_load_balance, int codec, int via_lists, genie_table** out);
void genie_table_destroy(genie_table* t);

int genie_table_sizes(genie_table* t, uint64_t* n_post, uint64_t* n_seg, uint64_t* n_words, uint32_t* n_dims,
                      uint32_t* n_objects, uint64_t* n_keyseg);
/* Export the layout: inv [n_post], inv_pos [n_seg+1] (uncompressed), words
 * [n_words] and cpos [n_seg+1] (compressed tables), per-dim bounds, and the
 * _inv_index_map entries of every dim for values lo..hi+1 concatenated
 * (n_keyseg entries).  Any pointer may be NULL. */
int genie_table_export(genie_table* t, int32_t* inv, int32_t* inv_pos, uint32_t* words, int32_t* cpos, int32_t* lo,
                       int32_t* hi, int32_t* keyseg);

/* Compile ranges of n_queries queries (range.query in [0, n_queries)) with
 * Query::attr then build / build_and_apply_load_balance / build_compressed
 * and dump, query by query (match.cu:161-210).  merge_runs (compressed
 * tables): one item per range spanning its consecutive segments instead of
 * one per segment; every form of the search takes it as the same work.
 * *n_out receives the item count; when it exceeds cap nothing is copied and
 * GENIE_E_INVALID returns. */
int genie_table_compile(genie_table* t, const genie_range_t* ranges, uint64_t n_ranges, uint32_t n_queries,
                        int use_load_balance, int32_t max_load, int merge_runs, genie_dim_t* out, uint64_t cap,
                        uint64_t* n_out);

/* knn_MT's admission round (genie::matching::admit_tables): tables from
 * `start` admitted against free_bytes; GENIE_E_OOM ("MEMORY NOT ENOUGH")
 * when table `start` does not fit alone.  Host-only. */
int genie_knn_mt_admit(const uint64_t* estimate, uint64_t n_tables, uint64_t start, uint64_t free_bytes,
                       uint64_t* finish);

/* knn_search_MT over n_tables tables (interface.cc:859-909): table i's
 * ranges are ranges[range_ptr[i] .. range_ptr[i + 1]) for n_queries[i]
 * queries, k results each; results are concatenated by table (table i:
 * n_queries[i] x k slots).  budget: device memory knn_MT admits against
 * (0 = the device's free memory). */
int genie_tables_knn_mt(genie_table* const* tables, uint32_t n_tables, const genie_range_t* ranges,
                        const uint64_t* range_ptr, const uint32_t* n_queries, uint32_t k, uint64_t budget,
                        int32_t* out_ids, int32_t* out_counts);

/* Engine handle of the table (cpy_data_to_gpu); owned by the table. */
int genie_table_device(genie_table* t, genie_index** out);

/* Save / load a built table (genie::utility::SaveTable / LoadTable,
 * src/genie/utility/serialization.cc:18-56; the file format is described in
 * genie/utility/serialization.h).  A loaded table is not yet on the device. */
int genie_table_save(genie_table* t, const char* path);
int genie_table_load(const char* path, genie_table** out);

/* The serving precedent's JSON documents (genie/interface/wire.h;
 * src/dist_genie/parser.cu:16-217).  Invalid documents return
 * GENIE_E_INVALID with the reference's message in genie_table_last_error(). */
int genie_wire_parse_query(const char* json, int32_t* topk, uint32_t* n_queries);
int genie_wire_parse_config(const char* json, int32_t* dim, int32_t* radius, int32_t* num_of_cluster,
                            int32_t* data_format, float* selectivity);
/* Search a query document over the clusters tables[0..n_tables) with the
 * configuration document's settings (ExecuteMultitableQuery + MergeResult,
 * search.cu:34-52, sorting.cc:20-64) and write GenerateOutput's "id,count"
 * lines (file.cu:42-53) into out (NUL-terminated, cap bytes); *needed gets
 * the bytes required.  GENIE_E_INVALID when out is too small. */
int genie_wire_search(genie_table* const* tables, uint32_t n_tables, const char* config_json, const char* query_json,
                      char* out, uint64_t cap, uint64_t* needed);

/* Legacy search: queries from ranges, genie::original::knn_search(table,
 * queries, ids, counts, config); k slots per query. */
int genie_table_knn(genie_table* t, const genie_range_t* ranges, uint64_t n_ranges, uint32_t n_queries, uint32_t k,
                    int use_load_balance, int32_t max_load, int32_t* out_ids, int32_t* out_counts);

const char* genie_table_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
