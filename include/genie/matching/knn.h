// Matching entry below the legacy interface (src/genie/matching/knn.h;
// knn.cu:106-183 knn, 185-283 knn_MT), answered by the MI355X engine.
#ifndef GENIE_MATCHING_KNN_H_
#define GENIE_MATCHING_KNN_H_

#include <cstdint>
#include <vector>

#include <genie/query/query.h>
#include <genie/table/inv_table.h>

namespace genie {
namespace matching {

// Top-k (ids, counts) of every query: queries.size() x queries[0].topk()
// slots, query-major, count descending then id ascending, padded (0, 0).
// hash_table_size and bitmap_bits are accepted for signature compatibility
// and unused; max_load drives load-balanced query compilation.
void knn(genie::table::inv_table& table, std::vector<genie::query::Query>& queries,
         std::vector<int>& h_top_indexes, std::vector<int>& h_top_count, int hash_table_size,
         int max_load, int bitmap_bits);

// Several tables, each with its own queries (knn.cu:185-283).  Tables are
// admitted in order while the device's free memory stays above their plans'
// estimated memory plus 50 MiB (knn.cu:207-243, estimate from
// genie_hip_plan_estimate); the admitted tables run concurrently, one
// automatic plan (the dense or set form when the batch qualifies, else the
// plain form) and one stream each, and are released before the next round.
// A table that does not fit alone throws gpu_bad_alloc("MEMORY NOT ENOUGH").
// Results per table: queries[i].size() x queries[i][0].topk() slots.
void knn_MT(std::vector<genie::table::inv_table*>& table, std::vector<std::vector<genie::query::Query> >& queries,
            std::vector<std::vector<int> >& h_top_indexes, std::vector<std::vector<int> >& h_top_count,
            std::vector<int>& hash_table_size, std::vector<int>& max_load, int bitmap_bits);

// knn_MT's admission: the tables from `start` on that fit one round given
// `free_bytes` (each table's estimate taken from it while more than
// estimate + 50 MiB is free); returns the end of the round.  Throws
// gpu_bad_alloc("MEMORY NOT ENOUGH") when table `start` does not fit alone.
size_t admit_tables(const std::vector<uint64_t>& estimate, size_t start, uint64_t free_bytes);

// Caps the device memory knn_MT admits against (0, the default: the free
// memory the device reports at each round).  No counterpart in the
// reference; for callers that share the device and for tests.
void set_knn_mt_memory_budget(uint64_t bytes);

// Compile the queries into work items (match.cu:161-210 build_queries);
// returns the largest number of items of one query.  merge_runs (compressed
// tables): one item per range over its consecutive segments
// (Query::build_compressed).
int build_queries(std::vector<genie::query::Query>& queries, genie::table::inv_table& table,
                  std::vector<genie::query::Query::dim>& dims, int max_load, bool merge_runs = false);

}  // namespace matching
}  // namespace genie

#endif
