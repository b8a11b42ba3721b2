// Matching entry below the legacy interface (src/genie/matching/knn.h;
// knn.cu:106-183 knn, 185-283 knn_MT), answered by the MI355X engine.
#ifndef GENIE_MATCHING_KNN_H_
#define GENIE_MATCHING_KNN_H_

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
// admitted while their estimated scratch fits in free device memory.
void knn_MT(std::vector<genie::table::inv_table*>& table, std::vector<std::vector<genie::query::Query> >& queries,
            std::vector<std::vector<int> >& h_top_indexes, std::vector<std::vector<int> >& h_top_count,
            std::vector<int>& hash_table_size, std::vector<int>& max_load, int bitmap_bits);

// Compile the queries into work items (match.cu:161-210 build_queries);
// returns the largest number of items of one query.  merge_runs (compressed
// tables): one item per range over its consecutive segments
// (Query::build_compressed).
int build_queries(std::vector<genie::query::Query>& queries, genie::table::inv_table& table,
                  std::vector<genie::query::Query::dim>& dims, int max_load, bool merge_runs = false);

}  // namespace matching
}  // namespace genie

#endif
