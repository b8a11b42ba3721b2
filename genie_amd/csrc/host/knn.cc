// Matching entry points of the legacy interface on the MI355X engine
// (src/genie/matching/knn.cu:106-283, match.cu:161-220).
#include <genie/exception/exception.h>
#include <genie/matching/knn.h>
#include <genie/table/inv_compr_table.h>

#include <algorithm>
#include <new>

#include "genie_hip.h"
#include "host_common.h"

namespace genie {
namespace matching {

using genie::query::Query;

int build_queries(std::vector<Query>& queries, genie::table::inv_table& table, std::vector<Query::dim>& dims,
                  int max_load, bool merge_runs)
{
    if (table.build_status() != genie::table::inv_table::builded)
        throw genie::exception::cpu_runtime_error("table not built!");
    const bool compressed = dynamic_cast<genie::table::inv_compr_table*>(&table) != nullptr;
    int max_count = -1;
    try {
        for (Query& q : queries) {
            if (q.ref_table() != &table)
                throw genie::exception::cpu_runtime_error(
                    "Can't build queries. Queries constructed for different table!");
            if (compressed)
                q.build_compressed(max_load, merge_runs);
            else if (q.use_load_balance)
                q.build_and_apply_load_balance(max_load);
            else
                q.build();
            const size_t before = dims.size();
            q.dump(dims);
            max_count = std::max(max_count, (int)(dims.size() - before));
        }
    } catch (std::bad_alloc& e) {
        throw genie::exception::cpu_bad_alloc(e.what());
    }
    return max_count;
}

// Work items use the position of each query in `queries` as the query id.
static void run_engine(genie::table::inv_table& table, std::vector<Query>& queries, std::vector<Query::dim>& dims,
                       uint32_t k, std::vector<int>& ids, std::vector<int>& counts)
{
    // the reference keys items by Query::index(); remap to batch positions
    std::vector<int> pos_of;
    {
        int max_index = 0;
        for (const Query& q : queries)
            max_index = std::max(max_index, q.index());
        pos_of.assign((size_t)max_index + 1, -1);
        for (size_t i = 0; i < queries.size(); ++i)
            if (queries[i].index() >= 0 && pos_of[(size_t)queries[i].index()] < 0)
                pos_of[(size_t)queries[i].index()] = (int)i;
    }
    std::vector<genie_dim_t> items;
    items.reserve(dims.size());
    for (const Query::dim& d : dims) {
        genie_dim_t it;
        it.query = (d.query >= 0 && (size_t)d.query < pos_of.size()) ? pos_of[(size_t)d.query] : -1;
        it.order = d.order;
        it.start_pos = d.start_pos;
        it.end_pos = d.end_pos;
        it.weight = d.weight;
        if (it.query < 0)
            throw genie::exception::cpu_runtime_error("work item refers to an unknown query");
        items.push_back(it);
    }
    std::vector<uint32_t> kq(queries.size(), k);
    ids.assign(queries.size() * (size_t)k, 0);
    counts.assign(queries.size() * (size_t)k, 0);
    genie_amd::check(genie_hip_match_topk(table.device_index(), items.data(), items.size(),
                                          (uint32_t)queries.size(), kq.data(), 0, ids.data(), counts.data()));
}

void knn(genie::table::inv_table& table, std::vector<Query>& queries, std::vector<int>& h_top_indexes,
         std::vector<int>& h_top_count, int hash_table_size, int max_load, int bitmap_bits)
{
    (void)hash_table_size;
    (void)bitmap_bits;
    if (queries.empty())
        throw genie::exception::cpu_runtime_error("Queries not loaded!");
    std::vector<Query::dim> dims;
    build_queries(queries, table, dims, max_load);
    // output slots per query follow the first query's k (knn.cu:158-159)
    const int k = queries[0].topk();
    if (k <= 0 || k > GENIE_MAX_K)
        throw genie::exception::cpu_runtime_error("k must be in [1, " + std::to_string(GENIE_MAX_K) + "]");
    run_engine(table, queries, dims, (uint32_t)k, h_top_indexes, h_top_count);
}

void knn_MT(std::vector<genie::table::inv_table*>& table, std::vector<std::vector<Query> >& queries,
            std::vector<std::vector<int> >& h_top_indexes, std::vector<std::vector<int> >& h_top_count,
            std::vector<int>& hash_table_size, std::vector<int>& max_load, int bitmap_bits)
{
    if (table.size() != queries.size())
        throw genie::exception::cpu_runtime_error("one query set per table expected");
    h_top_indexes.resize(table.size());
    h_top_count.resize(table.size());
    for (size_t i = 0; i < table.size(); ++i) {
        if (queries[i].empty()) {
            h_top_indexes[i].clear();
            h_top_count[i].clear();
            continue;
        }
        knn(*table[i], queries[i], h_top_indexes[i], h_top_count[i],
            i < hash_table_size.size() ? hash_table_size[i] : 0, i < max_load.size() ? max_load[i] : 0,
            bitmap_bits);
    }
}

}  // namespace matching
}  // namespace genie
