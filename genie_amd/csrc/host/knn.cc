// Matching entry points of the legacy interface on the MI355X engine
// (src/genie/matching/knn.cu:106-283, match.cu:161-220).
#include <genie/exception/exception.h>
#include <genie/matching/knn.h>
#include <genie/table/inv_compr_table.h>

#include <algorithm>
#include <atomic>
#include <new>
#include <string>

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

// Work items of the engine: the position of each query in `queries` is its
// query id (the reference keys items by Query::index()).
static std::vector<genie_dim_t> engine_items(const std::vector<Query>& queries, const std::vector<Query::dim>& dims)
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
    return items;
}

static void run_engine(genie::table::inv_table& table, std::vector<Query>& queries, std::vector<Query::dim>& dims,
                       uint32_t k, std::vector<int>& ids, std::vector<int>& counts)
{
    const std::vector<genie_dim_t> items = engine_items(queries, dims);
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

namespace {
constexpr uint64_t kAdmitTolerance = 50ull << 20;  // device memory always left free (knn.cu:209)
std::atomic<uint64_t> g_mt_budget{0};               // 0: the device's free memory
constexpr int kMtMode = 2;  // automatic plans: the dense or set form when the batch qualifies, else plain
}  // namespace

size_t admit_tables(const std::vector<uint64_t>& estimate, size_t start, uint64_t free_bytes)
{
    size_t finish = start;
    while (finish < estimate.size() && free_bytes > estimate[finish] + kAdmitTolerance) {
        free_bytes -= estimate[finish];
        ++finish;
    }
    if (finish == start && start < estimate.size())
        throw genie::exception::gpu_bad_alloc("MEMORY NOT ENOUGH");
    return finish;
}

void set_knn_mt_memory_budget(uint64_t bytes) { g_mt_budget.store(bytes); }

void knn_MT(std::vector<genie::table::inv_table*>& table, std::vector<std::vector<Query> >& queries,
            std::vector<std::vector<int> >& h_top_indexes, std::vector<std::vector<int> >& h_top_count,
            std::vector<int>& hash_table_size, std::vector<int>& max_load, int bitmap_bits)
{
    (void)hash_table_size;
    (void)bitmap_bits;
    if (table.size() != queries.size())
        throw genie::exception::cpu_runtime_error("one query set per table expected");
    const size_t T = table.size();
    h_top_indexes.assign(T, {});
    h_top_count.assign(T, {});
    // every table's work items and its plan's device memory
    std::vector<std::vector<genie_dim_t> > items(T);
    std::vector<uint32_t> kt(T, 0);
    std::vector<uint64_t> est(T, 0);
    for (size_t i = 0; i < T; ++i) {
        if (queries[i].empty())
            continue;
        std::vector<Query::dim> dims;
        build_queries(queries[i], *table[i], dims, i < max_load.size() ? max_load[i] : 0);
        const int k = queries[i][0].topk();  // output slots per query follow the first query (knn.cu:158-159)
        if (k <= 0 || k > GENIE_MAX_K)
            throw genie::exception::cpu_runtime_error("k must be in [1, " + std::to_string(GENIE_MAX_K) + "]");
        kt[i] = (uint32_t)k;
        items[i] = engine_items(queries[i], dims);
        genie_amd::check(genie_hip_plan_estimate(table[i]->device_index(), (uint32_t)queries[i].size(),
                                                 items[i].size(), kt[i], kMtMode, &est[i]));
    }
    // rounds of tables admitted while the device keeps kAdmitTolerance free
    // (knn.cu:207-243); a round's tables run concurrently, one plan and
    // stream each, and their memory is released before the next round
    for (size_t start = 0; start < T;) {
        uint64_t free_bytes = g_mt_budget.load();
        if (free_bytes == 0) {
            uint64_t total = 0;
            genie_amd::check(genie_hip_mem_info(&free_bytes, &total));
        }
        const size_t finish = admit_tables(est, start, free_bytes);
        std::vector<genie_plan*> plans(finish - start, nullptr);
        std::vector<void*> streams(finish - start, nullptr);
        auto release = [&] {
            for (size_t j = 0; j < plans.size(); ++j) {
                if (streams[j])
                    (void)genie_hip_stream_synchronize(streams[j]);
                genie_hip_plan_destroy(plans[j]);
                genie_hip_stream_destroy(streams[j]);
            }
        };
        try {
            for (size_t i = start; i < finish; ++i) {
                if (queries[i].empty())
                    continue;
                const size_t j = i - start;
                const uint32_t nq = (uint32_t)queries[i].size();
                std::vector<uint32_t> kq(nq, kt[i]);
                genie_amd::check(genie_hip_stream_create(&streams[j]));
                genie_amd::check(genie_hip_plan_create(table[i]->device_index(), nq, std::max<uint64_t>(1, items[i].size()),
                                                       kt[i], 0, kMtMode, &plans[j]));
                genie_amd::check(genie_hip_plan_upload(plans[j], items[i].data(), items[i].size(), nq, kq.data()));
                genie_amd::check(genie_hip_plan_run(plans[j], streams[j]));
            }
            for (size_t i = start; i < finish; ++i) {
                if (queries[i].empty())
                    continue;
                const size_t j = i - start;
                h_top_indexes[i].assign(queries[i].size() * (size_t)kt[i], 0);
                h_top_count[i].assign(queries[i].size() * (size_t)kt[i], 0);
                genie_amd::check(genie_hip_plan_download(plans[j], streams[j], h_top_indexes[i].data(),
                                                         h_top_count[i].data()));
            }
        } catch (...) {
            release();
            throw;
        }
        release();
        start = finish;
    }
}

}  // namespace matching
}  // namespace genie
