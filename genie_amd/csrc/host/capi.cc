// C entry points over the C++ host library (include/genie_table.h).
#include "genie_table.h"

#include <genie/matching/knn.h>
#include <genie/original/interface.h>
#include <genie/interface/wire.h>
#include <genie/utility/serialization.h>

#include <cstring>

#include <memory>
#include <new>
#include <string>

#include "host_common.h"

struct genie_table {
    std::shared_ptr<genie::table::inv_table> t;
};

namespace {

thread_local std::string g_msg;

template <class F>
int guard(F f)
{
    try {
        f();
        return GENIE_OK;
    } catch (genie::exception::gpu_bad_alloc& e) {
        g_msg = e.what();
        return GENIE_E_OOM;
    } catch (genie::exception::gpu_runtime_error& e) {
        g_msg = e.what();
        return GENIE_E_RUNTIME;
    } catch (std::bad_alloc& e) {
        g_msg = e.what();
        return GENIE_E_OOM;
    } catch (std::exception& e) {
        g_msg = e.what();
        return GENIE_E_INVALID;
    }
}

genie::table::inv_table* make_table(int codec, uint32_t max_length)
{
    if (codec == 0)
        return new genie::table::inv_table();
    auto* c = new genie::table::inv_compr_table();
    c->setCompression((genie::compression::COMPRESSION_TYPE)codec);
    c->setUncompressedPostingListMaxLength(max_length);
    return c;
}

std::vector<genie::query::Query> make_queries(genie::table::inv_table& t, const genie_range_t* ranges, uint64_t n,
                                              uint32_t n_queries, int use_lb, uint32_t k)
{
    std::vector<genie::query::Query> qs;
    qs.reserve(n_queries);
    for (uint32_t q = 0; q < n_queries; ++q) {
        qs.emplace_back(t, (int)q);
        qs.back().topk((int)k);
        qs.back().use_load_balance = use_lb != 0;
    }
    for (uint64_t i = 0; i < n; ++i) {
        const genie_range_t& r = ranges[i];
        if (r.query < 0 || (uint32_t)r.query >= n_queries)
            throw genie::exception::cpu_runtime_error("range refers to an unknown query");
        qs[(size_t)r.query].attr(r.dim, r.low, r.up, r.weight, r.order);
    }
    return qs;
}

}  // namespace

extern "C" {

const char* genie_table_last_error(void) { return g_msg.c_str(); }

int genie_table_build_rows(const int32_t* rows, uint32_t n, uint32_t d, uint32_t max_length, int use_load_balance,
                           int codec, int via_lists, genie_table** out)
{
    return guard([&] {
        std::unique_ptr<genie::table::inv_table> t(make_table(codec, max_length));
        if (via_lists) {
            genie::table::inv_list list;
            for (uint32_t j = 0; j < d; ++j) {
                std::vector<int> col(n);
                for (uint32_t i = 0; i < n; ++i)
                    col[i] = rows[(size_t)i * d + j];
                list.invert(col);
                t->append(list);
            }
            t->build(max_length, use_load_balance != 0);
        } else {
            t->build_from_rows(rows, n, d, max_length, use_load_balance != 0);
        }
        *out = new genie_table{std::move(t)};
    });
}

int genie_table_build_sets(const int64_t* row_ptr, const int32_t* vals, uint32_t n, uint32_t max_length,
                           int use_load_balance, int codec, int via_lists, genie_table** out)
{
    return guard([&] {
        std::unique_ptr<genie::table::inv_table> t(make_table(codec, max_length));
        if (via_lists) {
            std::vector<std::vector<int> > rows(n);
            for (uint32_t i = 0; i < n; ++i)
                rows[i].assign(vals + row_ptr[i], vals + row_ptr[i + 1]);
            genie::table::inv_list list;
            list.invert_bijectMap(rows);
            t->append(list);
            t->build(max_length, use_load_balance != 0);
        } else {
            t->build_from_sets(row_ptr, vals, n, max_length, use_load_balance != 0);
        }
        *out = new genie_table{std::move(t)};
    });
}

void genie_table_destroy(genie_table* t) { delete t; }

int genie_table_sizes(genie_table* tt, uint64_t* n_post, uint64_t* n_seg, uint64_t* n_words, uint32_t* n_dims,
                      uint32_t* n_objects, uint64_t* n_keyseg)
{
    return guard([&] {
        genie::table::inv_table& t = *tt->t;
        auto* c = dynamic_cast<genie::table::inv_compr_table*>(&t);
        const std::vector<int>& inv = c ? *c->uncompressedInv() : *t.inv();
        const std::vector<int>& pos = c ? *c->uncompressedInvPos() : *t.inv_pos();
        *n_post = inv.size();
        *n_seg = pos.size() - 1;
        *n_words = c ? c->compressedInv()->size() : 0;
        *n_dims = (uint32_t)t.m_size();
        *n_objects = (uint32_t)t.i_size();
        uint64_t ks = 0;
        for (int d = 0; d < t.m_size(); ++d)
            ks += (uint64_t)(t.get_upperbound_of_list(d) - t.get_lowerbound_of_list(d)) + 2;
        *n_keyseg = ks;
    });
}

int genie_table_export(genie_table* tt, int32_t* inv, int32_t* inv_pos, uint32_t* words, int32_t* cpos, int32_t* lo,
                       int32_t* hi, int32_t* keyseg)
{
    return guard([&] {
        genie::table::inv_table& t = *tt->t;
        auto* c = dynamic_cast<genie::table::inv_compr_table*>(&t);
        const std::vector<int>& vinv = c ? *c->uncompressedInv() : *t.inv();
        const std::vector<int>& vpos = c ? *c->uncompressedInvPos() : *t.inv_pos();
        if (inv)
            std::copy(vinv.begin(), vinv.end(), inv);
        if (inv_pos)
            std::copy(vpos.begin(), vpos.end(), inv_pos);
        if (c && words)
            std::copy(c->compressedInv()->begin(), c->compressedInv()->end(), words);
        if (c && cpos)
            std::copy(c->compressedInvPos()->begin(), c->compressedInvPos()->end(), cpos);
        uint64_t k = 0;
        for (int d = 0; d < t.m_size(); ++d) {
            const int l = t.get_lowerbound_of_list(d), h = t.get_upperbound_of_list(d);
            if (lo)
                lo[d] = l;
            if (hi)
                hi[d] = h;
            for (int v = l; v <= h + 1; ++v, ++k)
                if (keyseg)
                    keyseg[k] = t.inv_index_map()->at((size_t)((d << t.shifter()) + v - l));
        }
    });
}

int genie_table_compile(genie_table* tt, const genie_range_t* ranges, uint64_t n_ranges, uint32_t n_queries,
                        int use_load_balance, int32_t max_load, int merge_runs, genie_dim_t* out, uint64_t cap,
                        uint64_t* n_out)
{
    return guard([&] {
        std::vector<genie::query::Query> qs = make_queries(*tt->t, ranges, n_ranges, n_queries, use_load_balance, 1);
        std::vector<genie::query::Query::dim> dims;
        genie::matching::build_queries(qs, *tt->t, dims, max_load, merge_runs != 0);
        *n_out = dims.size();
        if (dims.size() > cap)
            throw genie::exception::cpu_runtime_error("output capacity too small");
        static_assert(sizeof(genie::query::Query::dim) == sizeof(genie_dim_t), "layout");
        std::copy(dims.begin(), dims.end(), reinterpret_cast<genie::query::Query::dim*>(out));
        *n_out = dims.size();
    });
}

int genie_table_save(genie_table* tt, const char* path)
{
    if (!tt || !tt->t || !path)
        return GENIE_E_INVALID;
    return guard([&] { genie::utility::SaveTable(path, tt->t); });
}

int genie_table_load(const char* path, genie_table** out)
{
    if (!path || !out)
        return GENIE_E_INVALID;
    *out = nullptr;
    return guard([&] {
        auto* g = new genie_table();
        try {
            g->t = genie::utility::LoadTable(path);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int genie_wire_parse_query(const char* json, int32_t* topk, uint32_t* n_queries)
{
    if (!json || !topk || !n_queries)
        return GENIE_E_INVALID;
    return guard([&] {
        genie::original::GPUGenie_Config config;
        genie::wire::QueryBatch batch;
        std::string err;
        if (!genie::wire::ParseQuery(json, config, batch, &err))
            throw genie::exception::cpu_runtime_error(err);
        *topk = batch.topk;
        *n_queries = (uint32_t)batch.content.size();
    });
}

int genie_wire_parse_config(const char* json, int32_t* dim, int32_t* radius, int32_t* num_of_cluster,
                            int32_t* data_format, float* selectivity)
{
    if (!json)
        return GENIE_E_INVALID;
    return guard([&] {
        genie::original::GPUGenie_Config config;
        genie::wire::DistConfig extra;
        std::string err;
        if (!genie::wire::ParseConfiguration(json, config, extra, &err))
            throw genie::exception::cpu_runtime_error(err);
        if (dim)
            *dim = config.dim;
        if (radius)
            *radius = config.query_radius;
        if (num_of_cluster)
            *num_of_cluster = extra.num_of_cluster;
        if (data_format)
            *data_format = extra.data_format;
        if (selectivity)
            *selectivity = config.selectivity;
    });
}

int genie_wire_search(genie_table* const* tables, uint32_t n_tables, const char* config_json, const char* query_json,
                      char* out, uint64_t cap, uint64_t* needed)
{
    if (!tables || !config_json || !query_json || !needed)
        return GENIE_E_INVALID;
    return guard([&] {
        genie::original::GPUGenie_Config config;
        genie::wire::DistConfig extra;
        genie::wire::QueryBatch batch;
        std::string err;
        if (!genie::wire::ParseConfiguration(config_json, config, extra, &err) ||
            !genie::wire::ParseQuery(query_json, config, batch, &err))
            throw genie::exception::cpu_runtime_error(err);
        std::vector<std::shared_ptr<genie::table::inv_table> > ts;
        for (uint32_t i = 0; i < n_tables; ++i) {
            if (!tables[i] || !tables[i]->t)
                throw genie::exception::cpu_runtime_error("null table");
            ts.push_back(tables[i]->t);
        }
        std::vector<genie::wire::Result> res = genie::wire::ExecuteMultitableQuery(config, ts, batch);
        const std::string text = genie::wire::GenerateOutput(res, config.num_of_topk);
        *needed = text.size() + 1;
        if (!out || cap < text.size() + 1)
            throw genie::exception::cpu_runtime_error("output buffer too small");
        std::memcpy(out, text.c_str(), text.size() + 1);
    });
}

int genie_table_device(genie_table* tt, genie_index** out)
{
    return guard([&] { *out = tt->t->device_index(); });
}

int genie_table_knn(genie_table* tt, const genie_range_t* ranges, uint64_t n_ranges, uint32_t n_queries, uint32_t k,
                    int use_load_balance, int32_t max_load, int32_t* out_ids, int32_t* out_counts)
{
    return guard([&] {
        std::vector<genie::query::Query> qs = make_queries(*tt->t, ranges, n_ranges, n_queries, use_load_balance, k);
        genie::original::GPUGenie_Config config;
        config.num_of_topk = (int)k;
        config.use_load_balance = use_load_balance != 0;
        if (max_load > 0) {
            config.posting_list_max_length = max_load;
            config.multiplier = 1.0f;
        }
        std::vector<int> ids, counts;
        genie::original::knn_search(*tt->t, qs, ids, counts, config);
        std::copy(ids.begin(), ids.end(), out_ids);
        std::copy(counts.begin(), counts.end(), out_counts);
    });
}

int genie_knn_mt_admit(const uint64_t* estimate, uint64_t n_tables, uint64_t start, uint64_t free_bytes,
                       uint64_t* finish)
{
    if ((!estimate && n_tables) || !finish)
        return GENIE_E_INVALID;
    return guard([&] {
        const std::vector<uint64_t> est(estimate, estimate + n_tables);
        *finish = genie::matching::admit_tables(est, (size_t)start, free_bytes);
    });
}

int genie_tables_knn_mt(genie_table* const* tables, uint32_t n_tables, const genie_range_t* ranges,
                        const uint64_t* range_ptr, const uint32_t* n_queries, uint32_t k, uint64_t budget,
                        int32_t* out_ids, int32_t* out_counts)
{
    if ((!tables || !range_ptr || !n_queries) && n_tables)
        return GENIE_E_INVALID;
    return guard([&] {
        std::vector<genie::table::inv_table*> raw(n_tables);
        std::vector<std::vector<genie::query::Query> > qs(n_tables);
        std::vector<genie::original::GPUGenie_Config> configs(std::max<uint32_t>(1, n_tables));
        for (uint32_t i = 0; i < n_tables; ++i) {
            if (!tables[i] || !tables[i]->t)
                throw genie::exception::cpu_runtime_error("null table");
            raw[i] = tables[i]->t.get();
            qs[i] = make_queries(*raw[i], ranges + range_ptr[i], range_ptr[i + 1] - range_ptr[i], n_queries[i], 0, k);
            configs[i].num_of_topk = (int)k;
        }
        std::vector<std::vector<int> > ids, counts;
        genie::matching::set_knn_mt_memory_budget(budget);
        try {
            genie::original::knn_search_MT(raw, qs, ids, counts, configs);
        } catch (...) {
            genie::matching::set_knn_mt_memory_budget(0);
            throw;
        }
        genie::matching::set_knn_mt_memory_budget(0);
        uint64_t off = 0;
        for (uint32_t i = 0; i < n_tables; ++i) {
            std::copy(ids[i].begin(), ids[i].end(), out_ids + off);
            std::copy(counts[i].begin(), counts[i].end(), out_counts + off);
            off += (uint64_t)n_queries[i] * k;
        }
    });
}

}  // extern "C"
