// The serving precedent's documents and multi-cluster search
// (src/dist_genie/parser.cu, search.cu, sorting.cc, file.cu, online.cu); see
// include/genie/interface/wire.h.
#include <genie/interface/wire.h>
#include <genie/utility/serialization.h>

#include <algorithm>
#include <functional>
#include <sstream>

#include "json.h"

namespace genie {
namespace wire {

using genie::original::GPUGenie_Config;
using genie::query::Query;
using genie::table::inv_table;
using genie_amd::Json;

namespace {

bool say(std::string* error, const std::string& what)
{
    if (error)
        *error = what;
    return false;
}

}  // namespace

bool ParseConfiguration(const std::string& text, GPUGenie_Config& config, DistConfig& extra, std::string* error)
{
    Json j;
    std::string perr;
    if (!genie_amd::parse_json(text, j, &perr))
        return say(error, "configuration is not a valid JSON document: " + perr);
    if (j.kind != Json::Object)
        return say(error, "configuration is not a JSON object");
    static const char* const kString[] = {"data_file"};
    static const char* const kInt[] = {"dim",           "count_threshold", "data_type", "search_type",
                                       "max_data_size", "num_of_cluster",  "data_format", "radius"};
    static const char* const kFloat[] = {"selectivity"};
    for (const char* e : kString) {
        const Json* v = j.get(e);
        if (!v)
            return say(error, std::string("Entry ") + e + " is missing");
        if (v->kind != Json::String)
            return say(error, std::string("Entry ") + e + " has wrong type");
    }
    for (const char* e : kInt) {
        const Json* v = j.get(e);
        if (!v)
            return say(error, std::string("Entry ") + e + " is missing");
        if (!v->is_int())
            return say(error, std::string("Entry ") + e + " has wrong type");
    }
    for (const char* e : kFloat) {
        const Json* v = j.get(e);
        if (!v)
            return say(error, std::string("Entry ") + e + " is missing");
        if (!v->is_float())
            return say(error, std::string("Entry ") + e + " has wrong type");
    }
    extra.data_file = j.get("data_file")->s;
    extra.num_of_cluster = (int)j.get("num_of_cluster")->i;
    extra.data_format = (int)j.get("data_format")->i;
    config.dim = (int)j.get("dim")->i;
    config.count_threshold = (int)j.get("count_threshold")->i;
    config.query_radius = (int)j.get("radius")->i;
    config.use_device = 0;
    config.selectivity = (float)j.get("selectivity")->d;
    config.use_adaptive_range = config.selectivity > 0.0f;
    config.use_load_balance = false;
    config.posting_list_max_length = 6400;
    config.multiplier = 1.5f;
    config.use_multirange = false;
    config.save_to_gpu = true;
    config.data_type = (int)j.get("data_type")->i;
    config.search_type = (int)j.get("search_type")->i;
    config.max_data_size = (unsigned int)j.get("max_data_size")->i;
    return true;
}

bool ParseQuery(const std::string& text, GPUGenie_Config& config, QueryBatch& batch, std::string* error)
{
    Json j;
    std::string perr;
    if (!genie_amd::parse_json(text, j, &perr))
        return say(error, "Received query is not a valid JSON document: " + perr);
    if (j.kind != Json::Object)
        return say(error, "Received query is not a JSON object");
    const Json* topk = j.get("topk");
    if (!topk)
        return say(error, "Entry 'topk' is missing");
    if (!topk->is_int())
        return say(error, "Entry 'topk' should be an interger");
    const Json* qs = j.get("queries");
    if (!qs)
        return say(error, "Entry 'queries' is missing");
    if (qs->kind != Json::Array)
        return say(error, "Entry 'queries' should be an array");
    for (const Json& q : qs->a) {
        const Json* c = q.kind == Json::Object ? q.get("content") : nullptr;
        if (!c)
            return say(error, "Some query misses the 'content' section");
        if (c->kind != Json::Array)
            return say(error, "Query's 'content' section should be an array");
        const Json* cl = q.get("clusters");
        if (!cl)
            return say(error, "Some query misses the 'clusters' section");
        if (cl->kind != Json::Array)
            return say(error, "Query's 'clusters' section should be an array");
        // the reference reads the elements with GetInt(); require integers
        for (const Json& v : c->a)
            if (!v.is_int())
                return say(error, "Query's 'content' values should be integers");
        for (const Json& v : cl->a)
            if (!v.is_int())
                return say(error, "Query's 'clusters' values should be integers");
    }
    batch = QueryBatch();
    batch.topk = (int)topk->i;
    for (const Json& q : qs->a) {
        std::vector<int> content, clusters;
        for (const Json& v : q.get("content")->a)
            content.push_back((int)v.i);
        for (const Json& v : q.get("clusters")->a)
            clusters.push_back((int)v.i);
        batch.content.push_back(std::move(content));
        batch.clusters.push_back(std::move(clusters));
    }
    config.num_of_topk = batch.topk;
    config.hashtable_size = (float)(config.num_of_topk * 1.5 * config.count_threshold);
    return true;
}

std::vector<std::shared_ptr<inv_table> > ReadData(GPUGenie_Config& config, const DistConfig& extra)
{
    std::vector<std::shared_ptr<inv_table> > tables((size_t)std::max(0, extra.num_of_cluster));
    std::vector<std::vector<int> > data;
    for (int i = 0; i < extra.num_of_cluster; ++i) {
        const std::string base = extra.data_file + "_" + std::to_string(i) + "_0";
        if (extra.data_format == 0) {
            data.clear();
            genie::utility::read_file(data, (base + ".csv").c_str(), -1);
            config.data_points = &data;
            inv_table* raw = nullptr;
            genie::original::preprocess_for_knn_csv(config, raw);
            tables[(size_t)i] = std::shared_ptr<inv_table>(raw, [](inv_table* p) { delete[] p; });
            config.data_points = nullptr;
        } else if (extra.data_format == 1) {
            tables[(size_t)i] = genie::utility::LoadTable(base + ".dat");
        } else {
            throw genie::exception::cpu_runtime_error("unknown data_format " + std::to_string(extra.data_format));
        }
        if (config.save_to_gpu)
            tables[(size_t)i]->cpy_data_to_gpu();
    }
    return tables;
}

std::vector<Result> ExecuteMultitableQuery(GPUGenie_Config& config, std::vector<std::shared_ptr<inv_table> >& tables,
                                           const QueryBatch& batch)
{
    const size_t nq = batch.content.size();
    const int topk = config.num_of_topk;
    if (topk <= 0)
        throw genie::exception::cpu_runtime_error("topk must be positive");
    // queries of every cluster (Cluster::m_queries / m_queries_id, parser.cu:196-211)
    std::vector<std::vector<std::vector<int> > > cq(tables.size());
    std::vector<std::vector<int> > cid(tables.size());
    for (size_t q = 0; q < nq; ++q)
        for (int c : batch.clusters[q]) {
            if (c < 0 || (size_t)c >= tables.size() || !tables[(size_t)c])
                throw genie::exception::cpu_runtime_error("query " + std::to_string(q) + " names cluster " +
                                                          std::to_string(c) + ", which is not loaded");
            cq[(size_t)c].push_back(batch.content[q]);
            cid[(size_t)c].push_back((int)q);
        }
    // global id offset of every cluster: the sizes of the clusters before it
    std::vector<int> offset(tables.size(), 0);
    for (size_t c = 1; c < tables.size(); ++c)
        offset[c] = offset[c - 1] + (tables[c - 1] ? tables[c - 1]->i_size() : 0);
    // one batch per cluster that has queries
    std::vector<inv_table*> raw;
    std::vector<std::vector<Query> > queries;
    std::vector<GPUGenie_Config> configs;
    std::vector<size_t> which;
    for (size_t c = 0; c < tables.size(); ++c) {
        if (cq[c].empty())
            continue;
        GPUGenie_Config cc = config;
        cc.num_of_queries = (int)cq[c].size();
        cc.query_points = &cq[c];
        std::vector<Query> qv;
        genie::original::load_query(*tables[c], qv, cc);
        raw.push_back(tables[c].get());
        queries.push_back(std::move(qv));
        configs.push_back(cc);
        which.push_back(c);
    }
    std::vector<std::vector<int> > h_topk, h_count;
    if (!raw.empty())
        genie::original::knn_search_MT(raw, queries, h_topk, h_count, configs);
    std::vector<Result> results(nq);
    for (size_t t = 0; t < which.size(); ++t) {
        const size_t c = which[t];
        for (size_t i = 0; i < cq[c].size(); ++i)
            for (int k = 0; k < topk; ++k) {
                const size_t slot = i * (size_t)topk + (size_t)k;
                const int count = h_count[t].at(slot);
                const int id = count > 0 ? offset[c] + h_topk[t].at(slot) : 0;
                results[(size_t)cid[c][i]].emplace_back(count, id);
            }
    }
    for (Result& r : results)
        std::sort(r.begin(), r.end(), std::greater<std::pair<int, int> >());
    return results;
}

std::string GenerateOutput(std::vector<Result>& results, int topk)
{
    std::ostringstream out;
    for (Result& r : results) {
        std::sort(r.begin(), r.end(), std::greater<std::pair<int, int> >());
        for (int i = 0; i < topk; ++i) {
            if ((size_t)i < r.size())
                out << r[(size_t)i].second << "," << r[(size_t)i].first << "\n";
            else
                out << "0,0\n";
        }
    }
    return out.str();
}

}  // namespace wire
}  // namespace genie
