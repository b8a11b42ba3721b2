// GENIE 1st/2nd-level API: Config, policies, Search / BuildTable /
// BuildQuery / Match, CSV loading (src/genie/interface/config.cc:10-106,
// execution_policy.cc:14-66, search.cc:11-42, io.cc:10-34;
// src/genie/execution_policy/single_value.cc:10-33, single_range.cc:11-45,
// validation.cc:7-23; table/interface.cc:8-18; query/interface.cc:7-24;
// matching/interface.cc:9-26).
#include <genie/genie.h>
#include <genie/original/interface.h>
#include <genie/utility/serialization.h>

#include <cstdio>
#include <mutex>

#include "genie_hip.h"
#include "host_common.h"

namespace genie_amd {

static std::mutex g_dev_mu;
static int g_dev = 0;
static bool g_dev_set = false;

void set_device(int id)
{
    std::lock_guard<std::mutex> lk(g_dev_mu);
    g_dev = id;
    check(genie_hip_init(id));
    g_dev_set = true;
}

void ensure_device()
{
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (!g_dev_set) {
        check(genie_hip_init(g_dev));
        g_dev_set = true;
    }
}

}  // namespace genie_amd

namespace genie {

using genie::exception::InvalidConfigurationException;

Config& Config::SetK(const uint32_t k)
{
    k_ = k;
    k_initialized_ = true;
    return *this;
}

Config& Config::SetQueryRange(const uint32_t r)
{
    query_range_ = r;
    query_range_initialized_ = true;
    return *this;
}

Config& Config::SetNumOfQueries(const uint32_t n)
{
    num_of_queries_ = n;
    num_of_queries_initialized_ = true;
    return *this;
}

Config& Config::SetSaveToGpu(const bool s)
{
    save_to_gpu_ = s;
    return *this;
}

Config& Config::SetGpuId(const uint8_t id)
{
    gpu_id_ = id;
    return *this;
}

Config& Config::LoadFromFile(const std::string&) { throw genie::exception::NotImplementedException(); }

void Config::Validate() const
{
    if (!k_initialized_)
        throw InvalidConfigurationException("K is not set.");
    if (!num_of_queries_initialized_)
        throw InvalidConfigurationException("Number of query is not set.");
}

void Config::DisplayConfiguration()
{
    std::printf("GENIE configuration:\nK: %u\nNumber of queries: %u\n", k_, num_of_queries_);
    if (IsQueryRangeSet())
        std::printf("Query range: %u\n", query_range_);
    std::printf("Save to GPU: %s\nGPU used: %u\n\n", save_to_gpu_ ? "true" : "false", (unsigned)gpu_id_);
}

namespace execution_policy {
namespace validation {

void ValidateK(uint32_t k)
{
    if (k == 0)
        throw InvalidConfigurationException("K should be larger than 0.");
}

void ValidateNumOfQueries(uint32_t n)
{
    if (n == 0)
        throw InvalidConfigurationException("Number of query should be larger than 0.");
}

void ValidateQueryRange(uint32_t r)
{
    if (r == 0)
        throw InvalidConfigurationException("Query range should be larger than 0.");
}

}  // namespace validation
}  // namespace execution_policy

void ExecutionPolicy::Validate()
{
    execution_policy::validation::ValidateK(k_);
    execution_policy::validation::ValidateNumOfQueries(num_of_queries_);
}

std::shared_ptr<ExecutionPolicy> MakePolicy(const Config& config)
{
    config.Validate();
    std::shared_ptr<ExecutionPolicy> policy;
    if (config.IsQueryRangeSet()) {
        auto* p = new execution_policy::SingleRangeExecutionPolicy();
        p->SetK(config.GetK());
        p->SetNumOfQueries(config.GetNumOfQueries());
        p->SetQueryRange(config.GetQueryRange());
        policy.reset(p);
    } else {
        auto* p = new execution_policy::SingleValueExecutionPolicy();
        p->SetK(config.GetK());
        p->SetNumOfQueries(config.GetNumOfQueries());
        policy.reset(p);
    }
    policy->Validate();
    genie_amd::set_device(config.GetGpuId());
    return policy;
}

namespace execution_policy {

// table/interface.cc:8-18 -> preprocess_for_knn_csv -> load_table
std::shared_ptr<genie::table::inv_table> SingleValueExecutionPolicy::BuildTable(const TableData& table_data)
{
    dim_ = (uint32_t)table_data.at(0).size();
    genie::original::GPUGenie_Config config;
    config.data_points = &const_cast<TableData&>(table_data);
    genie::table::inv_table* t = nullptr;
    genie::original::preprocess_for_knn_csv(config, t);
    return std::shared_ptr<genie::table::inv_table>(t, [](genie::table::inv_table* p) { delete[] p; });
}

static std::vector<genie::query::Query> load_queries(const std::shared_ptr<const genie::table::inv_table>& table,
                                                     const QueryData& query_data, uint32_t range, uint32_t k)
{
    // query/interface.cc:7-24
    std::vector<genie::query::Query> queries;
    genie::original::GPUGenie_Config config;
    config.query_points = &const_cast<QueryData&>(query_data);
    config.query_radius = (int)range;
    config.num_of_topk = (int)k;
    auto* t = const_cast<genie::table::inv_table*>(table.get());
    config.dim = t->m_size();
    genie::original::load_query_singlerange(*t, queries, config);
    return queries;
}

std::vector<genie::query::Query> SingleValueExecutionPolicy::BuildQuery(
    const std::shared_ptr<const genie::table::inv_table>& table, const QueryData& query_data)
{
    return load_queries(table, query_data, 0, k_);
}

std::vector<genie::query::Query> SingleRangeExecutionPolicy::BuildQuery(
    const std::shared_ptr<const genie::table::inv_table>& table, const QueryData& query_data)
{
    return load_queries(table, query_data, query_range_, k_);
}

void SingleRangeExecutionPolicy::Validate()
{
    ExecutionPolicy::Validate();
    validation::ValidateQueryRange(query_range_);
}

// matching/interface.cc:9-26
SearchResult SingleValueExecutionPolicy::Match(const std::shared_ptr<const genie::table::inv_table>& table,
                                               const std::vector<genie::query::Query>& queries)
{
    genie::original::GPUGenie_Config config;
    config.hashtable_size = (float)(dim_ * k_ * 1.5);
    config.count_threshold = (int)dim_;
    std::vector<int> ids, counts;
    genie::original::knn_search(*const_cast<genie::table::inv_table*>(table.get()),
                                const_cast<std::vector<genie::query::Query>&>(queries), ids, counts, config);
    return std::make_pair(std::move(ids), std::move(counts));
}

// The reference leaves multi-table Match unimplemented (single_value.cc:29-33);
// here every table is searched with its own queries and the results are
// concatenated in table order.
SearchResult SingleValueExecutionPolicy::Match(
    const std::vector<std::shared_ptr<const genie::table::inv_table> >& tables,
    const std::vector<std::vector<genie::query::Query> >& queries)
{
    if (tables.size() != queries.size())
        throw InvalidConfigurationException("one query set per table expected");
    SearchResult out;
    for (size_t i = 0; i < tables.size(); ++i) {
        if (queries[i].empty())
            continue;
        SearchResult r = Match(tables[i], queries[i]);
        out.first.insert(out.first.end(), r.first.begin(), r.first.end());
        out.second.insert(out.second.end(), r.second.begin(), r.second.end());
    }
    return out;
}

}  // namespace execution_policy

std::shared_ptr<genie::table::inv_table> LoadTableFromBinary(const std::string& filename)
{
    return genie::utility::LoadTable(filename);
}

void SaveTableToBinary(const std::string& filename, const std::shared_ptr<const genie::table::inv_table>& table)
{
    genie::utility::SaveTable(filename, table);
}

TableData LoadTableDataFromCsv(const std::string& filename)
{
    TableData d;
    genie::utility::read_file(d, filename.c_str(), -1);
    return d;
}

QueryData LoadQueryDataFromCsv(const std::string& filename, const std::shared_ptr<const ExecutionPolicy>& policy)
{
    QueryData d;
    genie::utility::read_file(d, filename.c_str(), (int)policy->GetNumOfQueries());
    return d;
}

SearchResult Search(std::shared_ptr<ExecutionPolicy>& policy, const std::string& table_filename,
                    const std::string& query_filename)
{
    TableData table_data = LoadTableDataFromCsv(table_filename);
    QueryData query_data = LoadQueryDataFromCsv(query_filename, policy);
    std::shared_ptr<genie::table::inv_table> table = BuildTable(policy, table_data);
    std::vector<genie::query::Query> queries = BuildQuery(policy, table, query_data);
    return Match(policy, table, queries);
}

std::shared_ptr<genie::table::inv_table> BuildTable(std::shared_ptr<ExecutionPolicy>& policy,
                                                    const TableData& table_data)
{
    return policy->BuildTable(table_data);
}

std::vector<genie::query::Query> BuildQuery(std::shared_ptr<ExecutionPolicy>& policy,
                                            const std::shared_ptr<genie::table::inv_table>& table,
                                            const QueryData& query_data)
{
    return policy->BuildQuery(table, query_data);
}

SearchResult Match(std::shared_ptr<ExecutionPolicy>& policy, const std::shared_ptr<genie::table::inv_table>& table,
                   const std::vector<genie::query::Query>& queries)
{
    return policy->Match(table, queries);
}

}  // namespace genie
