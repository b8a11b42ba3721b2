// Legacy GENIE interface on the MI355X engine
// (src/genie/original/interface.cc:150-909).
#include <genie/matching/knn.h>
#include <genie/original/interface.h>

#include <algorithm>
#include <map>
#include <new>

#include "genie_hip.h"
#include "host_common.h"

namespace genie {
namespace original {

using genie::query::Query;
using genie::table::inv_compr_table;
using genie::table::inv_list;
using genie::table::inv_table;

static inv_table* new_tables(size_t n, const GPUGenie_Config& config)
{
    if (config.compression == genie::compression::NO_COMPRESSION)
        return new inv_table[n];
    if (n != 1)
        throw genie::exception::cpu_runtime_error("Compression for multiple tables to yet supported");
    inv_compr_table* t = new inv_compr_table[1];
    t[0].setCompression(config.compression);
    t[0].setUncompressedPostingListMaxLength((size_t)config.posting_list_max_length);
    return t;
}

static void load_rows(inv_table& t, std::vector<std::vector<int> >& rows, GPUGenie_Config& config)
{
    switch (config.search_type) {
    case 0: load_table(t, rows, config); break;
    case 1: load_table_bijectMap(t, rows, config); break;
    case 2: throw genie::exception::cpu_runtime_error("sequence search is not supported");
    default: throw genie::exception::cpu_runtime_error("Unrecognised search type!");
    }
}

bool preprocess_for_knn_csv(GPUGenie_Config& config, inv_table*& _table)
{
    config.num_of_topk = (int)config.num_of_iteration * config.num_of_topk;
    std::vector<std::vector<int> >& data = *config.data_points;
    if (config.max_data_size >= data.size() || config.max_data_size == 0) {
        if (data.empty())
            throw genie::exception::cpu_runtime_error("no data input!");
        _table = new_tables(1, config);
        _table[0].set_table_index(0);
        _table[0].set_total_num_of_table(1);
        load_rows(_table[0], data, config);
        return true;
    }
    // multiload: split the rows into tables of max_data_size rows; the last
    // two share the remainder (interface.cc:197-278)
    const size_t n = data.size(), m = config.max_data_size;
    size_t table_num, cycle;
    if (n % m == 0) {
        table_num = n / m;
        cycle = table_num;
    } else {
        table_num = n / m + 1;
        cycle = table_num - 2;
    }
    if (config.compression != genie::compression::NO_COMPRESSION)
        throw genie::exception::cpu_runtime_error("Compression for multiple tables to yet supported");
    _table = new inv_table[table_num];
    for (size_t i = 0; i < cycle; ++i) {
        std::vector<std::vector<int> > part(data.begin() + i * m, data.begin() + (i + 1) * m);
        _table[i].set_table_index((int)i);
        _table[i].set_total_num_of_table((int)table_num);
        load_rows(_table[i], part, config);
    }
    if (table_num != cycle) {
        const size_t second_last = (n - cycle * m) / 2;
        std::vector<std::vector<int> > p1(data.begin() + cycle * m, data.begin() + cycle * m + second_last);
        std::vector<std::vector<int> > p2(data.begin() + cycle * m + second_last, data.end());
        _table[cycle].set_table_index((int)cycle);
        _table[cycle].set_total_num_of_table((int)table_num);
        _table[cycle + 1].set_table_index((int)cycle + 1);
        _table[cycle + 1].set_total_num_of_table((int)table_num);
        load_rows(_table[cycle], p1, config);
        load_rows(_table[cycle + 1], p2, config);
    }
    return true;
}

// Merge the per-table results of every query and keep the best
// topk / iterations (merge_knn_results_from_multiload, interface.cc:114-149).
// Ties are ordered by id, a valid instance of the reference's unstable sort.
static void merge_results(std::vector<std::vector<int> >& res, std::vector<std::vector<int> >& cnt,
                          std::vector<int>& result, std::vector<int>& result_count, size_t table_num, int topk,
                          size_t query_num, unsigned int iterations)
{
    const int individual = topk / (int)iterations;
    result.clear();
    result_count.clear();
    for (size_t i = 0; i < query_num; ++i) {
        std::vector<std::pair<int, int> > all;  // (count, id)
        for (size_t j = 0; j < table_num; ++j)
            for (int s = (int)(i * topk + (size_t)individual * (iterations - 1)); s < (int)((i + 1) * topk); ++s)
                all.emplace_back(cnt[j][(size_t)s], res[j][(size_t)s]);
        std::stable_sort(all.begin(), all.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
            if (a.first != b.first)
                return a.first > b.first;
            if (a.first == 0)
                return false;  // keep padding where it is
            return a.second < b.second;
        });
        for (int s = 0; s < individual; ++s) {
            result.push_back(all[(size_t)s].second);
            result_count.push_back(all[(size_t)s].first);
        }
    }
}

void knn_search_after_preprocess(GPUGenie_Config& config, inv_table*& _table, std::vector<int>& result,
                                 std::vector<int>& result_count)
{
    const size_t table_num = (size_t)_table[0].get_total_num_of_table();
    std::vector<std::vector<int> > res(table_num), cnt(table_num);
    std::vector<Query> queries;
    int accumulate = 0;
    for (size_t i = 0; i < table_num; ++i) {
        queries.clear();
        load_query(_table[i], queries, config);
        knn_search(_table[i], queries, res[i], cnt[i], config);
        if (i == 0)
            continue;
        accumulate += _table[i - 1].i_size();
        for (size_t j = 0; j < res[i].size(); ++j)
            if (cnt[i][j] > 0)
                res[i][j] += accumulate;
    }
    merge_results(res, cnt, result, result_count, table_num, config.num_of_topk, queries.size(),
                  config.num_of_iteration);
}

void load_table(inv_table& table, std::vector<std::vector<int> >& data_points, GPUGenie_Config& config)
{
    inv_list list;
    const size_t dims = data_points[0].size();
    for (size_t i = 0; i < dims; ++i) {
        std::vector<int> col;
        col.reserve(data_points.size());
        for (size_t j = 0; j < data_points.size(); ++j)
            col.push_back(data_points[j][i]);
        list.invert(col);
        table.append(list);
    }
    table.build((size_t)config.posting_list_max_length, config.use_load_balance);
    if (config.save_to_gpu)
        table.cpy_data_to_gpu();
    table.is_stored_in_gpu = config.save_to_gpu;
}

void load_table(inv_table& table, int* data, unsigned int item_num, unsigned int* index, unsigned int row_num,
                GPUGenie_Config& config)
{
    const unsigned int row_size = row_num == 1 ? item_num : index[1] - index[0];
    inv_list list;
    for (unsigned int i = 0; i < row_size; ++i) {
        std::vector<int> col;
        col.reserve(row_num);
        for (unsigned int j = 0; j < row_num; ++j)
            col.push_back(data[index[j] + i - index[0]]);
        list.invert(col);
        table.append(list);
    }
    table.build((size_t)config.posting_list_max_length, config.use_load_balance);
    if (config.save_to_gpu)
        table.cpy_data_to_gpu();
    table.is_stored_in_gpu = config.save_to_gpu;
}

void load_table_bijectMap(inv_table& table, std::vector<std::vector<int> >& data_points, GPUGenie_Config& config)
{
    if (config.use_subsequence_search)
        throw genie::exception::cpu_runtime_error("sub-sequence search is not supported");
    inv_list list;
    list.invert_bijectMap(data_points);
    table.append(list);
    table.build((size_t)config.posting_list_max_length, config.use_load_balance);
    if (config.save_to_gpu)
        table.cpy_data_to_gpu();
    table.is_stored_in_gpu = config.save_to_gpu;
}

void load_table_bijectMap(inv_table& table, int* data, unsigned int item_num, unsigned int* index,
                          unsigned int row_num, GPUGenie_Config& config)
{
    if (config.use_subsequence_search)
        throw genie::exception::cpu_runtime_error("sub-sequence search is not supported");
    inv_list list;
    list.invert_bijectMap(data, item_num, index, row_num);
    table.append(list);
    table.build((size_t)config.posting_list_max_length, config.use_load_balance);
    if (config.save_to_gpu)
        table.cpy_data_to_gpu();
    table.is_stored_in_gpu = config.save_to_gpu;
}

void load_query(inv_table& table, std::vector<Query>& queries, GPUGenie_Config& config)
{
    if (config.search_type == 2)
        throw genie::exception::cpu_runtime_error("sequence search is not supported");
    if (config.use_multirange)
        load_query_multirange(table, queries, config);
    else
        load_query_singlerange(table, queries, config);
}

// "qid dim value selectivity weight" lines, grouped by qid in id order
// (interface.cc:427-472).
void load_query_multirange(inv_table& table, std::vector<Query>& queries, GPUGenie_Config& config)
{
    queries.clear();
    if (!config.multirange_query_points)
        throw genie::exception::cpu_runtime_error("no multirange query points");
    std::map<int, Query> by_qid;
    for (const genie::utility::attr_t& a : *config.multirange_query_points) {
        auto it = by_qid.find(a.qid);
        if (it == by_qid.end()) {
            Query q(table, a.qid);
            q.topk(config.num_of_topk);
            if (config.selectivity > 0.0f)
                q.selectivity(config.selectivity);
            if (config.use_load_balance)
                q.use_load_balance = true;
            it = by_qid.emplace(a.qid, q).first;
        }
        it->second.attr(a.dim, a.value, a.weight, a.sel, it->second.count_ranges());
    }
    for (auto& kv : by_qid) {
        if (queries.size() >= (size_t)config.num_of_queries)
            break;
        kv.second.apply_adaptive_query_range();
        queries.push_back(kv.second);
    }
}

// One query per row: value +- radius on every attribute (on attribute 0 for
// bijectMap tables), weight 1 (interface.cc:473-517).
void load_query_singlerange(inv_table& table, std::vector<Query>& queries, GPUGenie_Config& config)
{
    std::vector<std::vector<int> >& points = *config.query_points;
    const int radius = config.query_radius;
    for (size_t i = 0; i < points.size(); ++i) {
        Query q(table, (int)i);
        for (size_t j = 0; j < points[i].size() && (config.search_type == 1 || j < (size_t)config.dim); ++j) {
            const int value = points[i][j];
            q.attr(config.search_type == 1 ? 0 : (int)j, value - radius, value + radius, GPUGENIE_DEFAULT_WEIGHT,
                   (int)j);
        }
        q.topk(config.num_of_topk);
        q.selectivity(config.selectivity);
        if (config.use_adaptive_range)
            q.apply_adaptive_query_range();
        if (config.use_load_balance)
            q.use_load_balance = true;
        queries.push_back(q);
    }
}

void knn_search_for_csv_data(std::vector<int>& result, std::vector<int>& result_count, GPUGenie_Config& config)
{
    inv_table* table = nullptr;
    preprocess_for_knn_csv(config, table);
    try {
        knn_search_after_preprocess(config, table, result, result_count);
    } catch (...) {
        delete[] table;
        throw;
    }
    delete[] table;
}

void knn_search(std::vector<int>& result, GPUGenie_Config& config)
{
    std::vector<int> result_count;
    knn_search(result, result_count, config);
}

void knn_search(std::vector<int>& result, std::vector<int>& result_count, GPUGenie_Config& config)
{
    // error translation of interface.cc:779-824
    try {
        switch (config.data_type) {
        case 0: knn_search_for_csv_data(result, result_count, config); break;
        case 1: throw genie::exception::cpu_runtime_error("Binary .dat data_type no longer supported\n");
        default: throw genie::exception::cpu_runtime_error("Please check data type in config\n");
        }
    } catch (genie::exception::gpu_bad_alloc&) {
        throw;
    } catch (genie::exception::gpu_runtime_error&) {
        throw;
    } catch (genie::exception::cpu_runtime_error&) {
        throw;
    } catch (std::bad_alloc& e) {
        throw genie::exception::cpu_bad_alloc(e.what());
    } catch (std::exception& e) {
        throw genie::exception::cpu_runtime_error(e.what());
    }
}

void knn_search(inv_table& table, std::vector<Query>& queries, std::vector<int>& h_topk,
                std::vector<int>& h_topk_count, GPUGenie_Config& config)
{
    const int max_load = (int)(config.multiplier * config.posting_list_max_length) + 1;
    const int hashtable_size = config.hashtable_size <= 2
                                   ? (int)(table.i_size() * config.hashtable_size) + 1
                                   : (int)config.hashtable_size;
    genie::matching::knn(table, queries, h_topk, h_topk_count, hashtable_size, max_load, config.count_threshold);
}

void knn_search_MT(std::vector<inv_table*>& tables, std::vector<std::vector<Query> >& queries,
                   std::vector<std::vector<int> >& h_topk, std::vector<std::vector<int> >& h_topk_count,
                   std::vector<GPUGenie_Config>& configs)
{
    std::vector<int> hs, ml;
    for (size_t i = 0; i < tables.size(); ++i) {
        const GPUGenie_Config& c = configs.at(i);
        hs.push_back(c.hashtable_size <= 2 ? (int)(tables[i]->i_size() * c.hashtable_size) + 1
                                           : (int)c.hashtable_size);
        ml.push_back((int)(c.multiplier * c.posting_list_max_length) + 1);
    }
    genie::matching::knn_MT(tables, queries, h_topk, h_topk_count, hs, ml, configs.at(0).count_threshold);
}

void reset_device() { (void)genie_hip_init(0); }

void init_genie(GPUGenie_Config& config) { genie_amd::set_device(config.use_device); }

}  // namespace original
}  // namespace genie
