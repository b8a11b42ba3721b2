// Legacy GENIE interface (src/genie/original/interface.h:20-160 config and
// defaults; interface.cc:150-909 functions), on the MI355X engine.
#ifndef GENIE_ORIGINAL_INTERFACE_H_
#define GENIE_ORIGINAL_INTERFACE_H_

#include <string>
#include <vector>

#include <genie/compression/codecs.h>
#include <genie/exception/exception.h>
#include <genie/query/query.h>
#include <genie/table/inv_compr_table.h>
#include <genie/table/inv_table.h>
#include <genie/utility/FileReader.h>

#define GPUGENIE_DEFAULT_TOPK 10
#define GPUGENIE_DEFAULT_RADIUS 0
#define GPUGENIE_DEFAULT_THRESHOLD 0
#define GPUGENIE_DEFAULT_HASHTABLE_SIZE 1.0f
#define GPUGENIE_DEFAULT_WEIGHT 1
#define GPUGENIE_DEFAULT_DEVICE 0
#define GPUGENIE_DEFAULT_USE_ADAPTIVE_RANGE false
#define GPUGENIE_DEFAULT_SELECTIVITY -1.0f
#define GPUGENIE_DEFAULT_POSTING_LIST_LENGTH 100000
#define GPUGENIE_DEFAULT_LOAD_MULTIPLIER 3.0f
#define GPUGENIE_DEFAULT_USE_LOAD_BALANCE false
#define GPUGENIE_DEFAULT_USE_MULTIRANGE true
#define GPUGENIE_DEFAULT_NUM_OF_QUERIES 0

namespace genie {
namespace original {

typedef struct _GPUGenie_Config {
    int num_of_topk = GPUGENIE_DEFAULT_TOPK;
    int query_radius = GPUGENIE_DEFAULT_RADIUS;
    // Bit width hint of the reference's bitmap counters; counts here are
    // sized from the query weights, so this field has no effect.
    int count_threshold = GPUGENIE_DEFAULT_THRESHOLD;
    // Size of the reference's per-query c-PQ hash table; no effect here.
    float hashtable_size = GPUGENIE_DEFAULT_HASHTABLE_SIZE;
    int use_device = GPUGENIE_DEFAULT_DEVICE;
    int dim = 0;
    bool use_adaptive_range = GPUGENIE_DEFAULT_USE_ADAPTIVE_RANGE;
    float selectivity = GPUGENIE_DEFAULT_SELECTIVITY;
    std::vector<std::vector<int> >* data_points = nullptr;
    int* data = nullptr;
    unsigned int* index = nullptr;
    unsigned int item_num = 0;
    unsigned int row_num = 0;
    int search_type = 0;  // 0 single value, 1 bijectMap (2, sequence, is not supported)
    int data_type = 0;    // 0 csv rows
    unsigned int max_data_size = 0;
    bool save_to_gpu = false;
    std::vector<std::vector<int> >* query_points = nullptr;
    std::vector<genie::utility::attr_t>* multirange_query_points = nullptr;
    int posting_list_max_length = GPUGENIE_DEFAULT_POSTING_LIST_LENGTH;
    float multiplier = GPUGENIE_DEFAULT_LOAD_MULTIPLIER;
    bool use_load_balance = GPUGENIE_DEFAULT_USE_LOAD_BALANCE;
    bool use_multirange = GPUGENIE_DEFAULT_USE_MULTIRANGE;
    int num_of_queries = GPUGENIE_DEFAULT_NUM_OF_QUERIES;
    bool use_subsequence_search = false;
    int data_gram_length = 3;
    unsigned int num_of_iteration = 1;
    genie::compression::COMPRESSION_TYPE compression = genie::compression::DEFAULT_COMPRESSION_TYPE;
} GPUGenie_Config;

bool preprocess_for_knn_csv(GPUGenie_Config& config, genie::table::inv_table*& _table);
void knn_search_after_preprocess(GPUGenie_Config& config, genie::table::inv_table*& _table,
                                 std::vector<int>& result, std::vector<int>& result_count);
void knn_search(std::vector<int>& result, std::vector<int>& result_count, GPUGenie_Config& config);
void knn_search(std::vector<int>& result, GPUGenie_Config& config);
void knn_search(genie::table::inv_table& table, std::vector<genie::query::Query>& queries,
                std::vector<int>& h_topk, std::vector<int>& h_topk_count, GPUGenie_Config& config);
void knn_search_MT(std::vector<genie::table::inv_table*>& tables,
                   std::vector<std::vector<genie::query::Query> >& queries,
                   std::vector<std::vector<int> >& h_topk, std::vector<std::vector<int> >& h_topk_count,
                   std::vector<GPUGenie_Config>& configs);
void knn_search_for_csv_data(std::vector<int>& result, std::vector<int>& result_count, GPUGenie_Config& config);

void load_table(genie::table::inv_table& table, std::vector<std::vector<int> >& data_points, GPUGenie_Config& config);
void load_table(genie::table::inv_table& table, int* data, unsigned int item_num, unsigned int* index,
                unsigned int row_num, GPUGenie_Config& config);
void load_table_bijectMap(genie::table::inv_table& table, std::vector<std::vector<int> >& data_points,
                          GPUGenie_Config& config);
void load_table_bijectMap(genie::table::inv_table& table, int* data, unsigned int item_num, unsigned int* index,
                          unsigned int row_num, GPUGenie_Config& config);
void load_query(genie::table::inv_table& table, std::vector<genie::query::Query>& queries, GPUGenie_Config& config);
void load_query_singlerange(genie::table::inv_table& table, std::vector<genie::query::Query>& queries,
                            GPUGenie_Config& config);
void load_query_multirange(genie::table::inv_table& table, std::vector<genie::query::Query>& queries,
                           GPUGenie_Config& config);

void reset_device();
void init_genie(GPUGenie_Config& config);

}  // namespace original
}  // namespace genie

#endif
