// The serving precedent's request path, without its transport: the JSON
// configuration and query documents of the multi-node server
// (src/dist_genie/parser.cu:16-217, README.md:115-133), the search of every
// query on the clusters it names with the per-cluster id offsets
// (search.cu:34-52, online.cu:80-105, sorting.cc:20-64) and the "id,count"
// output (file.cu:42-53).  TCP, MPI and multi-node placement are not part of
// this build; one process searches every cluster on one GPU.
#ifndef GENIE_INTERFACE_WIRE_H_
#define GENIE_INTERFACE_WIRE_H_

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include <genie/original/interface.h>

namespace genie {
namespace wire {

// distgenie::DistGenieConfig (config.h:6-14) fields the path uses.
struct DistConfig {
    std::string data_file;
    int num_of_cluster = 0;
    int data_format = 0;  // 0: csv "<data_file>_<cluster>_<rank>.csv"; 1: table file ".dat"
};

// A parsed query document: {"topk": k, "queries": [{"content": [...],
// "clusters": [...]}, ...]}.
struct QueryBatch {
    int topk = 0;
    std::vector<std::vector<int> > content;
    std::vector<std::vector<int> > clusters;
};

// (count, global id) pairs of one query, like distgenie::Result.
typedef std::vector<std::pair<int, int> > Result;

// ParseConfigurationFile + ValidateConfiguration (parser.cu:16-124): every
// entry present and of its type (data_file string; dim, count_threshold,
// data_type, search_type, max_data_size, num_of_cluster, data_format, radius
// integers; selectivity a real number), then the same settings of `config`
// (adaptive range on for selectivity > 0, load balance off, posting list
// length 6400, multiplier 1.5, single-range queries, resident tables).
// False with the reason in *error when the document is invalid.
bool ParseConfiguration(const std::string& json, genie::original::GPUGenie_Config& config, DistConfig& extra,
                        std::string* error);

// ValidateAndParseQuery (parser.cu:126-217): topk an integer; queries an
// array of objects each with a "content" array and a "clusters" array.  Sets
// config.num_of_topk and config.hashtable_size as the reference does.
bool ParseQuery(const std::string& json, genie::original::GPUGenie_Config& config, QueryBatch& batch,
                std::string* error);

// ReadData (file.cu:14-40) for one process (rank 0): every cluster's table
// from "<data_file>_<i>_0.csv" (built as preprocess_for_knn_csv builds) or
// "<data_file>_<i>_0.dat" (genie::utility::LoadTable).
std::vector<std::shared_ptr<genie::table::inv_table> > ReadData(genie::original::GPUGenie_Config& config,
                                                                const DistConfig& extra);

// ExecuteMultitableQuery + MergeResult: every query is searched on each
// cluster it names (one batch per cluster, knn_search_MT), each cluster's
// top-k ids are shifted by the cluster's id offset (the sizes of the
// clusters before it, CalculateIdOffset), and each query's pairs are sorted
// by (count, id) descending.  Padding slots (count 0) keep id 0.
std::vector<Result> ExecuteMultitableQuery(genie::original::GPUGenie_Config& config,
                                           std::vector<std::shared_ptr<genie::table::inv_table> >& tables,
                                           const QueryBatch& batch);

// GenerateOutput: topk lines "id,count" per query in query order, best
// first; a query with fewer pairs (no cluster named) is padded with "0,0".
std::string GenerateOutput(std::vector<Result>& results, int topk);

}  // namespace wire
}  // namespace genie

#endif
