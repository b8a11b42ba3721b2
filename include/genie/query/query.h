// Query model and compilation into posting spans (src/genie/query/query.h:
// 34-314, query.cc:87-113 attr, 141-191 build_and_apply_load_balance,
// 196-261 apply_adaptive_query_range, 273-333 build, 464-533
// build_compressed, 536-560 dump).
#ifndef GENIE_QUERY_QUERY_H_
#define GENIE_QUERY_QUERY_H_

#include <map>
#include <vector>

#include <genie/table/inv_table.h>

namespace genie {
namespace query {

typedef unsigned int u32;
typedef unsigned long long u64;

class Query {
public:
    struct range {
        int query;
        int order;
        int dim;
        int low;
        int up;
        float weight;
        int low_offset;
        int up_offset;
        float selectivity;
    };
    // Work item handed to the matching engine: 20 bytes, identical to
    // genie_dim_t of include/genie_hip.h.
    struct dim {
        int query;
        int order;
        int start_pos;
        int end_pos;
        float weight;
        dim() = default;
        dim(int query, int order, int start_pos, int end_pos, float weight)
            : query(query), order(order), start_pos(start_pos), end_pos(end_pos), weight(weight) {}
    };

    Query();
    Query(genie::table::inv_table* ref, int index);
    Query(genie::table::inv_table& ref, int index);

    genie::table::inv_table* ref_table() const { return _ref_table; }

    void attr(int index, int low, int up, float weight, int order);
    void attr(int index, int value, float weight, float selectivity, int order);
    void attr(int index, int low, int up, float weight, float selectivity, int order);
    void clear_dim(int index);

    void selectivity(float s) { _selectivity = s; }
    float selectivity() const { return _selectivity; }
    void apply_adaptive_query_range();

    void build();
    void build(std::vector<dim>& out);
    void build_and_apply_load_balance(int max_load);
    // One work item per compressed segment of each range (query.cc:464-533);
    // with merge_runs, one item per range spanning all of its segments, which
    // the engine takes as the same work with far fewer items.
    void build_compressed(int max_load, bool merge_runs = false);

    void topk(int k) { _topk = k; }
    int topk() const { return _topk; }

    int dump(std::vector<dim>& vout);
    int dump(std::vector<range>& ranges);

    int index() const { return _index; }
    int count_ranges() const { return _count; }

    bool is_load_balanced = false;
    bool use_load_balance = false;
    std::vector<dim> _dims;

private:
    genie::table::inv_table* _ref_table = nullptr;
    std::map<int, std::vector<range> > _attr_map;
    std::map<int, std::vector<dim> > _dim_map;
    int _topk = 1;
    float _selectivity = -1.0f;
    int _index = -1;
    int _count = -1;
};

}  // namespace query
}  // namespace genie

#endif
