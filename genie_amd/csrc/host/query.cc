// Query construction and compilation into posting spans
// (src/genie/query/query.cc:87-113, 141-191, 196-261, 273-333, 464-533, 536-560).
#include <genie/exception/exception.h>
#include <genie/query/query.h>
#include <genie/table/inv_compr_table.h>

#include <cmath>

namespace genie {
namespace query {

static_assert(sizeof(Query::dim) == 20, "Query::dim must keep the 20-byte work-item layout");

Query::Query() = default;

Query::Query(genie::table::inv_table* ref, int index) : _ref_table(ref), _index(index), _count(-1) {}

Query::Query(genie::table::inv_table& ref, int index) : _ref_table(&ref), _index(index), _count(0) {}

void Query::attr(int index, int low, int up, float weight, int order)
{
    attr(index, low, up, weight, -1.0f, order);
}

void Query::attr(int index, int value, float weight, float selectivity, int order)
{
    attr(index, value, value, weight, selectivity, order);
}

void Query::attr(int index, int low, int up, float weight, float selectivity, int order)
{
    if (!_ref_table || index < 0 || index >= _ref_table->m_size())
        return;  // ranges on unknown attributes are dropped (query.cc:94-95)
    range r;
    r.low = low;
    r.up = up;
    r.weight = weight;
    r.dim = index;
    r.query = _index;
    r.order = order;
    r.low_offset = 0;
    r.up_offset = 0;
    r.selectivity = selectivity;
    _attr_map[index].push_back(r);
    _count++;
}

void Query::clear_dim(int index)
{
    auto it = _attr_map.find(index);
    if (it == _attr_map.end())
        return;
    _count -= (int)it->second.size();
    _attr_map.erase(it);
}

void Query::apply_adaptive_query_range()
{
    genie::table::inv_table& table = *_ref_table;
    if (table.build_status() == genie::table::inv_table::not_builded)
        return;
    const u32 global_threshold = _selectivity > 0 ? (u32)std::ceil(_selectivity * table.i_size()) : 0;
    for (auto& kv : _attr_map) {
        const int index = kv.first;
        for (range& d : kv.second) {
            u32 local_threshold;
            if (d.selectivity > 0)
                local_threshold = (u32)std::ceil(d.selectivity * table.i_size());
            else if (_selectivity > 0)
                local_threshold = global_threshold;
            else
                return;  // no selectivity set (query.cc:222-226)
            unsigned int count = 0;
            for (int vi = d.low; vi <= d.up; ++vi)
                if (table.list_contain(index, vi))
                    count += table.get_posting_list_size(index, vi);
            // widen the range one value each side until enough postings
            while (count < local_threshold) {
                if (d.low > 0) {
                    d.low--;
                    if (!table.list_contain(index, d.low))
                        continue;
                    count += table.get_posting_list_size(index, d.low);
                }
                if (!table.list_contain(index, d.up + 1)) {
                    if (d.low == 0)
                        break;
                } else {
                    d.up++;
                    count += table.get_posting_list_size(index, d.up);
                }
            }
        }
    }
}

// Span of one clamped range: [inv_pos[map[low key]], inv_pos[map[up key + 1]]).
static bool span_of(genie::table::inv_table& t, int index, int low, int up, int* sp, int* ep, int* s0, int* s1)
{
    if (low > up || low > t.get_upperbound_of_list(index) || up < t.get_lowerbound_of_list(index))
        return false;
    const int lb = t.get_lowerbound_of_list(index);
    low = std::max(low, lb);
    up = std::min(up, t.get_upperbound_of_list(index));
    const int d = index << t.shifter();
    std::unordered_map<size_t, int>& map = *t.inv_index_map();
    const int a = map.at((size_t)(d + low - lb)), b = map.at((size_t)(d + up - lb + 1));
    std::vector<int>& pos = *t.inv_pos();
    *s0 = a;
    *s1 = b;
    *sp = pos[(size_t)a];
    *ep = pos[(size_t)b];
    return true;
}

void Query::build()
{
    genie::table::inv_table& table = *_ref_table;
    for (auto& kv : _attr_map) {
        const int index = kv.first;
        if (kv.second.empty())
            continue;
        std::vector<dim>& out = _dim_map[index];
        for (const range& r : kv.second) {
            int sp, ep, s0, s1;
            if (!span_of(table, index, r.low, r.up, &sp, &ep, &s0, &s1))
                continue;
            out.emplace_back(_index, r.order, sp, ep, r.weight);
        }
    }
}

void Query::build(std::vector<dim>& dims)
{
    genie::table::inv_table& table = *_ref_table;
    for (auto& kv : _attr_map)
        for (const range& r : kv.second) {
            int sp, ep, s0, s1;
            if (!span_of(table, kv.first, r.low, r.up, &sp, &ep, &s0, &s1))
                continue;
            dims.emplace_back(_index, r.order, sp, ep, r.weight);
        }
}

void Query::build_and_apply_load_balance(int max_load)
{
    build();
    if (max_load <= 0)
        return;
    _dims.clear();
    for (auto& kv : _dim_map)
        for (const dim& d : kv.second) {
            const unsigned int length = (unsigned int)(d.end_pos - d.start_pos);
            if ((unsigned int)max_load > length) {
                _dims.push_back(d);
                continue;
            }
            unsigned int j = 1;
            for (; (unsigned int)max_load * j <= length; ++j) {
                const int st = (int)(max_load * (j - 1)) + d.start_pos;
                _dims.emplace_back(d.query, d.order, st, st + max_load, d.weight);
            }
            if ((unsigned int)max_load * (j - 1) != length)
                _dims.emplace_back(d.query, d.order, (int)(max_load * (j - 1)) + d.start_pos, d.end_pos, d.weight);
        }
    is_load_balanced = true;
}

void Query::build_compressed(int max_load, bool merge_runs)
{
    (void)max_load;
    auto* table = dynamic_cast<genie::table::inv_compr_table*>(_ref_table);
    if (!table)
        throw genie::exception::cpu_runtime_error("build_compressed needs a compressed table");
    std::vector<int>& cpos = *table->compressedInvPos();
    for (auto& kv : _attr_map) {
        const int index = kv.first;
        if (kv.second.empty())
            continue;
        std::vector<dim>& out = _dim_map[index];
        for (const range& r : kv.second) {
            int sp, ep, s0, s1;
            if (!span_of(*table, index, r.low, r.up, &sp, &ep, &s0, &s1))
                continue;
            if (merge_runs) {  // the range's segments follow one another: one item
                if (s1 > s0)
                    out.emplace_back(_index, r.order, cpos[(size_t)s0], cpos[(size_t)s1], r.weight);
                continue;
            }
            // one work item per compressed segment of the range
            for (int s = s0; s < s1; ++s)
                out.emplace_back(_index, r.order, cpos[(size_t)s], cpos[(size_t)s + 1], r.weight);
        }
    }
}

int Query::dump(std::vector<dim>& vout)
{
    if (is_load_balanced) {
        vout.insert(vout.end(), _dims.begin(), _dims.end());
        return (int)_dims.size();
    }
    int count = 0;
    for (auto& kv : _dim_map) {
        count += (int)kv.second.size();
        vout.insert(vout.end(), kv.second.begin(), kv.second.end());
    }
    return count;
}

int Query::dump(std::vector<range>& ranges)
{
    ranges.clear();
    for (auto& kv : _attr_map)
        ranges.insert(ranges.end(), kv.second.begin(), kv.second.end());
    return (int)ranges.size();
}

}  // namespace query
}  // namespace genie
