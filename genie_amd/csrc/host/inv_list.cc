// Inverted lists of one attribute (src/genie/table/inv_list.cc:51-86,
// 295-320): bounds of the values, then ascending row ids per value.
#include <genie/table/inv_list.h>

#include <cstdlib>

namespace genie {
namespace table {

void inv_list::invert(std::vector<int>& vin)
{
    _size = (int)vin.size();
    _inv.clear();
    if (vin.empty()) {
        _bound = {0, -1};
        return;
    }
    int lo = vin[0], hi = vin[0];
    for (int v : vin) {
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    _bound = {lo, hi};
    // counting pass first so every list is allocated once
    std::vector<int> len((size_t)hi - lo + 1, 0);
    for (int v : vin)
        ++len[(size_t)(v - lo)];
    _inv.resize(len.size());
    for (size_t j = 0; j < len.size(); ++j)
        _inv[j].reserve(len[j]);
    for (size_t i = 0; i < vin.size(); ++i)
        _inv[(size_t)(vin[i] - lo)].push_back((int)i);
}

void inv_list::invert(std::vector<std::string>& vin)
{
    std::vector<int> v;
    v.reserve(vin.size());
    for (auto& s : vin)
        v.push_back(std::atoi(s.c_str()));
    invert(v);
}

void inv_list::invert_bijectMap(std::vector<std::vector<int> >& vin)
{
    _size = (int)vin.size();
    _inv.clear();
    bool any = false;
    int lo = 0, hi = -1;
    for (auto& row : vin)
        for (int v : row) {
            if (!any) {
                lo = hi = v;
                any = true;
            }
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
    _bound = {lo, hi};
    if (!any)
        return;
    _inv.resize((size_t)hi - lo + 1);
    for (size_t i = 0; i < vin.size(); ++i)
        for (int v : vin[i])
            _inv[(size_t)(v - lo)].push_back((int)i);
}

void inv_list::invert_bijectMap(int* data, unsigned int item_num, unsigned int* index, unsigned int row_num)
{
    std::vector<std::vector<int> > rows(row_num);
    for (unsigned int r = 0; r < row_num; ++r) {
        unsigned int b = index[r] - index[0];
        unsigned int e = r + 1 < row_num ? index[r + 1] - index[0] : item_num;
        rows[r].assign(data + b, data + e);
    }
    invert_bijectMap(rows);
}

std::vector<int>* inv_list::index(int value)
{
    if (!contains(value))
        return nullptr;
    return &_inv[(size_t)(value - _bound.first)];
}

const std::vector<int>* inv_list::index(int value) const
{
    if (!contains(value))
        return nullptr;
    return &_inv[(size_t)(value - _bound.first)];
}

}  // namespace table
}  // namespace genie
