// One attribute's inverted lists: for every value in [min, max] the ascending
// ids of the objects holding it (src/genie/table/inv_list.h, inv_list.cc:
// 51-86 invert_bijectMap, 295-320 invert).
#ifndef GENIE_TABLE_INV_LIST_H_
#define GENIE_TABLE_INV_LIST_H_

#include <string>
#include <utility>
#include <vector>

namespace genie {
namespace table {

class inv_list {
public:
    inv_list() : _size(0), _bound(0, -1) {}
    explicit inv_list(std::vector<int>& vin) : inv_list() { invert(vin); }
    explicit inv_list(std::vector<int>* vin) : inv_list() { invert(*vin); }

    // Single-value column: row i holds value vin[i].
    void invert(std::vector<int>& vin);
    void invert(std::vector<int>* vin) { invert(*vin); }
    void invert(std::vector<std::string>& vin);
    // One dimension, many values per row; a row id is repeated once per
    // occurrence of the value in the row.
    void invert_bijectMap(std::vector<std::vector<int> >& vin);
    void invert_bijectMap(int* data, unsigned int item_num, unsigned int* index, unsigned int row_num);

    int min() const { return _bound.first; }
    int max() const { return _bound.second; }
    int size() const { return _size; }
    bool contains(int value) const { return value >= _bound.first && value <= _bound.second; }
    // Ids holding `value`, or NULL when value is outside [min, max].
    std::vector<int>* index(int value);
    const std::vector<int>* index(int value) const;
    int value_range() const { return (int)_inv.size(); }

private:
    int _size;
    std::pair<int, int> _bound;
    std::vector<std::vector<int> > _inv;
};

}  // namespace table
}  // namespace genie

#endif
