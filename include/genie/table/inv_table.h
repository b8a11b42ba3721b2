// Inverted index with GENIE's flat layout (src/genie/table/inv_table.h:
// 110-174, inv_table.cu:107-128, 268-350):
//   key(d, v)        = (d << 16) + v - min_d          (shifter() == 16)
//   _inv             = every key's ascending id list, concatenated in key order
//   _inv_pos         = segment starts; a key's list opens a new segment every
//                      max_length ids when load balancing is on
//   _inv_index_map   = key -> first segment of the key, for v = min_d..max_d+1
// The table can be made resident in HBM (cpy_data_to_gpu) as a handle of the
// MI355X engine (include/genie_hip.h).
#ifndef GENIE_TABLE_INV_TABLE_H_
#define GENIE_TABLE_INV_TABLE_H_

#include <cstddef>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include <genie/table/inv_list.h>

struct genie_index;

namespace genie {
namespace table {

class inv_table {
public:
    enum status { not_builded, builded };

    inv_table() = default;
    inv_table(const inv_table&) = delete;
    inv_table& operator=(const inv_table&) = delete;
    virtual ~inv_table();

    // Kept for API compatibility: the engine handle owns the device copy.
    genie_index* d_inv_p = nullptr;
    bool is_stored_in_gpu = false;
    int shift_bits_sequence = 0;

    void append(inv_list& inv);
    void append(inv_list* inv) { if (inv) append(*inv); }
    // Flatten the appended lists (inv_table.cu:268-350).
    void build(size_t max_length, bool use_load_balance);
    // The same layout built straight from row-major single-value rows
    // [n][d] (one counting sort per attribute, attributes in parallel),
    // without keeping per-attribute inv_list copies.
    void build_from_rows(const int32_t* rows, uint32_t n, uint32_t d, size_t max_length, bool use_load_balance);
    // The same layout for set-valued rows (bijectMap, one attribute); CSR
    // row_ptr [n+1].
    void build_from_sets(const int64_t* row_ptr, const int32_t* vals, uint32_t n, size_t max_length,
                         bool use_load_balance);
    // Upload to HBM and keep resident; returns true.
    virtual bool cpy_data_to_gpu();
    virtual void clear_gpu_mem();
    void clear();

    bool empty() const { return _size == -1; }
    int m_size() const { return _dim_size; }
    int i_size() const { return _size <= -1 ? 0 : _size; }
    int shifter() const { return _shifter; }
    status build_status() const { return _build_status; }

    std::vector<inv_list>* inv_lists() { return &_inv_lists; }
    std::vector<int>* ck() { return &_ck; }
    virtual std::vector<int>* inv() { return &_inv; }
    virtual std::vector<int>* inv_pos() { return &_inv_pos; }
    std::unordered_map<size_t, int>* inv_index_map() { return &_inv_index_map; }

    int get_posting_list_size(int attr_index, int value) const;
    bool list_contain(int attr_index, int value) const;
    int get_upperbound_of_list(int attr_index) const;
    int get_lowerbound_of_list(int attr_index) const;

    void set_table_index(int i) { table_index = i; }
    void set_total_num_of_table(int n) { total_num_of_table = n; }
    int get_table_index() const { return table_index; }
    int get_total_num_of_table() const { return total_num_of_table; }

    // True when every object holds exactly one value per dim (built from
    // invert(), not invert_bijectMap()).
    bool single_value() const { return _single_value; }

    // Dense key numbering used by the engine: key_base(d) + v - min_d.
    uint32_t n_dense_keys() const { return _n_dense_keys; }
    // Start position in _inv of every dense key, n_dense_keys()+1 entries.
    const std::vector<uint64_t>& dense_key_start() const { return _dense_key_start; }
    const std::vector<uint32_t>& dense_key_dim() const { return _dense_key_dim; }

    // Engine handle, uploading on first use.
    genie_index* device_index();

    // Fields of the reference's sequence search (search_type 2,
    // inv_table.h:176-199, 230): carried through save/load so that a saved
    // table holds the reference's full field list; this build does not run
    // sequence search.
    unsigned int _shift_bits_subsequence() const { return shift_bits_subsequence; }
    void set_shift_bits_subsequence(unsigned int v) { shift_bits_subsequence = v; }
    int get_min_value_sequence() const { return min_value_sequence; }
    void set_min_value_sequence(int v) { min_value_sequence = v; }
    int get_max_value_sequence() const { return max_value_sequence; }
    void set_max_value_sequence(int v) { max_value_sequence = v; }
    int get_gram_length_sequence() const { return gram_length_sequence; }
    void set_gram_length_sequence(int v) { gram_length_sequence = v; }
    std::unordered_map<int, int>* get_distinct_map(int dim);

    // Persistence (genie::utility::SaveTable / LoadTable): the table's fields
    // in the order of the reference's archive (serialization.cc:68-146),
    // appended to / read from a byte buffer at `at`; finish_load validates the
    // layout and rebuilds the derived arrays.
    virtual void save_fields(std::string& out) const;
    virtual void load_fields(const std::string& in, size_t& at);
    void finish_load(bool single_value);

protected:
    status _build_status = not_builded;
    int _shifter = 16;
    int _size = -1;
    int _dim_size = 0;
    bool _single_value = true;
    int table_index = 0;
    int total_num_of_table = 1;
    std::vector<inv_list> _inv_lists;
    std::vector<int> inv_list_upperbound;
    std::vector<int> inv_list_lowerbound;
    std::vector<std::vector<int> > posting_list_size;
    std::vector<int> _ck;
    std::vector<int> _inv;
    std::unordered_map<size_t, int> _inv_index_map;
    std::vector<int> _inv_pos;
    uint32_t _n_dense_keys = 0;
    std::vector<uint64_t> _dense_key_start;
    std::vector<uint32_t> _dense_key_dim;
    unsigned int shift_bits_subsequence = 0;
    int min_value_sequence = 0;
    int max_value_sequence = 0;
    int gram_length_sequence = 1;
    std::vector<std::unordered_map<int, int> > _distinct_map;

    // _ck and the dense key arrays from the bounds, _inv_index_map and
    // _inv_pos (after a load).
    void rebuild_derived();

    virtual int upload(genie_index** out);
    // Runs after every build; compressed tables encode their segments here.
    virtual void after_build(size_t max_length) { (void)max_length; }
    // Fill _inv_index_map, bounds, dense keys and _inv_pos from per-dim
    // bounds and per-key lengths, with the ids written by `fill`.
    template <class Fill>
    void flatten_counts(const std::vector<int>& lo, const std::vector<int>& hi,
                        const std::vector<std::vector<int> >& len, size_t max_length, Fill fill);
};

}  // namespace table
}  // namespace genie

#endif
