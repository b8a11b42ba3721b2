// Flattened inverted index and its HBM-resident engine handle
// (src/genie/table/inv_table.cu:30-45, 107-128, 164-171, 268-350;
// inv_compr_table.cu:31-119).
#include <genie/exception/exception.h>
#include <genie/table/inv_compr_table.h>
#include <genie/table/inv_table.h>

#include <algorithm>
#include <string>

#include "genie_hip.h"
#include "host_common.h"

namespace genie {
namespace table {

inv_table::~inv_table() { clear_gpu_mem(); }

void inv_table::clear()
{
    _build_status = not_builded;
    _inv_lists.clear();
    _ck.clear();
    _inv.clear();
    clear_gpu_mem();
}

void inv_table::clear_gpu_mem()
{
    if (d_inv_p) {
        genie_hip_index_destroy(d_inv_p);
        d_inv_p = nullptr;
    }
    is_stored_in_gpu = false;
}

void inv_table::append(inv_list& inv)
{
    // rows must agree across attributes (inv_table.cu:109)
    if (_size != -1 && _size != inv.size())
        return;
    _build_status = not_builded;
    _size = inv.size();
    _inv_lists.push_back(inv);
    _dim_size = (int)_inv_lists.size();
    std::vector<int> line;
    line.reserve((size_t)std::max(0, inv.value_range()));
    bool one_value_per_row = true;
    size_t total = 0;
    for (int i = 0; i < inv.value_range(); ++i) {
        const std::vector<int>* l = inv.index(i + inv.min());
        line.push_back((int)l->size());
        total += l->size();
        for (size_t j = 1; j < l->size() && one_value_per_row; ++j)
            one_value_per_row = (*l)[j] != (*l)[j - 1];
    }
    if (total != (size_t)inv.size() || !one_value_per_row)
        _single_value = false;
    posting_list_size.push_back(line);
    inv_list_upperbound.push_back(inv.max());
    inv_list_lowerbound.push_back(inv.min());
}

int inv_table::get_posting_list_size(int attr_index, int value) const
{
    if ((size_t)attr_index < posting_list_size.size() && value >= inv_list_lowerbound[attr_index] &&
        value <= inv_list_upperbound[attr_index])
        return posting_list_size[attr_index][value - inv_list_lowerbound[attr_index]];
    return 0;
}

bool inv_table::list_contain(int attr_index, int value) const
{
    return value <= inv_list_upperbound[attr_index] && value >= inv_list_lowerbound[attr_index];
}

int inv_table::get_upperbound_of_list(int attr_index) const
{
    return (size_t)attr_index < inv_list_upperbound.size() ? inv_list_upperbound[attr_index] : -1;
}

int inv_table::get_lowerbound_of_list(int attr_index) const
{
    return (size_t)attr_index < inv_list_lowerbound.size() ? inv_list_lowerbound[attr_index] : -1;
}

void inv_table::build(size_t max_length, bool use_load_balance)
{
    if (!use_load_balance || max_length == 0)
        max_length = (size_t)-1;
    _ck.clear();
    _inv.clear();
    _inv_pos.clear();
    _inv_index_map.clear();
    clear_gpu_mem();
    size_t total = 0;
    for (size_t d = 0; d < _inv_lists.size(); ++d) {
        inv_list& l = _inv_lists[d];
        if (d + 1 < _inv_lists.size() && (long long)l.max() - l.min() + 1 >= (1ll << _shifter))
            throw genie::exception::cpu_runtime_error(
                "an attribute spans 65536 or more values; keys of neighbouring attributes would collide");
        for (int v = l.min(); v <= l.max(); ++v)
            total += l.index(v)->size();
    }
    if (total >= 0x7FFFFFFFull)
        throw genie::exception::cpu_runtime_error("index larger than 2^31-1 postings");
    _inv.reserve(total);
    _n_dense_keys = 0;
    _dense_key_start.clear();
    _dense_key_dim.clear();
    // first segment of every key; keys absent from a dim's range inherit the
    // next segment, like _inv_index in the reference
    std::vector<int> key_seg;
    for (size_t d = 0; d < _inv_lists.size(); ++d) {
        inv_list& l = _inv_lists[d];
        const int dim_key = (int)(d << _shifter);
        for (int v = l.min(); v <= l.max(); ++v) {
            const size_t key = (size_t)(dim_key + v - l.min());
            if (_ck.size() <= key) {
                size_t last = _ck.size();
                _ck.resize(key + 1);
                key_seg.resize(key + 1);
                for (; last <= key; ++last) {
                    _ck[last] = (int)_inv.size();
                    key_seg[last] = (int)_inv_pos.size();
                }
            }
            _dense_key_start.push_back(_inv.size());
            _dense_key_dim.push_back((uint32_t)d);
            ++_n_dense_keys;
            const std::vector<int>& ids = *l.index(v);
            for (size_t j = 0; j < ids.size(); ++j) {
                if (j % max_length == 0)
                    _inv_pos.push_back((int)_inv.size());
                _inv.push_back(ids[j]);
            }
            _ck[key] = (int)_inv.size();
        }
    }
    key_seg.push_back((int)_inv_pos.size());
    _inv_pos.push_back((int)_inv.size());
    _dense_key_start.push_back(_inv.size());
    for (size_t d = 0; d < _inv_lists.size(); ++d) {
        inv_list& l = _inv_lists[d];
        const int dim_key = (int)(d << _shifter);
        for (int v = l.min(); v <= l.max() + 1; ++v) {
            const size_t key = (size_t)(dim_key + v - l.min());
            _inv_index_map.emplace(key, key_seg.at(key));
        }
    }
    _build_status = builded;
    after_build(max_length);
}

template <class Fill>
void inv_table::flatten_counts(const std::vector<int>& lo, const std::vector<int>& hi,
                               const std::vector<std::vector<int> >& len, size_t max_length, Fill fill)
{
    if (max_length == 0)
        max_length = (size_t)-1;
    const size_t D = lo.size();
    _ck.clear();
    _inv.clear();
    _inv_pos.clear();
    _inv_index_map.clear();
    _inv_lists.clear();
    clear_gpu_mem();
    _dense_key_start.clear();
    _dense_key_dim.clear();
    posting_list_size.assign(len.begin(), len.end());
    inv_list_lowerbound.assign(lo.begin(), lo.end());
    inv_list_upperbound.assign(hi.begin(), hi.end());
    _dim_size = (int)D;
    size_t total = 0;
    for (size_t d = 0; d < D; ++d) {
        // key (d << 16) + range would reach the next attribute's keys
        if (d + 1 < D && (long long)hi[d] - lo[d] + 1 >= (1ll << _shifter))
            throw genie::exception::cpu_runtime_error(
                "an attribute spans 65536 or more values; keys of neighbouring attributes would collide");
        for (int c : len[d])
            total += (size_t)c;
    }
    if (total >= 0x7FFFFFFFull)
        throw genie::exception::cpu_runtime_error("index larger than 2^31-1 postings");
    // key starts and segments
    std::vector<std::vector<size_t> > start(D);
    size_t pos = 0;
    for (size_t d = 0; d < D; ++d) {
        const size_t nv = len[d].size();
        start[d].resize(nv);
        for (size_t v = 0; v < nv; ++v) {
            const size_t key = (d << _shifter) + v;
            if (_ck.size() <= key)
                _ck.resize(key + 1, (int)pos);
            start[d][v] = pos;
            _dense_key_start.push_back(pos);
            _dense_key_dim.push_back((uint32_t)d);
            _inv_index_map.emplace(key, (int)_inv_pos.size());
            for (size_t j = 0; j < (size_t)len[d][v]; j += max_length)
                _inv_pos.push_back((int)(pos + j));
            pos += (size_t)len[d][v];
            _ck[key] = (int)pos;
        }
        // key max+1 of the dim: the next segment
        _inv_index_map.emplace((d << _shifter) + nv, (int)_inv_pos.size());
    }
    _inv_pos.push_back((int)pos);
    _dense_key_start.push_back(pos);
    _n_dense_keys = (uint32_t)(_dense_key_start.size() - 1);
    _inv.resize(total);
    fill(start, _inv.data());
    _build_status = builded;
    after_build(max_length);
}

void inv_table::build_from_rows(const int32_t* rows, uint32_t n, uint32_t d, size_t max_length,
                                bool use_load_balance)
{
    if (n == 0 || d == 0)
        throw genie::exception::cpu_runtime_error("no data input!");
    if (!use_load_balance)
        max_length = (size_t)-1;
    std::vector<int> lo(d), hi(d);
    std::vector<std::vector<int> > len(d);
    genie_amd::parallel_for(d, [&](size_t j) {
        int a = rows[j], b = rows[j];
        for (uint32_t i = 0; i < n; ++i) {
            const int x = rows[(size_t)i * d + j];
            a = x < a ? x : a;
            b = x > b ? x : b;
        }
        lo[j] = a;
        hi[j] = b;
        if (j + 1 < d && (long long)b - a + 1 >= (1ll << 16))
            return;  // rejected by flatten_counts
        len[j].assign((size_t)(b - a + 1), 0);
        for (uint32_t i = 0; i < n; ++i)
            ++len[j][(size_t)(rows[(size_t)i * d + j] - a)];
    });
    _size = (int)n;
    _single_value = true;
    flatten_counts(lo, hi, len, max_length, [&](const std::vector<std::vector<size_t> >& start, int* inv) {
        genie_amd::parallel_for(d, [&](size_t j) {
            std::vector<size_t> cur(start[j]);
            for (uint32_t i = 0; i < n; ++i)
                inv[cur[(size_t)(rows[(size_t)i * d + j] - lo[j])]++] = (int)i;
        });
    });
}

void inv_table::build_from_sets(const int64_t* row_ptr, const int32_t* vals, uint32_t n, size_t max_length,
                                bool use_load_balance)
{
    if (n == 0 || row_ptr[n] <= 0)
        throw genie::exception::cpu_runtime_error("no data input!");
    if (!use_load_balance)
        max_length = (size_t)-1;
    int a = vals[0], b = vals[0];
    for (int64_t i = 0; i < row_ptr[n]; ++i) {
        a = vals[i] < a ? vals[i] : a;
        b = vals[i] > b ? vals[i] : b;
    }
    std::vector<int> lo(1, a), hi(1, b);
    std::vector<std::vector<int> > len(1);
    len[0].assign((size_t)((long long)b - a + 1), 0);
    for (int64_t i = 0; i < row_ptr[n]; ++i)
        ++len[0][(size_t)(vals[i] - a)];
    _size = (int)n;
    _single_value = false;
    flatten_counts(lo, hi, len, max_length, [&](const std::vector<std::vector<size_t> >& start, int* inv) {
        std::vector<size_t> cur(start[0]);
        for (uint32_t r = 0; r < n; ++r)
            for (int64_t i = row_ptr[r]; i < row_ptr[r + 1]; ++i)
                inv[cur[(size_t)(vals[i] - a)]++] = (int)r;
    });
}

int inv_table::upload(genie_index** out)
{
    genie_index_desc desc = {};
    desc.n_objects = (uint32_t)i_size();
    desc.n_dims = (uint32_t)m_size();
    desc.n_keys = _n_dense_keys;
    desc.n_postings = _inv.size();
    desc.postings = reinterpret_cast<const uint32_t*>(_inv.data());
    desc.key_start = _dense_key_start.data();
    desc.key_dim = _dense_key_dim.data();
    std::vector<uint32_t> upos(_inv_pos.begin(), _inv_pos.end());
    desc.n_segments = (uint32_t)(upos.size() - 1);
    desc.seg_upos = upos.data();
    desc.codec = 0;
    desc.single_value = _single_value ? 1 : 0;
    return genie_hip_index_create(&desc, out);
}

bool inv_table::cpy_data_to_gpu()
{
    if (_build_status != builded)
        throw genie::exception::cpu_runtime_error("table not built!");
    if (!d_inv_p) {
        genie_amd::ensure_device();
        genie_amd::check(upload(&d_inv_p));
    }
    is_stored_in_gpu = true;
    return true;
}

genie_index* inv_table::device_index()
{
    if (!d_inv_p)
        cpy_data_to_gpu();
    return d_inv_p;
}

void inv_compr_table::setCompression(genie::compression::COMPRESSION_TYPE c)
{
    if (build_status() == builded)
        return;  // the reference refuses to change a built table (inv_compr_table.cu:150-156)
    m_compression = c;
}

void inv_compr_table::after_build(size_t max_length)
{
    if (m_compression == genie::compression::NO_COMPRESSION)
        throw genie::exception::cpu_runtime_error("No compression codec available!");
    const std::vector<int>& inv = *inv_table::inv();
    const std::vector<int>& pos = *inv_table::inv_pos();
    const size_t cap = std::min(std::min(max_length, m_max_length), genie::compression::kMaxSegmentLength);
    m_comprInv.clear();
    m_comprInvPos.assign(1, 0);
    std::vector<uint32_t> enc;
    for (size_t s = 0; s + 1 < pos.size(); ++s) {
        const size_t a = (size_t)pos[s], b = (size_t)pos[s + 1];
        if (b - a == 0 || b - a > cap)
            throw genie::exception::cpu_runtime_error(
                "compressed tables need load balancing with segments of at most " + std::to_string(cap) + " ids");
        genie::compression::encode(m_compression, reinterpret_cast<const uint32_t*>(inv.data()) + a, b - a, enc);
        if (enc.size() > cap)
            throw genie::exception::cpu_runtime_error("compressed segment longer than the segment limit");
        m_comprInv.insert(m_comprInv.end(), enc.begin(), enc.end());
        m_comprInvPos.push_back((int)m_comprInv.size());
    }
}

double inv_compr_table::getCompressionRatio()
{
    if (build_status() != builded || uncompressedInv()->empty())
        return -1;
    return 32.0 * (double)m_comprInv.size() / (double)uncompressedInv()->size();
}

int inv_compr_table::upload(genie_index** out)
{
    const std::vector<int>& inv = *inv_table::inv();
    genie_index_desc desc = {};
    desc.n_objects = (uint32_t)i_size();
    desc.n_dims = (uint32_t)m_size();
    desc.n_keys = _n_dense_keys;
    desc.n_postings = inv.size();
    desc.postings = reinterpret_cast<const uint32_t*>(inv.data());
    desc.key_start = _dense_key_start.data();
    desc.key_dim = _dense_key_dim.data();
    std::vector<uint32_t> upos(inv_table::inv_pos()->begin(), inv_table::inv_pos()->end());
    std::vector<uint32_t> cpos(m_comprInvPos.begin(), m_comprInvPos.end());
    desc.n_segments = (uint32_t)(upos.size() - 1);
    desc.seg_upos = upos.data();
    desc.codec = (int32_t)m_compression;
    desc.n_words = m_comprInv.size();
    desc.words = m_comprInv.data();
    desc.seg_cpos = cpos.data();
    desc.single_value = _single_value ? 1 : 0;
    return genie_hip_index_create(&desc, out);
}

}  // namespace table
}  // namespace genie
