// Index persistence: SaveTable / LoadTable (names of the reference,
// src/genie/utility/serialization.cc:18-56) in a native binary format.
//
// The reference writes a Boost binary archive of the table
// (serialization.cc:68-146).  Boost is not part of this build, so the file is
// this build's own format.  It holds the same fields in the same order, so a
// converter is a field-by-field copy:
//
//   header   char[8] "GENIEIDX"   u32 version (1)   u32 kind (0 inv_table,
//            1 inv_compr_table)   u32 byte-order mark 0x01020304
//   fields   i32 _size, i32 _dim_size,
//            vec<i32> _inv, vec<i32> _inv_pos,
//            map _inv_index_map      (u64 n, then n x {u64 key, i32 value},
//                                     keys ascending),
//            vec<i32> inv_list_lowerbound, vec<i32> inv_list_upperbound,
//            u64 n + n x vec<i32> posting_list_size,
//            u32 shift_bits_subsequence, i32 min_value_sequence,
//            i32 max_value_sequence, i32 gram_length_sequence,
//            i32 shift_bits_sequence,
//            u64 n + n x map _distinct_map (u64 m, then m x {i32 key, i32 value},
//                                           keys ascending)
//   kind 1   u8 is_compressed, i32 compression (COMPRESSION_TYPE),
//            u64 uncompressed segment length, vec<u32> compressed words,
//            vec<i32> compressed word offsets
//   this     i32 shifter, u8 single_value
//   build
//   trailer  u64 FNV-1a (64-bit) of every byte before it
//
// vec<T> is a u64 count followed by the elements.  Everything is
// little-endian.  Loading checks the header, every count against the bytes
// left, the checksum and the layout invariants, then rebuilds the derived
// arrays (key ends and the engine's dense key numbering) from the bounds,
// _inv_index_map and _inv_pos.
#include <genie/exception/exception.h>
#include <genie/table/inv_compr_table.h>
#include <genie/table/inv_table.h>
#include <genie/utility/serialization.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

namespace genie {
namespace {

constexpr char kMagic[8] = {'G', 'E', 'N', 'I', 'E', 'I', 'D', 'X'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kByteOrder = 0x01020304u;

[[noreturn]] void bad(const std::string& what)
{
    throw genie::exception::cpu_runtime_error("Loading from file failed (" + what + ")");
}

// Writer: appends to a byte buffer.
struct Out {
    std::string& buf;
    explicit Out(std::string& b) : buf(b) {}
    template <class T>
    void put(T v)
    {
        buf.append(reinterpret_cast<const char*>(&v), sizeof(T));
    }
    template <class T>
    void vec(const std::vector<T>& v)
    {
        put<uint64_t>(v.size());
        if (!v.empty())
            buf.append(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
    }
};

// Reader over a loaded byte buffer; every read is bounds-checked.
struct In {
    const std::string& buf;
    size_t& at;
    In(const std::string& b, size_t& pos) : buf(b), at(pos) {}
    size_t left() const { return buf.size() - at; }
    template <class T>
    T get()
    {
        if (left() < sizeof(T))
            bad("truncated");
        T v;
        std::memcpy(&v, buf.data() + at, sizeof(T));
        at += sizeof(T);
        return v;
    }
    uint64_t count(size_t elem)
    {
        const uint64_t n = get<uint64_t>();
        if (elem && n > left() / elem)
            bad("count larger than the file");
        return n;
    }
    template <class T>
    void vec(std::vector<T>& v)
    {
        const uint64_t n = count(sizeof(T));
        v.resize((size_t)n);
        if (n)
            std::memcpy(v.data(), buf.data() + at, (size_t)n * sizeof(T));
        at += (size_t)n * sizeof(T);
    }
};

uint64_t fnv1a(const char* p, size_t n)
{
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= (unsigned char)p[i];
        h *= 1099511628211ull;
    }
    return h;
}

}  // namespace

namespace table {

std::unordered_map<int, int>* inv_table::get_distinct_map(int dim)
{
    if (dim < 0)
        return nullptr;
    if ((size_t)dim >= _distinct_map.size())
        _distinct_map.resize((size_t)dim + 1);
    return &_distinct_map[(size_t)dim];
}

void inv_table::save_fields(std::string& out) const
{
    Out o(out);
    o.put<int32_t>(_size);
    o.put<int32_t>(_dim_size);
    o.vec(_inv);
    o.vec(_inv_pos);
    std::vector<std::pair<size_t, int> > m(_inv_index_map.begin(), _inv_index_map.end());
    std::sort(m.begin(), m.end());
    o.put<uint64_t>(m.size());
    for (auto& kv : m) {
        o.put<uint64_t>(kv.first);
        o.put<int32_t>(kv.second);
    }
    o.vec(inv_list_lowerbound);
    o.vec(inv_list_upperbound);
    o.put<uint64_t>(posting_list_size.size());
    for (auto& v : posting_list_size)
        o.vec(v);
    o.put<uint32_t>(shift_bits_subsequence);
    o.put<int32_t>(min_value_sequence);
    o.put<int32_t>(max_value_sequence);
    o.put<int32_t>(gram_length_sequence);
    o.put<int32_t>(shift_bits_sequence);
    o.put<uint64_t>(_distinct_map.size());
    for (auto& dm : _distinct_map) {
        std::vector<std::pair<int, int> > e(dm.begin(), dm.end());
        std::sort(e.begin(), e.end());
        o.put<uint64_t>(e.size());
        for (auto& kv : e) {
            o.put<int32_t>(kv.first);
            o.put<int32_t>(kv.second);
        }
    }
}

void inv_table::load_fields(const std::string& buf, size_t& at)
{
    In in(buf, at);
    _size = in.get<int32_t>();
    _dim_size = in.get<int32_t>();
    in.vec(_inv);
    in.vec(_inv_pos);
    const uint64_t nm = in.count(12);
    _inv_index_map.clear();
    _inv_index_map.reserve((size_t)nm);
    for (uint64_t i = 0; i < nm; ++i) {
        const uint64_t key = in.get<uint64_t>();
        const int32_t v = in.get<int32_t>();
        _inv_index_map.emplace((size_t)key, v);
    }
    in.vec(inv_list_lowerbound);
    in.vec(inv_list_upperbound);
    const uint64_t np = in.count(8);
    posting_list_size.assign((size_t)np, {});
    for (auto& v : posting_list_size)
        in.vec(v);
    shift_bits_subsequence = in.get<uint32_t>();
    min_value_sequence = in.get<int32_t>();
    max_value_sequence = in.get<int32_t>();
    gram_length_sequence = in.get<int32_t>();
    shift_bits_sequence = in.get<int32_t>();
    const uint64_t nd = in.count(8);
    _distinct_map.assign((size_t)nd, {});
    for (auto& dm : _distinct_map) {
        const uint64_t ne = in.count(8);
        for (uint64_t i = 0; i < ne; ++i) {
            const int32_t k = in.get<int32_t>();
            const int32_t v = in.get<int32_t>();
            dm.emplace(k, v);
        }
    }
}

void inv_table::finish_load(bool single_value)
{
    _single_value = single_value;
    is_stored_in_gpu = false;
    table_index = 0;
    total_num_of_table = 1;
    rebuild_derived();
}

void inv_table::rebuild_derived()
{
    if (_size < 0 || _dim_size < 0 || inv_list_lowerbound.size() != (size_t)_dim_size ||
        inv_list_upperbound.size() != (size_t)_dim_size)
        bad("bounds do not match the dims");
    if (_inv_pos.empty() || _inv_pos.front() != 0 || (size_t)_inv_pos.back() != _inv.size())
        bad("segment offsets do not span the postings");
    for (size_t i = 1; i < _inv_pos.size(); ++i)
        if (_inv_pos[i] < _inv_pos[i - 1])
            bad("segment offsets decrease");
    for (int id : _inv)
        if (id < 0 || id >= _size)
            bad("posting id out of range");
    _ck.clear();
    _dense_key_start.clear();
    _dense_key_dim.clear();
    auto seg_of = [&](size_t key) -> int {
        auto it = _inv_index_map.find(key);
        if (it == _inv_index_map.end() || it->second < 0 || (size_t)it->second >= _inv_pos.size())
            bad("key map incomplete");
        return it->second;
    };
    uint64_t prev_end = 0;
    for (int d = 0; d < _dim_size; ++d) {
        const int lo = inv_list_lowerbound[(size_t)d], hi = inv_list_upperbound[(size_t)d];
        if (hi < lo || (d + 1 < _dim_size && (long long)hi - lo + 1 >= (1ll << _shifter)))
            bad("dim bounds");
        for (int v = lo; v <= hi; ++v) {
            const size_t key = ((size_t)d << _shifter) + (size_t)(v - lo);
            const uint64_t s = (uint64_t)_inv_pos[(size_t)seg_of(key)];
            const uint64_t e = (uint64_t)_inv_pos[(size_t)seg_of(key + 1)];
            if (s != prev_end || e < s)
                bad("key lists are not contiguous");
            if (_ck.size() <= key)
                _ck.resize(key + 1, (int)s);
            _ck[key] = (int)e;
            _dense_key_start.push_back(s);
            _dense_key_dim.push_back((uint32_t)d);
            prev_end = e;
        }
    }
    if (prev_end != _inv.size())
        bad("key lists do not cover the postings");
    _dense_key_start.push_back(_inv.size());
    _n_dense_keys = (uint32_t)(_dense_key_start.size() - 1);
    _build_status = builded;
}

void inv_compr_table::save_fields(std::string& out) const
{
    inv_table::save_fields(out);
    Out o(out);
    o.put<uint8_t>(1);  // built compressed tables are always compressed
    o.put<int32_t>((int32_t)m_compression);
    o.put<uint64_t>((uint64_t)m_max_length);
    o.vec(m_comprInv);
    o.vec(m_comprInvPos);
}

void inv_compr_table::load_fields(const std::string& buf, size_t& at)
{
    inv_table::load_fields(buf, at);
    In in(buf, at);
    if (in.get<uint8_t>() != 1)
        bad("compressed table not compressed");
    m_compression = (genie::compression::COMPRESSION_TYPE)in.get<int32_t>();
    m_max_length = (size_t)in.get<uint64_t>();
    in.vec(m_comprInv);
    in.vec(m_comprInvPos);
    if (m_comprInvPos.size() != inv_table::_inv_pos.size() || m_comprInvPos.empty() || m_comprInvPos.front() != 0 ||
        (size_t)m_comprInvPos.back() != m_comprInv.size())
        bad("compressed segment offsets");
    for (size_t i = 1; i < m_comprInvPos.size(); ++i)
        if (m_comprInvPos[i] < m_comprInvPos[i - 1])
            bad("compressed segment offsets decrease");
}

}  // namespace table

namespace utility {

// The table's own fields are written through save_fields; this build's
// extras (shifter, single-value flag) follow, then the checksum.
void SaveTable(const std::string& filename, const std::shared_ptr<const genie::table::inv_table>& table)
{
    if (!table)
        throw genie::exception::cpu_runtime_error("no table to save");
    if (table->get_table_index() != 0 || table->get_total_num_of_table() != 1)
        throw genie::exception::cpu_runtime_error("Saving multiple tables not supported");
    if (table->build_status() != genie::table::inv_table::builded)
        throw genie::exception::cpu_runtime_error("Cannot serialize inv::table that has not yet been built.");
    const bool compressed = dynamic_cast<const genie::table::inv_compr_table*>(table.get()) != nullptr;
    std::string bytes(kMagic, 8);
    Out o(bytes);
    o.put<uint32_t>(kVersion);
    o.put<uint32_t>(compressed ? 1u : 0u);
    o.put<uint32_t>(kByteOrder);
    table->save_fields(bytes);
    o.put<int32_t>(table->shifter());
    o.put<uint8_t>(table->single_value() ? 1 : 0);
    const uint64_t h = fnv1a(bytes.data(), bytes.size());
    std::ofstream ofs(filename, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!ofs.good())
        throw genie::exception::cpu_runtime_error("Saving to file failed (cannot open " + filename + ")");
    ofs.write(bytes.data(), (std::streamsize)bytes.size());
    ofs.write(reinterpret_cast<const char*>(&h), 8);
    if (!ofs.good())
        throw genie::exception::cpu_runtime_error("Saving to file failed (write error)");
}

std::shared_ptr<genie::table::inv_table> LoadTable(const std::string& filename)
{
    std::ifstream ifs(filename, std::ios::in | std::ios::binary);
    if (!ifs.good())
        throw genie::exception::cpu_runtime_error("Loading from file failed (fstream not good)");
    std::string bytes((std::istreambuf_iterator<char>(ifs)), std::istreambuf_iterator<char>());
    if (bytes.size() < 8 + 12 + 8 || std::memcmp(bytes.data(), kMagic, 8) != 0)
        bad("not a GENIE index file");
    uint64_t h;
    std::memcpy(&h, bytes.data() + bytes.size() - 8, 8);
    if (fnv1a(bytes.data(), bytes.size() - 8) != h)
        bad("checksum mismatch");
    uint32_t head[3];
    std::memcpy(head, bytes.data() + 8, sizeof(head));
    if (head[2] != kByteOrder)
        bad("byte order");
    if (head[0] != kVersion)
        bad("format version " + std::to_string(head[0]));
    if (head[1] > 1)
        bad("table kind");
    std::shared_ptr<genie::table::inv_table> t;
    if (head[1] == 1)
        t = std::make_shared<genie::table::inv_compr_table>();
    else
        t = std::make_shared<genie::table::inv_table>();
    const std::string body = bytes.substr(0, bytes.size() - 8);
    size_t at = 20;
    t->load_fields(body, at);
    In in(body, at);
    const int32_t shifter = in.get<int32_t>();
    const uint8_t single = in.get<uint8_t>();
    if (at != body.size())
        bad("trailing bytes");
    if (shifter != t->shifter())
        bad("key shifter " + std::to_string(shifter));
    t->finish_load(single != 0);
    return t;
}

}  // namespace utility
}  // namespace genie
