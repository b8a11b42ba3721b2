// Internal structures shared by the engine translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <vector>

#include "genie_hip.h"

namespace gh {

// Objects per directory tile.  Every workgroup tile is a multiple of it.
constexpr uint32_t kDirTile = 8192;
// Keys whose list is at most this long have no directory entry: a workgroup
// reads the whole list and keeps the ids inside its tile.
constexpr uint32_t kShortList = 256;
// Work items longer than this many postings are split into chunks on the
// device so that the waves of a workgroup share long lists evenly.
constexpr uint32_t kChunk = 32768;
constexpr uint32_t kNoDir = 0xFFFFFFFFu;

// Directory rows are graded: key k's row has an entry for every 2^s-th
// directory tile boundary (entry j describes boundary (j << s) * kDirTile,
// j = 0 .. ceil(n_dtiles / 2^s)), s chosen per key so that the row stays
// small next to the list it indexes.  dir_off[k] is the row's offset in
// words, a multiple of kDirAlign, with s in its low bits.
constexpr uint32_t kDirAlign = 16;
constexpr uint32_t kDirShiftMask = kDirAlign - 1;
// Tile-directory rows cost at most an eighth of their key's posting bytes;
// segment-directory rows at most 16 entries per segment of their key.
constexpr uint32_t kDirPostingsPerEntry = 8;
constexpr uint32_t kSdirEntriesPerSegment = 16;

__host__ __device__ inline uint32_t dir_row_len(uint32_t n_dtiles, uint32_t s)
{
    return ((n_dtiles + (1u << s) - 1) >> s) + 1;
}

// Finest grade whose row has at most max_entries entries; -1 when none has.
inline int dir_grade(uint32_t n_dtiles, uint64_t max_entries)
{
    for (uint32_t s = 0; s <= kDirShiftMask; ++s)
        if (dir_row_len(n_dtiles, s) <= max_entries)
            return (int)s;
    return -1;
}

// Lays out the rows of the keys for which grade_of(k) >= 0: off[k] gets the
// packed offset (kNoDir for the others), keys the row keys in order.
// Returns the directory's words, or 0 past the 32-bit offset range.
template <class GradeOf>
inline uint64_t dir_layout(uint32_t n_keys, uint32_t n_dtiles, GradeOf grade_of, std::vector<uint32_t>& off,
                           std::vector<uint32_t>& keys)
{
    off.assign(n_keys, kNoDir);
    keys.clear();
    uint64_t words = 0;
    for (uint32_t k = 0; k < n_keys; ++k) {
        const int s = grade_of(k);
        if (s < 0)
            continue;
        off[k] = (uint32_t)words | (uint32_t)s;
        keys.push_back(k);
        words += (dir_row_len(n_dtiles, (uint32_t)s) + kDirShiftMask) & ~(uint64_t)kDirShiftMask;
        if (words >= 0xFFFFFFF0ull)
            return 0;
    }
    return words ? words : kDirAlign;
}

// One compiled chunk of a work item: postings [sp, ep) covering keys k0..k1.
// kp is the position of the chunk's first (chunk, key) pair among all pairs
// of the batch (exclusive prefix of k1 - k0 + 1 over chunks).
struct CChunk {
    uint32_t k0, k1, sp, ep, w, q, kp, pad;
};

// Position -> key lookups over a non-decreasing array of key starts
// (starts[0..nk]): buckets of 2^shift positions, each holding the index of
// the first key that starts at or after the bucket, so a lookup searches the
// few keys of one bucket.
struct PosIndex {
    uint32_t shift = 0;
    std::vector<uint32_t> first;  // [n_buckets + 1]
    void build(const std::vector<uint64_t>& starts);
    // Last key whose start is <= pos (std::upper_bound - 1).
    uint32_t last_le(const uint64_t* starts, uint64_t pos) const;
    // First key whose start is >= pos (std::lower_bound).
    uint32_t first_ge(const uint64_t* starts, uint64_t pos) const;
};

// Segment view of a compressed index.
struct DevIndex {
    uint32_t n_objects = 0, n_dims = 0, n_keys = 0;
    uint64_t n_postings = 0;
    int32_t codec = 0;
    int32_t single_value = 1;
    uint32_t n_segments = 0;
    uint64_t n_words = 0;
    uint32_t n_dtiles = 0;
    // device
    uint32_t* post = nullptr;       // [n_postings]   (uncompressed)
    // key starts and tile directory in posting positions; a compressed index
    // has them too (built from its one decode at creation), for batches whose
    // segments are decoded once into a scratch of the uncompressed layout
    uint32_t* key_start = nullptr;  // [n_keys + 1]   positions (u32)
    uint32_t* dir_off = nullptr;    // [n_keys] packed row offset and grade, or kNoDir
    uint32_t* dir = nullptr;        // graded rows of the keys with a row
    uint32_t* words = nullptr;      // [n_words]      (compressed)
    uint32_t* seg_cpos = nullptr;   // [n_segments + 1]
    uint32_t* seg_upos = nullptr;   // [n_segments + 1]
    uint32_t* seg_key = nullptr;    // [n_segments]
    uint32_t* seg_first = nullptr;  // [n_segments] first id of each segment
    uint32_t* seg_last = nullptr;   // [n_segments] last id
    uint32_t* key_seg = nullptr;    // [n_keys + 1] first segment of each key
    uint4* seg_bph = nullptr;       // [n_segments] BP32 block header word offsets, 8 x u16 (BP32 codecs)
    uint32_t* sdir_off = nullptr;   // [n_keys] row of each key with more than 4 segments (compressed)
    uint32_t* sdir = nullptr;       // graded segment-directory rows (compressed)
    // forward codes of a single-value table (dense form of the batched search)
    bool dense = false;
    uint32_t dense_dpad = 0;        // dims rounded up to 16
    uint32_t dense_npad = 0;        // objects rounded up to 64 (row stride of codes)
    uint32_t* codes = nullptr;      // [dense_words(dpad)][dense_npad] packed 10-bit fields (code * 4)
    uint32_t* key_slot = nullptr;   // [n_keys] dim * 256 + code
    // forward lists of a set-valued table (set form of the batched search):
    // the keys of every object's occurrences, duplicates kept
    bool sets = false;
    uint32_t set_max_len = 0;       // most occurrences of one object
    uint32_t set_slots = 0;         // objects rounded up to 64
    uint32_t* obj_perm = nullptr;   // [set_slots] object of each slot (kSetEmpty: none)
    uint32_t* blk_off = nullptr;    // [set_slots / 64 + 1] first word of each 64-slot block
    uint32_t* blk_key = nullptr;    // [blk_off[last]] keys + 1 (0: padding), column-major per block
    uint64_t bytes = 0;
    // host copies used for validation and count bounds
    std::vector<uint64_t> h_key_start;
    std::vector<uint32_t> h_key_dim;
    std::vector<uint32_t> h_key_mult;  // max multiplicity of an id within each key's list
    std::vector<uint32_t> h_seg_cpos;
    std::vector<uint32_t> h_seg_key;
    std::vector<uint32_t> h_key_seg;
    std::vector<uint64_t> h_key_wstart;  // [n_keys + 1] first word of each key (compressed)
    PosIndex key_pos;                    // over h_key_start (uncompressed) or h_key_wstart (compressed)
    std::vector<uint32_t> h_seg_upos;    // [n_segments + 1] first posting of each segment (compressed)
    std::vector<uint64_t> h_seg_cpos64;  // h_seg_cpos widened, and its position index (compressed)
    PosIndex seg_pos;
};

// Candidates per query the dense and set forms' host total holds at least
// (a query whose counts crowd at its threshold asks for its ties).
constexpr uint64_t kDenseCapFloor = 16384;
// Initial factor of k that a plan gives each query's candidate space in the
// dense and set forms; the tile forms (plain and lane) take k per tile.
constexpr uint32_t kDenseCapFactor = 64;
// Fewest objects per tile of the tile forms (32-bit counters).
constexpr uint32_t kMinTileObjs = 16384;

// Candidate keys a new plan first gives one query with this k (the sizing of
// plan_size_candidates before any overflow): the dense and set forms hold
// max(64 k, kDenseCapFloor), at most one per object; the tile forms k per
// tile, counted here at the smallest tile size.
inline uint64_t initial_candidates(uint32_t k, bool dense, uint32_t n_objects)
{
    if (!dense)
        return (uint64_t)((n_objects + kMinTileObjs - 1) / kMinTileObjs) * k;
    const uint64_t c = (uint64_t)kDenseCapFactor * k;
    return std::min<uint64_t>(c > kDenseCapFloor ? c : kDenseCapFloor, n_objects);
}

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
// Checks a batch's items against the index: query ids, spans inside the
// posting (or compressed word) array, integral weights in [0, 65535],
// compressed spans on segment boundaries.
int validate_items(const DevIndex& d, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries);
// genie_hip_plan_upload for items already validated one request at a time
// (the server's batches): the same, without validating the items again.
int plan_upload_validated(genie_plan* p, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                          const uint32_t* k_per_query);
int plan_upload_staged(genie_plan* p, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                       const uint32_t* k_per_query, bool validate);

}  // namespace gh

struct genie_index {
    gh::DevIndex d;
};
