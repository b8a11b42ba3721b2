// MI355X (gfx950) match-count top-k engine: index upload, tile directory,
// device-side item compilation, the LDS-tiled scan/count/select kernel and
// the per-query merge.  C ABI in include/genie_hip.h.
//
// Data flow for one batch (plain form, one query per workgroup):
//   items (Query::dim, 20 B) --k_chunk_count/scan/k_chunk_write--> chunks
//   chunks --k_scan_tile<LB>--> per (tile, query): counts in LDS, exact
//        per-tile top-k under (count desc, id asc), pruned by the query's
//        adaptive threshold, appended to the query's candidate list
//   candidates --k_merge--> exact top-k per query, sorted, padded (0,0).
//
// The reference runs one 256-thread block per work item with counts in a
// global bitmap and a global c-PQ hash table (match.cu:43-158,
// match_device_utils.h:161-473) and then heap_count_topk
// (heap_count.cu:161-282).  Here the count table lives in LDS per object
// tile and the threshold / top-k are exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_set>
#include <vector>

#include "codecs_dev.h"
#include "engine.h"
#include "kernels.h"
#include "kernels_batched.h"
#include "kernels_dense.h"
#include "kernels_hash.h"
#include "wave.h"

namespace gh {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(e_ == hipErrorOutOfMemory ? GENIE_E_OOM : GENIE_E_RUNTIME,      \
                        std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

template <class T>
static int dev_alloc(T** p, size_t n, uint64_t* acct = nullptr)
{
    *p = nullptr;
    if (n == 0)
        n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(e == hipErrorOutOfMemory ? GENIE_E_OOM : GENIE_E_RUNTIME,
                    std::string("hipMalloc(") + std::to_string(n * sizeof(T)) +
                        " B): " + hipGetErrorString(e));
    }
    if (acct)
        *acct += n * sizeof(T);
    return GENIE_OK;
}

template <class T>
static void dev_free(T*& p)
{
    if (p)
        (void)hipFree(p);
    p = nullptr;
}

}  // namespace gh

using namespace gh;

// ===========================================================================
// Plan
// ===========================================================================

struct genie_plan {
    genie_index* ix = nullptr;
    int mode = 0;
    uint32_t max_queries = 0, max_k = 0, count_bound = 0;
    uint64_t max_items = 0;
    uint32_t lane_bits = 8;
    // batch shape of the last upload
    uint32_t n_queries = 0;
    uint64_t n_items = 0, n_chunks = 0, total_k = 0, total_cap = 0;
    // a compressed index's batch compiles into runs of contiguous segments,
    // or (decoded once into the full-layout scratch) like an uncompressed
    // batch: chunks of at most kChunk postings per item.  n_chunks is the
    // count of the compile path the next run takes; the chunk buffers hold
    // the larger of the two
    uint64_t n_chunks_runs = 0, n_chunks_full = 0;
    uint32_t cap_factor = 8;
    // dense and set forms: multiplier of the device-side candidate asks,
    // raised after a run whose asks were too small (a sample that
    // underestimated a query's count distribution)
    uint32_t dense_mult = 1;
    uint64_t logical_bytes = 0, compulsory_bytes = 0;
    // plain form: per query, the hash form (selective queries) or the
    // direct-mapped kernels; the direct queries' postings choose the sparse
    // select
    uint64_t est_postings = 0;  // postings of the direct queries
    int sparse_mode = -1;       // -1 automatic, 0 off, 1 on
    int hash_mode = -1;         // -1 automatic, 0 off, 1 every query whose counts fit
    std::vector<uint64_t> h_qpost;        // postings of every query
    std::vector<uint32_t> h_qbound;       // count bound of every query
    std::vector<uint8_t> h_force_direct;  // queries whose table filled up in an earlier run
    std::vector<uint32_t> h_hwork;        // (query, tile) of the hash kernel
    std::vector<uint4> h_hq;              // per query: directory tiles per tile, key bits, log2 slots, tiles
    std::vector<uint32_t> h_dq;           // queries of the direct-mapped kernels
    uint32_t max_hash_tiles = 0;
    bool plain_stats = false;             // h_qpost / h_qbound hold this batch's figures
    uint32_t* d_hwork = nullptr;
    uint64_t hwork_alloc = 0;
    uint4* d_hq = nullptr;
    uint32_t* d_dq = nullptr;
    uint32_t* d_hov = nullptr;
    uint32_t* d_hov_any = nullptr;
    // compressed index, hash form: the batch's segments decoded once into a
    // scratch posting array, with per-query span lists
    bool dec_spans = false;
    std::vector<Span> h_sspan;
    std::vector<uint32_t> h_sq_ptr;
    std::vector<uint2> h_djobs;
    Span* d_sspan = nullptr;
    uint32_t* d_sq_ptr = nullptr;
    uint2* d_djobs = nullptr;
    uint32_t* d_scratch = nullptr;
    uint64_t sspan_alloc = 0, djobs_alloc = 0, scratch_alloc = 0;
    // compressed index, plain form: the batch's distinct segments decoded
    // once into a scratch of the uncompressed layout (each segment at its
    // posting positions), the items translated to posting positions, and the
    // batch then counted like one on an uncompressed index
    int decode_mode = -1;                 // -1 automatic (when the scratch can be allocated), 0 off
    bool dec_full = false;
    std::vector<uint2> h_fjobs;           // (segment, first posting position)
    std::vector<genie_dim_t> h_uitems;    // items in posting positions
    uint2* d_fjobs = nullptr;
    uint64_t fjobs_alloc = 0;
    genie_dim_t* d_uitems = nullptr;      // [max_items]
    uint32_t* d_full = nullptr;           // [n_postings]
    // device buffers
    genie_dim_t* d_items = nullptr;
    uint32_t* d_item_qptr = nullptr;
    uint32_t* d_nchunks = nullptr;
    uint32_t* d_chunk_ptr = nullptr;
    uint32_t* d_block_sums = nullptr;
    uint32_t* d_total = nullptr;
    CChunk* d_chunks = nullptr;
    uint64_t chunks_cap = 0;
    uint32_t* d_qc_ptr = nullptr;
    uint32_t* d_qk_ptr = nullptr;
    uint32_t* d_nk = nullptr;        // per chunk keys, then pair prefix
    uint32_t* d_kptr = nullptr;
    uint32_t* d_chunk_sums = nullptr;
    uint32_t* d_total_pairs = nullptr;
    uint32_t* d_kq = nullptr;
    uint32_t* d_gthr = nullptr;
    unsigned long long* d_gkey = nullptr;  // hash form: each query's largest tile k-th key
    uint32_t* d_cand_cnt = nullptr;
    uint32_t* d_cand_cap = nullptr;
    uint64_t* d_cand_base = nullptr;
    uint64_t* d_cand = nullptr;
    uint64_t cand_alloc = 0;
    uint32_t* d_overflow = nullptr;
    uint32_t* d_flag_any = nullptr;  // d_overflow | d_hov_any, written by the merge
    uint64_t* d_out_off = nullptr;
    int32_t* d_out_ids = nullptr;
    int32_t* d_out_counts = nullptr;
    uint64_t out_alloc = 0, outc_alloc = 0;
    std::vector<uint32_t> h_kq;
    std::vector<uint64_t> h_out_off;
    // batched form (mode 1)
    uint32_t block = 512;            // threads per workgroup
    uint32_t tile_b = 8192;          // objects per tile (LDS counter capacity / 8 B)
    uint32_t seglen = 65536;         // longest posting span one wave takes at a time
    uint32_t lanes = 8, n_groups = 0, n_work = 0;
    uint64_t n_pairs = 0, pairs_alloc = 0;
    GPair* d_pairs = nullptr;
    uint32_t* d_gp_ptr = nullptr;
    int32_t* d_gq = nullptr;
    uint32_t* d_queue = nullptr;
    uint64_t gq_alloc = 0, gp_alloc = 0;
    std::vector<genie_dim_t> h_items;  // the uploaded batch sorted by query (work queue, byte figures)
    bool bytes_valid = false;
    std::vector<uint32_t> h_qptr;
    double host_build_ms = 0;
    bool lanes_built = false;        // work queue of the lane form built for this batch
    bool batched_plain = false;      // batched plan: the plain form where the dense and set forms do not apply
    // dense form of the batched search (single-value index, whole-key unit-weight batch)
    bool dense_on = true;            // use it when the batch allows
    int dense_nt = 0;                // threads per workgroup of its kernels (0: 1,024 when placed, else 512)
    bool dense_ok = false;           // the uploaded batch allows it
    uint32_t dense_wp = 1;           // its weight planes per query (1: unit weights; 2-4: bits of the largest)
    bool sets_ok = false;            // ... or its set form (set-valued tables)
    std::vector<uint32_t> h_skeys;   // set form: every group's sorted keys,
    std::vector<uint4> h_smask;      // their weight planes (two uint4 per key)
    std::vector<uint32_t> h_sgptr;   // and the groups' entry ranges
    std::vector<uint32_t> h_sshift;  // per query: sample histogram shift
    uint32_t* d_skeys = nullptr;
    uint4* d_smask = nullptr;
    uint32_t* d_sgptr = nullptr;
    uint32_t* d_sshift = nullptr;
    uint64_t sent_alloc = 0, smask_alloc = 0;
    uint32_t dn_groups = 0;
    uint32_t* d_dtab = nullptr;      // [dn_groups][dpad * 256] query masks
    uint64_t dtab_alloc = 0;
    uint32_t* d_dhist = nullptr;     // [max_queries][kDenseBins]
    uint32_t* d_dlist = nullptr;     // sample tiles
    uint32_t n_dlist = 0, dlist_alloc = 0, n_dsample = 0;
    uint32_t* d_gsafe = nullptr;     // [max_queries] safe threshold
    uint32_t* d_dkeys = nullptr;     // [max_items][2] key range of every item
    std::vector<uint32_t> h_dkeys;
    uint32_t* d_redo = nullptr;      // [groups] second count of a group
    uint32_t* d_want = nullptr;      // [max_queries] candidate space asked for at spec / at safe
    uint32_t* d_want_safe = nullptr;
    unsigned long long* d_cap_used = nullptr;  // end of the assigned candidate space
    // uploads: what genie_hip_plan_upload copies to the device is gathered in
    // one pinned buffer and copied asynchronously by the next run, on its
    // stream; ev_up marks the end of those copies.  A staged copy names the
    // plan's pointer field, read when the copies are issued (a buffer grown
    // meanwhile is copied to where it lives)
    struct Staged {
        void** dst;
        uint64_t off, bytes;
    };
    uint8_t* h_stage = nullptr;
    uint64_t stage_cap = 0, stage_used = 0;
    std::vector<Staged> staged;
    bool staging = false, up_pending = false;
    hipEvent_t ev_up = nullptr;
    // graph replay
    bool graph_on = false, graph_valid = false;
    hipStream_t own_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // timing
    bool timing = false;
    hipEvent_t ev[8] = {};
    double ms_compile = 0, ms_scan = 0, ms_merge = 0, ms_sample = 0;
    bool timed_dense = false;
    bool timed_decode = false;  // the last timed run decoded segments into a scratch (ev[5] after it)
    bool ran = false;
};

static uint32_t bits_for(uint32_t bound)
{
    return bound <= 255 ? 8u : (bound <= 65535 ? 16u : 32u);
}

static uint32_t tile_objs_for(uint32_t lane_bits) { return (uint32_t)kCntWords * (32u / lane_bits); }

static void plan_drop_graph(genie_plan* p)
{
    if (p->gexec)
        (void)hipGraphExecDestroy(p->gexec);
    if (p->graph)
        (void)hipGraphDestroy(p->graph);
    p->gexec = nullptr;
    p->graph = nullptr;
    p->graph_valid = false;
}

static void plan_free_buffers(genie_plan* p)
{
    plan_drop_graph(p);
    dev_free(p->d_pairs);
    dev_free(p->d_gp_ptr);
    dev_free(p->d_gq);
    dev_free(p->d_queue);
    dev_free(p->d_items);
    dev_free(p->d_item_qptr);
    dev_free(p->d_nchunks);
    dev_free(p->d_chunk_ptr);
    dev_free(p->d_block_sums);
    dev_free(p->d_total);
    dev_free(p->d_chunks);
    dev_free(p->d_qc_ptr);
    dev_free(p->d_qk_ptr);
    dev_free(p->d_nk);
    dev_free(p->d_kptr);
    dev_free(p->d_chunk_sums);
    dev_free(p->d_total_pairs);
    dev_free(p->d_kq);
    dev_free(p->d_gthr);
    dev_free(p->d_gkey);
    dev_free(p->d_cand_cnt);
    dev_free(p->d_cand_cap);
    dev_free(p->d_cand_base);
    dev_free(p->d_cand);
    dev_free(p->d_overflow);
    dev_free(p->d_flag_any);
    dev_free(p->d_out_off);
    dev_free(p->d_out_ids);
    dev_free(p->d_out_counts);
    dev_free(p->d_dtab);
    dev_free(p->d_dhist);
    dev_free(p->d_dlist);
    dev_free(p->d_gsafe);
    dev_free(p->d_redo);
    dev_free(p->d_want);
    dev_free(p->d_want_safe);
    dev_free(p->d_cap_used);
    dev_free(p->d_dkeys);
    dev_free(p->d_hwork);
    dev_free(p->d_hq);
    dev_free(p->d_dq);
    dev_free(p->d_hov);
    dev_free(p->d_hov_any);
    dev_free(p->d_skeys);
    dev_free(p->d_smask);
    dev_free(p->d_sgptr);
    dev_free(p->d_sshift);
    dev_free(p->d_sspan);
    dev_free(p->d_sq_ptr);
    dev_free(p->d_djobs);
    dev_free(p->d_scratch);
    dev_free(p->d_fjobs);
    dev_free(p->d_uitems);
    dev_free(p->d_full);
}

// Count bound of every query: per (query, order) group the largest weight
// times how often one object can occur in the group's spans (1 per distinct
// dim for single-value tables; the sum of per-key multiplicities for
// bijectMap tables), summed over groups.  Items are bucketed by query and
// each query's entries sorted on their own (one entry per item when the item
// lies in one dim of a single-value table, one per key otherwise).
static int compute_bounds(const DevIndex& d, const genie_dim_t* items, uint64_t n_items,
                          uint32_t n_queries, uint32_t* out)
{
    std::vector<uint32_t> qptr(n_queries + 1, 0), idx(n_items);
    for (uint64_t i = 0; i < n_items; ++i)
        qptr[items[i].query + 1]++;
    for (uint32_t q = 0; q < n_queries; ++q)
        qptr[q + 1] += qptr[q];
    {
        std::vector<uint32_t> fill(qptr.begin(), qptr.end() - 1);
        for (uint64_t i = 0; i < n_items; ++i)
            idx[fill[items[i].query]++] = (uint32_t)i;
    }
    const uint64_t* ks = d.h_key_start.data();
    const uint32_t nk = d.n_keys;
    (void)nk;
    auto key_of_pos = [&](uint64_t pos) -> uint32_t { return d.key_pos.last_le(ks, pos); };
    auto one_query = [&](uint32_t q) {
        struct E {
            int32_t order;
            uint32_t x, w;
        };
        std::vector<E> v;
        for (uint32_t j = qptr[q]; j < qptr[q + 1]; ++j) {
            const genie_dim_t& it = items[idx[j]];
            if (it.end_pos <= it.start_pos)
                continue;
            uint32_t k0, k1;
            if (d.codec == 0) {
                k0 = key_of_pos((uint64_t)it.start_pos);
                k1 = key_of_pos((uint64_t)it.end_pos - 1);
            } else {
                const uint32_t* cp = d.h_seg_cpos.data();
                const uint32_t s0 = (uint32_t)(std::upper_bound(cp, cp + d.n_segments + 1, (uint32_t)it.start_pos) - cp - 1);
                const uint32_t s1 = (uint32_t)(std::upper_bound(cp, cp + d.n_segments + 1, (uint32_t)it.end_pos - 1) - cp - 1);
                k0 = d.h_seg_key[s0];
                k1 = d.h_seg_key[s1];
            }
            const uint32_t w = (uint32_t)it.weight;
            if (d.single_value && d.h_key_dim[k0] == d.h_key_dim[k1])
                v.push_back({it.order, d.h_key_dim[k0], w});
            else
                for (uint32_t k = k0; k <= k1; ++k)
                    v.push_back({it.order, d.single_value ? d.h_key_dim[k] : k, w});
        }
        std::sort(v.begin(), v.end(), [](const E& a, const E& b) {
            return a.order != b.order ? a.order < b.order : a.x < b.x;
        });
        uint64_t acc = 0;
        for (size_t i = 0; i < v.size();) {
            size_t j = i;
            uint32_t wmax = 0;
            uint64_t mult = 0;
            for (; j < v.size() && v[j].order == v[i].order; ++j) {
                wmax = std::max(wmax, v[j].w);
                if (j > i && v[j].x == v[j - 1].x)
                    continue;  // the same dim (key) again
                mult += d.single_value ? 1u : d.h_key_mult[v[j].x];
            }
            acc += (uint64_t)wmax * mult;
            i = j;
        }
        out[q] = (uint32_t)std::min<uint64_t>(acc, 0xFFFFFFFFull);
    };
    if (n_items < 65536) {
        for (uint32_t q = 0; q < n_queries; ++q)
            one_query(q);
    } else {
        const uint32_t T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        std::atomic<uint32_t> next(0);
        for (uint32_t t = 0; t < T; ++t)
            th.emplace_back([&] {
                for (uint32_t q; (q = next.fetch_add(64)) < n_queries;)
                    for (uint32_t e = q; e < std::min(n_queries, q + 64); ++e)
                        one_query(e);
            });
        for (auto& x : th)
            x.join();
    }
    return GENIE_OK;
}

void gh::PosIndex::build(const std::vector<uint64_t>& starts)
{
    const uint64_t nk = starts.size() - 1, span = starts.back() + 1;
    shift = 0;  // about two buckets per key
    while ((span >> shift) > 2 * nk + 1)
        ++shift;
    const uint64_t nb = (span >> shift) + 1;
    first.resize(nb + 1);
    uint64_t k = 0;
    for (uint64_t b = 0; b <= nb; ++b) {
        while (k <= nk && starts[k] < (b << shift))
            ++k;
        first[b] = (uint32_t)k;
    }
}

uint32_t gh::PosIndex::last_le(const uint64_t* starts, uint64_t pos) const
{
    const uint64_t b = std::min<uint64_t>(pos >> shift, first.size() - 2);
    const uint64_t* it = std::upper_bound(starts + first[b], starts + first[b + 1], pos);
    return (uint32_t)(it - starts) - 1;
}

uint32_t gh::PosIndex::first_ge(const uint64_t* starts, uint64_t pos) const
{
    const uint64_t b = std::min<uint64_t>(pos >> shift, first.size() - 2);
    return (uint32_t)(std::lower_bound(starts + first[b], starts + first[b + 1], pos) - starts);
}

int gh::validate_items(const DevIndex& d, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries)
{
    const uint64_t limit = d.codec == 0 ? d.n_postings : d.n_words;
    for (uint64_t i = 0; i < n_items; ++i) {
        const genie_dim_t& it = items[i];
        if (it.query < 0 || (uint32_t)it.query >= n_queries)
            return fail(GENIE_E_INVALID, "item " + std::to_string(i) + ": query id out of range");
        if (it.start_pos < 0 || it.end_pos < 0 || (uint64_t)it.end_pos > limit ||
            (it.end_pos > it.start_pos && (uint64_t)it.start_pos >= limit))
            return fail(GENIE_E_INVALID, "item " + std::to_string(i) + ": span outside the index");
        if (!(it.weight >= 0.0f) || it.weight != (float)(uint32_t)it.weight || it.weight > 65535.0f)
            return fail(GENIE_E_INVALID, "item " + std::to_string(i) +
                                             ": weight must be a non-negative integer <= 65535");
        if (d.codec != 0 && it.end_pos > it.start_pos) {
            // compressed items cover whole segments
            const uint64_t* cp = d.h_seg_cpos64.data();
            const uint32_t a = d.seg_pos.first_ge(cp, (uint64_t)it.start_pos);
            const uint32_t b = d.seg_pos.first_ge(cp, (uint64_t)it.end_pos);
            if (a > d.n_segments || cp[a] != (uint64_t)it.start_pos || b > d.n_segments ||
                cp[b] != (uint64_t)it.end_pos)
                return fail(GENIE_E_INVALID, "item " + std::to_string(i) +
                                                 ": compressed span does not align with segments");
        }
    }
    return GENIE_OK;
}

static bool dense_shape(const genie_index_desc* desc, std::vector<uint32_t>& key_base, std::vector<uint32_t>& per_dim);
static int build_sets(const genie_index_desc* desc, DevIndex& d);
static int build_dense(const genie_index_desc* desc, DevIndex& d, const uint32_t* post, const uint32_t* key_start);

// Compressed index: words, segment offsets, key -> segment map, per-segment
// first/last ids (one device decode pass that also validates every segment)
// and the segment directory of long keys.
static int create_compressed(const genie_index_desc* desc, genie_index** out)
{
    const uint32_t S = desc->n_segments, K = desc->n_keys;
    const uint32_t* up = desc->seg_upos;
    const uint32_t* cp = desc->seg_cpos;
    if (up[0] != 0 || up[S] != desc->n_postings || cp[0] != 0 || cp[S] != desc->n_words)
        return fail(GENIE_E_INVALID, "segment offsets must run from 0 to the array sizes");
    for (uint32_t i = 0; i < S; ++i)
        if (up[i + 1] <= up[i] || up[i + 1] - up[i] > kSegMax || cp[i + 1] <= cp[i] || cp[i + 1] - cp[i] > kSegMax)
            return fail(GENIE_E_INVALID, "segment " + std::to_string(i) + " is empty or longer than " +
                                             std::to_string(kSegMax) + " ids or words");
    // BP32 codecs: every segment's block header positions (word offsets
    // within the segment, at most 8 blocks of 128 values), so that decoders
    // find all miniblocks at once instead of walking the headers; the device
    // validation pass below rejects any segment these do not describe
    CodecShape cs{false, false, kCopy};
    if (!codec_shape(desc->codec, &cs))
        return fail(GENIE_E_INVALID, "unknown codec " + std::to_string(desc->codec));
    std::vector<uint4> bph;
    if (cs.composite || cs.leaf == kBp32) {
        bph.assign(S, make_uint4(0, 0, 0, 0));
        for (uint32_t sg = 0; sg < S; ++sg) {
            const uint32_t* w = desc->words + cp[sg];
            const uint32_t nw = cp[sg + 1] - cp[sg];
            const uint32_t o = cs.composite ? 1u : 0u;  // the BP32 part's length word
            if (o >= nw || (cs.composite && w[0] == 0))
                continue;
            const uint32_t n = std::min<uint32_t>(w[o], kSegMax);
            uint32_t hp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            uint32_t p = o + 1;
            for (uint32_t j = 0; j < (n + 127) / 128 && p < nw; ++j) {
                hp[j] = p;
                const uint32_t h = w[p];
                p += 1 + (h >> 24) + ((h >> 16) & 255u) + ((h >> 8) & 255u) + (h & 255u);
            }
            bph[sg] = make_uint4(hp[0] | hp[1] << 16, hp[2] | hp[3] << 16, hp[4] | hp[5] << 16, hp[6] | hp[7] << 16);
        }
    }
    // every key starts on a segment boundary
    std::vector<uint32_t> key_seg(K + 1), seg_key(S);
    for (uint32_t k = 0; k <= K; ++k) {
        const uint64_t ks = desc->key_start[k];
        const uint32_t sidx = (uint32_t)(std::lower_bound(up, up + S + 1, (uint32_t)ks) - up);
        if (sidx > S || (sidx < S && up[sidx] != ks) || (sidx == S && ks != desc->n_postings))
            return fail(GENIE_E_INVALID, "key " + std::to_string(k) + " does not start on a segment boundary");
        key_seg[k] = sidx;
    }
    for (uint32_t k = 0; k < K; ++k)
        for (uint32_t sg = key_seg[k]; sg < key_seg[k + 1]; ++sg)
            seg_key[sg] = k;

    genie_index* ix = new genie_index();
    DevIndex& d = ix->d;
    d.n_objects = desc->n_objects;
    d.n_dims = desc->n_dims;
    d.n_keys = K;
    d.n_postings = desc->n_postings;
    d.codec = desc->codec;
    d.single_value = desc->single_value;
    d.n_segments = S;
    d.n_words = desc->n_words;
    d.n_dtiles = (desc->n_objects + kDirTile - 1) / kDirTile;
    d.h_key_start.assign(desc->key_start, desc->key_start + K + 1);
    d.h_key_dim.assign(desc->key_dim, desc->key_dim + K);
    d.h_key_mult.assign(K, 1);
    d.h_seg_cpos.assign(cp, cp + S + 1);
    d.h_seg_upos.assign(up, up + S + 1);
    d.h_seg_key = seg_key;
    d.h_key_seg = key_seg;
    d.h_key_wstart.resize(K + 1);
    for (uint32_t k = 0; k <= K; ++k)
        d.h_key_wstart[k] = cp[key_seg[k]];
    d.key_pos.build(d.h_key_wstart);
    d.h_seg_cpos64.assign(cp, cp + S + 1);
    d.seg_pos.build(d.h_seg_cpos64);
    if (desc->postings) {
        for (uint32_t k = 0; k < K; ++k) {
            uint32_t best = 1, run = 1;
            for (uint64_t p = desc->key_start[k] + 1; p < desc->key_start[k + 1]; ++p) {
                run = desc->postings[p] == desc->postings[p - 1] ? run + 1 : 1;
                best = std::max(best, run);
            }
            d.h_key_mult[k] = best;
        }
    }
    // segment directory: rows for keys of more than 4 segments, graded to at
    // most 16 entries per segment of the key
    std::vector<uint32_t> sdir_off, dir_keys;
    const uint64_t sdir_words = dir_layout(K, d.n_dtiles, [&](uint32_t k) {
        const uint64_t ns = key_seg[k + 1] - key_seg[k];
        return ns > 4 ? dir_grade(d.n_dtiles, ns * kSdirEntriesPerSegment) : -1;
    }, sdir_off, dir_keys);
    // tile directory of the posting positions: rows for keys longer than
    // kShortList, graded to at most an eighth of the key's posting bytes (an
    // index of many short lists, such as LSH buckets, gets coarse rows)
    std::vector<uint32_t> tdir_off, tdir_keys;
    const uint64_t tdir_words = dir_layout(K, d.n_dtiles, [&](uint32_t k) {
        const uint64_t len = desc->key_start[k + 1] - desc->key_start[k];
        return len > kShortList ? dir_grade(d.n_dtiles, len / kDirPostingsPerEntry) : -1;
    }, tdir_off, tdir_keys);
    if (!sdir_words || !tdir_words) {
        delete ix;
        return fail(GENIE_E_INVALID, "segment or tile directory too large");
    }
    // the whole table is decoded once into a temporary posting array: the
    // tile directory and, when the table can have the dense form, its
    // forward codes are built from it
    std::vector<uint32_t> kbase, kper;
    const bool want_dense = dense_shape(desc, kbase, kper);
    std::vector<uint32_t> ks32(K + 1);
    for (uint32_t k = 0; k <= K; ++k)
        ks32[k] = (uint32_t)desc->key_start[k];
    int rc = GENIE_OK;
    uint32_t *d_upos = nullptr, *d_dir_keys = nullptr, *d_err = nullptr, *d_post = nullptr, *d_tdir_keys = nullptr;
    do {
        if ((rc = dev_alloc(&d_post, d.n_postings))) break;
        if ((rc = dev_alloc(&d.key_start, K + 1, &d.bytes))) break;
        if ((rc = dev_alloc(&d.dir_off, K, &d.bytes))) break;
        if ((rc = dev_alloc(&d.dir, tdir_words, &d.bytes))) break;
        if ((rc = dev_alloc(&d_tdir_keys, tdir_keys.size()))) break;
        if (hipError_t e = hipMemcpy(d.key_start, ks32.data(), (K + 1) * 4ull, hipMemcpyHostToDevice)) {
            rc = fail(GENIE_E_RUNTIME, std::string("index upload: ") + hipGetErrorString(e));
            break;
        }
        if ((rc = dev_alloc(&d.words, d.n_words, &d.bytes))) break;
        if ((rc = dev_alloc(&d.seg_cpos, S + 1, &d.bytes))) break;
        if ((rc = dev_alloc(&d.seg_key, S, &d.bytes))) break;
        if ((rc = dev_alloc(&d.seg_first, S, &d.bytes))) break;
        if ((rc = dev_alloc(&d.seg_last, S, &d.bytes))) break;
        if ((rc = dev_alloc(&d.seg_bph, bph.empty() ? 1 : S, &d.bytes))) break;
        if ((rc = dev_alloc(&d.key_seg, K + 1, &d.bytes))) break;
        if ((rc = dev_alloc(&d.sdir_off, K, &d.bytes))) break;
        if ((rc = dev_alloc(&d.sdir, sdir_words, &d.bytes))) break;
        if ((rc = dev_alloc(&d_upos, S + 1))) break;
        if ((rc = dev_alloc(&d_dir_keys, dir_keys.size()))) break;
        if ((rc = dev_alloc(&d_err, 1))) break;
        hipError_t e;
        const uint32_t zero = 0;
        if ((e = hipMemcpy(d.words, desc->words, d.n_words * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d.seg_cpos, cp, (S + 1) * 4, hipMemcpyHostToDevice)) ||
            (!bph.empty() && (e = hipMemcpy(d.seg_bph, bph.data(), (uint64_t)S * 16, hipMemcpyHostToDevice))) ||
            (e = hipMemcpy(d_upos, up, (S + 1) * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d.seg_key, seg_key.data(), S * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d.key_seg, key_seg.data(), (K + 1) * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d.sdir_off, sdir_off.data(), K * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d_dir_keys, dir_keys.data(), dir_keys.size() * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d.dir_off, tdir_off.data(), K * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d_tdir_keys, tdir_keys.data(), tdir_keys.size() * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d_err, &zero, 4, hipMemcpyHostToDevice))) {
            rc = fail(GENIE_E_RUNTIME, std::string("index upload: ") + hipGetErrorString(e));
            break;
        }
        const uint32_t blocks = std::min<uint32_t>((S + kDecodeWaves - 1) / kDecodeWaves, 65535);
        k_seg_bounds<<<blocks, kDecodeWaves * kWave>>>(d.codec, d.words, d.seg_cpos, d_upos, S, d.n_objects,
                                                        d.seg_first, d.seg_last, d_err, d_post);
        uint32_t err = 0;
        if ((e = hipGetLastError()) || (e = hipDeviceSynchronize()) ||
            (e = hipMemcpy(&err, d_err, 4, hipMemcpyDeviceToHost))) {
            rc = fail(GENIE_E_RUNTIME, std::string("segment decode: ") + hipGetErrorString(e));
            break;
        }
        if (err) {
            rc = fail(GENIE_E_INVALID, "a compressed segment does not decode to its ascending ids");
            break;
        }
        std::vector<uint32_t> first(S), last(S);
        if ((e = hipMemcpy(first.data(), d.seg_first, S * 4, hipMemcpyDeviceToHost)) ||
            (e = hipMemcpy(last.data(), d.seg_last, S * 4, hipMemcpyDeviceToHost))) {
            rc = fail(GENIE_E_RUNTIME, std::string("segment bounds: ") + hipGetErrorString(e));
            break;
        }
        for (uint32_t sg = 0; sg + 1 < S && !rc; ++sg)
            if (seg_key[sg] == seg_key[sg + 1] && last[sg] > first[sg + 1])
                rc = fail(GENIE_E_INVALID, "segments of key " + std::to_string(seg_key[sg]) + " are not ascending");
        if (rc)
            break;
        if (!dir_keys.empty()) {
            const uint32_t g = (uint32_t)std::min<uint64_t>((dir_keys.size() + 3) / 4, 65536);
            k_build_sdir<<<g, 256>>>(d.seg_last, d.key_seg, d_dir_keys, (uint32_t)dir_keys.size(), d.sdir_off,
                                      d.n_dtiles, d.sdir);
            if ((e = hipGetLastError()) || (e = hipDeviceSynchronize())) {
                rc = fail(GENIE_E_RUNTIME, std::string("segment directory: ") + hipGetErrorString(e));
                break;
            }
        }
        if (!tdir_keys.empty()) {
            const uint32_t g = (uint32_t)std::min<uint64_t>((tdir_keys.size() + 3) / 4, 65536);
            k_build_dir<<<g, 256>>>(d_post, d.key_start, d_tdir_keys, (uint32_t)tdir_keys.size(), d.dir_off,
                                     d.n_dtiles, d.dir);
            if ((e = hipGetLastError()) || (e = hipDeviceSynchronize())) {
                rc = fail(GENIE_E_RUNTIME, std::string("tile directory: ") + hipGetErrorString(e));
                break;
            }
        }
        if (want_dense)
            rc = build_dense(desc, d, d_post, d.key_start);
        if (!rc)
            rc = build_sets(desc, d);
    } while (0);
    dev_free(d_upos);
    dev_free(d_dir_keys);
    dev_free(d_tdir_keys);
    dev_free(d_err);
    dev_free(d_post);
    if (rc) {
        genie_hip_index_destroy(ix);
        return rc;
    }
    *out = ix;
    return GENIE_OK;
}

static uint32_t plan_n_tiles(const genie_plan* p);
static ScanArgs scan_args(genie_plan* p);

// ---- per-batch buffers that grow ------------------------------------------

constexpr uint64_t kGrowSlackBytes = 256ull << 20;  // most a growing buffer takes beyond its need

// Grows *buf (holding *alloc elements) to at least `need` elements; the
// contents are not kept.  Freeing the old buffer waits for the whole device
// (every stream, every executor of a server), so a buffer grows to room for a
// full batch at this batch's size per query, at least twice its old size, at
// most kGrowSlackBytes beyond the need (the exact need when that fails): a
// serving loop whose batches grow step by step then stalls a few times at
// most instead of at every new largest batch.  `worst` (elements, optional)
// is what the plan's largest batch can need; it is taken within the same
// limit.
template <class T>
static int grow_buf(const genie_plan* p, T** buf, uint64_t* alloc, uint64_t need, uint64_t worst = 0)
{
    if (need <= *alloc)
        return GENIE_OK;
    const uint64_t old = *alloc;
    dev_free(*buf);
    *alloc = 0;
    const uint64_t nq = std::max<uint32_t>(1, p->n_queries);
    const uint64_t full = (need + nq - 1) / nq * std::max<uint64_t>(nq, p->max_queries);
    const uint64_t n = std::max(need, std::min(std::max({full, 2 * old, worst}), need + kGrowSlackBytes / sizeof(T)));
    if (n > need) {
        if (hipMalloc((void**)buf, n * sizeof(T)) == hipSuccess) {
            *alloc = n;
            return GENIE_OK;
        }
        *buf = nullptr;
        (void)hipGetLastError();  // the exact size is tried next
    }
    if (int rc = dev_alloc(buf, need))
        return rc;
    *alloc = need;
    return GENIE_OK;
}

// ---- batched form: host-side work queue ---------------------------------

// Group consecutive queries into lanes of one u64 counter word, resolve every
// item into per-key posting spans (segment ranges for a compressed index),
// merge spans a group holds more than once (one read, increments summed over
// the lanes), split long spans for wave balance, upload.
static int build_batched(genie_plan* p)
{
    const auto t0 = std::chrono::steady_clock::now();
    const DevIndex& d = p->ix->d;
    const uint32_t LB = p->lane_bits, lanes = 64 / LB, nq = p->n_queries;
    p->lanes = lanes;
    p->n_groups = (nq + lanes - 1) / lanes;
    const uint32_t G = p->n_groups;
    std::vector<int32_t> gq((size_t)G * lanes, -1);
    for (uint32_t q = 0; q < nq; ++q)
        gq[q] = (int32_t)q;  // query q -> group q / lanes, lane q % lanes
    std::vector<std::vector<GPair>> per(G);
    const uint64_t* ks = d.h_key_start.data();
    const uint32_t* cp = d.h_seg_cpos.data();
    auto group_pairs = [&](uint32_t g) {
        std::vector<GPair>& out = per[g];
        for (uint32_t l = 0; l < lanes; ++l) {
            const uint32_t q = g * lanes + l;
            if (q >= nq)
                break;
            const uint64_t unit = 1ull << (LB * l);
            for (uint32_t i = p->h_qptr[q]; i < p->h_qptr[q + 1]; ++i) {
                const genie_dim_t& it = p->h_items[i];
                if (it.end_pos <= it.start_pos)
                    continue;
                const uint64_t inc = (uint64_t)(uint32_t)it.weight * unit;
                if (d.codec == 0) {
                    const uint32_t sp = (uint32_t)it.start_pos, ep = (uint32_t)it.end_pos;
                    const uint32_t k0 = (uint32_t)(std::upper_bound(ks, ks + d.n_keys + 1, (uint64_t)sp) - ks - 1);
                    const uint32_t k1 = (uint32_t)(std::upper_bound(ks, ks + d.n_keys + 1, (uint64_t)ep - 1) - ks - 1);
                    for (uint32_t k = k0; k <= k1; ++k) {
                        const uint32_t lo = std::max<uint32_t>(sp, (uint32_t)ks[k]);
                        const uint32_t hi = std::min<uint32_t>(ep, (uint32_t)ks[k + 1]);
                        if (lo < hi)
                            out.push_back({k, lo, hi, 0, inc});
                    }
                } else {
                    const uint32_t s0 = (uint32_t)(std::upper_bound(cp, cp + d.n_segments + 1, (uint32_t)it.start_pos) - cp - 1);
                    const uint32_t s1 = (uint32_t)(std::upper_bound(cp, cp + d.n_segments + 1, (uint32_t)it.end_pos - 1) - cp);
                    for (uint32_t sg = s0; sg < s1;) {
                        const uint32_t k = d.h_seg_key[sg];
                        const uint32_t e = std::min(s1, d.h_key_seg[k + 1]);
                        out.push_back({k, sg, e, 0, inc});
                        sg = e;
                    }
                }
            }
        }
        std::sort(out.begin(), out.end(), [](const GPair& a, const GPair& b) {
            if (a.key != b.key) return a.key < b.key;
            if (a.lo != b.lo) return a.lo < b.lo;
            return a.hi < b.hi;
        });
        size_t m = 0;
        for (size_t i = 0; i < out.size(); ++i) {
            if (m && out[m - 1].key == out[i].key && out[m - 1].lo == out[i].lo && out[m - 1].hi == out[i].hi)
                out[m - 1].inc += out[i].inc;  // lanes are disjoint bit fields; repeats add within a lane
            else
                out[m++] = out[i];
        }
        out.resize(m);
        if (d.codec == 0 && p->seglen) {
            std::vector<GPair> split;
            split.reserve(out.size());
            for (const GPair& x : out)
                for (uint32_t a = x.lo; a < x.hi; a += p->seglen)
                    split.push_back({x.key, a, std::min(x.hi, a + p->seglen), 0, x.inc});
            out.swap(split);
        }
    };
    {
        const uint32_t T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        std::atomic<uint32_t> next(0);
        for (uint32_t t = 0; t < std::min(T, G); ++t)
            th.emplace_back([&] {
                for (uint32_t g; (g = next.fetch_add(1)) < G;)
                    group_pairs(g);
            });
        for (auto& x : th)
            x.join();
    }
    std::vector<uint32_t> gp(G + 1, 0);
    for (uint32_t g = 0; g < G; ++g)
        gp[g + 1] = gp[g] + (uint32_t)per[g].size();
    const uint64_t np = gp[G];
    std::vector<GPair> all(np);
    for (uint32_t g = 0; g < G; ++g)
        std::copy(per[g].begin(), per[g].end(), all.begin() + gp[g]);
    // sized apart: a batch of a wider lane width has more groups in the same G x lanes
    if (int rc = grow_buf(p, &p->d_pairs, &p->pairs_alloc, np))
        return rc;
    if (int rc = grow_buf(p, &p->d_gp_ptr, &p->gp_alloc, G + 1ull))
        return rc;
    if (int rc = grow_buf(p, &p->d_gq, &p->gq_alloc, (uint64_t)G * lanes))
        return rc;
    if (np)
        HIP_TRY(hipMemcpy(p->d_pairs, all.data(), np * sizeof(GPair), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->d_gp_ptr, gp.data(), gp.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->d_gq, gq.data(), gq.size() * 4, hipMemcpyHostToDevice));
    p->n_pairs = np;
    p->n_work = G * plan_n_tiles(p);
    p->host_build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return GENIE_OK;
}

// *placed = false (nothing launched) when the occupancy query places no
// workgroup of NT threads.
template <int LB, bool CMP, int NT>
static int launch_batched_t(genie_plan* p, const BatchArgs& b, hipStream_t s, bool* placed)
{
    *placed = false;
    const size_t lds = (size_t)(p->tile_b + kWave) * 8;
    auto kern = k_scan_batched<LB, CMP, NT>;
    hipFuncAttributes fa;
    HIP_TRY(hipFuncGetAttributes(&fa, (const void*)kern));
    if (fa.sharedSizeBytes + lds > 160 * 1024)
        return fail(GENIE_E_INVALID, "batched kernel with " + std::to_string(NT) + " threads and tile " +
                                         std::to_string(p->tile_b) + " does not fit in LDS");
    HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0, n_cu = 0, dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, lds));
    if (per_cu < 1)
        return GENIE_OK;
    const uint32_t grid = std::max(1u, std::min<uint32_t>(b.n_work, (uint32_t)(per_cu * n_cu)));
    kern<<<grid, NT, lds, s>>>(b);
    *placed = true;
    return GENIE_OK;
}

// The plan's block size, or the next smaller one where the occupancy query
// places none (some runtimes place no 1,024-thread workgroup of more than 64
// VGPRs); every block size gives the same results.
template <int LB>
static int launch_batched_lb(genie_plan* p, const BatchArgs& b, hipStream_t s)
{
    const bool c = p->ix->d.codec != 0;
    bool placed = false;
    if (p->block == 1024) {
        if (int rc = c ? launch_batched_t<LB, true, 1024>(p, b, s, &placed)
                       : launch_batched_t<LB, false, 1024>(p, b, s, &placed))
            return rc;
    }
    if (!placed && p->block >= 512) {
        if (int rc = c ? launch_batched_t<LB, true, 512>(p, b, s, &placed)
                       : launch_batched_t<LB, false, 512>(p, b, s, &placed))
            return rc;
    }
    if (!placed) {
        if (int rc = c ? launch_batched_t<LB, true, 256>(p, b, s, &placed)
                       : launch_batched_t<LB, false, 256>(p, b, s, &placed))
            return rc;
    }
    if (!placed)
        return fail(GENIE_E_INVALID, "batched kernel does not fit a CU with tile " + std::to_string(p->tile_b));
    return GENIE_OK;
}

static int launch_batched(genie_plan* p, hipStream_t s)
{
    BatchArgs b;
    b.s = scan_args(p);
    b.s.tile_objs = p->tile_b;
    b.s.n_tiles = plan_n_tiles(p);
    b.s.dt_per_tile = p->tile_b / kDirTile;
    b.pairs = p->d_pairs;
    b.gp_ptr = p->d_gp_ptr;
    b.gq = p->d_gq;
    b.n_groups = p->n_groups;
    b.lanes = p->lanes;
    b.n_work = p->n_work;
    b.queue = p->d_queue;
    if (b.n_work == 0)
        return GENIE_OK;
    switch (p->lane_bits) {
    case 8: return launch_batched_lb<8>(p, b, s);
    case 16: return launch_batched_lb<16>(p, b, s);
    default: return launch_batched_lb<32>(p, b, s);
    }
}

// Whether an index can have the dense form, from its key table alone: a
// single-value table of at most kDenseMaxDims dims with at most 256 keys per
// dim, keys ordered by dim, and N postings in every dim.  Fills the first key
// and the key count of every dim.
static bool dense_shape(const genie_index_desc* desc, std::vector<uint32_t>& key_base, std::vector<uint32_t>& per_dim)
{
    const uint32_t D = desc->n_dims, N = desc->n_objects, K = desc->n_keys;
    if (!desc->single_value || D == 0 || D > kDenseMaxDims)
        return false;
    key_base.assign(D, 0xFFFFFFFFu);
    per_dim.assign(D, 0);
    std::vector<uint64_t> dim_post(D, 0);
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t dm = desc->key_dim[k];
        if (dm >= D || (k && dm < desc->key_dim[k - 1]))
            return false;
        if (key_base[dm] == 0xFFFFFFFFu)
            key_base[dm] = k;
        if (++per_dim[dm] > kDenseKeysPerDim)
            return false;
        dim_post[dm] += desc->key_start[k + 1] - desc->key_start[k];
    }
    for (uint32_t dm = 0; dm < D; ++dm)
        if (dim_post[dm] != N)
            return false;
    return true;
}

// Forward codes for the dense form, from the uncompressed posting lists on
// the device (`post`, `key_start` as 32-bit positions), for an index that
// passes dense_shape and in which (checked on the device) every object
// appears exactly once per dim.  Any other index simply has no dense form;
// only device errors fail.
static int build_dense(const genie_index_desc* desc, DevIndex& d, const uint32_t* post, const uint32_t* key_start)
{
    const uint32_t D = desc->n_dims, N = desc->n_objects, K = desc->n_keys;
    std::vector<uint32_t> key_base, per_dim;
    if (!post || !dense_shape(desc, key_base, per_dim))
        return GENIE_OK;
    // slot of a key within its dim: rank by list length, longest first, so
    // the 32 most frequent keys of a dim sit in 32 different LDS banks
    std::vector<uint32_t> slot(K);
    for (uint32_t dm = 0; dm < D; ++dm) {
        std::vector<uint32_t> ks(per_dim[dm]);
        std::iota(ks.begin(), ks.end(), key_base[dm]);
        std::stable_sort(ks.begin(), ks.end(), [&](uint32_t a, uint32_t b) {
            return desc->key_start[a + 1] - desc->key_start[a] > desc->key_start[b + 1] - desc->key_start[b];
        });
        for (uint32_t r = 0; r < ks.size(); ++r)
            slot[ks[r]] = dm * kDenseKeysPerDim + r;
    }
    const uint32_t dpad = (D + 15) / 16 * 16, npad = (N + 63) / 64 * 64;
    const uint32_t nw = (uint32_t)dense_words((int)dpad);
    if ((uint64_t)nw * npad * 4 >= 0x80000000ull)
        return GENIE_OK;  // 32-bit buffer offsets
    std::vector<uint32_t> dim_word(dpad, 0), dim_shift(dpad, 0);
    for (uint32_t w = 0; w < nw; ++w)
        for (int f = 0; f < 3; ++f) {
            const int dd = dense_field_dim((int)dpad, (int)w, f);
            if (dd >= 0) {
                dim_word[dd] = w;
                dim_shift[dd] = (uint32_t)dense_field_shift(f);
            }
        }
    const uint32_t bm_words = (D + 31) / 32;
    uint32_t *codes = nullptr, *key_slot = nullptr, *bm = nullptr, *flag = nullptr, *dim = nullptr;
    uint32_t *d_word = nullptr, *d_shift = nullptr;
    uint64_t bytes = 0;
    int rc = GENIE_OK;
    uint32_t h_flag = 0;
    do {
        if ((rc = dev_alloc(&codes, (uint64_t)nw * npad, &bytes))) break;
        if ((rc = dev_alloc(&key_slot, K, &bytes))) break;
        if ((rc = dev_alloc(&bm, (uint64_t)N * bm_words))) break;
        if ((rc = dev_alloc(&flag, 1))) break;
        if ((rc = dev_alloc(&dim, K))) break;
        if ((rc = dev_alloc(&d_word, dpad))) break;
        if ((rc = dev_alloc(&d_shift, dpad))) break;
        hipError_t e;
        if ((e = hipMemset(codes, 0, (uint64_t)nw * npad * 4)) || (e = hipMemset(bm, 0, (uint64_t)N * bm_words * 4)) ||
            (e = hipMemset(flag, 0, 4)) ||
            (e = hipMemcpy(key_slot, slot.data(), K * 4ull, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(dim, desc->key_dim, K * 4ull, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d_word, dim_word.data(), dpad * 4ull, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d_shift, dim_shift.data(), dpad * 4ull, hipMemcpyHostToDevice))) {
            rc = fail(GENIE_E_RUNTIME, std::string("dense codes: ") + hipGetErrorString(e));
            break;
        }
        k_dense_codes<<<K, 256>>>(post, key_start, dim, key_slot, d_word, d_shift, npad, bm_words, codes, bm,
                                  flag);
        k_dense_holes<<<(N + 255) / 256, 256>>>(bm, N, D, bm_words, flag);
        if ((e = hipGetLastError()) || (e = hipMemcpy(&h_flag, flag, 4, hipMemcpyDeviceToHost))) {
            rc = fail(GENIE_E_RUNTIME, std::string("dense codes: ") + hipGetErrorString(e));
            break;
        }
    } while (0);
    dev_free(d_word);
    dev_free(d_shift);
    dev_free(bm);
    dev_free(flag);
    dev_free(dim);
    if (rc || h_flag) {
        dev_free(codes);
        dev_free(key_slot);
        return rc;
    }
    d.dense = true;
    d.dense_dpad = dpad;
    d.dense_npad = npad;
    d.codes = codes;
    d.key_slot = key_slot;
    d.bytes += bytes;
    return GENIE_OK;
}

// Forward lists for the set form: for a set-valued table whose postings are
// given and whose objects hold at most 255 occurrences, every object's keys
// (an object listed twice in one key's list appears twice).  Laid out for
// waves: within each tile of kDenseTile objects the objects are ordered by
// list length, longest first (ties by id; `obj_perm` maps a slot to its
// object), and each block of 64 slots stores its lists column-major -- key i
// of slot j, plus one, at blk_off[b] + 64 i + j, 0 past the end of a list --
// so that a wave reads 64 consecutive words per step and loops as often as
// its block's longest list.  Built on the host; any other index simply has no
// set form; only device errors fail.
static int build_sets(const genie_index_desc* desc, DevIndex& d)
{
    if (desc->single_value || !desc->postings || desc->n_postings == 0 || desc->n_keys >= (1u << 24))
        return GENIE_OK;
    const uint32_t N = desc->n_objects, K = desc->n_keys;
    std::vector<uint32_t> ptr(N + 1, 0);
    for (uint64_t i = 0; i < desc->n_postings; ++i)
        ptr[desc->postings[i] + 1]++;
    uint32_t mx = 0;
    for (uint32_t o = 0; o < N; ++o) {
        mx = std::max(mx, ptr[o + 1]);
        ptr[o + 1] += ptr[o];
    }
    if (mx > 255)
        return GENIE_OK;
    std::vector<uint32_t> keys(desc->n_postings), fill(ptr.begin(), ptr.end() - 1);
    for (uint32_t k = 0; k < K; ++k)
        for (uint64_t i = desc->key_start[k]; i < desc->key_start[k + 1]; ++i)
            keys[fill[desc->postings[i]]++] = k;
    // slot order: per tile, by length descending, then id
    const uint32_t NB = (N + 63) / 64, NS = NB * 64;
    std::vector<uint32_t> perm(NS, kSetEmpty);
    for (uint32_t t0 = 0; t0 < N; t0 += kDenseTile) {
        const uint32_t t1 = std::min(N, t0 + kDenseTile);
        for (uint32_t o = t0; o < t1; ++o)
            perm[o] = o;
        std::stable_sort(perm.begin() + t0, perm.begin() + t1, [&](uint32_t x, uint32_t y) {
            return ptr[x + 1] - ptr[x] > ptr[y + 1] - ptr[y];
        });
    }
    std::vector<uint32_t> boff(NB + 1, 0);
    for (uint32_t b = 0; b < NB; ++b) {
        uint32_t len = 0;
        for (uint32_t j = 0; j < 64; ++j) {
            const uint32_t o = perm[b * 64 + j];
            if (o != kSetEmpty)
                len = std::max(len, ptr[o + 1] - ptr[o]);
        }
        if ((uint64_t)boff[b] + 64ull * len >= 0xFFFFFFFFull)
            return GENIE_OK;  // too large for 32-bit offsets: no set form
        boff[b + 1] = boff[b] + 64 * len;
    }
    // keys + 1, so that 0 pads the shorter lists of a block (and the kernel
    // reads 0 past a block's last column)
    std::vector<uint32_t> bkey(std::max<uint32_t>(1, boff[NB]), 0u);
    for (uint32_t b = 0; b < NB; ++b)
        for (uint32_t j = 0; j < 64; ++j) {
            const uint32_t o = perm[b * 64 + j];
            if (o == kSetEmpty)
                continue;
            for (uint32_t i = 0; i < ptr[o + 1] - ptr[o]; ++i)
                bkey[boff[b] + 64 * i + j] = keys[ptr[o] + i] + 1u;
        }
    uint32_t *d_perm = nullptr, *d_boff = nullptr, *d_bkey = nullptr;
    uint64_t bytes = 0;
    int rc = dev_alloc(&d_perm, NS, &bytes);
    if (!rc)
        rc = dev_alloc(&d_boff, NB + 1ull, &bytes);
    if (!rc)
        rc = dev_alloc(&d_bkey, bkey.size(), &bytes);
    hipError_t e = hipSuccess;
    if (!rc && ((e = hipMemcpy(d_perm, perm.data(), NS * 4ull, hipMemcpyHostToDevice)) ||
                (e = hipMemcpy(d_boff, boff.data(), (NB + 1ull) * 4, hipMemcpyHostToDevice)) ||
                (e = hipMemcpy(d_bkey, bkey.data(), bkey.size() * 4, hipMemcpyHostToDevice))))
        rc = fail(GENIE_E_RUNTIME, std::string("forward lists: ") + hipGetErrorString(e));
    if (rc) {
        dev_free(d_perm);
        dev_free(d_boff);
        dev_free(d_bkey);
        return rc;
    }
    d.sets = true;
    d.set_max_len = mx;
    d.set_slots = NS;
    d.obj_perm = d_perm;
    d.blk_off = d_boff;
    d.blk_key = d_bkey;
    d.bytes += bytes;
    return GENIE_OK;
}

static int static_scratch_init();

extern "C" {

const char* genie_hip_last_error(void) { return g_err.c_str(); }

int genie_hip_stream_synchronize(void* stream)
{
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return GENIE_OK;
}

int genie_hip_stream_create(void** stream)
{
    if (!stream)
        return fail(GENIE_E_INVALID, "null argument");
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return GENIE_OK;
}

int genie_hip_stream_destroy(void* stream)
{
    if (stream)
        HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return GENIE_OK;
}

int genie_hip_mem_info(uint64_t* free_bytes, uint64_t* total_bytes)
{
    if (!free_bytes || !total_bytes)
        return fail(GENIE_E_INVALID, "null argument");
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return GENIE_OK;
}

int genie_hip_copy_probe(uint64_t bytes)
{
    uint8_t *a = nullptr, *b = nullptr;
    if (int rc = dev_alloc(&a, bytes))
        return rc;
    if (int rc = dev_alloc(&b, bytes)) {
        dev_free(a);
        return rc;
    }
    hipError_t e = hipMemset(a, 1, bytes);
    if (e == hipSuccess)
        e = hipMemcpy(b, a, bytes, hipMemcpyDeviceToDevice);
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    dev_free(a);
    dev_free(b);
    if (e != hipSuccess)
        return fail(GENIE_E_RUNTIME, std::string("copy probe: ") + hipGetErrorString(e));
    return GENIE_OK;
}

int genie_hip_device_synchronize(void)
{
    HIP_TRY(hipDeviceSynchronize());
    return GENIE_OK;
}

int genie_hip_init(int device)
{
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (n <= 0)
        return fail(GENIE_E_RUNTIME, "no HIP device");
    if (device < 0 || device >= n)
        device = 0;  // Init falls back to device 0 (init.cu:21-28)
    HIP_TRY(hipSetDevice(device));
    // set-up of the device's static scratch (genie_hip_match_topk_static)
    return static_scratch_init();
}

int genie_hip_index_create(const genie_index_desc* desc, genie_index** out)
{
    if (!desc || !out)
        return fail(GENIE_E_INVALID, "null argument");
    *out = nullptr;
    if (desc->n_keys == 0 || desc->n_objects == 0 || !desc->key_start || !desc->key_dim)
        return fail(GENIE_E_INVALID, "empty index");
    if (desc->codec != 0) {
        CodecShape cs;
        if (!codec_shape(desc->codec, &cs))
            return fail(GENIE_E_INVALID, "unknown codec " + std::to_string(desc->codec));
        if (!desc->words || !desc->seg_cpos || !desc->seg_upos || desc->n_segments == 0)
            return fail(GENIE_E_INVALID, "compressed index needs words, seg_cpos and seg_upos");
        if (!desc->postings && !desc->single_value)
            return fail(GENIE_E_INVALID, "compressed set-valued index needs the uncompressed postings too");
    } else if (!desc->postings && desc->n_postings) {
        return fail(GENIE_E_INVALID, "missing postings");
    }
    if (desc->n_postings >= 0xFFFFFFFFull)
        return fail(GENIE_E_INVALID, "index too large: more than 2^32-1 postings");
    if (desc->n_objects >= 0xFFFFFFFFu)
        return fail(GENIE_E_INVALID, "too many objects");
    if (desc->key_start[0] != 0 || desc->key_start[desc->n_keys] != desc->n_postings)
        return fail(GENIE_E_INVALID, "key_start must run from 0 to n_postings");
    for (uint32_t k = 0; k < desc->n_keys; ++k)
        if (desc->key_start[k] > desc->key_start[k + 1])
            return fail(GENIE_E_INVALID, "key_start must be non-decreasing");
    if (desc->postings)
        for (uint64_t p = 0; p < desc->n_postings; ++p)
            if (desc->postings[p] >= desc->n_objects)
                return fail(GENIE_E_INVALID, "posting id out of range at position " + std::to_string(p));
    if (desc->codec != 0)
        return create_compressed(desc, out);

    genie_index* ix = new genie_index();
    DevIndex& d = ix->d;
    d.n_objects = desc->n_objects;
    d.n_dims = desc->n_dims;
    d.n_keys = desc->n_keys;
    d.n_postings = desc->n_postings;
    d.codec = desc->codec;
    d.single_value = desc->single_value;
    d.n_dtiles = (desc->n_objects + kDirTile - 1) / kDirTile;
    d.h_key_start.assign(desc->key_start, desc->key_start + desc->n_keys + 1);
    d.key_pos.build(d.h_key_start);
    d.h_key_dim.assign(desc->key_dim, desc->key_dim + desc->n_keys);
    d.h_key_mult.assign(desc->n_keys, 1);
    std::vector<uint32_t> ks32(desc->n_keys + 1), dir_off, dir_keys;
    for (uint32_t k = 0; k <= desc->n_keys; ++k)
        ks32[k] = (uint32_t)desc->key_start[k];
    for (uint32_t k = 0; k < desc->n_keys; ++k) {
        const uint64_t s = desc->key_start[k], e = desc->key_start[k + 1];
        uint32_t best = e > s ? 1 : 0, run = 1;
        for (uint64_t p = s + 1; p < e; ++p) {
            if (desc->postings[p] < desc->postings[p - 1]) {
                delete ix;
                return fail(GENIE_E_INVALID, "posting list of key " + std::to_string(k) + " is not sorted");
            }
            run = desc->postings[p] == desc->postings[p - 1] ? run + 1 : 1;
            best = std::max(best, run);
        }
        d.h_key_mult[k] = best;
    }
    // tile directory: rows for keys longer than kShortList, graded to at most
    // an eighth of the key's posting bytes
    const uint64_t dir_words = dir_layout(d.n_keys, d.n_dtiles, [&](uint32_t k) {
        const uint64_t len = desc->key_start[k + 1] - desc->key_start[k];
        return len > kShortList ? dir_grade(d.n_dtiles, len / kDirPostingsPerEntry) : -1;
    }, dir_off, dir_keys);
    if (!dir_words) {
        delete ix;
        return fail(GENIE_E_INVALID, "tile directory too large");
    }
    int rc = GENIE_OK;
    uint32_t* d_dir_keys = nullptr;
    do {
        if ((rc = dev_alloc(&d.post, d.n_postings, &d.bytes))) break;
        if ((rc = dev_alloc(&d.key_start, d.n_keys + 1, &d.bytes))) break;
        if ((rc = dev_alloc(&d.dir_off, d.n_keys, &d.bytes))) break;
        if ((rc = dev_alloc(&d.dir, dir_words, &d.bytes))) break;
        if ((rc = dev_alloc(&d_dir_keys, dir_keys.size()))) break;
        hipError_t e;
        if ((e = hipMemcpy(d.post, desc->postings, d.n_postings * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d.key_start, ks32.data(), ks32.size() * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d.dir_off, dir_off.data(), dir_off.size() * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d_dir_keys, dir_keys.data(), dir_keys.size() * 4, hipMemcpyHostToDevice))) {
            rc = fail(GENIE_E_RUNTIME, std::string("index upload: ") + hipGetErrorString(e));
            break;
        }
        if (!dir_keys.empty()) {
            const uint32_t blocks = (uint32_t)std::min<uint64_t>((dir_keys.size() + 3) / 4, 65536);
            k_build_dir<<<blocks, 256>>>(d.post, d.key_start, d_dir_keys, (uint32_t)dir_keys.size(), d.dir_off,
                                         d.n_dtiles, d.dir);
            if ((e = hipGetLastError()) || (e = hipDeviceSynchronize())) {
                rc = fail(GENIE_E_RUNTIME, std::string("directory build: ") + hipGetErrorString(e));
                break;
            }
        }
    } while (0);
    dev_free(d_dir_keys);
    if (!rc)
        rc = build_dense(desc, d, d.post, d.key_start);
    if (!rc)
        rc = build_sets(desc, d);
    if (rc) {
        genie_hip_index_destroy(ix);
        return rc;
    }
    *out = ix;
    return GENIE_OK;
}

int genie_hip_index_destroy(genie_index* ix)
{
    if (!ix)
        return GENIE_OK;
    DevIndex& d = ix->d;
    dev_free(d.post);
    dev_free(d.key_start);
    dev_free(d.dir_off);
    dev_free(d.dir);
    dev_free(d.words);
    dev_free(d.seg_cpos);
    dev_free(d.seg_upos);
    dev_free(d.seg_key);
    dev_free(d.seg_first);
    dev_free(d.seg_last);
    dev_free(d.seg_bph);
    dev_free(d.key_seg);
    dev_free(d.sdir_off);
    dev_free(d.sdir);
    dev_free(d.codes);
    dev_free(d.key_slot);
    dev_free(d.obj_perm);
    dev_free(d.blk_off);
    dev_free(d.blk_key);
    delete ix;
    return GENIE_OK;
}

uint64_t genie_hip_index_bytes(const genie_index* ix) { return ix ? ix->d.bytes : 0; }

int genie_hip_count_bounds(const genie_index* ix, const genie_dim_t* items, uint64_t n_items,
                           uint32_t n_queries, uint32_t* out_bounds)
{
    if (!ix || (!items && n_items) || !out_bounds)
        return fail(GENIE_E_INVALID, "null argument");
    int rc = validate_items(ix->d, items, n_items, n_queries);
    if (rc)
        return rc;
    return compute_bounds(ix->d, items, n_items, n_queries, out_bounds);
}

int genie_hip_plan_create(genie_index* ix, uint32_t max_queries, uint64_t max_items, uint32_t max_k,
                          uint32_t count_bound, int mode, genie_plan** out)
{
    if (!ix || !out)
        return fail(GENIE_E_INVALID, "null argument");
    if (max_queries == 0 || max_k == 0 || max_k > GENIE_MAX_K)
        return fail(GENIE_E_INVALID, "max_queries must be > 0 and max_k in [1, GENIE_MAX_K]");
    if (mode != 0 && mode != 1 && mode != 2)
        return fail(GENIE_E_INVALID, "unknown plan mode");
    genie_plan* p = new genie_plan();
    p->ix = ix;
    p->mode = mode;
    p->max_queries = max_queries;
    p->max_items = max_items;
    p->max_k = max_k;
    p->count_bound = count_bound;
    p->lane_bits = bits_for(count_bound);
    int rc = GENIE_OK;
    uint64_t nb = (max_items + 1023) / 1024 + 1;
    do {
        if ((rc = dev_alloc(&p->d_items, max_items))) break;
        if ((rc = dev_alloc(&p->d_item_qptr, max_queries + 1))) break;
        if ((rc = dev_alloc(&p->d_nchunks, max_items))) break;
        if ((rc = dev_alloc(&p->d_chunk_ptr, max_items))) break;
        if ((rc = dev_alloc(&p->d_block_sums, nb))) break;
        if ((rc = dev_alloc(&p->d_total, 1))) break;
        if ((rc = dev_alloc(&p->d_qc_ptr, max_queries + 1))) break;
        if ((rc = dev_alloc(&p->d_qk_ptr, max_queries + 1))) break;
        if ((rc = dev_alloc(&p->d_total_pairs, 1))) break;
        if ((rc = dev_alloc(&p->d_kq, max_queries))) break;
        if ((rc = dev_alloc(&p->d_gthr, max_queries))) break;
        if ((rc = dev_alloc(&p->d_gkey, max_queries))) break;
        if ((rc = dev_alloc(&p->d_cand_cnt, max_queries))) break;
        if ((rc = dev_alloc(&p->d_cand_cap, max_queries))) break;
        if ((rc = dev_alloc(&p->d_cand_base, max_queries))) break;
        if ((rc = dev_alloc(&p->d_overflow, 1))) break;
        if ((rc = dev_alloc(&p->d_flag_any, 1))) break;
        if ((rc = dev_alloc(&p->d_out_off, max_queries + 1))) break;
        if ((rc = dev_alloc(&p->d_queue, 1))) break;
        if ((rc = dev_alloc(&p->d_hq, max_queries))) break;
        if ((rc = dev_alloc(&p->d_dq, max_queries))) break;
        if ((rc = dev_alloc(&p->d_hov, max_queries))) break;
        if ((rc = dev_alloc(&p->d_hov_any, 1))) break;
        if (hipStreamCreateWithFlags(&p->own_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_up, hipEventDisableTiming) != hipSuccess) {
            rc = fail(GENIE_E_RUNTIME, "hipStreamCreate");
            break;
        }
        for (auto& e : p->ev)
            if (hipEventCreate(&e) != hipSuccess) {
                rc = fail(GENIE_E_RUNTIME, "hipEventCreate");
                break;
            }
    } while (0);
    if (rc) {
        genie_hip_plan_destroy(p);
        return rc;
    }
    *out = p;
    return GENIE_OK;
}

int genie_hip_plan_estimate(const genie_index* ix, uint32_t n_queries, uint64_t n_items, uint32_t max_k, int mode,
                            uint64_t* bytes)
{
    if (!ix || !bytes)
        return fail(GENIE_E_INVALID, "null argument");
    if (mode != 0 && mode != 1 && mode != 2)
        return fail(GENIE_E_INVALID, "unknown plan mode");
    const DevIndex& d = ix->d;
    const uint64_t nq = n_queries, ni = n_items;
    uint64_t b = 1ull << 20;                      // small fixed buffers, events, rounding
    b += ni * (sizeof(genie_dim_t) + 4 + 4 + 8);  // items, chunk counts and offsets, dense key ranges
    b += ni * (sizeof(CChunk) + 8);               // compiled chunks (one per item; spans past kChunk add more)
    b += nq * 96;                                 // per-query thresholds, counters, offsets
    const bool dense = (d.dense || d.sets) && mode != 0;
    b += nq * initial_candidates(max_k, dense, d.n_objects) * 8;  // candidate keys
    b += nq * (uint64_t)max_k * 8;                                // output rows
    if (dense)  // query-mask tables per group (8 queries at the fewest) and the sample histograms
        b += (nq + 7) / 8 * (uint64_t)std::max<uint32_t>(d.dense_dpad, 1) * kDenseKeysPerDim * 4 +
             nq * kDenseBins * 4ull;
    if (d.codec != 0 && mode != 1)  // full-layout decode scratch of the plain form
        b += d.n_postings * 4;
    *bytes = b;
    return GENIE_OK;
}

int genie_hip_plan_destroy(genie_plan* p)
{
    if (!p)
        return GENIE_OK;
    if (p->up_pending)
        (void)hipEventSynchronize(p->ev_up);  // issued copies land before their buffers go
    plan_free_buffers(p);
    for (auto& e : p->ev)
        if (e)
            (void)hipEventDestroy(e);
    if (p->ev_up)
        (void)hipEventDestroy(p->ev_up);
    if (p->own_stream)
        (void)hipStreamDestroy(p->own_stream);
    if (p->h_stage)
        (void)hipHostFree(p->h_stage);
    delete p;
    return GENIE_OK;
}

}  // extern "C"

// ---- uploads ---------------------------------------------------------------

// Opens the staging of an upload: the copies of the previous one (issued by a
// run) have left the pinned buffer before it is written again.
static int stage_begin(genie_plan* p)
{
    if (p->up_pending) {
        HIP_TRY(hipEventSynchronize(p->ev_up));
        p->up_pending = false;
    }
    p->stage_used = 0;
    p->staged.clear();
    p->staging = true;
    return GENIE_OK;
}

// Issues the staged copies on stream s, in order (the next run on s reads
// them); ev_up marks their end.
static int stage_issue(genie_plan* p, hipStream_t s)
{
    if (p->staged.empty())
        return GENIE_OK;
    for (const genie_plan::Staged& c : p->staged)
        HIP_TRY(hipMemcpyAsync(*c.dst, p->h_stage + c.off, c.bytes, hipMemcpyHostToDevice, s));
    p->staged.clear();
    HIP_TRY(hipEventRecord(p->ev_up, s));
    p->up_pending = true;
    return GENIE_OK;
}

// Host-to-device copy of `bytes` from `src` to the buffer *dst: staged while
// an upload is open (src may go away at once; the copies go out with the next
// run, on its stream), else synchronous after the staged ones.
template <class T>
static int h2d(genie_plan* p, T** dst, const void* src, uint64_t bytes)
{
    if (bytes == 0)
        return GENIE_OK;
    if (!p->staging) {
        if (!p->staged.empty()) {
            if (int rc = stage_issue(p, p->own_stream))
                return rc;
        }
        if (p->up_pending) {
            HIP_TRY(hipEventSynchronize(p->ev_up));
            p->up_pending = false;
        }
        HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return GENIE_OK;
    }
    const uint64_t off = (p->stage_used + 255) & ~255ull;
    if (off + bytes > p->stage_cap) {
        // first time: room for a batch of max_items items (their 20 B, key
        // ranges and per-query words), so that a serving loop seldom grows it
        // (freeing pinned memory waits for the whole device)
        const uint64_t est = p->stage_cap ? 0 : std::min<uint64_t>(64ull << 20, p->max_items * 40 + p->max_queries * 64ull);
        const uint64_t cap = std::max<uint64_t>({off + bytes, 2 * p->stage_cap, 1ull << 20, est});
        uint8_t* nb = nullptr;
        if (hipHostMalloc((void**)&nb, cap, hipHostMallocDefault) != hipSuccess)
            return fail(GENIE_E_OOM, "pinned upload buffer of " + std::to_string(cap) + " bytes");
        if (p->stage_used)
            std::memcpy(nb, p->h_stage, p->stage_used);
        if (p->h_stage)
            (void)hipHostFree(p->h_stage);
        p->h_stage = nb;
        p->stage_cap = cap;
    }
    std::memcpy(p->h_stage + off, src, bytes);
    p->staged.push_back({reinterpret_cast<void**>(dst), off, bytes});
    p->stage_used = off + bytes;
    return GENIE_OK;
}

// Mode 1 (batched) and mode 2 (automatic) run the dense form when the batch
// allows it; otherwise mode 1 runs the lane form and mode 2 the plain form.
static bool dense_active(const genie_plan* p) { return p->mode != 0 && p->dense_on && (p->dense_ok || p->sets_ok); }

// Queries of one group of the dense or set form: the bits of a mask word
// over the weight planes per query.
static uint32_t dense_qpg(const genie_plan* p) { return p->dense_ok ? 32u / p->dense_wp : 32u; }

// Whether the plan's batches that the dense and set forms do not take run the
// lane form (batched plans, unless the autotuner or the caller chose the
// plain form for them); otherwise they run the plain form.
static bool lane_form(const genie_plan* p) { return p->mode == 1 && !p->batched_plain; }

static uint32_t plan_tile_objs(const genie_plan* p)
{
    if (dense_active(p))
        return kDenseTile;
    return lane_form(p) ? p->tile_b : tile_objs_for(p->lane_bits);
}

static uint32_t plan_n_tiles(const genie_plan* p)
{
    const uint32_t t = plan_tile_objs(p);
    return (p->ix->d.n_objects + t - 1) / t;
}

// Largest useful cap_factor: the tile forms emit at most k per tile, the
// dense form at most every object once per query.
static uint32_t max_cap_factor(const genie_plan* p)
{
    if (!dense_active(p))
        return std::max(plan_n_tiles(p), lane_form(p) ? 0u : p->max_hash_tiles);
    uint32_t kmin = 0xFFFFFFFFu;
    for (uint32_t q = 0; q < p->n_queries; ++q)
        kmin = std::min(kmin, p->h_kq[q]);
    return (p->ix->d.n_objects + kmin - 1) / kmin;
}

// The tile forms start with the most the batch can emit (k per tile and
// query): on a large index a batch of a few queries runs every tile at once,
// and each emits its k before any threshold is known, so a smaller space
// overflowed on almost every batch and ran it again.
static uint32_t initial_cap_factor(const genie_plan* p)
{
    return dense_active(p) ? kDenseCapFactor : max_cap_factor(p);
}

// Smallest dense_mult at which every device-side ask is the whole index: an
// ask is at least 1,024 x dense_mult before it is capped at n_objects
// (dense_want).
static uint32_t dense_mult_max(const genie_plan* p) { return (p->ix->d.n_objects + 1023) / 1024; }

// Candidate capacity per query: cap_factor x k, at most what the form can emit.
static int plan_size_candidates(genie_plan* p)
{
    const bool dn = dense_active(p);
    const uint32_t n_tiles = plan_n_tiles(p);
    const uint32_t f = std::min<uint32_t>(p->cap_factor, max_cap_factor(p));
    std::vector<uint32_t> cap(p->n_queries);
    std::vector<uint64_t> base(p->n_queries);
    uint64_t tot = 0;
    const bool plain = !dn && !lane_form(p);
    for (uint32_t q = 0; q < p->n_queries; ++q) {
        // the dense form emits each object at most once; the tile forms at
        // most k per tile (a hash-form query: per tile of its own)
        uint64_t tiles = n_tiles;
        if (plain && q < p->h_hq.size() && p->h_hq[q].x) {
            const uint64_t tobj = (uint64_t)p->h_hq[q].x * kDirTile;
            tiles = (p->ix->d.n_objects + tobj - 1) / tobj;
        }
        const uint64_t limit = dn ? (uint64_t)p->ix->d.n_objects : tiles * p->h_kq[q];
        // the dense form's device sizing lays each query's ask out in this
        // total; a query whose counts crowd at its threshold asks for its
        // ties, so small k gets a floor of kDenseCapFloor (memory only)
        const uint64_t want = (uint64_t)f * p->h_kq[q];
        cap[q] = (uint32_t)std::min<uint64_t>(dn ? std::max<uint64_t>(want, kDenseCapFloor) : want, limit);
        base[q] = tot;
        tot += cap[q];
    }
    // the largest batch: max_queries queries of max_k, sized as above
    const uint64_t kw = p->max_k;
    const uint64_t qworst = std::min<uint64_t>(dn ? std::max<uint64_t>((uint64_t)f * kw, kDenseCapFloor) : (uint64_t)f * kw,
                                               dn ? (uint64_t)p->ix->d.n_objects : (uint64_t)n_tiles * kw);
    if (int rc = grow_buf(p, &p->d_cand, &p->cand_alloc, tot, qworst * p->max_queries))
        return rc;
    p->total_cap = tot;
    // the dense and set forms lay every query's space out in this total on
    // the device (k_cap_prefix, each run, before anything reads it); a switch
    // to another form sizes the candidates again and uploads its own
    if (dn)
        return GENIE_OK;
    if (int rc = h2d(p, &p->d_cand_cap, cap.data(), cap.size() * 4))
        return rc;
    return h2d(p, &p->d_cand_base, base.data(), base.size() * 8);
}

// Whether the uploaded batch (kept on the host by mode-1 plans) can run the
// dense form: every item with postings and a nonzero weight has an integer
// weight of at most 15 (all 1: unit masks, 32 queries a group; else as many
// weight planes a query as the largest weight has bits, WP = 2-4, and 32 / WP
// queries a group), a query's items never overlap,
// and each run of a query's items that follow one another without a gap
// covers whole keys (a run is one item of a plain batch, the pieces of a
// load-balanced range, or the segments of a range of a compressed index).
// The key range of every run is kept for the device-side mask build on the
// run's first item (h_dkeys: first key | the query's bit position << 24, end
// key | weight << 16; empty for every other item), and for a weighted batch
// every query's sample histogram shift (n_dims x its largest weight >> shift
// < 256).
static bool dense_batch_ok(genie_plan* p, uint32_t n_queries)
{
    const DevIndex& d = p->ix->d;
    if (p->mode == 0 || !d.dense)
        return false;
    // items of a compressed index span words: keys start at their first
    // segment's first word
    const uint64_t* ks = d.codec == 0 ? d.h_key_start.data() : d.h_key_wstart.data();
    const uint32_t nk = d.n_keys;
    uint32_t wmax = 1;
    for (const genie_dim_t& it : p->h_items) {
        if (it.end_pos <= it.start_pos || it.weight == 0.0f)
            continue;
        if (!(it.weight >= 1.0f && it.weight <= 15.0f) || it.weight != (float)(uint32_t)it.weight)
            return false;
        wmax = std::max(wmax, (uint32_t)it.weight);
    }
    const uint32_t wp = wmax == 1 ? 1u : (wmax <= 3 ? 2u : (wmax <= 7 ? 3u : 4u)), qpg = 32u / wp;
    p->dense_wp = wp;
    p->h_dkeys.assign(2 * p->h_items.size(), 0);
    if (wp > 1)
        p->h_sshift.assign(n_queries, 0);
    struct Span {
        uint64_t sp, ep;
        uint32_t item;
    };
    std::vector<Span> iv;
    for (uint32_t q = 0; q < n_queries; ++q) {
        iv.clear();
        uint32_t qw = 1;
        for (uint32_t i = p->h_qptr[q]; i < p->h_qptr[q + 1]; ++i) {
            const genie_dim_t& it = p->h_items[i];
            if (it.end_pos <= it.start_pos || it.weight == 0.0f)
                continue;
            if (wp == 1 && it.weight != 1.0f)
                return false;
            qw = std::max(qw, (uint32_t)it.weight);
            iv.push_back({(uint64_t)it.start_pos, (uint64_t)it.end_pos, i});
        }
        // compiled batches list a query's items in position order already
        bool ordered = true;
        for (size_t i = 1; i < iv.size() && ordered; ++i)
            ordered = iv[i - 1].sp <= iv[i].sp;
        if (!ordered)
            std::sort(iv.begin(), iv.end(), [](const Span& x, const Span& y) { return x.sp < y.sp; });
        for (size_t r = 0; r < iv.size();) {
            size_t e = r + 1;
            uint64_t end = iv[r].ep;
            const uint32_t w = (uint32_t)p->h_items[iv[r].item].weight;
            // a run joins items of one weight (a weighted mask holds one
            // weight per key)
            while (e < iv.size() && iv[e].sp == end && (uint32_t)p->h_items[iv[e].item].weight == w)
                end = iv[e++].ep;
            if (e < iv.size() && iv[e].sp < end)
                return false;  // overlapping items
            const uint64_t sp = iv[r].sp;
            const uint32_t ka = d.key_pos.last_le(ks, sp);
            const uint32_t kb = d.key_pos.first_ge(ks, end);
            if (ks[ka] != sp || kb > nk || ks[kb] != end)
                return false;
            // keys < 2^15 in a dense index (<= 128 dims of <= 256 keys)
            p->h_dkeys[2 * iv[r].item] = ka | ((q % qpg) * wp) << 24;
            p->h_dkeys[2 * iv[r].item + 1] = kb | (wp > 1 ? w << 16 : 0u);
            r = e;
        }
        if (wp > 1) {  // counts <= n_dims x 15 < kSetNoQuery
            uint32_t sh = 0;
            while ((d.n_dims * qw) >> sh >= 256)
                ++sh;
            p->h_sshift[q] = sh;
        }
    }
    return true;
}

// Whether the uploaded batch can run the set form (set-valued tables with
// forward lists): per query, the sum of its items' weights over the
// positions (items may repeat a key, or split it into segments or chunks) is
// constant over every key it touches, i.e. a weight per key, at most
// kSetMaxWeight; the longest forward list times the query's largest weight is
// below kSetNoQuery (that threshold marks "no query"); a group of 32 queries
// holds at most kSetEntries keys.  Builds every group's entries (keys
// ascending, bit j of plane b = bit b of query j's weight for the key) and
// every query's sample histogram shift (its count bound >> shift < 256).
static bool sets_batch_ok(genie_plan* p, uint32_t n_queries)
{
    const DevIndex& d = p->ix->d;
    if (p->mode == 0 || !d.sets)
        return false;
    const uint64_t* ks = d.codec == 0 ? d.h_key_start.data() : d.h_key_wstart.data();
    struct Ev {
        uint64_t pos;
        int64_t dw;
    };
    struct KW {
        uint32_t key, q, w;
    };
    std::vector<Ev> ev;
    std::vector<KW> gk;  // the group's (key, query in group, weight)
    const uint32_t G = (n_queries + 31) / 32;
    p->h_skeys.clear();
    p->h_smask.clear();
    p->h_sgptr.assign(G + 1, 0);
    p->h_sshift.assign(n_queries, 0);
    for (uint32_t g = 0; g < G; ++g) {
        gk.clear();
        for (uint32_t q = g * 32; q < std::min(n_queries, g * 32 + 32); ++q) {
            // the query's weight as a step function of the position: items add
            // their weight over their span (a key named twice adds twice)
            ev.clear();
            for (uint32_t i = p->h_qptr[q]; i < p->h_qptr[q + 1]; ++i) {
                const genie_dim_t& it = p->h_items[i];
                if (it.end_pos <= it.start_pos || it.weight == 0.0f)
                    continue;
                if (it.weight > (float)kSetMaxWeight)
                    return false;
                ev.push_back({(uint64_t)it.start_pos, (int64_t)it.weight});
                ev.push_back({(uint64_t)it.end_pos, -(int64_t)it.weight});
            }
            std::sort(ev.begin(), ev.end(), [](const Ev& x, const Ev& y) { return x.pos < y.pos; });
            // every piece of constant nonzero weight covers whole keys
            uint32_t wmax = 0;
            const size_t first = gk.size();
            int64_t w = 0;
            uint64_t from = 0;
            for (size_t i = 0; i < ev.size();) {
                const uint64_t pos = ev[i].pos;
                int64_t nw = w;
                for (; i < ev.size() && ev[i].pos == pos; ++i)
                    nw += ev[i].dw;
                if (nw == w)
                    continue;
                if (w > 0) {
                    if (w > (int64_t)kSetMaxWeight)
                        return false;
                    const uint32_t ka = d.key_pos.last_le(ks, from);
                    const uint32_t kb = d.key_pos.first_ge(ks, pos);
                    if (ks[ka] != from || kb > d.n_keys || ks[kb] != pos)
                        return false;
                    for (uint32_t k = ka; k < kb; ++k)
                        if (ks[k + 1] > ks[k])  // keys with postings
                            gk.push_back({k, q - g * 32, (uint32_t)w});
                    if (gk.size() - first > kSetEntries)
                        return false;  // more keys than one group's table holds
                    wmax = std::max(wmax, (uint32_t)w);
                }
                w = nw;
                from = pos;
            }
            const uint64_t bound = (uint64_t)d.set_max_len * wmax;
            if (bound >= kSetNoQuery)
                return false;
            uint32_t sh = 0;
            while ((bound >> sh) > 255)
                ++sh;
            p->h_sshift[q] = sh;
        }
        std::sort(gk.begin(), gk.end(), [](const KW& x, const KW& y) { return x.key < y.key; });
        for (size_t i = 0; i < gk.size();) {
            uint32_t m[8] = {};
            const uint32_t key = gk[i].key;
            for (; i < gk.size() && gk[i].key == key; ++i)
                for (int b = 0; b < kSetWeightPlanes; ++b)
                    m[b] |= ((gk[i].w >> b) & 1u) << gk[i].q;
            p->h_skeys.push_back(key);
            p->h_smask.push_back(make_uint4(m[0], m[1], m[2], m[3]));
            p->h_smask.push_back(make_uint4(m[4], m[5], m[6], m[7]));
        }
        if (p->h_skeys.size() - p->h_sgptr[g] > kSetEntries)
            return false;
        p->h_sgptr[g + 1] = (uint32_t)p->h_skeys.size();
    }
    return true;
}

// Device buffers of the dense form for the uploaded batch: the query mask
// tables, the sample histogram and the sample tile list (evenly spread tiles,
// at least 16 and at least 1/16 of all tiles).
static int prepare_dense(genie_plan* p)
{
    const DevIndex& d = p->ix->d;
    p->dn_groups = (p->n_queries + dense_qpg(p) - 1) / dense_qpg(p);
    const uint64_t tw = (uint64_t)d.dense_dpad * kDenseKeysPerDim;
    if (int rc = grow_buf(p, &p->d_dtab, &p->dtab_alloc, p->dn_groups * tw))
        return rc;
    if (!p->d_dhist)
        if (int rc = dev_alloc(&p->d_dhist, (uint64_t)p->max_queries * kDenseBins))
            return rc;
    if (p->dense_ok) {
        if (!p->d_dkeys)
            if (int rc = dev_alloc(&p->d_dkeys, 2 * std::max<uint64_t>(1, p->max_items)))
                return rc;
        if (p->n_items)
            if (int rc = h2d(p, &p->d_dkeys, p->h_dkeys.data(), p->n_items * 8))
                return rc;
        if (p->dense_wp > 1) {  // weighted: per-query sample histogram shifts
            if (!p->d_sshift)
                if (int rc = dev_alloc(&p->d_sshift, p->max_queries))
                    return rc;
            if (int rc = h2d(p, &p->d_sshift, p->h_sshift.data(), p->h_sshift.size() * 4))
                return rc;
        }
    } else {  // set form: the groups' key entries
        const uint64_t ne = std::max<uint64_t>(1, p->h_skeys.size());
        if (int rc = grow_buf(p, &p->d_skeys, &p->sent_alloc, ne))
            return rc;
        if (int rc = grow_buf(p, &p->d_smask, &p->smask_alloc, 2 * ne))
            return rc;
        if (!p->d_sgptr)
            if (int rc = dev_alloc(&p->d_sgptr, (p->max_queries + 31) / 32 + 1))
                return rc;
        if (!p->d_sshift)
            if (int rc = dev_alloc(&p->d_sshift, p->max_queries))
                return rc;
        if (!p->h_skeys.empty()) {
            if (int rc = h2d(p, &p->d_skeys, p->h_skeys.data(), p->h_skeys.size() * 4))
                return rc;
            if (int rc = h2d(p, &p->d_smask, p->h_smask.data(), p->h_smask.size() * sizeof(uint4)))
                return rc;
        }
        if (int rc = h2d(p, &p->d_sgptr, p->h_sgptr.data(), p->h_sgptr.size() * 4))
            return rc;
        if (int rc = h2d(p, &p->d_sshift, p->h_sshift.data(), p->h_sshift.size() * 4))
            return rc;
    }
    if (!p->d_gsafe)
        if (int rc = dev_alloc(&p->d_gsafe, p->max_queries))
            return rc;
    if (!p->d_redo)  // groups of 8 queries at the fewest
        if (int rc = dev_alloc(&p->d_redo, (p->max_queries + 7) / 8))
            return rc;
    if (!p->d_want) {
        if (int rc = dev_alloc(&p->d_want, p->max_queries))
            return rc;
        if (int rc = dev_alloc(&p->d_want_safe, p->max_queries))
            return rc;
        if (int rc = dev_alloc(&p->d_cap_used, 1))
            return rc;
    }
    const uint32_t n_tiles = (d.n_objects + kDenseTile - 1) / kDenseTile;
    const uint32_t ns = std::min(n_tiles, std::max(8u, (n_tiles + 31) / 32));
    std::vector<uint32_t> list(ns);
    uint32_t n_sample = 0;
    for (uint32_t i = 0; i < ns; ++i) {
        list[i] = (uint32_t)((uint64_t)i * n_tiles / ns);
        n_sample += std::min(kDenseTile, d.n_objects - list[i] * kDenseTile);
    }
    p->n_dsample = n_sample;
    if (ns > p->dlist_alloc) {
        dev_free(p->d_dlist);
        p->dlist_alloc = 0;
        if (int rc = dev_alloc(&p->d_dlist, ns))
            return rc;
        p->dlist_alloc = ns;
    }
    if (int rc = h2d(p, &p->d_dlist, list.data(), ns * 4ull))
        return rc;
    p->n_dlist = ns;
    return GENIE_OK;
}

// The lane form's work queue is built when the lane form is first needed.
static int ensure_lanes(genie_plan* p)
{
    if (p->lanes_built)
        return GENIE_OK;
    if (int rc = build_batched(p))
        return rc;
    p->lanes_built = true;
    return GENIE_OK;
}

// Per-query figures of the plain form: count bounds (unless the plan has a
// fixed bound) and postings (from the segment offsets when compressed).
static void plain_query_stats(genie_plan* p)
{
    if (p->plain_stats)
        return;
    const DevIndex& d = p->ix->d;
    const uint32_t nq = p->n_queries;
    if (p->count_bound != 0) {
        p->h_qbound.assign(nq, p->count_bound);
    } else if (p->h_qbound.size() != nq || p->dense_ok) {  // upload skips a dense batch's bounds
        p->h_qbound.resize(nq);
        compute_bounds(d, p->h_items.data(), p->n_items, nq, p->h_qbound.data());
    }
    p->h_qpost.assign(nq, 0);
    const uint32_t* cp = d.h_seg_cpos.data();
    for (uint32_t q = 0; q < nq; ++q) {
        uint64_t P = 0;
        for (uint32_t i = p->h_qptr[q]; i < p->h_qptr[q + 1]; ++i) {
            const genie_dim_t& it = p->h_items[i];
            if (it.end_pos <= it.start_pos)
                continue;
            if (d.codec == 0) {
                P += (uint64_t)(it.end_pos - it.start_pos);
            } else {  // compressed spans start and end on segment boundaries
                const uint32_t s0 = (uint32_t)(std::lower_bound(cp, cp + d.n_segments + 1, (uint32_t)it.start_pos) - cp);
                const uint32_t s1 = (uint32_t)(std::lower_bound(cp, cp + d.n_segments + 1, (uint32_t)it.end_pos) - cp);
                P += d.h_seg_upos[s1] - d.h_seg_upos[s0];
            }
        }
        p->h_qpost[q] = P;
    }
    p->plain_stats = true;
}

// Compressed index, hash form: every item of a hash-form query becomes a
// span of the batch's scratch posting array (its segments' ids, in item
// order), decoded once per batch by k_decode_jobs; the hash kernel then
// counts them like uncompressed postings, keeping the ids of its tile.  A
// batch whose decoded postings pass 2^31 keeps the in-kernel decode.
static int plan_decoded_spans(genie_plan* p)
{
    const DevIndex& d = p->ix->d;
    p->dec_spans = false;
    p->h_djobs.clear();
    if (d.codec == 0 || p->h_hwork.empty())
        return GENIE_OK;
    const uint32_t nq = p->n_queries;
    const uint32_t* cp = d.h_seg_cpos.data();
    const uint32_t* up = d.h_seg_upos.data();
    p->h_sspan.clear();
    p->h_djobs.clear();
    p->h_sq_ptr.assign(nq + 1, 0);
    uint64_t off = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        p->h_sq_ptr[q] = (uint32_t)p->h_sspan.size();
        if (!p->h_hq[q].x)
            continue;
        for (uint32_t i = p->h_qptr[q]; i < p->h_qptr[q + 1]; ++i) {
            const genie_dim_t& it = p->h_items[i];
            if (it.end_pos <= it.start_pos || it.weight == 0.0f)
                continue;
            const uint32_t s0 = (uint32_t)(std::lower_bound(cp, cp + d.n_segments + 1, (uint32_t)it.start_pos) - cp);
            const uint32_t s1 = (uint32_t)(std::lower_bound(cp, cp + d.n_segments + 1, (uint32_t)it.end_pos) - cp);
            const uint64_t n = up[s1] - up[s0];
            if (off + n >= 0x80000000ull)
                return GENIE_OK;  // too large for one batch: in-kernel decode
            p->h_sspan.push_back(Span{(uint32_t)off, (uint32_t)(off + n), (uint32_t)it.weight << 1 | 1u});
            for (uint32_t sg = s0; sg < s1; ++sg)
                p->h_djobs.push_back(make_uint2(sg, (uint32_t)(off + up[sg] - up[s0])));
            off += n;
        }
    }
    p->h_sq_ptr[nq] = (uint32_t)p->h_sspan.size();
    auto grow = [](auto** ptr, uint64_t& alloc, uint64_t need) -> int {
        if (need <= alloc)
            return GENIE_OK;
        dev_free(*ptr);
        alloc = 0;
        if (int rc = dev_alloc(ptr, need))
            return rc;
        alloc = need;
        return GENIE_OK;
    };
    if (int rc = grow(&p->d_sspan, p->sspan_alloc, std::max<uint64_t>(1, p->h_sspan.size())))
        return rc;
    if (int rc = grow(&p->d_djobs, p->djobs_alloc, std::max<uint64_t>(1, p->h_djobs.size())))
        return rc;
    if (int rc = grow(&p->d_scratch, p->scratch_alloc, std::max<uint64_t>(1, off)))
        return rc;
    if (!p->d_sq_ptr)
        if (int rc = dev_alloc(&p->d_sq_ptr, p->max_queries + 1))
            return rc;
    if (!p->h_sspan.empty())
        if (int rc = h2d(p, &p->d_sspan, p->h_sspan.data(), p->h_sspan.size() * sizeof(Span)))
            return rc;
    if (!p->h_djobs.empty())
        if (int rc = h2d(p, &p->d_djobs, p->h_djobs.data(), p->h_djobs.size() * sizeof(uint2)))
            return rc;
    if (int rc = h2d(p, &p->d_sq_ptr, p->h_sq_ptr.data(), (nq + 1) * 4ull))
        return rc;
    p->dec_spans = true;
    return GENIE_OK;
}

// Compressed index, plain form: the segments any item of the batch names,
// each once, become decode jobs into a scratch array of the uncompressed
// layout, and every item is translated from word to posting positions (items
// start and end on segment boundaries).  The batch then runs the
// uncompressed kernels over the scratch: a segment shared by many queries and
// tiles is decoded once per batch instead of once per (query, tile).  Off
// (the in-kernel decode stays) when the scratch cannot be allocated.
static int plan_full_decode(genie_plan* p)
{
    const DevIndex& d = p->ix->d;
    p->dec_full = false;
    p->n_chunks = p->n_chunks_runs;
    if (d.codec == 0 || p->decode_mode == 0 || !d.key_start || p->n_items == 0)
        return GENIE_OK;
    if (!p->d_full) {
        if (dev_alloc(&p->d_full, std::max<uint64_t>(1, d.n_postings)) != GENIE_OK)
            return GENIE_OK;  // no room for the scratch: in-kernel decode
    }
    if (!p->d_uitems)
        if (int rc = dev_alloc(&p->d_uitems, std::max<uint64_t>(1, p->max_items)))
            return rc;
    const uint64_t* cp = d.h_seg_cpos64.data();
    const uint32_t* up = d.h_seg_upos.data();
    std::vector<uint2> runs;
    runs.reserve(p->n_items);
    p->h_uitems.resize(p->n_items);
    for (uint64_t i = 0; i < p->n_items; ++i) {
        genie_dim_t u = p->h_items[i];
        if (u.end_pos > u.start_pos) {
            const uint32_t s0 = d.seg_pos.first_ge(cp, (uint64_t)u.start_pos);
            const uint32_t s1 = d.seg_pos.first_ge(cp, (uint64_t)u.end_pos);
            u.start_pos = (int32_t)up[s0];
            u.end_pos = (int32_t)up[s1];
            runs.push_back(make_uint2(s0, s1));
        } else {
            u.start_pos = u.end_pos = 0;
        }
        p->h_uitems[i] = u;
    }
    std::sort(runs.begin(), runs.end(), [](const uint2& a, const uint2& b) { return a.x < b.x; });
    p->h_fjobs.clear();
    uint32_t next = 0;  // runs are sorted by start: segments below `next` are jobs already
    for (const uint2& r : runs) {
        for (uint32_t sg = std::max(r.x, next); sg < r.y; ++sg)
            p->h_fjobs.push_back(make_uint2(sg, up[sg]));
        next = std::max(next, r.y);
    }
    if (p->h_fjobs.size() > p->fjobs_alloc) {
        dev_free(p->d_fjobs);
        p->fjobs_alloc = 0;
        if (int rc = dev_alloc(&p->d_fjobs, p->h_fjobs.size()))
            return rc;
        p->fjobs_alloc = p->h_fjobs.size();
    }
    if (!p->h_fjobs.empty())
        if (int rc = h2d(p, &p->d_fjobs, p->h_fjobs.data(), p->h_fjobs.size() * sizeof(uint2)))
            return rc;
    if (int rc = h2d(p, &p->d_uitems, p->h_uitems.data(), p->n_items * sizeof(genie_dim_t)))
        return rc;
    // the run compiles the translated items like an uncompressed batch; the
    // chunk buffers were sized for it at upload
    if (p->n_chunks_full > p->chunks_cap)
        return fail(GENIE_E_RUNTIME, "chunk buffers smaller than the decoded batch needs");
    p->n_chunks = p->n_chunks_full;
    p->dec_full = true;
    return GENIE_OK;
}

// The plain form's split of the batch: a query with at most N / 16 postings
// (every query whose counts fit, when the hash form is forced) runs the hash
// kernel over tiles of whole directory tiles holding about kHashTarget of its
// postings each, with 2x that many table slots; the others, and queries
// whose table filled up in an earlier run, run the direct-mapped kernels.
// Key bits: key = tile-local id + 1, so tiles shrink until the counts fit in
// the remaining bits.
static int plan_split_plain(genie_plan* p)
{
    plain_query_stats(p);
    const DevIndex& d = p->ix->d;
    const uint32_t nq = p->n_queries;
    const uint64_t N = d.n_objects;
    p->h_hwork.clear();
    p->h_dq.clear();
    p->h_hq.assign(nq, make_uint4(0, 0, 0, 0));
    p->max_hash_tiles = 0;
    p->est_postings = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        const uint64_t P = p->h_qpost[q];
        bool hashed = false;
        if (p->hash_mode != 0 && !p->h_force_direct[q] && (p->hash_mode == 1 || P * 16 <= N)) {
            const uint64_t nt0 = std::max<uint64_t>(1, (P + kHashTarget - 1) / kHashTarget);
            uint64_t dtpt = std::max<uint64_t>(1, ((N + nt0 - 1) / nt0 + kDirTile - 1) / kDirTile);
            // tiles start on multiples of 16 directory tiles (of a power of two
            // below that), so that directory rows up to grade 4 clip the spans
            // exactly
            if (dtpt >= 16) {
                dtpt = (dtpt + 15) / 16 * 16;
            } else {
                uint64_t pow2 = 1;
                while (pow2 < dtpt)
                    pow2 <<= 1;
                dtpt = pow2;
            }
            uint32_t kb = 0;
            for (;;) {
                const uint64_t tobj = dtpt * kDirTile;
                kb = 64u - (uint32_t)__builtin_clzll(tobj);  // bits of tobj (keys are 1..tobj)
                if (kb < 32 && p->h_qbound[q] < (1u << (32 - kb)))
                    break;
                if (dtpt == 1) {
                    dtpt = 0;
                    break;
                }
                dtpt = (dtpt + 1) / 2;
            }
            if (dtpt) {
                const uint64_t tobj = dtpt * kDirTile;
                const uint64_t nt = (N + tobj - 1) / tobj;
                const uint64_t per = (P + nt - 1) / nt;
                uint32_t lhs = 9;
                while (lhs < (uint32_t)kHashLogMax && (1ull << lhs) < 2 * per)
                    ++lhs;
                p->h_hq[q] = make_uint4((uint32_t)dtpt, kb, lhs, (uint32_t)nt);
                p->max_hash_tiles = std::max(p->max_hash_tiles, (uint32_t)nt);
                hashed = true;
            }
        }
        if (!hashed) {
            p->h_dq.push_back(q);
            p->est_postings += P;
        }
    }
    // the hash work tile-major (every query's first tile, then every second
    // one, ...), so that a query's later tiles mostly start after its first
    // has published its k-th key, and select from fewer eligible slots
    for (uint32_t t = 0; t < p->max_hash_tiles; ++t)
        for (uint32_t q = 0; q < nq; ++q)
            if (t < p->h_hq[q].w) {
                p->h_hwork.push_back(q);
                p->h_hwork.push_back(t);
            }
    if (int rc = plan_full_decode(p))
        return rc;
    if (p->dec_full)
        p->dec_spans = false;
    else if (int rc = plan_decoded_spans(p))
        return rc;
    if (p->h_hwork.size() > p->hwork_alloc) {
        dev_free(p->d_hwork);
        p->hwork_alloc = 0;
        if (int rc = dev_alloc(&p->d_hwork, p->h_hwork.size()))
            return rc;
        p->hwork_alloc = p->h_hwork.size();
    }
    if (!p->h_hwork.empty())
        if (int rc = h2d(p, &p->d_hwork, p->h_hwork.data(), p->h_hwork.size() * 4))
            return rc;
    if (int rc = h2d(p, &p->d_hq, p->h_hq.data(), nq * sizeof(uint4)))
        return rc;
    if (!p->h_dq.empty())
        if (int rc = h2d(p, &p->d_dq, p->h_dq.data(), p->h_dq.size() * 4))
            return rc;
    return GENIE_OK;
}

// After a change of form or configuration: the active form's buffers and a
// fresh candidate space.
static int plan_settle_form(genie_plan* p)
{
    plan_drop_graph(p);
    if (p->mode == 0 || p->n_queries == 0)
        return GENIE_OK;
    if (dense_active(p)) {
        if (int rc = prepare_dense(p))
            return rc;
    } else if (lane_form(p)) {
        if (int rc = ensure_lanes(p))
            return rc;
    } else {
        if (int rc = plan_split_plain(p))
            return rc;
    }
    p->cap_factor = initial_cap_factor(p);
    p->dense_mult = 1;
    return plan_size_candidates(p);
}

static int plan_upload(genie_plan* p, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                       const uint32_t* k_per_query, bool validate)
{
    if (n_queries == 0 || n_queries > p->max_queries || n_items > p->max_items)
        return fail(GENIE_E_INVALID, "batch larger than the plan");
    const DevIndex& d = p->ix->d;
    if (validate)
        if (int rc = validate_items(d, items, n_items, n_queries))
            return rc;
    for (uint32_t q = 0; q < n_queries; ++q)
        if (k_per_query[q] == 0 || k_per_query[q] > p->max_k)
            return fail(GENIE_E_INVALID, "k of query " + std::to_string(q) + " outside [1, max_k]");
    // items grouped by query (stable), CSR pointer
    std::vector<uint32_t> qptr(n_queries + 1, 0);
    for (uint64_t i = 0; i < n_items; ++i)
        qptr[items[i].query + 1]++;
    for (uint32_t q = 0; q < n_queries; ++q)
        qptr[q + 1] += qptr[q];
    std::vector<genie_dim_t> sorted(n_items);
    {
        std::vector<uint32_t> fillp(qptr.begin(), qptr.end() - 1);
        for (uint64_t i = 0; i < n_items; ++i)
            sorted[fillp[items[i].query]++] = items[i];
    }
    // the batch stays on the host: the lane form's work queue and the byte
    // figures of genie_hip_plan_bytes are built from it when needed
    p->h_items.swap(sorted);
    p->h_qptr.swap(qptr);
    p->bytes_valid = false;
    p->dense_ok = dense_batch_ok(p, n_queries);
    p->sets_ok = !p->dense_ok && sets_batch_ok(p, n_queries);
    p->plain_stats = false;
    p->h_force_direct.assign(n_queries, 0);
    p->h_hwork.clear();
    p->h_dq.clear();
    p->h_hq.clear();
    p->max_hash_tiles = 0;
    // count bound of this batch
    if (p->count_bound == 0) {
        if (p->dense_ok && p->dense_wp == 1) {
            p->lane_bits = 8;  // dense, unit weights: each dim once, counts <= 128
        } else {  // weighted dense batches count up to 128 x 15
            p->h_qbound.resize(n_queries);
            compute_bounds(d, p->h_items.data(), n_items, n_queries, p->h_qbound.data());
            uint32_t mx = 0;
            for (uint32_t v : p->h_qbound)
                mx = std::max(mx, v);
            p->lane_bits = bits_for(mx);
        }
    }
    // chunks of the plain and lane forms
    uint64_t nch = 0, nch_full = 0, span_total = 0;
    bool prev_ok = false;
    for (uint64_t i = 0; i < n_items; ++i) {
        const genie_dim_t& it = p->h_items[i];
        if (it.end_pos > it.start_pos) {
            uint64_t len = (uint64_t)(it.end_pos - it.start_pos);
            span_total += len;
            if (d.codec == 0) {
                nch += (len + kChunk - 1) / kChunk;
            } else {
                // runs of contiguous segments (k_crun_flags uses the same rule)
                const genie_dim_t& pv = p->h_items[i - (i ? 1 : 0)];
                const bool starts = i == 0 || !prev_ok || pv.query != it.query || pv.weight != it.weight ||
                                    pv.end_pos != it.start_pos;
                nch += starts ? 1 : 0;
                // the item's postings, chunked as k_chunk_count does after
                // the full-layout decode translates it (plan_full_decode)
                const uint64_t* cp = d.h_seg_cpos64.data();
                const uint32_t s0 = d.seg_pos.first_ge(cp, (uint64_t)it.start_pos);
                const uint32_t s1 = d.seg_pos.first_ge(cp, (uint64_t)it.end_pos);
                const uint64_t plen = d.h_seg_upos[s1] - d.h_seg_upos[s0];
                nch_full += (plen + kChunk - 1) / kChunk;
            }
        }
        prev_ok = it.end_pos > it.start_pos;
    }
    const uint64_t nch_cap = std::max(nch, nch_full);
    if (nch_cap > p->chunks_cap) {
        dev_free(p->d_chunks);
        dev_free(p->d_nk);
        dev_free(p->d_kptr);
        dev_free(p->d_chunk_sums);
        p->chunks_cap = 0;
        if (int rc = dev_alloc(&p->d_chunks, nch_cap))
            return rc;
        if (int rc = dev_alloc(&p->d_nk, nch_cap))
            return rc;
        if (int rc = dev_alloc(&p->d_kptr, nch_cap))
            return rc;
        if (int rc = dev_alloc(&p->d_chunk_sums, (nch_cap + 1023) / 1024 + 1))
            return rc;
        p->chunks_cap = nch_cap;
    }
    p->n_queries = n_queries;
    p->n_items = n_items;
    p->n_chunks = nch;
    p->n_chunks_runs = nch;
    p->n_chunks_full = d.codec == 0 ? nch : nch_full;
    p->dec_full = false;
    p->est_postings = d.codec == 0 || d.n_words == 0
                          ? span_total
                          : (uint64_t)((double)span_total * (double)d.n_postings / (double)d.n_words);
    p->h_kq.assign(k_per_query, k_per_query + n_queries);
    p->h_out_off.assign(n_queries + 1, 0);
    for (uint32_t q = 0; q < n_queries; ++q)
        p->h_out_off[q + 1] = p->h_out_off[q] + k_per_query[q];
    p->total_k = p->h_out_off[n_queries];
    const uint64_t out_worst = (uint64_t)p->max_queries * p->max_k;
    if (int rc = grow_buf(p, &p->d_out_ids, &p->out_alloc, p->total_k, out_worst))
        return rc;
    if (int rc = grow_buf(p, &p->d_out_counts, &p->outc_alloc, p->total_k, out_worst))
        return rc;
    if (n_items)
        if (int rc = h2d(p, &p->d_items, p->h_items.data(), n_items * sizeof(genie_dim_t)))
            return rc;
    if (int rc = h2d(p, &p->d_item_qptr, p->h_qptr.data(), p->h_qptr.size() * 4))
        return rc;
    if (int rc = h2d(p, &p->d_kq, k_per_query, n_queries * 4ull))
        return rc;
    if (int rc = h2d(p, &p->d_out_off, p->h_out_off.data(), p->h_out_off.size() * 8))
        return rc;
    plan_drop_graph(p);
    if (p->mode != 0) {
        p->lanes_built = false;
        return plan_settle_form(p);
    }
    if (int rc = plan_split_plain(p))
        return rc;
    p->cap_factor = initial_cap_factor(p);
    return plan_size_candidates(p);
}

// The batch's device copies are staged in pinned memory and go out with the
// next run, on its stream.  A failed upload leaves the plan without a batch:
// it may have replaced the host copy of the previous batch or some of its
// device buffers, so run and download refuse until the next upload succeeds.
int gh::plan_upload_staged(genie_plan* p, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                           const uint32_t* k_per_query, bool validate)
{
    if (int rc = stage_begin(p))
        return rc;
    const int rc = plan_upload(p, items, n_items, n_queries, k_per_query, validate);
    p->staging = false;
    if (rc) {
        p->staged.clear();  // a failed upload sends nothing
        p->n_queries = 0;
        p->n_items = 0;
        p->bytes_valid = false;
        p->ran = false;
        p->lanes_built = false;
        plan_drop_graph(p);
    }
    return rc;
}

int gh::plan_upload_validated(genie_plan* p, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                              const uint32_t* k_per_query)
{
    if (!p || (!items && n_items) || !k_per_query)
        return fail(GENIE_E_INVALID, "null argument");
    return plan_upload_staged(p, items, n_items, n_queries, k_per_query, false);
}

extern "C" {

int genie_hip_plan_upload(genie_plan* p, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                          const uint32_t* k_per_query)
{
    if (!p || (!items && n_items) || !k_per_query)
        return fail(GENIE_E_INVALID, "null argument");
    return gh::plan_upload_staged(p, items, n_items, n_queries, k_per_query, true);
}

}  // extern "C"

static ScanArgs scan_args(genie_plan* p)
{
    const DevIndex& d = p->ix->d;
    ScanArgs a;
    a.codec = d.codec;
    a.words = d.words;
    a.seg_cpos = d.seg_cpos;
    a.seg_first = d.seg_first;
    a.seg_last = d.seg_last;
    a.key_seg = d.key_seg;
    a.seg_bph = (d.codec > 0 && codec_has_bp32(d.codec)) ? d.seg_bph : nullptr;
    a.sdir_off = d.sdir_off;
    a.sdir = d.sdir;
    a.post = d.post;
    a.key_start = d.key_start;
    a.dir_off = d.dir_off;
    a.dir = d.dir;
    a.chunks = p->d_chunks;
    a.qc_ptr = p->d_qc_ptr;
    a.qk_ptr = p->d_qk_ptr;
    a.kq = p->d_kq;
    a.qlist = nullptr;
    a.n_qlist = 0;
    a.n_objects = d.n_objects;
    a.n_queries = p->n_queries;
    a.tile_objs = tile_objs_for(p->lane_bits);
    a.n_tiles = (d.n_objects + a.tile_objs - 1) / a.tile_objs;
    a.dt_per_tile = a.tile_objs / kDirTile;
    a.n_dtiles = d.n_dtiles;
    a.gthr = p->d_gthr;
    a.cand_cnt = p->d_cand_cnt;
    a.cand_cap = p->d_cand_cap;
    a.cand_base = p->d_cand_base;
    a.cand = p->d_cand;
    a.overflow = p->d_overflow;
    return a;
}

// Whether the plain form marks touched counter words and selects sparsely:
// on average fewer postings per (query, tile) than a quarter of the tile's
// counter words (each tile still falls back to the dense select when it
// turns out busier, kSparseWords).
static bool plain_sparse(const genie_plan* p, uint32_t n_tiles)
{
    if (p->sparse_mode >= 0)
        return p->sparse_mode != 0;
    if (p->h_dq.empty())
        return false;
    const double per = (double)p->est_postings / ((double)p->h_dq.size() * (double)n_tiles);
    return per <= (double)kCntWords / 4;
}

template <bool SPARSE>
static void launch_scan_t(uint32_t lane_bits, const ScanArgs& a, uint32_t blocks, hipStream_t s)
{
    if (a.codec == 0) {
        switch (lane_bits) {
        case 8: k_scan_tile<8, SPARSE><<<blocks, kScanThreads, 0, s>>>(a); break;
        case 16: k_scan_tile<16, SPARSE><<<blocks, kScanThreads, 0, s>>>(a); break;
        default: k_scan_tile<32, SPARSE><<<blocks, kScanThreads, 0, s>>>(a); break;
        }
    } else {
        switch (lane_bits) {
        case 8: k_scan_tile_c<8, SPARSE><<<blocks, kScanThreads, 0, s>>>(a); break;
        case 16: k_scan_tile_c<16, SPARSE><<<blocks, kScanThreads, 0, s>>>(a); break;
        default: k_scan_tile_c<32, SPARSE><<<blocks, kScanThreads, 0, s>>>(a); break;
        }
    }
}

// k_decode_jobs instantiated for the index's codec.
static int launch_decode_jobs(const DevIndex& d, const uint2* jobs, uint32_t nj, uint32_t* out, hipStream_t s)
{
    const uint32_t g = std::min<uint32_t>((nj + kDecodeWaves - 1) / kDecodeWaves, 65535);
    const dim3 b(kDecodeWaves * kWave);
    switch (d.codec) {
    case 1: k_decode_jobs<1><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 2: k_decode_jobs<2><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 3: k_decode_jobs<3><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 4: k_decode_jobs<4><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 5: k_decode_jobs<5><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 6: k_decode_jobs<6><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 7: k_decode_jobs<7><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 8: k_decode_jobs<8><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 9: k_decode_jobs<9><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 10: k_decode_jobs<10><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 11: k_decode_jobs<11><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 12: k_decode_jobs<12><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    case 13: k_decode_jobs<13><<<g, b, 0, s>>>(d.words, d.seg_cpos, d.seg_bph, jobs, nj, out); break;
    default: return fail(GENIE_E_INVALID, "unknown codec " + std::to_string(d.codec));
    }
    return GENIE_OK;
}

// The plain form's scan: the hash kernel over the (query, tile) work of the
// selective queries, then the direct-mapped kernel over the other queries.
static int launch_scan(genie_plan* p, hipStream_t s)
{
    ScanArgs a = scan_args(p);
    const uint64_t nwork = p->h_hwork.size() / 2;
    const uint64_t ndq = p->h_dq.size();
    const uint64_t blocks = (uint64_t)a.n_tiles * ndq;
    if (blocks >= 0x7FFFFFFFull || nwork >= 0x7FFFFFFFull)
        return fail(GENIE_E_INVALID, "batch too large for one launch");
    if (p->dec_full) {
        // the batch's segments, each once, into the full-layout scratch; the
        // kernels below then read it as an uncompressed posting array
        const DevIndex& d = p->ix->d;
        const uint32_t nj = (uint32_t)p->h_fjobs.size();
        if (nj)
            if (int rc = launch_decode_jobs(d, p->d_fjobs, nj, p->d_full, s))
                return rc;
        if (p->timing) {
            HIP_TRY(hipEventRecord(p->ev[5], s));
            p->timed_decode = true;
        }
        a.codec = 0;
        a.post = p->d_full;
    }
    if (nwork) {
        HashArgs h;
        h.s = a;
        h.work = p->d_hwork;
        h.hq = p->d_hq;
        h.hov = p->d_hov;
        h.hov_any = p->d_hov_any;
        h.gkey = p->d_gkey;
        h.spans = nullptr;
        h.sq_ptr = nullptr;
        const DevIndex& d = p->ix->d;
        if (a.codec == 0) {
            k_scan_hash<false><<<(uint32_t)nwork, kScanThreads, 0, s>>>(h);
        } else if (p->dec_spans) {
            const uint32_t nj = (uint32_t)p->h_djobs.size();
            if (nj)
                if (int rc = launch_decode_jobs(d, p->d_djobs, nj, p->d_scratch, s))
                    return rc;
            if (p->timing) {
                HIP_TRY(hipEventRecord(p->ev[5], s));
                p->timed_decode = true;
            }
            h.s.post = p->d_scratch;
            h.spans = p->d_sspan;
            h.sq_ptr = p->d_sq_ptr;
            k_scan_hash<false><<<(uint32_t)nwork, kScanThreads, 0, s>>>(h);
        } else {
            k_scan_hash<true><<<(uint32_t)nwork, kScanThreads, 0, s>>>(h);
        }
    }
    if (ndq) {
        a.qlist = ndq == p->n_queries ? nullptr : p->d_dq;
        a.n_qlist = (uint32_t)ndq;
        if (plain_sparse(p, a.n_tiles))
            launch_scan_t<true>(p->lane_bits, a, (uint32_t)blocks, s);
        else
            launch_scan_t<false>(p->lane_bits, a, (uint32_t)blocks, s);
    }
    return GENIE_OK;
}

static int launch_merge(genie_plan* p, hipStream_t s);

// Grid of the dense kernels: every CU's share of resident workgroups, a
// multiple of the 8 XCDs the workgroup ids are dealt to.  The 1,024-thread
// kernel when the occupancy query places it, else the 512-thread one (some
// runtimes place no 1,024-thread workgroup of more than 64 VGPRs); a kernel
// the query does not place is never launched, since it would never run.
template <int NBLK, bool SAMPLE, bool SETS = false, int WP = 1>
static int launch_dense_t(const DenseArgs& a, hipStream_t s, int nt)
{
    auto k1 = k_dense_count<NBLK, SAMPLE, SETS, kDenseThreads, WP>;  // static LDS
    auto k2 = k_dense_count<NBLK, SAMPLE, SETS, kDenseThreads / 2, WP>;
    int per_cu = 0, n_cu = 0, dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    if (nt != kDenseThreads / 2)
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k1, kDenseThreads, 0));
    // the grid, and the tile chunks per group for its workgroups per XCD
    DenseArgs b = a;
    uint32_t n_wg = 0;
    auto grid = [&](int per) {
        n_wg = std::max(8u, (uint32_t)(per * n_cu) / 8 * 8);
        b.chunks = dense_chunks(a.G, n_wg / 8, !SAMPLE && !SETS);
    };
    if (per_cu >= 1) {
        grid(per_cu);
        k1<<<n_wg, kDenseThreads, 0, s>>>(b);
        return GENIE_OK;
    }
    if (nt != kDenseThreads)
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k2, kDenseThreads / 2, 0));
    if (per_cu >= 1) {
        grid(per_cu);
        k2<<<n_wg, kDenseThreads / 2, 0, s>>>(b);
        return GENIE_OK;
    }
    hipFuncAttributes fa{};
    hipDeviceProp_t prop{};
    HIP_TRY(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(k2)));
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    return fail(GENIE_E_INVALID, "dense kernel does not fit a CU at 1024 or 512 threads: " +
                                     std::to_string(fa.sharedSizeBytes) + " B of LDS, " +
                                     std::to_string(fa.numRegs) + " VGPRs at 512; device " + prop.gcnArchName +
                                     ": " + std::to_string(prop.maxSharedMemoryPerMultiProcessor) +
                                     " B of LDS per CU, " + std::to_string(prop.maxThreadsPerMultiProcessor) +
                                     " threads per CU");
}

template <bool SAMPLE, int WP>
static int launch_dense_w(uint32_t nblk, const DenseArgs& a, hipStream_t s, int nt)
{
    switch (nblk) {
    case 1: return launch_dense_t<1, SAMPLE, false, WP>(a, s, nt);
    case 2: return launch_dense_t<2, SAMPLE, false, WP>(a, s, nt);
    case 3: return launch_dense_t<3, SAMPLE, false, WP>(a, s, nt);
    case 4: return launch_dense_t<4, SAMPLE, false, WP>(a, s, nt);
    case 5: return launch_dense_t<5, SAMPLE, false, WP>(a, s, nt);
    case 6: return launch_dense_t<6, SAMPLE, false, WP>(a, s, nt);
    case 7: return launch_dense_t<7, SAMPLE, false, WP>(a, s, nt);
    case 8: return launch_dense_t<8, SAMPLE, false, WP>(a, s, nt);
    }
    return fail(GENIE_E_INVALID, "dense form supports at most 128 dims");
}

// nblk 0: the set form; wp: weight planes per query of a dense batch (1:
// unit weights; 2, 3, 4: weights below 4, 8, 16)
template <bool SAMPLE>
static int launch_dense(uint32_t nblk, const DenseArgs& a, hipStream_t s, int nt, uint32_t wp)
{
    if (nblk == 0)
        return launch_dense_t<0, SAMPLE, true>(a, s, nt);
    switch (wp) {
    case 2: return launch_dense_w<SAMPLE, 2>(nblk, a, s, nt);
    case 3: return launch_dense_w<SAMPLE, 3>(nblk, a, s, nt);
    case 4: return launch_dense_w<SAMPLE, 4>(nblk, a, s, nt);
    }
    return launch_dense_w<SAMPLE, 1>(nblk, a, s, nt);
}

// Dense form: query masks, sample pass + thresholds, main pass, merge.
static int run_dense(genie_plan* p, hipStream_t s)
{
    const DevIndex& d = p->ix->d;
    const uint32_t nq = p->n_queries;
    // the sample histogram was cleared with the run's other resets
    if (p->timing)
        HIP_TRY(hipEventRecord(p->ev[1], s));
    // the sample pass builds every group's mask table in LDS from the items
    // and writes it out for the main pass
    DenseArgs a;
    a.items = p->d_items;
    a.item_qptr = p->d_item_qptr;
    a.dkeys = p->d_dkeys;
    a.key_slot = d.key_slot;
    a.codes = d.codes;
    a.n_pad = d.dense_npad;
    a.tab = p->d_dtab;
    a.D_pad = d.dense_dpad;
    a.G = p->dn_groups;
    a.n_queries = nq;
    a.n_objects = d.n_objects;
    a.n_tiles = (d.n_objects + kDenseTile - 1) / kDenseTile;
    a.list = p->d_dlist;
    a.n_list = p->n_dlist;
    a.gthr = p->d_gthr;
    a.redo = nullptr;
    a.hist = p->d_dhist;
    a.cand_cnt = p->d_cand_cnt;
    a.cand_cap = p->d_cand_cap;
    a.cand_base = p->d_cand_base;
    a.cand = p->d_cand;
    a.overflow = p->d_overflow;
    a.obj_perm = d.obj_perm;
    a.blk_off = d.blk_off;
    a.blk_key = d.blk_key;
    a.n_slots = d.set_slots;
    a.set_keys = p->d_skeys;
    a.set_mask = p->d_smask;
    a.set_gptr = p->d_sgptr;
    a.hshift = (p->dense_ok && p->dense_wp == 1) ? nullptr : p->d_sshift;
    const uint32_t nblk = p->dense_ok ? d.dense_dpad / 16 : 0;  // 0: the set form
    if (int rc = launch_dense<true>(nblk, a, s, p->dense_nt, p->dense_wp))
        return rc;
    // thresholds, then every query's candidate space from its sample
    const uint32_t mult = p->dense_mult;
    const uint32_t ratio16 = (uint32_t)std::min<uint64_t>(
        ((uint64_t)d.n_objects << 16) / std::max<uint32_t>(1, p->n_dsample) + 1, 0xFFFFFFFFull);
    k_dense_thresh<<<(nq + 3) / 4, 256, 0, s>>>(p->d_dhist, p->d_kq, nq, p->n_dsample, d.n_objects, ratio16, mult,
                                                a.hshift, p->d_gthr, p->d_gsafe, p->d_want, p->d_want_safe);
    k_cap_prefix<<<1, 1024, 0, s>>>(p->d_want, p->d_want_safe, mult > 1 ? 1u : 0u, nq, p->total_cap, p->d_cand_cap,
                                    p->d_cand_base, p->d_cap_used);
    if (p->timing)
        HIP_TRY(hipEventRecord(p->ev[4], s));
    a.list = nullptr;
    a.n_list = 0;
    if (int rc = launch_dense<false>(nblk, a, s, p->dense_nt, p->dense_wp))
        return rc;
    // queries the speculative threshold left short: their groups again
    k_dense_check<<<p->dn_groups, 64, 0, s>>>(p->d_kq, nq, p->d_gthr, p->d_gsafe, p->d_cand_cnt, p->d_redo,
                                              p->d_want_safe, p->d_cand_cap, p->d_cand_base, p->d_cap_used,
                                              p->total_cap,
                                              (p->dense_ok && p->dense_wp == 1) ? kDenseNoQuery : kSetNoQuery,
                                              dense_qpg(p));
    a.redo = p->d_redo;
    if (int rc = launch_dense<false>(nblk, a, s, p->dense_nt, p->dense_wp))
        return rc;
    if (p->timing)
        HIP_TRY(hipEventRecord(p->ev[2], s));
    return launch_merge(p, s);
}

static int plan_run_once(genie_plan* p, hipStream_t s);

// Candidate overflow after a completed run on stream s: grow the capacity
// and run again (the reference retries on hash-table overflow in the same
// way, match.cu:356-421).
// A hash-form query whose table filled up is run on the direct-mapped form
// instead, in the same way.
static int settle_overflow(genie_plan* p, hipStream_t s)
{
    for (int attempt = 0;; ++attempt) {
        uint32_t ov = 0, hv = 0;
        HIP_TRY(hipMemcpy(&ov, p->d_overflow, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&hv, p->d_hov_any, 4, hipMemcpyDeviceToHost));
        if (!ov && !hv)
            return GENIE_OK;
        if (attempt > 12)
            return fail(GENIE_E_RUNTIME, "candidate overflow at full capacity");
        plan_drop_graph(p);
        if (hv) {
            std::vector<uint32_t> flags(p->n_queries);
            HIP_TRY(hipMemcpy(flags.data(), p->d_hov, p->n_queries * 4ull, hipMemcpyDeviceToHost));
            for (uint32_t q = 0; q < p->n_queries; ++q)
                if (flags[q])
                    p->h_force_direct[q] = 1;
            if (int rc = plan_split_plain(p))
                return rc;
        }
        if (ov) {
            const uint32_t fmax = max_cap_factor(p);
            if (dense_active(p)) {
                // two limits: the host total (cap_factor) and each query's
                // ask on the device (dense_mult); each grows until it gives
                // every query room for every object
                const uint32_t mmax = dense_mult_max(p);
                if (p->cap_factor >= fmax && p->dense_mult >= mmax)
                    return fail(GENIE_E_RUNTIME, "candidate overflow at full capacity");
                p->dense_mult = (uint32_t)std::min<uint64_t>((uint64_t)p->dense_mult * 4, mmax);
            } else if (p->cap_factor >= fmax) {
                return fail(GENIE_E_RUNTIME, "candidate overflow at full capacity");
            }
            p->cap_factor = (uint32_t)std::min<uint64_t>((uint64_t)p->cap_factor * 4, fmax);
        }
        if (int rc = plan_size_candidates(p))
            return rc;
        if (int rc = plan_run_once(p, s))
            return rc;
        HIP_TRY(hipStreamSynchronize(s));
    }
}

static int plan_run_once(genie_plan* p, hipStream_t s)
{
    const DevIndex& d = p->ix->d;
    const uint32_t nq = p->n_queries;
    const uint32_t ni = (uint32_t)p->n_items;
    if (p->timing)
        HIP_TRY(hipEventRecord(p->ev[0], s));
    // per-run resets in one launch (fill kernels, not memsets: graph-safe)
    {
        ZeroList z{};
        z.p[0] = p->d_gthr, z.n[0] = nq;
        z.p[1] = p->d_cand_cnt, z.n[1] = nq;
        z.p[2] = p->d_overflow, z.n[2] = 1;
        z.p[3] = p->d_hov, z.n[3] = nq;
        z.p[4] = p->d_hov_any, z.n[4] = 1;
        if (dense_active(p))
            z.p[5] = p->d_dhist, z.n[5] = (uint64_t)nq * kDenseBins;
        else
            z.p[5] = reinterpret_cast<uint32_t*>(p->d_gkey), z.n[5] = 2ull * nq;
        const uint64_t most = std::max<uint64_t>(nq, z.n[5]);
        k_zero_many<<<(uint32_t)std::min<uint64_t>((most + 255) / 256, 1024), 256, 0, s>>>(z);
    }
    p->timed_dense = dense_active(p);
    p->timed_decode = false;
    if (p->timed_dense)
        return run_dense(p, s);
    if (lane_form(p)) {
        // lane form: the work queue was built at upload
        fill_u32(p->d_queue, 0, 1, s);
        if (p->timing)
            HIP_TRY(hipEventRecord(p->ev[1], s));
        if (int rc = launch_batched(p, s))
            return rc;
        if (p->timing)
            HIP_TRY(hipEventRecord(p->ev[2], s));
        return launch_merge(p, s);
    }
    // compile items into chunks (uncompressed, or a compressed batch decoded
    // into the full-layout scratch: its items in posting positions) or
    // segment runs (compressed); the kernels below write n_chunks chunks
    if (p->n_chunks > p->chunks_cap)
        return fail(GENIE_E_RUNTIME, "batch compiles into more chunks than the plan holds");
    if (ni && (d.codec == 0 || p->dec_full)) {
        const genie_dim_t* items = p->dec_full ? p->d_uitems : p->d_items;
        k_chunk_count<<<(ni + 255) / 256, 256, 0, s>>>(items, ni, p->d_nchunks);
        const uint32_t nb = (ni + 1023) / 1024;
        k_scan_blocks<<<nb, 256, 0, s>>>(p->d_nchunks, p->d_chunk_ptr, ni, p->d_block_sums);
        k_scan_sums<<<1, 256, 0, s>>>(p->d_block_sums, nb, p->d_total);
        k_scan_add<<<nb, 256, 0, s>>>(p->d_chunk_ptr, ni, p->d_block_sums);
        k_chunk_write<<<(ni + 255) / 256, 256, 0, s>>>(items, ni, p->d_chunk_ptr, d.key_start,
                                                      d.n_keys, p->d_chunks);
    } else if (ni) {
        k_crun_flags<<<(ni + 255) / 256, 256, 0, s>>>(p->d_items, ni, d.seg_cpos, d.n_segments, p->d_nchunks);
        const uint32_t nb = (ni + 1023) / 1024;
        k_scan_blocks<<<nb, 256, 0, s>>>(p->d_nchunks, p->d_chunk_ptr, ni, p->d_block_sums);
        k_scan_sums<<<1, 256, 0, s>>>(p->d_block_sums, nb, p->d_total);
        k_scan_add<<<nb, 256, 0, s>>>(p->d_chunk_ptr, ni, p->d_block_sums);
        k_crun_write<<<(ni + 255) / 256, 256, 0, s>>>(p->d_items, ni, d.seg_cpos, d.n_segments, d.seg_key,
                                                     p->d_chunk_ptr, p->d_chunks);
    } else {
        fill_u32(p->d_total, 0, 1, s);
    }
    k_query_chunk_ptr<<<(nq + 1 + 255) / 256, 256, 0, s>>>(p->d_item_qptr, nq, p->d_chunk_ptr, ni,
                                                          p->d_total, p->d_qc_ptr);
    // pair prefix: every (chunk, key) pair gets a batch-wide position
    if (p->n_chunks) {
        const uint32_t nc = (uint32_t)p->n_chunks;
        const uint32_t g = std::min<uint32_t>((nc + 255) / 256, 4096);
        k_chunk_nkeys<<<g, 256, 0, s>>>(p->d_chunks, p->d_total, p->d_nk);
        const uint32_t nb = (nc + 1023) / 1024;
        k_scan_blocks<<<nb, 256, 0, s>>>(p->d_nk, p->d_kptr, nc, p->d_chunk_sums);
        k_scan_sums<<<1, 256, 0, s>>>(p->d_chunk_sums, nb, p->d_total_pairs);
        k_scan_add<<<nb, 256, 0, s>>>(p->d_kptr, nc, p->d_chunk_sums);
        k_chunk_set_kp<<<g, 256, 0, s>>>(p->d_chunks, p->d_total, p->d_kptr);
    } else {
        fill_u32(p->d_total_pairs, 0, 1, s);
    }
    k_query_pair_ptr<<<(nq + 1 + 255) / 256, 256, 0, s>>>(p->d_qc_ptr, nq, p->d_kptr, p->d_total,
                                                         p->d_total_pairs, p->d_qk_ptr);
    if (p->timing)
        HIP_TRY(hipEventRecord(p->ev[1], s));
    if (int rc = launch_scan(p, s))
        return rc;
    if (p->timing)
        HIP_TRY(hipEventRecord(p->ev[2], s));
    return launch_merge(p, s);
}

static int launch_merge(genie_plan* p, hipStream_t s)
{
    const uint32_t nq = p->n_queries;
    MergeArgs m;
    m.cand = p->d_cand;
    m.cand_base = p->d_cand_base;
    m.cand_cnt = p->d_cand_cnt;
    m.cand_cap = p->d_cand_cap;
    m.kq = p->d_kq;
    m.out_off = p->d_out_off;
    m.out_ids = p->d_out_ids;
    m.out_counts = p->d_out_counts;
    m.key_bits = 32 + p->lane_bits;
    m.flag_a = p->d_overflow;
    m.flag_b = p->d_hov_any;
    m.flag_any = p->d_flag_any;
    k_merge<<<nq, kMergeThreads, 0, s>>>(m);
    if (p->timing)
        HIP_TRY(hipEventRecord(p->ev[3], s));
    HIP_TRY(hipGetLastError());
    p->ran = true;
    return GENIE_OK;
}

extern "C" {

int genie_hip_plan_run(genie_plan* p, void* stream)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    if (p->n_queries == 0)
        return fail(GENIE_E_NOT_BUILT, "no batch uploaded");
    hipStream_t s = (hipStream_t)stream;
    if (int rc = stage_issue(p, s))
        return rc;
    if (!p->graph_on || p->timing)
        return plan_run_once(p, s);
    if (!p->graph_valid) {
        // capture the whole run (fill kernels, no memset nodes) once
        plan_drop_graph(p);
        HIP_TRY(hipStreamBeginCapture(p->own_stream, hipStreamCaptureModeThreadLocal));
        int rc = plan_run_once(p, p->own_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(p->own_stream, &g);
        if (rc) {
            if (g)
                (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess)
            return fail(GENIE_E_RUNTIME, std::string("graph capture: ") + hipGetErrorString(e));
        p->graph = g;
        HIP_TRY(hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0));
        p->graph_valid = true;
    }
    HIP_TRY(hipGraphLaunch(p->gexec, s));
    p->ran = true;
    return GENIE_OK;
}

int genie_hip_plan_set_graph(genie_plan* p, int on)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    p->graph_on = on != 0;
    if (!p->graph_on)
        plan_drop_graph(p);
    return GENIE_OK;
}

int genie_hip_plan_set_config(genie_plan* p, uint32_t block, uint32_t tile_objs, uint32_t seglen)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    if (block != 256 && block != 512 && block != 1024)
        return fail(GENIE_E_INVALID, "block must be 256, 512 or 1024");
    if (tile_objs == 0 || tile_objs % kDirTile || (uint64_t)(tile_objs + kWave) * 8 > 160 * 1024)
        return fail(GENIE_E_INVALID, "tile must be a multiple of " + std::to_string(kDirTile) +
                                         " objects and fit in LDS");
    if (seglen < 1024)
        return fail(GENIE_E_INVALID, "segment length must be at least 1024 postings");
    const bool rebuild = seglen != p->seglen || tile_objs != p->tile_b;
    p->block = block;
    p->tile_b = tile_objs;
    p->seglen = seglen;
    plan_drop_graph(p);
    if (rebuild)
        p->lanes_built = false;
    if (lane_form(p) && p->n_queries && rebuild && !dense_active(p))
        return plan_settle_form(p);
    return GENIE_OK;
}

int genie_hip_plan_set_dense(genie_plan* p, int on)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    if (p->mode == 0)
        return fail(GENIE_E_INVALID, "the dense form belongs to batched and automatic plans");
    if (p->dense_on == (on != 0))
        return GENIE_OK;
    p->dense_on = on != 0;
    return plan_settle_form(p);
}

int genie_hip_plan_set_dense_threads(genie_plan* p, int threads)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    if (threads != 0 && threads != kDenseThreads && threads != kDenseThreads / 2)
        return fail(GENIE_E_INVALID, "dense threads must be 0, 512 or 1024");
    p->dense_nt = threads;
    plan_drop_graph(p);
    return GENIE_OK;
}

int genie_hip_plan_set_sparse(genie_plan* p, int mode)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    if (mode < -1 || mode > 1)
        return fail(GENIE_E_INVALID, "sparse select mode must be -1, 0 or 1");
    plan_drop_graph(p);
    p->sparse_mode = mode;
    return GENIE_OK;
}

int genie_hip_plan_set_hash(genie_plan* p, int mode)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    if (mode < -1 || mode > 1)
        return fail(GENIE_E_INVALID, "hash form mode must be -1, 0 or 1");
    plan_drop_graph(p);
    p->hash_mode = mode;
    if (p->n_queries && !lane_form(p) && !dense_active(p)) {
        std::fill(p->h_force_direct.begin(), p->h_force_direct.end(), 0);
        if (int rc = plan_split_plain(p))
            return rc;
        return plan_size_candidates(p);
    }
    return GENIE_OK;
}

int genie_hip_plan_set_decode(genie_plan* p, int mode)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    if (mode != -1 && mode != 0)
        return fail(GENIE_E_INVALID, "decode mode must be -1 or 0");
    plan_drop_graph(p);
    p->decode_mode = mode;
    if (p->n_queries && !lane_form(p) && !dense_active(p)) {
        if (int rc = plan_split_plain(p))
            return rc;
        return plan_size_candidates(p);
    }
    return GENIE_OK;
}

int genie_hip_plan_decode_active(genie_plan* p, int* active)
{
    if (!p || !active)
        return fail(GENIE_E_INVALID, "null argument");
    *active = p->n_queries && !lane_form(p) && !dense_active(p) && p->dec_full ? 1 : 0;
    return GENIE_OK;
}

int genie_hip_plan_hash_queries(genie_plan* p, uint32_t* n)
{
    if (!p || !n)
        return fail(GENIE_E_INVALID, "null argument");
    uint32_t c = 0;
    if (!lane_form(p) && !dense_active(p))
        for (const uint4& h : p->h_hq)
            c += h.x != 0;
    *n = c;
    return GENIE_OK;
}

int genie_hip_plan_sparse_active(genie_plan* p, int* active)
{
    if (!p || !active)
        return fail(GENIE_E_INVALID, "null argument");
    const DevIndex& d = p->ix->d;
    const uint32_t to = tile_objs_for(p->lane_bits);
    *active = p->n_queries && plain_sparse(p, (d.n_objects + to - 1) / to) ? 1 : 0;
    return GENIE_OK;
}

int genie_hip_plan_dense_active(genie_plan* p, int* active)
{
    if (!p || !active)
        return fail(GENIE_E_INVALID, "null argument");
    *active = dense_active(p) ? 1 : 0;
    return GENIE_OK;
}

int genie_hip_plan_get_config(genie_plan* p, uint32_t* block, uint32_t* tile_objs, uint32_t* seglen,
                              uint32_t* lane_bits)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    if (block)
        *block = p->block;
    if (tile_objs)
        *tile_objs = plan_tile_objs(p);
    if (seglen)
        *seglen = lane_form(p) ? p->seglen : kChunk;
    if (lane_bits)
        *lane_bits = p->lane_bits;
    return GENIE_OK;
}

int genie_hip_plan_host_ms(genie_plan* p, double* ms)
{
    if (!p || !ms)
        return fail(GENIE_E_INVALID, "null argument");
    *ms = p->host_build_ms;
    return GENIE_OK;
}

// Autotuner of the batched form: every (block, tile, segment length)
// candidate runs the uploaded batch (one warm-up, `reps` timed runs on the
// plan's stream); the fastest is kept.  Results go to out_ms (one entry per
// candidate, in the order block-major, then tile, then segment length) when
// given; *n_out receives the candidate count.
int genie_hip_plan_autotune(genie_plan* p, int reps, double* out_ms, uint32_t* n_out)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    if (p->mode != 1 || p->n_queries == 0)
        return fail(GENIE_E_INVALID, "autotune needs a batched plan with an uploaded batch");
    const uint32_t blocks[] = {256, 512, 1024};
    const uint32_t tiles[] = {8192, 16384};
    const uint32_t segs[] = {16384, 65536};
    if (int rc = stage_issue(p, p->own_stream))
        return rc;
    const bool was_graph = p->graph_on, try_dense = p->dense_on && (p->dense_ok || p->sets_ok);
    p->graph_on = false;
    double best = 1e30;
    uint32_t bb = p->block, bt = p->tile_b, bs = p->seglen, n = 0;
    bool best_dense = false, best_plain = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // one settled run (the first may overflow the candidate space), then reps
    auto time_it = [&](double* ms) -> int {
        *ms = -1;
        if (int rc = plan_run_once(p, p->own_stream))
            return rc;
        HIP_TRY(hipStreamSynchronize(p->own_stream));
        if (int rc = settle_overflow(p, p->own_stream))
            return rc;
        HIP_TRY(hipEventRecord(e0, p->own_stream));
        for (int r = 0; r < std::max(1, reps); ++r)
            if (int rc = plan_run_once(p, p->own_stream))
                return rc;
        HIP_TRY(hipEventRecord(e1, p->own_stream));
        HIP_TRY(hipEventSynchronize(e1));
        float f = 0;
        HIP_TRY(hipEventElapsedTime(&f, e0, e1));
        *ms = f / std::max(1, reps);
        return GENIE_OK;
    };
    int rc = GENIE_OK;
    p->batched_plain = false;
    if (try_dense)
        p->dense_on = false;
    rc = plan_settle_form(p);
    for (uint32_t blk : blocks)
        for (uint32_t tl : tiles)
            for (uint32_t sg : segs) {
                if (rc)
                    break;
                double ms = -1;
                if ((rc = genie_hip_plan_set_config(p, blk, tl, sg)) == GENIE_OK)
                    rc = time_it(&ms);
                if (rc == GENIE_E_INVALID)  // a candidate that does not fit this device: skip it
                    rc = GENIE_OK;
                if (rc)
                    break;
                if (out_ms)
                    out_ms[n] = ms;
                ++n;
                if (ms >= 0 && ms < best) {
                    best = ms;
                    bb = blk;
                    bt = tl;
                    bs = sg;
                }
            }
    if (!rc && try_dense) {
        // then the dense (or set) form
        double ms = -1;
        p->dense_on = true;
        if ((rc = plan_settle_form(p)) == GENIE_OK)
            rc = time_it(&ms);
        if (!rc) {
            if (out_ms)
                out_ms[n] = ms;
            ++n;
            if (ms >= 0 && ms < best) {
                best = ms;
                best_dense = true;
            }
        }
    }
    if (!rc) {
        // and the plain form last: a batched plan never runs slower than it
        double ms = -1;
        p->dense_on = false;
        p->batched_plain = true;
        if ((rc = plan_settle_form(p)) == GENIE_OK)
            rc = time_it(&ms);
        if (!rc) {
            if (out_ms)
                out_ms[n] = ms;
            ++n;
            if (ms >= 0 && ms < best) {
                best = ms;
                best_plain = true;
                best_dense = false;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (n_out)
        *n_out = n;
    p->graph_on = was_graph;
    if (rc)
        return rc;
    p->batched_plain = best_plain;
    p->dense_on = best_dense;
    p->block = bb;
    p->tile_b = bt;
    p->seglen = bs;
    p->lanes_built = false;  // the lane form's queue is rebuilt for the chosen configuration
    return plan_settle_form(p);
}

int genie_hip_plan_thresholds(genie_plan* p, void* stream, uint32_t* out)
{
    if (!p || !out)
        return fail(GENIE_E_INVALID, "null argument");
    if (!p->ran)
        return fail(GENIE_E_NOT_BUILT, "plan has not run");
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    HIP_TRY(hipMemcpy(out, p->d_gthr, p->n_queries * 4ull, hipMemcpyDeviceToHost));
    return GENIE_OK;
}

int genie_hip_plan_set_form(genie_plan* p, int form)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    if (p->mode != 1)
        return fail(GENIE_E_INVALID, "the form of a batched plan only");
    if (form != 0 && form != 1)
        return fail(GENIE_E_INVALID, "form must be 0 (plain) or 1 (lane)");
    if (p->batched_plain == (form == 0))
        return GENIE_OK;
    p->batched_plain = form == 0;
    return plan_settle_form(p);
}

int genie_hip_plan_form(genie_plan* p, int* form)
{
    if (!p || !form)
        return fail(GENIE_E_INVALID, "null argument");
    if (dense_active(p))
        *form = p->dense_ok ? 2 : 3;
    else
        *form = lane_form(p) ? 1 : 0;
    return GENIE_OK;
}

int genie_hip_plan_download(genie_plan* p, void* stream, int32_t* out_ids, int32_t* out_counts)
{
    if (!p || !out_ids || !out_counts)
        return fail(GENIE_E_INVALID, "null argument");
    if (!p->ran)
        return fail(GENIE_E_NOT_BUILT, "plan has not run");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipStreamSynchronize(s));
    if (int rc = settle_overflow(p, s))
        return rc;
    if (p->timing) {
        float a = 0, b = 0, c = 0, e = 0;
        HIP_TRY(hipEventElapsedTime(&a, p->ev[0], p->ev[1]));
        if (p->timed_dense) {
            HIP_TRY(hipEventElapsedTime(&e, p->ev[1], p->ev[4]));
            HIP_TRY(hipEventElapsedTime(&b, p->ev[4], p->ev[2]));
        } else {
            HIP_TRY(hipEventElapsedTime(&b, p->ev[1], p->ev[2]));
        }
        HIP_TRY(hipEventElapsedTime(&c, p->ev[2], p->ev[3]));
        p->ms_compile = a;
        p->ms_sample = e;
        p->ms_scan = b;
        p->ms_merge = c;
    }
    // on the run's stream; the buffers may be host or device memory
    HIP_TRY(hipMemcpyAsync(out_ids, p->d_out_ids, p->total_k * 4, hipMemcpyDefault, s));
    HIP_TRY(hipMemcpyAsync(out_counts, p->d_out_counts, p->total_k * 4, hipMemcpyDefault, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GENIE_OK;
}

int genie_hip_plan_download_async(genie_plan* p, void* stream, int32_t* out_ids, int32_t* out_counts,
                                  uint32_t* overflow)
{
    if (!p || !out_ids || !out_counts || !overflow)
        return fail(GENIE_E_INVALID, "null argument");
    if (!p->ran)
        return fail(GENIE_E_NOT_BUILT, "plan has not run");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(out_ids, p->d_out_ids, p->total_k * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(out_counts, p->d_out_counts, p->total_k * 4, hipMemcpyDeviceToHost, s));
    // nonzero when the candidate space overflowed or a hash-form table filled
    // up: either way genie_hip_plan_download settles the batch
    HIP_TRY(hipMemcpyAsync(overflow, p->d_flag_any, 4, hipMemcpyDeviceToHost, s));
    return GENIE_OK;
}

int genie_hip_host_alloc(void** ptr, uint64_t bytes)
{
    if (!ptr)
        return fail(GENIE_E_INVALID, "null argument");
    HIP_TRY(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return GENIE_OK;
}

int genie_hip_host_free(void* ptr)
{
    if (ptr)
        HIP_TRY(hipHostFree(ptr));
    return GENIE_OK;
}

int genie_hip_plan_set_timing(genie_plan* p, int on)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    p->timing = on != 0;
    return GENIE_OK;
}

int genie_hip_plan_kernel_ms(genie_plan* p, const char* name, double* ms)
{
    if (!p || !name || !ms)
        return fail(GENIE_E_INVALID, "null argument");
    if (p->timing && p->ran) {
        // refresh from the events of the last run
        float v = 0;
        const hipEvent_t scan0 = p->timed_dense ? p->ev[4] : p->ev[1];
        if (!strcmp(name, "compile") && hipEventElapsedTime(&v, p->ev[0], p->ev[1]) == hipSuccess) { *ms = v; return GENIE_OK; }
        if (!strcmp(name, "sample")) {
            if (!p->timed_dense) { *ms = 0; return GENIE_OK; }
            if (hipEventElapsedTime(&v, p->ev[1], p->ev[4]) == hipSuccess) { *ms = v; return GENIE_OK; }
        }
        if (!strcmp(name, "scan") && hipEventElapsedTime(&v, scan0, p->ev[2]) == hipSuccess) { *ms = v; return GENIE_OK; }
        if (!strcmp(name, "decode")) {  // part of "scan": the segments decoded into a scratch
            if (!p->timed_decode) { *ms = 0; return GENIE_OK; }
            if (hipEventElapsedTime(&v, p->ev[1], p->ev[5]) == hipSuccess) { *ms = v; return GENIE_OK; }
        }
        if (!strcmp(name, "merge") && hipEventElapsedTime(&v, p->ev[2], p->ev[3]) == hipSuccess) { *ms = v; return GENIE_OK; }
    }
    return fail(GENIE_E_INVALID, std::string("no timing for kernel ") + name);
}

int genie_hip_plan_bytes(genie_plan* p, uint64_t* logical, uint64_t* compulsory)
{
    if (!p || !logical || !compulsory)
        return fail(GENIE_E_INVALID, "null argument");
    if (!p->bytes_valid) {
        uint64_t logical_b = 0;
        std::vector<std::pair<uint32_t, uint32_t>> spans;
        spans.reserve(p->n_items);
        for (uint64_t i = 0; i < p->n_items; ++i) {
            const genie_dim_t& it = p->h_items[i];
            if (it.end_pos > it.start_pos) {
                logical_b += 4ull * (uint32_t)(it.end_pos - it.start_pos);
                spans.push_back({(uint32_t)it.start_pos, (uint32_t)it.end_pos});
            }
        }
        std::sort(spans.begin(), spans.end());
        uint64_t comp = 0, cur_e = 0;
        for (auto& sp : spans) {  // union of position ranges
            const uint64_t a = std::max<uint64_t>(sp.first, cur_e);
            if (sp.second > a)
                comp += 4 * (sp.second - a);
            cur_e = std::max<uint64_t>(cur_e, sp.second);
        }
        p->logical_bytes = logical_b;
        p->compulsory_bytes = comp;
        p->bytes_valid = true;
    }
    *logical = p->logical_bytes;
    *compulsory = p->compulsory_bytes;
    return GENIE_OK;
}

int genie_hip_match_topk(genie_index* ix, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                         const uint32_t* k_per_query, uint32_t count_bound, int32_t* out_ids,
                         int32_t* out_counts)
{
    return genie_hip_match_topk_stream(ix, items, n_items, n_queries, k_per_query, count_bound, out_ids, out_counts,
                                       nullptr);
}

int genie_hip_match_topk_stream(genie_index* ix, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                                const uint32_t* k_per_query, uint32_t count_bound, int32_t* out_ids,
                                int32_t* out_counts, void* stream)
{
    if (!ix || (!items && n_items) || !k_per_query || !out_ids || !out_counts)
        return fail(GENIE_E_INVALID, "null argument");
    if (n_queries == 0)
        return GENIE_OK;
    uint32_t max_k = 0;
    for (uint32_t q = 0; q < n_queries; ++q)
        max_k = std::max(max_k, k_per_query[q]);
    if (max_k == 0 || max_k > GENIE_MAX_K)
        return fail(GENIE_E_INVALID, "k outside [1, GENIE_MAX_K]");
    genie_plan* p = nullptr;
    int rc = genie_hip_plan_create(ix, n_queries, n_items, max_k, count_bound, 2, &p);
    if (rc)
        return rc;
    if (!(rc = genie_hip_plan_upload(p, items, n_items, n_queries, k_per_query)) &&
        !(rc = genie_hip_plan_run(p, stream)))
        rc = genie_hip_plan_download(p, stream, out_ids, out_counts);
    genie_hip_plan_destroy(p);
    return rc;
}

int genie_hip_decode(int codec, const uint32_t* words, const uint32_t* cpos, uint32_t n_arrays,
                     uint32_t out_stride, uint32_t* out, uint32_t* out_len)
{
    CodecShape cs;
    if (!codec_shape(codec, &cs))
        return fail(GENIE_E_INVALID, "unknown codec " + std::to_string(codec));
    if (n_arrays == 0)
        return GENIE_OK;
    if (!words || !cpos || !out || !out_len)
        return fail(GENIE_E_INVALID, "null argument");
    const uint64_t nw = cpos[n_arrays];
    for (uint32_t i = 0; i < n_arrays; ++i)
        if (cpos[i + 1] < cpos[i])
            return fail(GENIE_E_INVALID, "array offsets must be non-decreasing");
    uint32_t *d_w = nullptr, *d_c = nullptr, *d_o = nullptr, *d_l = nullptr;
    int rc = GENIE_OK;
    do {
        if ((rc = dev_alloc(&d_w, nw + 4))) break;
        if ((rc = dev_alloc(&d_c, n_arrays + 1))) break;
        if ((rc = dev_alloc(&d_o, (uint64_t)n_arrays * out_stride))) break;
        if ((rc = dev_alloc(&d_l, n_arrays))) break;
        hipError_t e;
        if ((e = hipMemcpy(d_w, words, nw * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d_c, cpos, (n_arrays + 1) * 4, hipMemcpyHostToDevice))) {
            rc = fail(GENIE_E_RUNTIME, hipGetErrorString(e));
            break;
        }
        const uint32_t blocks = std::min<uint32_t>((n_arrays + kDecodeWaves - 1) / kDecodeWaves, 65535);
        k_decode_arrays<<<blocks, kDecodeWaves * kWave>>>(codec, d_w, d_c, n_arrays, out_stride, d_o, d_l);
        if ((e = hipGetLastError()) || (e = hipDeviceSynchronize()) ||
            (e = hipMemcpy(out, d_o, (uint64_t)n_arrays * out_stride * 4, hipMemcpyDeviceToHost)) ||
            (e = hipMemcpy(out_len, d_l, n_arrays * 4, hipMemcpyDeviceToHost))) {
            rc = fail(GENIE_E_RUNTIME, hipGetErrorString(e));
            break;
        }
    } while (0);
    dev_free(d_w);
    dev_free(d_c);
    dev_free(d_o);
    dev_free(d_l);
    return rc;
}

}  // extern "C"

// ---- no-allocation entry point ----------------------------------------------
// genie_hip_match_topk_static: the plain per-batch search (direct-mapped
// kernel, no hash form) with every device buffer carved out of one
// file-scope array instead of hipMalloc, for callers that cannot allocate.
// Each GPU the library runs on holds its own copy of the array until the
// process ends.

constexpr uint64_t kStaticScratchBytes = 32ull << 20;
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

namespace {

struct StaticScratch {
    std::mutex mu;          // calls on one device run one at a time
    uint8_t* ptr = nullptr; // this device's copy of `scratch`
};
std::mutex g_static_mu;
std::vector<std::unique_ptr<StaticScratch> > g_static;  // by device id

// The device buffers of one call, all inside the array.
struct StaticBufs {
    genie_dim_t* d_items;
    uint32_t *d_item_qptr, *d_nchunks, *d_chunk_ptr, *d_block_sums, *d_total;
    CChunk* d_chunks;
    uint32_t *d_nk, *d_kptr, *d_chunk_sums, *d_qc_ptr, *d_qk_ptr, *d_total_pairs;
    uint32_t *d_kq, *d_gthr, *d_cand_cnt, *d_cand_cap;
    uint64_t *d_cand_base, *d_cand;
    uint32_t *d_overflow, *d_hov_any, *d_flag_any;
    uint64_t* d_out_off;
    int32_t *d_out_ids, *d_out_counts;
};

// Bump allocation inside the array (256-byte aligned pieces); with a null
// base it only measures.
struct Carve {
    uint8_t* base = nullptr;
    uint64_t off = 0;
    template <class T>
    T* take(uint64_t n)
    {
        off = (off + 255) & ~255ull;
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += std::max<uint64_t>(n, 1) * sizeof(T);
        return p;
    }
};

}  // namespace

// This device's address of `scratch`, looked up once per device (with the
// device current); host code uses it, never the array's name.
static int static_scratch(StaticScratch** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_static_mu);
    if ((size_t)dev >= g_static.size())
        g_static.resize((size_t)dev + 1);
    if (!g_static[(size_t)dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GENIE_E_RUNTIME, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        auto s = std::make_unique<StaticScratch>();
        s->ptr = static_cast<uint8_t*>(p);
        g_static[(size_t)dev] = std::move(s);
    }
    *out = g_static[(size_t)dev].get();
    return GENIE_OK;
}

static int static_scratch_init()
{
    StaticScratch* ss = nullptr;
    return static_scratch(&ss);
}

extern "C" {

int genie_hip_static_scratch_bytes(uint64_t* bytes)
{
    if (!bytes)
        return fail(GENIE_E_INVALID, "null argument");
    *bytes = kStaticScratchBytes;
    return GENIE_OK;
}

int genie_hip_match_topk_static(genie_index* ix, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                                const uint32_t* k_per_query, uint32_t count_bound, int32_t* out_ids,
                                int32_t* out_counts, void* stream)
{
    if (!ix || (!items && n_items) || !k_per_query || !out_ids || !out_counts)
        return fail(GENIE_E_INVALID, "null argument");
    if (n_queries == 0)
        return GENIE_OK;
    const DevIndex& d = ix->d;
    for (uint32_t q = 0; q < n_queries; ++q)
        if (k_per_query[q] == 0 || k_per_query[q] > GENIE_MAX_K)
            return fail(GENIE_E_INVALID, "k of query " + std::to_string(q) + " outside [1, GENIE_MAX_K]");
    if (int rc = validate_items(d, items, n_items, n_queries))
        return rc;
    if (n_items >= 0xFFFFFFFFull)
        return fail(GENIE_E_INVALID, "too many work items");
    // items grouped by query (stable), CSR pointer
    const uint32_t nq = n_queries;
    std::vector<uint32_t> qptr(nq + 1, 0);
    for (uint64_t i = 0; i < n_items; ++i)
        qptr[items[i].query + 1]++;
    for (uint32_t q = 0; q < nq; ++q)
        qptr[q + 1] += qptr[q];
    std::vector<genie_dim_t> sorted(n_items);
    {
        std::vector<uint32_t> fillp(qptr.begin(), qptr.end() - 1);
        for (uint64_t i = 0; i < n_items; ++i)
            sorted[fillp[items[i].query]++] = items[i];
    }
    uint32_t bound = count_bound;
    if (bound == 0) {
        std::vector<uint32_t> qb(nq);
        compute_bounds(d, sorted.data(), n_items, nq, qb.data());
        for (uint32_t v : qb)
            bound = std::max(bound, v);
    }
    const uint32_t lane_bits = bits_for(bound);
    // chunks, as the compile kernels will write them
    uint64_t nch = 0, span_total = 0;
    for (uint64_t i = 0; i < n_items; ++i) {
        const genie_dim_t& it = sorted[i];
        if (it.end_pos <= it.start_pos)
            continue;
        const uint64_t len = (uint64_t)(it.end_pos - it.start_pos);
        span_total += len;
        if (d.codec == 0) {
            nch += (len + kChunk - 1) / kChunk;
        } else {
            const genie_dim_t* pv = i ? &sorted[i - 1] : nullptr;
            const bool starts = !pv || pv->end_pos <= pv->start_pos || pv->query != it.query ||
                                pv->weight != it.weight || pv->end_pos != it.start_pos;
            nch += starts ? 1 : 0;
        }
    }
    // candidate space: the direct-mapped select emits at most k per tile and
    // query, so this space never overflows
    const uint32_t tile_objs = tile_objs_for(lane_bits);
    const uint64_t n_tiles = (d.n_objects + tile_objs - 1) / tile_objs;
    if (n_tiles * nq >= 0x7FFFFFFFull)
        return fail(GENIE_E_INVALID, "batch too large for one launch");
    std::vector<uint32_t> cap(nq);
    std::vector<uint64_t> cbase(nq), out_off(nq + 1, 0);
    uint64_t cand_total = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        const uint64_t c = n_tiles * k_per_query[q];
        if (c >= 0xFFFFFFFFull)
            return fail(GENIE_E_INVALID, "candidate space of a query too large");
        cap[q] = (uint32_t)c;
        cbase[q] = cand_total;
        cand_total += c;
        out_off[q + 1] = out_off[q] + k_per_query[q];
    }
    const uint64_t total_k = out_off[nq];
    const uint64_t nb = (n_items + 1023) / 1024 + 1, ncb = (nch + 1023) / 1024 + 1;
    // the layout, measured first, then carved out of the array
    auto layout = [&](Carve& c, StaticBufs& v) {
        v.d_items = c.take<genie_dim_t>(n_items);
        v.d_item_qptr = c.take<uint32_t>(nq + 1);
        v.d_nchunks = c.take<uint32_t>(n_items);
        v.d_chunk_ptr = c.take<uint32_t>(n_items);
        v.d_block_sums = c.take<uint32_t>(nb);
        v.d_total = c.take<uint32_t>(1);
        v.d_chunks = c.take<CChunk>(nch);
        v.d_nk = c.take<uint32_t>(nch);
        v.d_kptr = c.take<uint32_t>(nch);
        v.d_chunk_sums = c.take<uint32_t>(ncb);
        v.d_qc_ptr = c.take<uint32_t>(nq + 1);
        v.d_qk_ptr = c.take<uint32_t>(nq + 1);
        v.d_total_pairs = c.take<uint32_t>(1);
        v.d_kq = c.take<uint32_t>(nq);
        v.d_gthr = c.take<uint32_t>(nq);
        v.d_cand_cnt = c.take<uint32_t>(nq);
        v.d_cand_cap = c.take<uint32_t>(nq);
        v.d_cand_base = c.take<uint64_t>(nq);
        v.d_cand = c.take<uint64_t>(cand_total);
        v.d_overflow = c.take<uint32_t>(1);
        v.d_hov_any = c.take<uint32_t>(1);
        v.d_flag_any = c.take<uint32_t>(1);
        v.d_out_off = c.take<uint64_t>(nq + 1);
        v.d_out_ids = c.take<int32_t>(total_k);
        v.d_out_counts = c.take<int32_t>(total_k);
    };
    StaticBufs v{};
    Carve probe;
    layout(probe, v);
    if (probe.off > kStaticScratchBytes)
        return fail(GENIE_E_INVALID, "batch needs " + std::to_string(probe.off) + " bytes of device scratch; the static "
                                     "scratch holds " + std::to_string(kStaticScratchBytes) +
                                     " (genie_hip_match_topk allocates what a batch needs)");
    StaticScratch* ss = nullptr;
    if (int rc = static_scratch(&ss))
        return rc;
    std::lock_guard<std::mutex> lk(ss->mu);
    Carve c;
    c.base = ss->ptr;
    layout(c, v);
    hipStream_t s = (hipStream_t)stream;
    if (n_items)
        HIP_TRY(hipMemcpyAsync(v.d_items, sorted.data(), n_items * sizeof(genie_dim_t), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(v.d_item_qptr, qptr.data(), (nq + 1) * 4ull, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(v.d_kq, k_per_query, nq * 4ull, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(v.d_cand_cap, cap.data(), nq * 4ull, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(v.d_cand_base, cbase.data(), nq * 8ull, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(v.d_out_off, out_off.data(), (nq + 1) * 8ull, hipMemcpyHostToDevice, s));
    // what must start at zero: the array keeps the previous call's contents
    {
        ZeroList z{};
        z.p[0] = v.d_gthr, z.n[0] = nq;
        z.p[1] = v.d_cand_cnt, z.n[1] = nq;
        z.p[2] = v.d_overflow, z.n[2] = 1;
        z.p[3] = v.d_hov_any, z.n[3] = 1;
        k_zero_many<<<(uint32_t)std::min<uint64_t>((nq + 255) / 256, 1024), 256, 0, s>>>(z);
    }
    const uint32_t ni = (uint32_t)n_items;
    if (ni) {
        const uint32_t nbk = (ni + 1023) / 1024;
        if (d.codec == 0)
            k_chunk_count<<<(ni + 255) / 256, 256, 0, s>>>(v.d_items, ni, v.d_nchunks);
        else
            k_crun_flags<<<(ni + 255) / 256, 256, 0, s>>>(v.d_items, ni, d.seg_cpos, d.n_segments, v.d_nchunks);
        k_scan_blocks<<<nbk, 256, 0, s>>>(v.d_nchunks, v.d_chunk_ptr, ni, v.d_block_sums);
        k_scan_sums<<<1, 256, 0, s>>>(v.d_block_sums, nbk, v.d_total);
        k_scan_add<<<nbk, 256, 0, s>>>(v.d_chunk_ptr, ni, v.d_block_sums);
        if (d.codec == 0)
            k_chunk_write<<<(ni + 255) / 256, 256, 0, s>>>(v.d_items, ni, v.d_chunk_ptr, d.key_start, d.n_keys,
                                                          v.d_chunks);
        else
            k_crun_write<<<(ni + 255) / 256, 256, 0, s>>>(v.d_items, ni, d.seg_cpos, d.n_segments, d.seg_key,
                                                         v.d_chunk_ptr, v.d_chunks);
    } else {
        fill_u32(v.d_total, 0, 1, s);
    }
    k_query_chunk_ptr<<<(nq + 1 + 255) / 256, 256, 0, s>>>(v.d_item_qptr, nq, v.d_chunk_ptr, ni, v.d_total,
                                                          v.d_qc_ptr);
    if (nch) {
        const uint32_t ncu = (uint32_t)nch;
        const uint32_t g = std::min<uint32_t>((ncu + 255) / 256, 4096);
        k_chunk_nkeys<<<g, 256, 0, s>>>(v.d_chunks, v.d_total, v.d_nk);
        const uint32_t nbk = (ncu + 1023) / 1024;
        k_scan_blocks<<<nbk, 256, 0, s>>>(v.d_nk, v.d_kptr, ncu, v.d_chunk_sums);
        k_scan_sums<<<1, 256, 0, s>>>(v.d_chunk_sums, nbk, v.d_total_pairs);
        k_scan_add<<<nbk, 256, 0, s>>>(v.d_kptr, ncu, v.d_chunk_sums);
        k_chunk_set_kp<<<g, 256, 0, s>>>(v.d_chunks, v.d_total, v.d_kptr);
    } else {
        fill_u32(v.d_total_pairs, 0, 1, s);
    }
    k_query_pair_ptr<<<(nq + 1 + 255) / 256, 256, 0, s>>>(v.d_qc_ptr, nq, v.d_kptr, v.d_total, v.d_total_pairs,
                                                         v.d_qk_ptr);
    // the direct-mapped kernel over every query
    ScanArgs a{};
    a.codec = d.codec;
    a.words = d.words;
    a.seg_cpos = d.seg_cpos;
    a.seg_first = d.seg_first;
    a.seg_last = d.seg_last;
    a.key_seg = d.key_seg;
    a.seg_bph = (d.codec > 0 && codec_has_bp32(d.codec)) ? d.seg_bph : nullptr;
    a.sdir_off = d.sdir_off;
    a.sdir = d.sdir;
    a.post = d.post;
    a.key_start = d.key_start;
    a.dir_off = d.dir_off;
    a.dir = d.dir;
    a.chunks = v.d_chunks;
    a.qc_ptr = v.d_qc_ptr;
    a.qk_ptr = v.d_qk_ptr;
    a.kq = v.d_kq;
    a.qlist = nullptr;
    a.n_qlist = nq;
    a.n_objects = d.n_objects;
    a.n_queries = nq;
    a.tile_objs = tile_objs;
    a.n_tiles = (uint32_t)n_tiles;
    a.dt_per_tile = tile_objs / kDirTile;
    a.n_dtiles = d.n_dtiles;
    a.gthr = v.d_gthr;
    a.cand_cnt = v.d_cand_cnt;
    a.cand_cap = v.d_cand_cap;
    a.cand_base = v.d_cand_base;
    a.cand = v.d_cand;
    a.overflow = v.d_overflow;
    const uint64_t est = d.codec == 0 || d.n_words == 0
                             ? span_total
                             : (uint64_t)((double)span_total * (double)d.n_postings / (double)d.n_words);
    const bool sparse = (double)est / ((double)nq * (double)n_tiles) <= (double)kCntWords / 4;
    const uint32_t blocks = (uint32_t)(n_tiles * nq);
    if (sparse)
        launch_scan_t<true>(lane_bits, a, blocks, s);
    else
        launch_scan_t<false>(lane_bits, a, blocks, s);
    MergeArgs m;
    m.cand = v.d_cand;
    m.cand_base = v.d_cand_base;
    m.cand_cnt = v.d_cand_cnt;
    m.cand_cap = v.d_cand_cap;
    m.kq = v.d_kq;
    m.out_off = v.d_out_off;
    m.out_ids = v.d_out_ids;
    m.out_counts = v.d_out_counts;
    m.key_bits = 32 + lane_bits;
    m.flag_a = v.d_overflow;
    m.flag_b = v.d_hov_any;
    m.flag_any = v.d_flag_any;
    k_merge<<<nq, kMergeThreads, 0, s>>>(m);
    HIP_TRY(hipGetLastError());
    uint32_t ovf = 0;
    HIP_TRY(hipMemcpyAsync(out_ids, v.d_out_ids, total_k * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(out_counts, v.d_out_counts, total_k * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&ovf, v.d_flag_any, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (ovf)
        return fail(GENIE_E_RUNTIME, "candidate overflow in the static scratch");
    return GENIE_OK;
}

}  // extern "C"
