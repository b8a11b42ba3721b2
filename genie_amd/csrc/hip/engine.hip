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
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_set>
#include <vector>

#include "codecs_dev.h"
#include "engine.h"
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

// ---------------------------------------------------------------------------
// Small utility kernels
// ---------------------------------------------------------------------------

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s)
{
    if (n == 0)
        return;
    uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    k_fill_u32<<<blocks, 256, 0, s>>>(p, v, n);
}

// Exclusive scan of n u32 values, 1024 per block (256 threads x 4).
__global__ void __launch_bounds__(256) k_scan_blocks(const uint32_t* in, uint32_t* out, uint32_t n,
                                                     uint32_t* block_sums)
{
    __shared__ uint32_t s_w[4];
    const uint32_t base = blockIdx.x * 1024u + threadIdx.x * 4u;
    uint32_t v[4], t = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v[j] = (base + j < n) ? in[base + j] : 0u;
        t += v[j];
    }
    uint32_t incl = wave_incl_scan(t);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 63)
        s_w[w] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (int i = 0; i < w; ++i)
        woff += s_w[i];
    uint32_t run = woff + incl - t;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (base + j < n)
            out[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 255)
        block_sums[blockIdx.x] = woff + incl;
}

__global__ void __launch_bounds__(256) k_scan_sums(uint32_t* sums, uint32_t nb, uint32_t* total)
{
    __shared__ uint32_t s_w[4];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0)
        s_carry = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < nb; b0 += 256) {
        uint32_t i = b0 + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t incl = wave_incl_scan(v);
        const int w = threadIdx.x >> 6;
        if (lane_id() == 63)
            s_w[w] = incl;
        __syncthreads();
        uint32_t woff = s_carry;
        for (int j = 0; j < w; ++j)
            woff += s_w[j];
        if (i < nb)
            sums[i] = woff + incl - v;
        __syncthreads();
        if (threadIdx.x == 255)
            s_carry = woff + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *total = s_carry;
}

__global__ void k_scan_add(uint32_t* out, uint32_t n, const uint32_t* sums)
{
    uint32_t i = blockIdx.x * 1024u + threadIdx.x;
    uint32_t add = sums[blockIdx.x];
    for (int j = 0; j < 4; ++j, i += 256)
        if (i < n)
            out[i] += add;
}

// ---------------------------------------------------------------------------
// Index directory: for every long key and every directory-tile boundary b,
// the position inside the key's list of the first id >= b * kDirTile.
// ---------------------------------------------------------------------------

__global__ void k_build_dir(const uint32_t* __restrict__ post, const uint32_t* __restrict__ key_start,
                            const uint32_t* __restrict__ dir_keys, uint32_t n_dir_keys,
                            uint32_t n_dtiles, uint32_t* __restrict__ dir)
{
    const uint64_t per = (uint64_t)n_dtiles + 1;
    const uint64_t total = (uint64_t)n_dir_keys * per;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t dk = (uint32_t)(i / per), b = (uint32_t)(i % per);
        const uint32_t k = dir_keys[dk];
        const uint32_t s = key_start[k], e = key_start[k + 1];
        const uint64_t target = (uint64_t)b * kDirTile;
        uint32_t lo = s, hi = e;  // first position with post[p] >= target
        while (lo < hi) {
            uint32_t mid = lo + ((hi - lo) >> 1);
            if ((uint64_t)post[mid] < target)
                lo = mid + 1;
            else
                hi = mid;
        }
        dir[i] = lo - s;
    }
}

// ---------------------------------------------------------------------------
// Compressed index: segment bounds and segment directory
// ---------------------------------------------------------------------------

constexpr int kDecodeWaves = 4;

// Decode every segment once: first and last id, decoded length check
// (against the uncompressed segment positions) and ascending order.
__global__ void __launch_bounds__(kDecodeWaves * kWave)
    k_seg_bounds(int codec, const uint32_t* __restrict__ words, const uint32_t* __restrict__ cpos,
                 const uint32_t* __restrict__ upos, uint32_t n_seg, uint32_t n_objects,
                 uint32_t* __restrict__ first, uint32_t* __restrict__ last, uint32_t* __restrict__ err)
{
    __shared__ DecodeScratch S[kDecodeWaves];
    const int wv = threadIdx.x / kWave, lane = lane_id();
    for (uint32_t sgi = blockIdx.x * kDecodeWaves + wv; sgi < n_seg; sgi += gridDim.x * kDecodeWaves) {
        const uint32_t w0 = cpos[sgi], nw = cpos[sgi + 1] - w0;
        const uint32_t want = upos[sgi + 1] - upos[sgi];
        if (nw > kSegMax || want > kSegMax) {
            if (lane == 0)
                atomicOr(err, 1u);
            continue;
        }
        stage_words(S[wv], words, w0, nw);
        uint32_t v[kPerLane];
        const uint32_t n = wave_decode(codec, S[wv], nw, v);
        bool bad = n != want || n == 0;
        uint32_t prev = lane ? 0u : 0u;
        // ascending within the segment and inside the object range
        uint32_t lastv = (uint32_t)__shfl_up((int)v[kPerLane - 1], 1, kWave);
        if (lane == 0)
            lastv = 0;
        for (int j = 0; j < kPerLane; ++j) {
            const uint32_t i = lane * kPerLane + j;
            if (i < n) {
                const uint32_t before = j ? v[j - 1] : lastv;
                if ((i > 0 && v[j] < before) || v[j] >= n_objects)
                    bad = true;
            }
        }
        (void)prev;
        if (ballot(bad) != 0) {
            if (lane == 0)
                atomicOr(err, 1u);
            wave_sync();
            continue;
        }
        if (!bad && n) {
            if (lane == 0)
                first[sgi] = v[0];
            if (lane == (int)((n - 1) / kPerLane))
                last[sgi] = v[(n - 1) % kPerLane];
        }
        wave_sync();
    }
}

// For every long key and directory-tile boundary b: the first segment of the
// key (relative) whose last id is >= b * kDirTile.
__global__ void k_build_sdir(const uint32_t* __restrict__ seg_last, const uint32_t* __restrict__ key_seg,
                             const uint32_t* __restrict__ dir_keys, uint32_t n_dir_keys, uint32_t n_dtiles,
                             uint32_t* __restrict__ sdir)
{
    const uint64_t per = (uint64_t)n_dtiles + 1;
    const uint64_t total = (uint64_t)n_dir_keys * per;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = dir_keys[i / per], b = (uint32_t)(i % per);
        const uint32_t s = key_seg[k], e = key_seg[k + 1];
        const uint64_t target = (uint64_t)b * kDirTile;
        uint32_t lo = s, hi = e;
        while (lo < hi) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if ((uint64_t)seg_last[mid] < target)
                lo = mid + 1;
            else
                hi = mid;
        }
        sdir[i] = lo - s;
    }
}

// Stand-alone decode of n arrays (tests, measurement).
__global__ void __launch_bounds__(kDecodeWaves * kWave)
    k_decode_arrays(int codec, const uint32_t* __restrict__ words, const uint32_t* __restrict__ cpos,
                    uint32_t n_arrays, uint32_t stride, uint32_t* __restrict__ out, uint32_t* __restrict__ out_len)
{
    __shared__ DecodeScratch S[kDecodeWaves];
    const int wv = threadIdx.x / kWave, lane = lane_id();
    for (uint32_t a = blockIdx.x * kDecodeWaves + wv; a < n_arrays; a += gridDim.x * kDecodeWaves) {
        const uint32_t w0 = cpos[a], nw = cpos[a + 1] - w0;
        if (nw > kSegMax) {
            if (lane == 0)
                out_len[a] = 0xFFFFFFFFu;
            continue;
        }
        stage_words(S[wv], words, w0, nw);
        uint32_t v[kPerLane];
        const uint32_t n = wave_decode(codec, S[wv], nw, v);
        if (lane == 0)
            out_len[a] = n;
        if (n != 0xFFFFFFFFu) {
            const uint32_t lim = min(n, stride);
            for (int j = 0; j < kPerLane; ++j) {
                const uint32_t i = lane * kPerLane + j;
                if (i < lim)
                    out[(uint64_t)a * stride + i] = v[j];
            }
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// Device-side item compilation
// ---------------------------------------------------------------------------

// Last key whose first position is <= pos (key_start is non-decreasing).
__device__ __forceinline__ uint32_t key_of(const uint32_t* key_start, uint32_t n_keys, uint32_t pos)
{
    uint32_t lo = 0, hi = n_keys;  // answer in [0, n_keys)
    while (hi - lo > 1) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (key_start[mid] <= pos)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ void k_chunk_count(const genie_dim_t* __restrict__ items, uint32_t n_items,
                              uint32_t* __restrict__ nchunks)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items)
        return;
    const genie_dim_t it = items[i];
    uint32_t len = it.end_pos > it.start_pos ? (uint32_t)(it.end_pos - it.start_pos) : 0u;
    nchunks[i] = (len + kChunk - 1) / kChunk;
}

__global__ void k_chunk_write(const genie_dim_t* __restrict__ items, uint32_t n_items,
                              const uint32_t* __restrict__ chunk_ptr,
                              const uint32_t* __restrict__ key_start, uint32_t n_keys,
                              CChunk* __restrict__ chunks)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items)
        return;
    const genie_dim_t it = items[i];
    if (it.end_pos <= it.start_pos)
        return;
    const uint32_t w = (uint32_t)it.weight;
    uint32_t c = chunk_ptr[i];
    for (uint32_t sp = (uint32_t)it.start_pos; sp < (uint32_t)it.end_pos; sp += kChunk, ++c) {
        uint32_t ep = min((uint32_t)it.end_pos, sp + kChunk);
        CChunk ch;
        ch.k0 = key_of(key_start, n_keys, sp);
        ch.k1 = key_of(key_start, n_keys, ep - 1);
        ch.sp = sp;
        ch.ep = ep;
        ch.w = w;
        ch.q = (uint32_t)it.query;
        ch.kp = 0;
        ch.pad = 0;
        chunks[c] = ch;
    }
}

// Number of keys each chunk covers (for the pair prefix).
__global__ void k_chunk_nkeys(const CChunk* __restrict__ chunks, const uint32_t* __restrict__ total,
                              uint32_t* __restrict__ nk)
{
    const uint32_t n = *total;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        nk[c] = chunks[c].k1 - chunks[c].k0 + 1;
}

__global__ void k_chunk_set_kp(CChunk* __restrict__ chunks, const uint32_t* __restrict__ total,
                               const uint32_t* __restrict__ kptr)
{
    const uint32_t n = *total;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        chunks[c].kp = kptr[c];
}

// qk_ptr[q] = first pair of query q.
__global__ void k_query_pair_ptr(const uint32_t* __restrict__ qc_ptr, uint32_t n_queries,
                                 const uint32_t* __restrict__ kptr, const uint32_t* __restrict__ total_chunks,
                                 const uint32_t* __restrict__ total_pairs, uint32_t* __restrict__ qk_ptr)
{
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > n_queries)
        return;
    uint32_t c = qc_ptr[q];
    qk_ptr[q] = c >= *total_chunks ? *total_pairs : kptr[c];
}

// Compressed items cover whole segments; consecutive items of one query with
// the same weight and adjacent segments merge into one run (a chunk whose
// sp/ep are segment indices), so a workgroup jumps over a long list with the
// segment directory instead of visiting one item per segment.
__device__ __forceinline__ uint32_t seg_at(const uint32_t* cpos, uint32_t n_seg, uint32_t pos)
{
    uint32_t lo = 0, hi = n_seg;  // last s with cpos[s] <= pos
    while (hi - lo > 1) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (cpos[mid] <= pos)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ bool citem_segs(const genie_dim_t& it, const uint32_t* cpos, uint32_t n_seg,
                                           uint32_t* s0, uint32_t* s1)
{
    if (it.end_pos <= it.start_pos)
        return false;
    *s0 = seg_at(cpos, n_seg, (uint32_t)it.start_pos);
    *s1 = seg_at(cpos, n_seg, (uint32_t)it.end_pos - 1) + 1;
    return true;
}

__device__ __forceinline__ bool crun_starts(const genie_dim_t* items, uint32_t i, const uint32_t* cpos,
                                            uint32_t n_seg)
{
    uint32_t a0, a1, b0, b1;
    if (!citem_segs(items[i], cpos, n_seg, &a0, &a1))
        return false;
    if (i == 0 || !citem_segs(items[i - 1], cpos, n_seg, &b0, &b1))
        return true;
    return items[i].query != items[i - 1].query || items[i].weight != items[i - 1].weight || a0 != b1;
}

__global__ void k_crun_flags(const genie_dim_t* __restrict__ items, uint32_t n_items,
                             const uint32_t* __restrict__ cpos, uint32_t n_seg, uint32_t* __restrict__ flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_items)
        flag[i] = crun_starts(items, i, cpos, n_seg) ? 1u : 0u;
}

__global__ void k_crun_write(const genie_dim_t* __restrict__ items, uint32_t n_items,
                             const uint32_t* __restrict__ cpos, uint32_t n_seg,
                             const uint32_t* __restrict__ seg_key, const uint32_t* __restrict__ rptr,
                             CChunk* __restrict__ runs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items)
        return;
    uint32_t s0, s1;
    if (!citem_segs(items[i], cpos, n_seg, &s0, &s1))
        return;
    const bool start = crun_starts(items, i, cpos, n_seg);
    const uint32_t r = start ? rptr[i] : rptr[i] - 1;
    if (start) {
        runs[r].sp = s0;
        runs[r].k0 = seg_key[s0];
        runs[r].w = (uint32_t)items[i].weight;
        runs[r].q = (uint32_t)items[i].query;
        runs[r].kp = 0;
        runs[r].pad = 0;
    }
    if (i + 1 == n_items || crun_starts(items, i + 1, cpos, n_seg) ||
        items[i + 1].end_pos <= items[i + 1].start_pos) {
        runs[r].ep = s1;
        runs[r].k1 = seg_key[s1 - 1];
    }
}

// qc_ptr[q] = first chunk of query q (items sorted by query, item_qptr CSR).
__global__ void k_query_chunk_ptr(const uint32_t* __restrict__ item_qptr, uint32_t n_queries,
                                  const uint32_t* __restrict__ chunk_ptr, uint32_t n_items,
                                  const uint32_t* __restrict__ total, uint32_t* __restrict__ qc_ptr)
{
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > n_queries)
        return;
    uint32_t ip = item_qptr[q];
    qc_ptr[q] = ip >= n_items ? *total : chunk_ptr[ip];
}

// ---------------------------------------------------------------------------
// Scan + count + per-tile select (plain form)
// ---------------------------------------------------------------------------

constexpr int kScanThreads = 512;
constexpr int kScanWaves = kScanThreads / kWave;
constexpr int kCntWords = 16384;  // 64 KiB of counters per workgroup
constexpr uint32_t kEmitFail = 0xFFFFFFFFu;

struct ScanArgs {
    // compressed index (codec != 0)
    int codec;
    const uint32_t* words;
    const uint32_t* seg_cpos;
    const uint32_t* seg_first;
    const uint32_t* seg_last;
    const uint32_t* key_seg;
    const uint32_t* sdir_off;
    const uint32_t* sdir;
    // uncompressed index
    const uint32_t* post;
    const uint32_t* key_start;
    const uint32_t* dir_off;
    const uint32_t* dir;
    const CChunk* chunks;
    const uint32_t* qc_ptr;
    const uint32_t* qk_ptr;
    const uint32_t* kq;
    uint32_t n_objects, n_queries, n_tiles, tile_objs, dt_per_tile, n_dtiles;
    uint32_t* gthr;       // per query adaptive threshold (max of finished tiles' k-th counts)
    uint32_t* cand_cnt;   // per query candidates appended
    const uint32_t* cand_cap;
    const uint64_t* cand_base;
    uint64_t* cand;       // keys: count << 32 | ~id
    uint32_t* overflow;
};

template <int LB>
__device__ __forceinline__ uint32_t lane_count(uint32_t v, int j)
{
    if constexpr (LB == 32)
        return v;
    else
        return (v >> (LB * j)) & ((1u << LB) - 1u);
}

__device__ __forceinline__ uint64_t make_key(uint32_t count, uint32_t id)
{
    return ((uint64_t)count << 32) | (uint32_t)(~id);
}

// Count postings of one key sub-span [a, b) into the tile counters.  Lanes
// read 16-byte groups of ids, four groups per lane issued back to back (the
// loads carry no branches, so all four are in flight together: indexes past
// the end are clamped to the last group and their ids are not counted);
// the unaligned head and tail go through single-id loads.  FILTER: the span
// may hold ids outside the tile (short lists without a directory entry).
template <int LB, bool FILTER>
__device__ __forceinline__ void count_one(uint32_t id, uint32_t obj0, uint32_t tile_objs, uint32_t w,
                                          uint32_t* cnt)
{
    constexpr int PER = 32 / LB;
    constexpr int SH = PER == 4 ? 2 : (PER == 2 ? 1 : 0);
    const uint32_t loc = id - obj0;
    // Ids outside the tile (and lanes with nothing to add, w == 0) go to a
    // per-lane scratch word past the counters: no branch, no shared address.
    const bool ok = !FILTER || loc < tile_objs;
    const uint32_t word = ok ? (loc >> SH) : (uint32_t)(kCntWords + lane_id());
    atomicAdd(&cnt[word], ok ? w << ((loc & (PER - 1)) * LB) : 0u);
}

template <int LB, bool FILTER>
__device__ __forceinline__ void count_span(const uint32_t* __restrict__ post, uint32_t a, uint32_t b,
                                           uint32_t obj0, uint32_t tile_objs, uint32_t w, uint32_t* cnt)
{
    const int lane = lane_id();
    const uint32_t a4 = (a + 3u) & ~3u;        // first 16-byte aligned position
    const uint32_t b4 = b & ~3u;               // end of the aligned body
    if (a4 >= b4) {                            // short span: one id per lane
        for (uint32_t p = a + lane; p < b; p += kWave)
            count_one<LB, FILTER>(post[p], obj0, tile_objs, w, cnt);
        return;
    }
    if (lane < (int)(a4 - a))
        count_one<LB, FILTER>(post[a + lane], obj0, tile_objs, w, cnt);
    if (lane < (int)(b - b4))
        count_one<LB, FILTER>(post[b4 + lane], obj0, tile_objs, w, cnt);
    const uint4* body = reinterpret_cast<const uint4*>(post + a4);
    const uint32_t ng = (b4 - a4) >> 2;        // 16-byte groups, >= 1
    for (uint32_t g0 = 0; g0 < ng; g0 += 4 * kWave) {
        uint4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            v[u] = body[min(g0 + u * kWave + lane, ng - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            // lanes past the end count into the scratch word instead of
            // branching, which would let the compiler sink their load
            const bool in = g0 + u * kWave + lane < ng;
            const uint32_t past = obj0 + tile_objs;  // an id just outside the tile
            count_one<LB, true>(in ? v[u].x : past, obj0, tile_objs, w, cnt);
            count_one<LB, true>(in ? v[u].y : past, obj0, tile_objs, w, cnt);
            count_one<LB, true>(in ? v[u].z : past, obj0, tile_objs, w, cnt);
            count_one<LB, true>(in ? v[u].w : past, obj0, tile_objs, w, cnt);
        }
    }
}

// Shared scratch of the select passes.
struct SelShared {
    uint32_t hist[64];
    uint32_t wv[kScanWaves];
    uint32_t res[8];
};

// Exact per-tile top-k under (count desc, id asc), pruned by the query's
// adaptive threshold, appended to the query's candidate list.  Called by all
// threads of the workgroup after the counters are complete.
template <int LB>
__device__ __forceinline__ void tile_select(const ScanArgs& a, const uint32_t* cnt, uint32_t q, uint32_t obj0,
                                            SelShared& S)
{
    constexpr int PER = 32 / LB;
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
    uint32_t* s_hist = S.hist;
    uint32_t* s_wv = S.wv;
    uint32_t* s_res = S.res;
    // ---- select: exact tile top-k under (count desc, id asc) ------------
    const uint32_t k = a.kq[q];
    const uint32_t wpw = kCntWords / kScanWaves;  // words per wave (contiguous)
    const uint32_t wbeg = wv * wpw;

    uint32_t m = 0;
    for (uint32_t i = lane; i < wpw; i += kWave) {
        uint32_t v = cnt[wbeg + i];
#pragma unroll
        for (int j = 0; j < PER; ++j)
            m = max(m, lane_count<LB>(v, j));
    }
    m = wave_max(m);
    if (lane == 0)
        s_wv[wv] = m;
    if (tid == 0)  // one read, shared, so every thread takes the same branches
        s_res[6] = __hip_atomic_load(&a.gthr[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    uint32_t M = 0;
    for (int i = 0; i < kScanWaves; ++i)
        M = max(M, s_wv[i]);
    const uint32_t lower = max(1u, s_res[6]);
    if (M < lower)
        return;  // nothing in this tile can reach the query's top-k

    // Threshold search in 64-wide count windows from the top.
    uint32_t top = M, cum = 0, ct = lower, ngt = 0, neq = 0, found = 0;
    for (;;) {
        const uint32_t wlo = top >= lower + 63 ? top - 63 : lower;
        __syncthreads();
        if (tid < 64)
            s_hist[tid] = 0;
        __syncthreads();
        for (uint32_t i = lane; i < wpw; i += kWave) {
            uint32_t v = cnt[wbeg + i];
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                uint32_t c = lane_count<LB>(v, j);
                hist_add_aggregated(s_hist, c - wlo, c >= wlo && c <= top);
            }
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t f = 0, cc = wlo, g = cum, e = 0;
            for (uint32_t c = top + 1; c-- > wlo;) {
                uint32_t h = s_hist[c - wlo];
                if (g + h >= k) {
                    f = 1;
                    cc = c;
                    e = h;
                    break;
                }
                g += h;
            }
            s_res[0] = f;
            s_res[1] = cc;
            s_res[2] = g;
            s_res[3] = e;
        }
        __syncthreads();
        found = s_res[0];
        cum = s_res[2];
        if (found) {
            ct = s_res[1];
            ngt = cum;
            neq = s_res[3];
            break;
        }
        if (wlo == lower) {
            ct = lower;  // fewer than k objects reach `lower`: emit them all
            ngt = cum;
            neq = 0;
            break;
        }
        top = wlo - 1;
    }

    const uint32_t need = found ? k - ngt : 0;
    const bool trim = found && neq > need;
    const uint32_t nemit = found ? ngt + min(neq, need) : ngt;
    if (tid == 0) {
        uint32_t base = atomicAdd(&a.cand_cnt[q], nemit);
        if ((uint64_t)base + nemit > a.cand_cap[q]) {
            atomicOr(a.overflow, 1u);
            base = kEmitFail;
        }
        if (found)
            atomicMax(&a.gthr[q], ct);
        s_res[4] = base;
        s_res[5] = 0;  // unordered slot counter
    }
    __syncthreads();
    const uint32_t base = s_res[4];
    if (base == kEmitFail || nemit == 0)
        return;
    uint64_t* out = a.cand + a.cand_base[q] + base;

    // Unordered emission: objects strictly above ct (and, when nothing has to
    // be trimmed, those equal to ct as well).
    const uint32_t cut = trim ? ct + 1 : ct;
    for (uint32_t i = lane; i < wpw; i += kWave) {
        uint32_t v = cnt[wbeg + i];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            uint32_t c = lane_count<LB>(v, j);
            bool take = c >= cut;
            uint64_t bm = ballot(take);
            if (bm) {
                uint32_t wbase = 0;
                if (lane == __builtin_ctzll(bm))
                    wbase = atomicAdd(&s_res[5], (uint32_t)__popcll(bm));
                wbase = (uint32_t)__shfl((int)wbase, __builtin_ctzll(bm), kWave);
                if (take) {
                    uint32_t id = obj0 + (wbeg + i) * PER + j;
                    out[wbase + lane_prefix(bm)] = make_key(c, id);
                }
            }
        }
    }
    if (!trim)
        return;

    // Ordered emission of ties at ct: the `need` smallest ids.  Pass A counts
    // ties per wave range, pass B ranks them in id order.
    uint32_t e = 0;
    for (uint32_t i = lane; i < wpw; i += kWave) {
        uint32_t v = cnt[wbeg + i];
#pragma unroll
        for (int j = 0; j < PER; ++j)
            e += lane_count<LB>(v, j) == ct;
    }
    e = wave_sum(e);
    __syncthreads();
    if (lane == 0)
        s_wv[wv] = e;
    __syncthreads();
    uint32_t run = 0;
    for (int i = 0; i < wv; ++i)
        run += s_wv[i];
    for (uint32_t i0 = 0; i0 < wpw && run < need; i0 += kWave) {
        const uint32_t i = i0 + lane;
        uint32_t v = i < wpw ? cnt[wbeg + i] : 0u;
        uint32_t mask = 0;
#pragma unroll
        for (int j = 0; j < PER; ++j)
            mask |= (uint32_t)(i < wpw && lane_count<LB>(v, j) == ct) << j;
        const uint32_t pc = (uint32_t)__popc(mask);
        const uint32_t incl = wave_incl_scan(pc);
        uint32_t r = run + incl - pc;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (mask & (1u << j)) {
                if (r < need) {
                    uint32_t id = obj0 + (wbeg + i) * PER + j;
                    out[ngt + r] = make_key(ct, id);
                }
                ++r;
            }
        }
        run += (uint32_t)__shfl((int)incl, 63, kWave);
    }
}

// Sub-span of one (chunk, key) pair inside the workgroup's tile.
struct Span {
    uint32_t lo, hi;   // posting positions
    uint32_t wf;       // weight << 1 | filter
};
constexpr int kSpanCap = 1024;

template <int LB>
__global__ void __launch_bounds__(kScanThreads) k_scan_tile(ScanArgs a)
{
    __shared__ uint4 s_cnt4[kCntWords / 4 + kWave / 4];  // counters + per-lane scratch
    uint32_t* cnt = reinterpret_cast<uint32_t*>(s_cnt4);
    __shared__ Span s_span[kSpanCap];
    __shared__ uint32_t s_next;
    __shared__ SelShared s_sel;

    const uint32_t t = blockIdx.x / a.n_queries;
    const uint32_t q = blockIdx.x % a.n_queries;
    const uint32_t obj0 = t * a.tile_objs;
    const uint32_t dt0 = t * a.dt_per_tile;
    const uint32_t dt1 = min(dt0 + a.dt_per_tile, a.n_dtiles);
    const int tid = threadIdx.x, lane = lane_id();

    for (int i = tid; i < kCntWords / 4; i += kScanThreads)
        s_cnt4[i] = make_uint4(0, 0, 0, 0);

    // ---- count ------------------------------------------------------------
    // Rounds of up to kSpanCap (chunk, key) pairs: every thread resolves the
    // pairs of its chunks inside the round's window (independent loads, so
    // the metadata latency is paid once per round), then the waves take the
    // resolved spans dynamically.
    const uint32_t c0 = a.qc_ptr[q], c1 = a.qc_ptr[q + 1];
    const uint32_t pbase = a.qk_ptr[q], npairs = a.qk_ptr[q + 1] - pbase;
    for (uint32_t w0 = 0; w0 < npairs; w0 += kSpanCap) {
        const uint32_t wn = min((uint32_t)kSpanCap, npairs - w0);
        if (tid == 0)
            s_next = 0;
        for (uint32_t c = c0 + tid; c < c1; c += kScanThreads) {
            const CChunk ch = a.chunks[c];
            const uint32_t cp = ch.kp - pbase;                  // chunk's first pair, query-relative
            const uint32_t nk = ch.k1 - ch.k0 + 1;
            if (cp + nk <= w0 || cp >= w0 + wn)
                continue;
            const uint32_t j0 = cp < w0 ? w0 - cp : 0;
            const uint32_t j1 = min(nk, w0 + wn - cp);
            for (uint32_t j = j0; j < j1; ++j) {
                const uint32_t k = ch.k0 + j;
                const uint32_t ks = a.key_start[k], ke = a.key_start[k + 1];
                uint32_t lo = max(ch.sp, ks), hi = min(ch.ep, ke);
                const uint32_t doff = a.dir_off[k];
                uint32_t filter = 1;
                if (doff != kNoDir) {
                    const uint32_t* d = a.dir + doff;
                    lo = max(lo, ks + d[dt0]);
                    hi = min(hi, ks + d[dt1]);
                    filter = 0;
                }
                Span sp;
                sp.lo = lo;
                sp.hi = max(lo, hi);
                sp.wf = ch.w << 1 | filter;
                s_span[cp + j - w0] = sp;
            }
        }
        __syncthreads();
        for (;;) {
            uint32_t si = 0;
            if (lane == 0)
                si = atomicAdd(&s_next, 1u);
            si = bcast_first(__shfl((int)si, 0, kWave));
            if (si >= wn)
                break;
            const Span sp = s_span[si];
            if (sp.lo < sp.hi) {
                if (sp.wf & 1)
                    count_span<LB, true>(a.post, sp.lo, sp.hi, obj0, a.tile_objs, sp.wf >> 1, cnt);
                else
                    count_span<LB, false>(a.post, sp.lo, sp.hi, obj0, a.tile_objs, sp.wf >> 1, cnt);
            }
        }
        __syncthreads();
    }
    if (npairs == 0)
        __syncthreads();  // the zeroing above must finish before the select passes

    tile_select<LB>(a, cnt, q, obj0, s_sel);
}

// Compressed form of k_scan_tile: pairs resolve to ranges of compressed
// segments overlapping the tile (segment directory for long keys), the waves
// take single segments, stage and decode them (codecs_dev.h) and count the
// decoded ids that fall inside the tile.
template <int LB>
__global__ void __launch_bounds__(kScanThreads) k_scan_tile_c(ScanArgs a)
{
    __shared__ uint4 s_cnt4[kCntWords / 4 + kWave / 4];
    uint32_t* cnt = reinterpret_cast<uint32_t*>(s_cnt4);
    __shared__ Span s_span[kSpanCap];
    __shared__ uint32_t s_pre[kSpanCap + 1];   // prefix of segments per span
    __shared__ DecodeScratch s_dec[kScanWaves];
    __shared__ uint32_t s_next;
    __shared__ uint32_t s_wsum[kScanWaves];
    __shared__ SelShared s_sel;

    const uint32_t t = blockIdx.x / a.n_queries;
    const uint32_t q = blockIdx.x % a.n_queries;
    const uint32_t obj0 = t * a.tile_objs;
    const uint32_t obj1 = obj0 + a.tile_objs;
    const uint32_t dt0 = t * a.dt_per_tile;
    const uint32_t dt1 = min(dt0 + a.dt_per_tile, a.n_dtiles);
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;

    for (int i = tid; i < kCntWords / 4; i += kScanThreads)
        s_cnt4[i] = make_uint4(0, 0, 0, 0);

    const uint32_t c0 = a.qc_ptr[q], c1 = a.qc_ptr[q + 1];
    const uint32_t pbase = a.qk_ptr[q], npairs = a.qk_ptr[q + 1] - pbase;
    for (uint32_t w0 = 0; w0 < npairs; w0 += kSpanCap) {
        const uint32_t wn = min((uint32_t)kSpanCap, npairs - w0);
        if (tid == 0)
            s_next = 0;
        for (uint32_t c = c0 + tid; c < c1; c += kScanThreads) {
            const CChunk ch = a.chunks[c];
            const uint32_t cp = ch.kp - pbase;
            const uint32_t nk = ch.k1 - ch.k0 + 1;
            if (cp + nk <= w0 || cp >= w0 + wn)
                continue;
            const uint32_t j0 = cp < w0 ? w0 - cp : 0;
            const uint32_t j1 = min(nk, w0 + wn - cp);
            for (uint32_t j = j0; j < j1; ++j) {
                const uint32_t k = ch.k0 + j;
                const uint32_t ks = a.key_seg[k], ke = a.key_seg[k + 1];
                uint32_t lo = max(ch.sp, ks), hi = min(ch.ep, ke);
                const uint32_t doff = a.sdir_off[k];
                if (doff != kNoDir) {
                    const uint32_t* d = a.sdir + doff;
                    const uint32_t mid = ks + d[dt1];
                    const uint32_t end = mid + (mid < ke && a.seg_first[mid] < obj1 ? 1u : 0u);
                    lo = max(lo, ks + d[dt0]);
                    hi = min(hi, end);
                }
                Span sp;
                sp.lo = lo;
                sp.hi = max(lo, hi);
                sp.wf = ch.w << 1 | 1u;
                s_span[cp + j - w0] = sp;
            }
        }
        __syncthreads();
        // prefix of segment counts over the window's spans (2 per thread)
        {
            const uint32_t i0 = 2 * tid, i1 = 2 * tid + 1;
            const uint32_t n0 = i0 < wn ? s_span[i0].hi - s_span[i0].lo : 0u;
            const uint32_t n1 = i1 < wn ? s_span[i1].hi - s_span[i1].lo : 0u;
            const uint32_t incl = wave_incl_scan(n0 + n1);
            if (lane == kWave - 1)
                s_wsum[wv] = incl;
            __syncthreads();
            uint32_t off = 0;
            for (int i = 0; i < wv; ++i)
                off += s_wsum[i];
            const uint32_t ex = off + incl - (n0 + n1);
            if (i0 <= wn)
                s_pre[i0] = ex;
            if (i1 <= wn)
                s_pre[i1] = ex + n0;
            if (i1 + 1 == wn)
                s_pre[wn] = ex + n0 + n1;  // the total, also when wn is the full window
        }
        __syncthreads();
        const uint32_t nsegs = s_pre[wn];
        DecodeScratch& D = s_dec[wv];
        for (;;) {
            uint32_t g = 0;
            if (lane == 0)
                g = atomicAdd(&s_next, 1u);
            g = bcast_first(__shfl((int)g, 0, kWave));
            if (g >= nsegs)
                break;
            uint32_t lo = 0, hi = wn;  // span with s_pre[span] <= g < s_pre[span + 1]
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_pre[mid] <= g)
                    lo = mid;
                else
                    hi = mid;
            }
            const Span sp = s_span[lo];
            const uint32_t sg = sp.lo + (g - s_pre[lo]);
            if (a.seg_last[sg] < obj0 || a.seg_first[sg] >= obj1)
                continue;  // segment entirely outside the tile
            const uint32_t cw0 = a.seg_cpos[sg], nw = a.seg_cpos[sg + 1] - cw0;
            stage_words(D, a.words, cw0, nw);
            uint32_t v[kPerLane];
            const uint32_t n = wave_decode(a.codec, D, nw, v);
            const uint32_t w = sp.wf >> 1;
#pragma unroll
            for (int j = 0; j < kPerLane; ++j) {
                const uint32_t i = lane * kPerLane + j;
                count_one<LB, true>(i < n ? v[j] : obj1, obj0, a.tile_objs, w, cnt);
            }
            wave_sync();  // the staging buffer is reused by the next segment
        }
        __syncthreads();
    }
    if (npairs == 0)
        __syncthreads();
    tile_select<LB>(a, cnt, q, obj0, s_sel);
}

// ---------------------------------------------------------------------------
// Merge: exact top-k of each query's candidate keys, sorted descending.
// ---------------------------------------------------------------------------

constexpr int kMergeThreads = 256;
constexpr int kMergeMaxK = GENIE_MAX_K;

struct MergeArgs {
    const uint64_t* cand;
    const uint64_t* cand_base;
    const uint32_t* cand_cnt;
    const uint32_t* cand_cap;
    const uint32_t* kq;
    const uint64_t* out_off;
    int32_t* out_ids;
    int32_t* out_counts;
    uint32_t key_bits;  // significant bits of a key (32 + bits of the count bound)
};

__global__ void __launch_bounds__(kMergeThreads) k_merge(MergeArgs a)
{
    __shared__ uint64_t s_keys[kMergeMaxK];
    __shared__ uint32_t s_hist[256];
    __shared__ uint32_t s_res[4];
    const uint32_t q = blockIdx.x;
    const int tid = threadIdx.x;
    const uint32_t k = a.kq[q];
    const uint32_t M = min(a.cand_cnt[q], a.cand_cap[q]);
    const uint64_t* keys = a.cand + a.cand_base[q];

    uint64_t kth = 0;  // every key >= kth is in the result
    if (M > k) {
        uint64_t prefix = 0, mask = 0;
        uint32_t remaining = k;
        int shift = ((int)((a.key_bits + 7) / 8) - 1) * 8;
        for (; shift >= 0; shift -= 8) {
            if (tid < 256)
                s_hist[tid] = 0;
            __syncthreads();
            for (uint32_t i0 = 0; i0 < M; i0 += kMergeThreads) {
                const uint32_t i = i0 + tid;
                uint64_t key = i < M ? keys[i] : 0;
                bool take = i < M && (key & mask) == prefix;
                hist_add_aggregated(s_hist, (uint32_t)(key >> shift) & 255u, take);
            }
            __syncthreads();
            if (tid == 0) {
                uint32_t cum = 0, b = 255;
                for (int d = 255; d >= 0; --d) {
                    if (cum + s_hist[d] >= remaining) {
                        b = (uint32_t)d;
                        break;
                    }
                    cum += s_hist[d];
                }
                s_res[0] = b;
                s_res[1] = remaining - cum;
            }
            __syncthreads();
            prefix |= (uint64_t)s_res[0] << shift;
            mask |= (uint64_t)255 << shift;
            remaining = s_res[1];
            __syncthreads();
        }
        kth = prefix;
    }
    // collect the selected keys
    if (tid == 0)
        s_res[2] = 0;
    uint32_t n2 = 1;
    const uint32_t nsel = min(M, k);
    while (n2 < nsel)
        n2 <<= 1;
    for (uint32_t i = tid; i < n2; i += kMergeThreads)
        s_keys[i] = 0;
    __syncthreads();
    for (uint32_t i0 = 0; i0 < M; i0 += kMergeThreads) {
        const uint32_t i = i0 + tid;
        uint64_t key = i < M ? keys[i] : 0;
        bool take = i < M && key >= kth;
        uint64_t bm = ballot(take);
        if (bm) {
            uint32_t wb = 0;
            if (lane_id() == __builtin_ctzll(bm))
                wb = atomicAdd(&s_res[2], (uint32_t)__popcll(bm));
            wb = (uint32_t)__shfl((int)wb, __builtin_ctzll(bm), kWave);
            if (take) {
                uint32_t slot = wb + lane_prefix(bm);
                if (slot < n2)
                    s_keys[slot] = key;
            }
        }
    }
    __syncthreads();
    // bitonic sort, descending
    for (uint32_t size = 2; size <= n2; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = tid; i < n2; i += kMergeThreads) {
                uint32_t j = i ^ stride;
                if (j > i) {
                    bool desc = (i & size) == 0;
                    uint64_t x = s_keys[i], y = s_keys[j];
                    if (desc ? (x < y) : (x > y)) {
                        s_keys[i] = y;
                        s_keys[j] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    const uint64_t o = a.out_off[q];
    for (uint32_t i = tid; i < k; i += kMergeThreads) {
        uint64_t key = i < nsel ? s_keys[i] : 0;
        if (key == 0) {
            a.out_ids[o + i] = 0;
            a.out_counts[o + i] = 0;
        } else {
            a.out_ids[o + i] = (int32_t)(~(uint32_t)key);
            a.out_counts[o + i] = (int32_t)(key >> 32);
        }
    }
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
    uint32_t cap_factor = 8;
    uint64_t logical_bytes = 0, compulsory_bytes = 0;
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
    uint32_t* d_cand_cnt = nullptr;
    uint32_t* d_cand_cap = nullptr;
    uint64_t* d_cand_base = nullptr;
    uint64_t* d_cand = nullptr;
    uint64_t cand_alloc = 0;
    uint32_t* d_overflow = nullptr;
    uint64_t* d_out_off = nullptr;
    int32_t* d_out_ids = nullptr;
    int32_t* d_out_counts = nullptr;
    uint64_t out_alloc = 0;
    std::vector<uint32_t> h_kq;
    std::vector<uint64_t> h_out_off;
    // timing
    bool timing = false;
    hipEvent_t ev[8] = {};
    double ms_compile = 0, ms_scan = 0, ms_merge = 0;
    bool ran = false;
};

static uint32_t bits_for(uint32_t bound)
{
    return bound <= 255 ? 8u : (bound <= 65535 ? 16u : 32u);
}

static uint32_t tile_objs_for(uint32_t lane_bits) { return (uint32_t)kCntWords * (32u / lane_bits); }

static void plan_free_buffers(genie_plan* p)
{
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
    dev_free(p->d_cand_cnt);
    dev_free(p->d_cand_cap);
    dev_free(p->d_cand_base);
    dev_free(p->d_cand);
    dev_free(p->d_overflow);
    dev_free(p->d_out_off);
    dev_free(p->d_out_ids);
    dev_free(p->d_out_counts);
}

// Count bound of every query: per (query, order) group the largest weight
// times how often one object can occur in the group's spans (1 per distinct
// dim for single-value tables; the sum of per-key multiplicities for
// bijectMap tables), summed over groups.
static int compute_bounds(const DevIndex& d, const genie_dim_t* items, uint64_t n_items,
                          uint32_t n_queries, uint32_t* out)
{
    struct G {
        uint32_t q;
        int32_t order;
        uint32_t w;
        uint32_t key;
    };
    std::vector<G> v;
    v.reserve(n_items * 2);
    const uint64_t* ks = d.h_key_start.data();
    const uint32_t nk = d.n_keys;
    auto key_of_pos = [&](uint64_t pos) -> uint32_t {
        return (uint32_t)(std::upper_bound(ks, ks + nk + 1, pos) - ks - 1);
    };
    for (uint64_t i = 0; i < n_items; ++i) {
        const genie_dim_t& it = items[i];
        if (it.end_pos <= it.start_pos)
            continue;
        uint32_t k0, k1;
        if (d.codec == 0) {
            k0 = key_of_pos((uint64_t)it.start_pos);
            k1 = key_of_pos((uint64_t)it.end_pos - 1);
        } else {
            const uint32_t* cp = d.h_seg_cpos.data();
            uint32_t s0 = (uint32_t)(std::upper_bound(cp, cp + d.n_segments + 1, (uint32_t)it.start_pos) - cp - 1);
            uint32_t s1 = (uint32_t)(std::upper_bound(cp, cp + d.n_segments + 1, (uint32_t)it.end_pos - 1) - cp - 1);
            k0 = d.h_seg_key[s0];
            k1 = d.h_seg_key[s1];
        }
        for (uint32_t k = k0; k <= k1; ++k)
            v.push_back({(uint32_t)it.query, it.order, (uint32_t)it.weight, k});
    }
    std::sort(v.begin(), v.end(), [](const G& a, const G& b) {
        if (a.q != b.q) return a.q < b.q;
        if (a.order != b.order) return a.order < b.order;
        return a.key < b.key;
    });
    std::vector<uint64_t> acc(n_queries, 0);
    for (size_t i = 0; i < v.size();) {
        size_t j = i;
        uint32_t wmax = 0;
        uint64_t mult = 0;
        std::unordered_set<uint32_t> dims;
        uint32_t last_key = 0xFFFFFFFFu;
        for (; j < v.size() && v[j].q == v[i].q && v[j].order == v[i].order; ++j) {
            wmax = std::max(wmax, v[j].w);
            if (v[j].key == last_key)
                continue;
            last_key = v[j].key;
            if (d.single_value)
                dims.insert(d.h_key_dim[v[j].key]);
            else
                mult += d.h_key_mult[v[j].key];
        }
        if (d.single_value)
            mult = dims.size();
        acc[v[i].q] += (uint64_t)wmax * mult;
        i = j;
    }
    for (uint32_t q = 0; q < n_queries; ++q)
        out[q] = (uint32_t)std::min<uint64_t>(acc[q], 0xFFFFFFFFull);
    return GENIE_OK;
}

static int validate_items(const DevIndex& d, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries)
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
            const uint32_t* cp = d.h_seg_cpos.data();
            if (!std::binary_search(cp, cp + d.n_segments + 1, (uint32_t)it.start_pos) ||
                !std::binary_search(cp, cp + d.n_segments + 1, (uint32_t)it.end_pos))
                return fail(GENIE_E_INVALID, "item " + std::to_string(i) +
                                                 ": compressed span does not align with segments");
        }
    }
    return GENIE_OK;
}

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
    d.h_seg_key = seg_key;
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
    std::vector<uint32_t> sdir_off(K, kNoDir), dir_keys;
    for (uint32_t k = 0; k < K; ++k)
        if (key_seg[k + 1] - key_seg[k] > 4) {
            sdir_off[k] = (uint32_t)(dir_keys.size() * (uint64_t)(d.n_dtiles + 1));
            dir_keys.push_back(k);
        }
    if ((uint64_t)dir_keys.size() * (d.n_dtiles + 1) >= 0xFFFFFFFFull) {
        delete ix;
        return fail(GENIE_E_INVALID, "segment directory too large");
    }
    int rc = GENIE_OK;
    uint32_t *d_upos = nullptr, *d_dir_keys = nullptr, *d_err = nullptr;
    do {
        if ((rc = dev_alloc(&d.words, d.n_words, &d.bytes))) break;
        if ((rc = dev_alloc(&d.seg_cpos, S + 1, &d.bytes))) break;
        if ((rc = dev_alloc(&d.seg_key, S, &d.bytes))) break;
        if ((rc = dev_alloc(&d.seg_first, S, &d.bytes))) break;
        if ((rc = dev_alloc(&d.seg_last, S, &d.bytes))) break;
        if ((rc = dev_alloc(&d.key_seg, K + 1, &d.bytes))) break;
        if ((rc = dev_alloc(&d.dir_off, K, &d.bytes))) break;  // segment directory offsets
        if ((rc = dev_alloc(&d.sdir, dir_keys.size() * (uint64_t)(d.n_dtiles + 1), &d.bytes))) break;
        if ((rc = dev_alloc(&d_upos, S + 1))) break;
        if ((rc = dev_alloc(&d_dir_keys, dir_keys.size()))) break;
        if ((rc = dev_alloc(&d_err, 1))) break;
        hipError_t e;
        const uint32_t zero = 0;
        if ((e = hipMemcpy(d.words, desc->words, d.n_words * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d.seg_cpos, cp, (S + 1) * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d_upos, up, (S + 1) * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d.seg_key, seg_key.data(), S * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d.key_seg, key_seg.data(), (K + 1) * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d.dir_off, sdir_off.data(), K * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d_dir_keys, dir_keys.data(), dir_keys.size() * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(d_err, &zero, 4, hipMemcpyHostToDevice))) {
            rc = fail(GENIE_E_RUNTIME, std::string("index upload: ") + hipGetErrorString(e));
            break;
        }
        const uint32_t blocks = std::min<uint32_t>((S + kDecodeWaves - 1) / kDecodeWaves, 65535);
        k_seg_bounds<<<blocks, kDecodeWaves * kWave>>>(d.codec, d.words, d.seg_cpos, d_upos, S, d.n_objects,
                                                        d.seg_first, d.seg_last, d_err);
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
            const uint64_t total = dir_keys.size() * (uint64_t)(d.n_dtiles + 1);
            const uint32_t g = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
            k_build_sdir<<<g, 256>>>(d.seg_last, d.key_seg, d_dir_keys, (uint32_t)dir_keys.size(), d.n_dtiles,
                                      d.sdir);
            if ((e = hipGetLastError()) || (e = hipDeviceSynchronize())) {
                rc = fail(GENIE_E_RUNTIME, std::string("segment directory: ") + hipGetErrorString(e));
                break;
            }
        }
    } while (0);
    dev_free(d_upos);
    dev_free(d_dir_keys);
    dev_free(d_err);
    if (rc) {
        genie_hip_index_destroy(ix);
        return rc;
    }
    *out = ix;
    return GENIE_OK;
}

extern "C" {

const char* genie_hip_last_error(void) { return g_err.c_str(); }

int genie_hip_stream_synchronize(void* stream)
{
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
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
    return GENIE_OK;
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
    d.h_key_dim.assign(desc->key_dim, desc->key_dim + desc->n_keys);
    d.h_key_mult.assign(desc->n_keys, 1);
    std::vector<uint32_t> ks32(desc->n_keys + 1), dir_off(desc->n_keys, kNoDir), dir_keys;
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
        if (e - s > kShortList) {
            dir_off[k] = (uint32_t)(dir_keys.size() * (uint64_t)(d.n_dtiles + 1));
            dir_keys.push_back(k);
        }
    }
    if ((uint64_t)dir_keys.size() * (d.n_dtiles + 1) >= 0xFFFFFFFFull) {
        delete ix;
        return fail(GENIE_E_INVALID, "tile directory too large");
    }
    int rc = GENIE_OK;
    uint32_t* d_dir_keys = nullptr;
    do {
        if ((rc = dev_alloc(&d.post, d.n_postings, &d.bytes))) break;
        if ((rc = dev_alloc(&d.key_start, d.n_keys + 1, &d.bytes))) break;
        if ((rc = dev_alloc(&d.dir_off, d.n_keys, &d.bytes))) break;
        if ((rc = dev_alloc(&d.dir, dir_keys.size() * (uint64_t)(d.n_dtiles + 1), &d.bytes))) break;
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
            uint64_t total = dir_keys.size() * (uint64_t)(d.n_dtiles + 1);
            uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
            k_build_dir<<<blocks, 256>>>(d.post, d.key_start, d_dir_keys, (uint32_t)dir_keys.size(),
                                         d.n_dtiles, d.dir);
            if ((e = hipGetLastError()) || (e = hipDeviceSynchronize())) {
                rc = fail(GENIE_E_RUNTIME, std::string("directory build: ") + hipGetErrorString(e));
                break;
            }
        }
    } while (0);
    dev_free(d_dir_keys);
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
    dev_free(d.key_seg);
    dev_free(d.sdir);
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
    if (mode != 0)
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
        if ((rc = dev_alloc(&p->d_cand_cnt, max_queries))) break;
        if ((rc = dev_alloc(&p->d_cand_cap, max_queries))) break;
        if ((rc = dev_alloc(&p->d_cand_base, max_queries))) break;
        if ((rc = dev_alloc(&p->d_overflow, 1))) break;
        if ((rc = dev_alloc(&p->d_out_off, max_queries + 1))) break;
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

int genie_hip_plan_destroy(genie_plan* p)
{
    if (!p)
        return GENIE_OK;
    plan_free_buffers(p);
    for (auto& e : p->ev)
        if (e)
            (void)hipEventDestroy(e);
    delete p;
    return GENIE_OK;
}

// Candidate capacity per query: cap_factor tiles' worth of k, at most all tiles.
static int plan_size_candidates(genie_plan* p)
{
    const DevIndex& d = p->ix->d;
    const uint32_t n_tiles = (d.n_objects + tile_objs_for(p->lane_bits) - 1) / tile_objs_for(p->lane_bits);
    const uint32_t f = std::min<uint32_t>(p->cap_factor, n_tiles);
    std::vector<uint32_t> cap(p->n_queries);
    std::vector<uint64_t> base(p->n_queries);
    uint64_t tot = 0;
    for (uint32_t q = 0; q < p->n_queries; ++q) {
        cap[q] = f * p->h_kq[q];
        base[q] = tot;
        tot += cap[q];
    }
    if (tot > p->cand_alloc) {
        dev_free(p->d_cand);
        p->cand_alloc = 0;
        if (int rc = dev_alloc(&p->d_cand, tot))
            return rc;
        p->cand_alloc = tot;
    }
    p->total_cap = tot;
    HIP_TRY(hipMemcpy(p->d_cand_cap, cap.data(), cap.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->d_cand_base, base.data(), base.size() * 8, hipMemcpyHostToDevice));
    return GENIE_OK;
}

int genie_hip_plan_upload(genie_plan* p, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                          const uint32_t* k_per_query)
{
    if (!p || (!items && n_items) || !k_per_query)
        return fail(GENIE_E_INVALID, "null argument");
    if (n_queries == 0 || n_queries > p->max_queries || n_items > p->max_items)
        return fail(GENIE_E_INVALID, "batch larger than the plan");
    const DevIndex& d = p->ix->d;
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
    // count bound of this batch
    if (p->count_bound == 0) {
        std::vector<uint32_t> b(n_queries);
        compute_bounds(d, sorted.data(), n_items, n_queries, b.data());
        uint32_t mx = 0;
        for (uint32_t v : b)
            mx = std::max(mx, v);
        p->lane_bits = bits_for(mx);
    }
    // chunks and algorithmic bytes
    uint64_t nch = 0, logical = 0;
    std::vector<std::pair<uint32_t, uint32_t>> spans;
    spans.reserve(n_items);
    bool prev_ok = false;
    for (uint64_t i = 0; i < n_items; ++i) {
        const genie_dim_t& it = sorted[i];
        if (it.end_pos > it.start_pos) {
            uint64_t len = (uint64_t)(it.end_pos - it.start_pos);
            if (d.codec == 0) {
                nch += (len + kChunk - 1) / kChunk;
            } else {
                // runs of contiguous segments (k_crun_flags uses the same rule)
                const genie_dim_t& pv = sorted[i - (i ? 1 : 0)];
                const bool starts = i == 0 || !prev_ok || pv.query != it.query || pv.weight != it.weight ||
                                    pv.end_pos != it.start_pos;
                nch += starts ? 1 : 0;
            }
            logical += 4 * len;
            spans.push_back({(uint32_t)it.start_pos, (uint32_t)it.end_pos});
        }
        prev_ok = it.end_pos > it.start_pos;
    }
    std::sort(spans.begin(), spans.end());
    uint64_t compulsory = 0, cur_e = 0;
    for (auto& s : spans) {  // union of position ranges
        uint64_t a = std::max<uint64_t>(s.first, cur_e);
        if (s.second > a)
            compulsory += 4 * (s.second - a);
        cur_e = std::max<uint64_t>(cur_e, s.second);
    }
    p->logical_bytes = logical;
    p->compulsory_bytes = compulsory;
    if (nch > p->chunks_cap) {
        dev_free(p->d_chunks);
        dev_free(p->d_nk);
        dev_free(p->d_kptr);
        dev_free(p->d_chunk_sums);
        p->chunks_cap = 0;
        if (int rc = dev_alloc(&p->d_chunks, nch))
            return rc;
        if (int rc = dev_alloc(&p->d_nk, nch))
            return rc;
        if (int rc = dev_alloc(&p->d_kptr, nch))
            return rc;
        if (int rc = dev_alloc(&p->d_chunk_sums, (nch + 1023) / 1024 + 1))
            return rc;
        p->chunks_cap = nch;
    }
    p->n_queries = n_queries;
    p->n_items = n_items;
    p->n_chunks = nch;
    p->h_kq.assign(k_per_query, k_per_query + n_queries);
    p->h_out_off.assign(n_queries + 1, 0);
    for (uint32_t q = 0; q < n_queries; ++q)
        p->h_out_off[q + 1] = p->h_out_off[q] + k_per_query[q];
    p->total_k = p->h_out_off[n_queries];
    if (p->total_k > p->out_alloc) {
        dev_free(p->d_out_ids);
        dev_free(p->d_out_counts);
        p->out_alloc = 0;
        if (int rc = dev_alloc(&p->d_out_ids, p->total_k))
            return rc;
        if (int rc = dev_alloc(&p->d_out_counts, p->total_k))
            return rc;
        p->out_alloc = p->total_k;
    }
    if (n_items)
        HIP_TRY(hipMemcpy(p->d_items, sorted.data(), n_items * sizeof(genie_dim_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->d_item_qptr, qptr.data(), qptr.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->d_kq, k_per_query, n_queries * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->d_out_off, p->h_out_off.data(), p->h_out_off.size() * 8, hipMemcpyHostToDevice));
    p->cap_factor = 8;
    return plan_size_candidates(p);
}

static int launch_scan(genie_plan* p, hipStream_t s)
{
    const DevIndex& d = p->ix->d;
    ScanArgs a;
    a.codec = d.codec;
    a.words = d.words;
    a.seg_cpos = d.seg_cpos;
    a.seg_first = d.seg_first;
    a.seg_last = d.seg_last;
    a.key_seg = d.key_seg;
    a.sdir_off = d.dir_off;
    a.sdir = d.sdir;
    a.post = d.post;
    a.key_start = d.key_start;
    a.dir_off = d.dir_off;
    a.dir = d.dir;
    a.chunks = p->d_chunks;
    a.qc_ptr = p->d_qc_ptr;
    a.qk_ptr = p->d_qk_ptr;
    a.kq = p->d_kq;
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
    const uint64_t blocks = (uint64_t)a.n_tiles * p->n_queries;
    if (blocks >= 0x7FFFFFFFull)
        return fail(GENIE_E_INVALID, "batch too large for one launch");
    if (d.codec == 0) {
        switch (p->lane_bits) {
        case 8: k_scan_tile<8><<<(uint32_t)blocks, kScanThreads, 0, s>>>(a); break;
        case 16: k_scan_tile<16><<<(uint32_t)blocks, kScanThreads, 0, s>>>(a); break;
        default: k_scan_tile<32><<<(uint32_t)blocks, kScanThreads, 0, s>>>(a); break;
        }
    } else {
        switch (p->lane_bits) {
        case 8: k_scan_tile_c<8><<<(uint32_t)blocks, kScanThreads, 0, s>>>(a); break;
        case 16: k_scan_tile_c<16><<<(uint32_t)blocks, kScanThreads, 0, s>>>(a); break;
        default: k_scan_tile_c<32><<<(uint32_t)blocks, kScanThreads, 0, s>>>(a); break;
        }
    }
    return GENIE_OK;
}

static int plan_run_once(genie_plan* p, hipStream_t s)
{
    const DevIndex& d = p->ix->d;
    const uint32_t nq = p->n_queries;
    const uint32_t ni = (uint32_t)p->n_items;
    if (p->timing)
        HIP_TRY(hipEventRecord(p->ev[0], s));
    fill_u32(p->d_gthr, 0, nq, s);
    fill_u32(p->d_cand_cnt, 0, nq, s);
    fill_u32(p->d_overflow, 0, 1, s);
    // compile items into chunks (uncompressed) or segment runs (compressed)
    if (ni && d.codec == 0) {
        k_chunk_count<<<(ni + 255) / 256, 256, 0, s>>>(p->d_items, ni, p->d_nchunks);
        const uint32_t nb = (ni + 1023) / 1024;
        k_scan_blocks<<<nb, 256, 0, s>>>(p->d_nchunks, p->d_chunk_ptr, ni, p->d_block_sums);
        k_scan_sums<<<1, 256, 0, s>>>(p->d_block_sums, nb, p->d_total);
        k_scan_add<<<nb, 256, 0, s>>>(p->d_chunk_ptr, ni, p->d_block_sums);
        k_chunk_write<<<(ni + 255) / 256, 256, 0, s>>>(p->d_items, ni, p->d_chunk_ptr, d.key_start,
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
    k_merge<<<nq, kMergeThreads, 0, s>>>(m);
    if (p->timing)
        HIP_TRY(hipEventRecord(p->ev[3], s));
    HIP_TRY(hipGetLastError());
    p->ran = true;
    return GENIE_OK;
}

int genie_hip_plan_run(genie_plan* p, void* stream)
{
    if (!p)
        return fail(GENIE_E_INVALID, "null plan");
    if (p->n_queries == 0)
        return fail(GENIE_E_NOT_BUILT, "no batch uploaded");
    return plan_run_once(p, (hipStream_t)stream);
}

int genie_hip_plan_download(genie_plan* p, void* stream, int32_t* out_ids, int32_t* out_counts)
{
    if (!p || !out_ids || !out_counts)
        return fail(GENIE_E_INVALID, "null argument");
    if (!p->ran)
        return fail(GENIE_E_NOT_BUILT, "plan has not run");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipStreamSynchronize(s));
    // candidate overflow: grow the capacity and run again (the reference
    // retries on hash-table overflow in the same way, match.cu:356-421)
    for (int attempt = 0;; ++attempt) {
        uint32_t ov = 0;
        HIP_TRY(hipMemcpy(&ov, p->d_overflow, 4, hipMemcpyDeviceToHost));
        if (!ov)
            break;
        const DevIndex& d = p->ix->d;
        const uint32_t n_tiles = (d.n_objects + tile_objs_for(p->lane_bits) - 1) / tile_objs_for(p->lane_bits);
        if (p->cap_factor >= n_tiles || attempt > 8)
            return fail(GENIE_E_RUNTIME, "candidate overflow at full capacity");
        p->cap_factor = std::min<uint32_t>(p->cap_factor * 4, n_tiles);
        if (int rc = plan_size_candidates(p))
            return rc;
        if (int rc = plan_run_once(p, s))
            return rc;
        HIP_TRY(hipStreamSynchronize(s));
    }
    if (p->timing) {
        float a = 0, b = 0, c = 0;
        HIP_TRY(hipEventElapsedTime(&a, p->ev[0], p->ev[1]));
        HIP_TRY(hipEventElapsedTime(&b, p->ev[1], p->ev[2]));
        HIP_TRY(hipEventElapsedTime(&c, p->ev[2], p->ev[3]));
        p->ms_compile = a;
        p->ms_scan = b;
        p->ms_merge = c;
    }
    HIP_TRY(hipMemcpy(out_ids, p->d_out_ids, p->total_k * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(out_counts, p->d_out_counts, p->total_k * 4, hipMemcpyDeviceToHost));
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
        if (!strcmp(name, "compile") && hipEventElapsedTime(&v, p->ev[0], p->ev[1]) == hipSuccess) { *ms = v; return GENIE_OK; }
        if (!strcmp(name, "scan") && hipEventElapsedTime(&v, p->ev[1], p->ev[2]) == hipSuccess) { *ms = v; return GENIE_OK; }
        if (!strcmp(name, "merge") && hipEventElapsedTime(&v, p->ev[2], p->ev[3]) == hipSuccess) { *ms = v; return GENIE_OK; }
    }
    return fail(GENIE_E_INVALID, std::string("no timing for kernel ") + name);
}

int genie_hip_plan_bytes(genie_plan* p, uint64_t* logical, uint64_t* compulsory)
{
    if (!p || !logical || !compulsory)
        return fail(GENIE_E_INVALID, "null argument");
    *logical = p->logical_bytes;
    *compulsory = p->compulsory_bytes;
    return GENIE_OK;
}

int genie_hip_match_topk(genie_index* ix, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                         const uint32_t* k_per_query, uint32_t count_bound, int32_t* out_ids,
                         int32_t* out_counts)
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
    int rc = genie_hip_plan_create(ix, n_queries, n_items, max_k, count_bound, 0, &p);
    if (rc)
        return rc;
    if (!(rc = genie_hip_plan_upload(p, items, n_items, n_queries, k_per_query)) &&
        !(rc = genie_hip_plan_run(p, nullptr)))
        rc = genie_hip_plan_download(p, nullptr, out_ids, out_counts);
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
