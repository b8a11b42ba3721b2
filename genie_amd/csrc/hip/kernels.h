// Device kernels of the match-count engine (plain form) and the shared
// count / select / merge building blocks.  Included by engine.hip only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "codecs_dev.h"
#include "engine.h"
#include "wave.h"

namespace gh {

// ---------------------------------------------------------------------------
// Small utility kernels
// ---------------------------------------------------------------------------

// Zeroes up to 6 arrays in one launch (the per-run resets of a plan).
struct ZeroList {
    uint32_t* p[6];
    uint64_t n[6];
};

__global__ void k_zero_many(ZeroList z)
{
    const uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
#pragma unroll
    for (int a = 0; a < 6; ++a)
        for (uint64_t i = i0; i < z.n[a]; i += st)
            z.p[a][i] = 0;
}

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
// Index directory: for every key with a row of grade s and every row entry j,
// the position inside the key's list of the first id >= (j << s) * kDirTile.
// One wave per row.
// ---------------------------------------------------------------------------

__global__ void k_build_dir(const uint32_t* __restrict__ post, const uint32_t* __restrict__ key_start,
                            const uint32_t* __restrict__ dir_keys, uint32_t n_dir_keys,
                            const uint32_t* __restrict__ dir_off, uint32_t n_dtiles, uint32_t* __restrict__ dir)
{
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t dk = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; dk < n_dir_keys; dk += waves) {
        const uint32_t k = dir_keys[dk];
        const uint32_t doff = dir_off[k], sh = doff & kDirShiftMask;
        uint32_t* row = dir + (doff & ~kDirShiftMask);
        const uint32_t n = dir_row_len(n_dtiles, sh);
        const uint32_t s = key_start[k], e = key_start[k + 1];
        for (uint32_t j = lane_id(); j < n; j += kWave) {
            const uint64_t target = ((uint64_t)j << sh) * kDirTile;
            uint32_t lo = s, hi = e;  // first position with post[p] >= target
            while (lo < hi) {
                const uint32_t mid = lo + ((hi - lo) >> 1);
                if ((uint64_t)post[mid] < target)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            row[j] = lo - s;
        }
    }
}

// Clips postings [lo, hi) of a key whose list starts at ks to the directory
// tiles [dt0, dt1) through the key's row (packed offset doff).  Returns true
// when the clipped span holds ids of those tiles only (both ends fall on row
// entries); otherwise the span also covers parts of the neighbouring tiles
// of the row's grade and its ids need the tile filter.
__device__ __forceinline__ bool dir_clip(const uint32_t* __restrict__ dir, uint32_t doff, uint32_t ks,
                                         uint32_t dt0, uint32_t dt1, uint32_t n_dtiles, uint32_t& lo, uint32_t& hi)
{
    const uint32_t s = doff & kDirShiftMask, m = (1u << s) - 1;
    const uint32_t* d = dir + (doff & ~kDirShiftMask);
    lo = max(lo, ks + d[dt0 >> s]);
    hi = min(hi, ks + d[(dt1 + m) >> s]);
    return (dt0 & m) == 0 && ((dt1 & m) == 0 || dt1 == n_dtiles);
}

// Clips segments [lo, hi) of a key whose segments are [ks, ke) to those that
// can hold ids of the directory tiles [dt0, dt1) (objects below obj1) through
// the key's segment-directory row: from the first segment whose last id
// reaches the row entry at or below dt0, to the first segment whose last id
// reaches the row entry at or above dt1, that one included when it starts
// below obj1 (later segments start at or after its last id).
__device__ __forceinline__ void sdir_clip(const uint32_t* __restrict__ sdir, uint32_t doff, uint32_t ks, uint32_t ke,
                                          uint32_t dt0, uint32_t dt1, uint32_t obj1,
                                          const uint32_t* __restrict__ seg_first, uint32_t& lo, uint32_t& hi)
{
    const uint32_t s = doff & kDirShiftMask, m = (1u << s) - 1;
    const uint32_t* d = sdir + (doff & ~kDirShiftMask);
    const uint32_t mid = ks + d[(dt1 + m) >> s];
    const uint32_t end = mid + (mid < ke && seg_first[mid] < obj1 ? 1u : 0u);
    lo = max(lo, ks + d[dt0 >> s]);
    hi = min(hi, end);
}

// ---------------------------------------------------------------------------
// Compressed index: segment bounds and segment directory
// ---------------------------------------------------------------------------

constexpr int kDecodeWaves = 4;

// Decode every segment once: first and last id, decoded length check
// (against the uncompressed segment positions) and ascending order.  With
// `post` set, the decoded ids of a valid segment are also written to
// post[upos[s] .. upos[s + 1]) (the uncompressed posting array).
__global__ void __launch_bounds__(kDecodeWaves * kWave)
    k_seg_bounds(int codec, const uint32_t* __restrict__ words, const uint32_t* __restrict__ cpos,
                 const uint32_t* __restrict__ upos, uint32_t n_seg, uint32_t n_objects,
                 uint32_t* __restrict__ first, uint32_t* __restrict__ last, uint32_t* __restrict__ err,
                 uint32_t* __restrict__ post)
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
            if (post) {
                const uint32_t base = upos[sgi];
                for (int j = 0; j < kPerLane; ++j) {
                    const uint32_t i = lane * kPerLane + j;
                    if (i < n)
                        post[base + i] = v[j];
                }
            }
        }
        wave_sync();
    }
}

// For every key with a segment-directory row of grade s and every row entry
// j: the first segment of the key (relative) whose last id is >=
// (j << s) * kDirTile.  One wave per row.
__global__ void k_build_sdir(const uint32_t* __restrict__ seg_last, const uint32_t* __restrict__ key_seg,
                             const uint32_t* __restrict__ dir_keys, uint32_t n_dir_keys,
                             const uint32_t* __restrict__ sdir_off, uint32_t n_dtiles, uint32_t* __restrict__ sdir)
{
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t dk = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; dk < n_dir_keys; dk += waves) {
        const uint32_t k = dir_keys[dk];
        const uint32_t doff = sdir_off[k], sh = doff & kDirShiftMask;
        uint32_t* row = sdir + (doff & ~kDirShiftMask);
        const uint32_t n = dir_row_len(n_dtiles, sh);
        const uint32_t s = key_seg[k], e = key_seg[k + 1];
        for (uint32_t j = lane_id(); j < n; j += kWave) {
            const uint64_t target = ((uint64_t)j << sh) * kDirTile;
            uint32_t lo = s, hi = e;
            while (lo < hi) {
                const uint32_t mid = lo + ((hi - lo) >> 1);
                if ((uint64_t)seg_last[mid] < target)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            row[j] = lo - s;
        }
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
    const uint4* seg_bph;     // BP32 block header positions per segment (null: walk the headers)
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
    const uint32_t* qlist;  // queries of the direct-mapped kernels (null: all)
    uint32_t n_qlist;
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
// Touched-word bitmap of the sparse select: counter word c of the tile is
// bit (c >> 6) & 31 of bitmap word (c >> 11) * 64 + (c & 63), so that wave v
// lane l of the select owns bitmap word 64 v + l, i.e. the counter words
// 2048 v + 64 i + l (i = 0..31): the words the dense passes give that lane,
// always in LDS bank l.
constexpr int kTouchWords = kCntWords / 32;
static_assert(kTouchWords == kScanThreads, "one bitmap word per select thread");
// A tile with at most this many touched counter words takes the sparse
// select; a busier one (a query with long lists in a sparse batch) the dense.
constexpr uint32_t kSparseWords = kCntWords / 8;

__device__ __forceinline__ void mark_touched(uint32_t* bm, uint32_t word)
{
    atomicOr(&bm[((word >> 11) << 6) | (word & 63u)], 1u << ((word >> 6) & 31u));
}

// MARK: the first increment of a counter word also sets its bit in `bm`.
template <int LB, bool FILTER, bool MARK = false>
__device__ __forceinline__ void count_one(uint32_t id, uint32_t obj0, uint32_t tile_objs, uint32_t w,
                                          uint32_t* cnt, uint32_t* bm = nullptr)
{
    constexpr int PER = 32 / LB;
    constexpr int SH = PER == 4 ? 2 : (PER == 2 ? 1 : 0);
    const uint32_t loc = id - obj0;
    // Ids outside the tile (and lanes with nothing to add, w == 0) go to a
    // per-lane scratch word past the counters: no branch, no shared address.
    const bool ok = !FILTER || loc < tile_objs;
    const uint32_t word = ok ? (loc >> SH) : (uint32_t)(kCntWords + lane_id());
    const uint32_t inc = ok ? w << ((loc & (PER - 1)) * LB) : 0u;
    if constexpr (MARK) {
        const uint32_t old = atomicAdd(&cnt[word], inc);
        if (old == 0 && inc != 0 && word < (uint32_t)kCntWords)
            mark_touched(bm, word);
    } else {
        atomicAdd(&cnt[word], inc);
    }
}

template <int LB, bool FILTER, bool MARK = false>
__device__ __forceinline__ void count_span(const uint32_t* __restrict__ post, uint32_t a, uint32_t b,
                                           uint32_t obj0, uint32_t tile_objs, uint32_t w, uint32_t* cnt,
                                           uint32_t* bm = nullptr)
{
    const int lane = lane_id();
    const uint32_t a4 = (a + 3u) & ~3u;        // first 16-byte aligned position
    const uint32_t b4 = b & ~3u;               // end of the aligned body
    if (a4 >= b4) {                            // short span: one id per lane
        for (uint32_t p = a + lane; p < b; p += kWave)
            count_one<LB, FILTER, MARK>(post[p], obj0, tile_objs, w, cnt, bm);
        return;
    }
    if (lane < (int)(a4 - a))
        count_one<LB, FILTER, MARK>(post[a + lane], obj0, tile_objs, w, cnt, bm);
    if (lane < (int)(b - b4))
        count_one<LB, FILTER, MARK>(post[b4 + lane], obj0, tile_objs, w, cnt, bm);
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
            count_one<LB, true, MARK>(in ? v[u].x : past, obj0, tile_objs, w, cnt, bm);
            count_one<LB, true, MARK>(in ? v[u].y : past, obj0, tile_objs, w, cnt, bm);
            count_one<LB, true, MARK>(in ? v[u].z : past, obj0, tile_objs, w, cnt, bm);
            count_one<LB, true, MARK>(in ? v[u].w : past, obj0, tile_objs, w, cnt, bm);
        }
    }
}

// Shared scratch of the select passes.
constexpr int kMaxWaves = 16;
struct SelShared {
    uint32_t hist[64];
    uint32_t wv[kMaxWaves];
    uint32_t res[8];
};

// Count accessors of the select passes.  A "word" holds PER objects.
// Plain form: LB-bit counters packed PER = 32 / LB to a u32 word.
template <int LB>
struct PackedAcc {
    static constexpr int PER = 32 / LB;
    const uint32_t* cnt;
    uint32_t nwords;
    __device__ __forceinline__ uint32_t word(uint32_t i) const { return cnt[i]; }
    __device__ __forceinline__ static uint32_t count(uint32_t v, int j) { return lane_count<LB>(v, j); }
};

// Batched form: one u64 word per object holding every query's LB-bit lane;
// `shift` selects the lane of the query being selected.
template <int LB>
struct LaneAcc {
    static constexpr int PER = 1;
    const uint64_t* cnt;
    uint32_t nwords;
    uint32_t shift;
    __device__ __forceinline__ uint32_t word(uint32_t i) const
    {
        if constexpr (LB == 32)
            return (uint32_t)(cnt[i] >> shift);
        else
            return (uint32_t)(cnt[i] >> shift) & ((1u << LB) - 1u);
    }
    __device__ __forceinline__ static uint32_t count(uint32_t v, int) { return v; }
};

// Exact per-tile top-k under (count desc, id asc), pruned by the query's
// adaptive threshold, appended to the query's candidate list.  Called by all
// threads of the workgroup after the counters are complete.
template <class Acc>
__device__ __forceinline__ void tile_select(const ScanArgs& a, const Acc& acc, uint32_t q, uint32_t obj0,
                                            SelShared& S)
{
    constexpr int PER = Acc::PER;
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
    const int nwaves = blockDim.x / kWave;
    uint32_t* s_hist = S.hist;
    uint32_t* s_wv = S.wv;
    uint32_t* s_res = S.res;
    // ---- select: exact tile top-k under (count desc, id asc) ------------
    const uint32_t k = a.kq[q];
    const uint32_t wpw = acc.nwords / nwaves;  // words per wave (contiguous)
    const uint32_t wbeg = wv * wpw;

    uint32_t m = 0;
    for (uint32_t i = lane; i < wpw; i += kWave) {
        uint32_t v = acc.word(wbeg + i);
#pragma unroll
        for (int j = 0; j < PER; ++j)
            m = max(m, Acc::count(v, j));
    }
    m = wave_max(m);
    if (lane == 0)
        s_wv[wv] = m;
    if (tid == 0)  // one read, shared, so every thread takes the same branches
        s_res[6] = __hip_atomic_load(&a.gthr[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    uint32_t M = 0;
    for (int i = 0; i < nwaves; ++i)
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
            uint32_t v = acc.word(wbeg + i);
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                uint32_t c = Acc::count(v, j);
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
        uint32_t v = acc.word(wbeg + i);
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            uint32_t c = Acc::count(v, j);
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
        uint32_t v = acc.word(wbeg + i);
#pragma unroll
        for (int j = 0; j < PER; ++j)
            e += Acc::count(v, j) == ct;
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
        uint32_t v = i < wpw ? acc.word(wbeg + i) : 0u;
        uint32_t mask = 0;
#pragma unroll
        for (int j = 0; j < PER; ++j)
            mask |= (uint32_t)(i < wpw && Acc::count(v, j) == ct) << j;
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

// Sparse form of tile_select for the plain kernels' packed counters: the same
// threshold search and the same emitted keys, but every pass visits only the
// counter words marked in `bm` (kTouchWords bitmap words; the lane owning
// bitmap word 64 v + l handles counter words 2048 v + 64 i + l).  Unmarked
// words hold only zero counts, which no pass would use (every window, cut
// and tie count is >= 1).  Passes over a lane's marked words run per lane;
// the ordered tie emission keeps the dense passes' word order.
template <int LB>
__device__ __forceinline__ void tile_select_sparse(const ScanArgs& a, const uint32_t* cnt, const uint32_t* bm,
                                                   uint32_t q, uint32_t obj0, SelShared& S)
{
    constexpr int PER = 32 / LB;
    constexpr uint32_t kWpw = kCntWords / kScanWaves;  // counter words per wave
    static_assert(kWpw == 64 * 32, "one bitmap word per lane covers 32 counter words");
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
    uint32_t* s_hist = S.hist;
    uint32_t* s_wv = S.wv;
    uint32_t* s_res = S.res;
    const uint32_t k = a.kq[q];
    const uint32_t mine = bm[tid];
    const uint32_t* cw = cnt + wv * kWpw + lane;  // word i of this lane: cw[64 * i]

    uint32_t m = 0;
    for (uint32_t b = mine; b; b &= b - 1) {
        const uint32_t v = cw[64 * __builtin_ctz(b)];
#pragma unroll
        for (int j = 0; j < PER; ++j)
            m = max(m, lane_count<LB>(v, j));
    }
    m = wave_max(m);
    if (lane == 0)
        s_wv[wv] = m;
    if (tid == 0)
        s_res[6] = __hip_atomic_load(&a.gthr[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    uint32_t M = 0;
    for (int i = 0; i < kScanWaves; ++i)
        M = max(M, s_wv[i]);
    const uint32_t lower = max(1u, s_res[6]);
    if (M < lower)
        return;

    uint32_t top = M, cum = 0, ct = lower, ngt = 0, neq = 0, found = 0;
    for (;;) {
        const uint32_t wlo = top >= lower + 63 ? top - 63 : lower;
        __syncthreads();
        if (tid < 64)
            s_hist[tid] = 0;
        __syncthreads();
        for (uint32_t b = mine; b; b &= b - 1) {
            const uint32_t v = cw[64 * __builtin_ctz(b)];
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const uint32_t c = lane_count<LB>(v, j);
                if (c >= wlo && c <= top)
                    atomicAdd(&s_hist[c - wlo], 1u);
            }
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t f = 0, cc = wlo, g = cum, e = 0;
            for (uint32_t c = top + 1; c-- > wlo;) {
                const uint32_t h = s_hist[c - wlo];
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
            ct = lower;
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
        s_res[5] = 0;
    }
    __syncthreads();
    const uint32_t base = s_res[4];
    if (base == kEmitFail || nemit == 0)
        return;
    uint64_t* out = a.cand + a.cand_base[q] + base;

    // unordered emission of the keys at or above the cut
    const uint32_t cut = trim ? ct + 1 : ct;
    for (uint32_t b = mine; b; b &= b - 1) {
        const uint32_t i = (uint32_t)__builtin_ctz(b);
        const uint32_t v = cw[64 * i];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t c = lane_count<LB>(v, j);
            if (c >= cut) {
                const uint32_t slot = atomicAdd(&s_res[5], 1u);
                const uint32_t id = obj0 + (wv * kWpw + 64 * i + lane) * PER + j;
                out[slot] = make_key(c, id);
            }
        }
    }
    if (!trim)
        return;

    // Ordered emission of the `need` smallest ids tied at ct, in the word
    // order of the dense passes (wave ranges, then i, then lane).
    uint32_t e = 0;
    for (uint32_t b = mine; b; b &= b - 1) {
        const uint32_t v = cw[64 * __builtin_ctz(b)];
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
    uint32_t any = mine;  // iterations where some lane of the wave has a marked word
    for (int off = 32; off > 0; off >>= 1)
        any |= (uint32_t)__shfl_xor((int)any, off, kWave);
    for (uint32_t bb = any; bb && run < need; bb &= bb - 1) {
        const uint32_t i = (uint32_t)__builtin_ctz(bb);
        const uint32_t v = (mine >> i) & 1u ? cw[64 * i] : 0u;
        uint32_t mask = 0;
#pragma unroll
        for (int j = 0; j < PER; ++j)
            mask |= (uint32_t)(lane_count<LB>(v, j) == ct) << j;
        const uint32_t pc = (uint32_t)__popc(mask);
        const uint32_t incl = wave_incl_scan(pc);
        uint32_t r = run + incl - pc;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (mask & (1u << j)) {
                if (r < need) {
                    const uint32_t id = obj0 + (wv * kWpw + 64 * i + lane) * PER + j;
                    out[ngt + r] = make_key(ct, id);
                }
                ++r;
            }
        }
        run += (uint32_t)__shfl((int)incl, 63, kWave);
    }
}

// Select of a plain kernel: sparse when few counter words were touched.
template <int LB, bool SPARSE>
__device__ __forceinline__ void plain_select(const ScanArgs& a, const uint32_t* cnt, const uint32_t* bm, uint32_t q,
                                             uint32_t obj0, SelShared& S)
{
    if constexpr (SPARSE) {
        const int lane = lane_id(), wv = threadIdx.x >> 6;
        const uint32_t pc = wave_sum((uint32_t)__popc(bm[threadIdx.x]));
        if (lane == 0)
            S.wv[wv] = pc;
        __syncthreads();
        uint32_t touched = 0;
        for (int i = 0; i < kScanWaves; ++i)
            touched += S.wv[i];
        __syncthreads();  // S.wv is reused by the select
        if (touched <= kSparseWords) {
            tile_select_sparse<LB>(a, cnt, bm, q, obj0, S);
            return;
        }
    }
    tile_select(a, PackedAcc<LB>{cnt, (uint32_t)kCntWords}, q, obj0, S);
}

// Sub-span of one (chunk, key) pair inside the workgroup's tile.
struct Span {
    uint32_t lo, hi;   // posting positions
    uint32_t wf;       // weight << 1 | filter
};
constexpr int kSpanCap = 1024;

// Resolves the window [w0, w0 + wn) of a query's (chunk, key) pairs into
// posting spans inside directory tiles [dt0, dt1): every thread takes the
// pairs of its chunks (independent loads, so the metadata latency is paid
// once per window).  Spans of short keys (no directory entry) cover the whole
// list and carry the filter flag.  Called by all threads.
__device__ __forceinline__ void resolve_spans(const ScanArgs& a, uint32_t c0, uint32_t c1, uint32_t pbase, uint32_t w0,
                                              uint32_t wn, uint32_t dt0, uint32_t dt1, Span* s_span)
{
    for (uint32_t c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
        const CChunk ch = a.chunks[c];
        const uint32_t cp = ch.kp - pbase;  // chunk's first pair, query-relative
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
            if (doff != kNoDir)
                filter = dir_clip(a.dir, doff, ks, dt0, dt1, a.n_dtiles, lo, hi) ? 0u : 1u;
            Span sp;
            sp.lo = lo;
            sp.hi = max(lo, hi);
            sp.wf = ch.w << 1 | filter;
            s_span[cp + j - w0] = sp;
        }
    }
}

// Flattens a resolved window of wn <= 2 * blockDim.x spans: each span's end
// is replaced by the flat index of its first posting (a prefix of the span
// lengths, 2 spans per thread); returns the window's posting count.  Called
// by all threads after the spans are complete; ends with a barrier.
__device__ __forceinline__ uint32_t flat_prefix(Span* s_span, uint32_t wn, uint32_t* s_wsum)
{
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6, nw = blockDim.x / kWave;
    const uint32_t i0 = 2 * tid, i1 = 2 * tid + 1;
    const uint32_t n0 = i0 < wn ? s_span[i0].hi - s_span[i0].lo : 0u;
    const uint32_t n1 = i1 < wn ? s_span[i1].hi - s_span[i1].lo : 0u;
    const uint32_t incl = wave_incl_scan(n0 + n1);
    if (lane == kWave - 1)
        s_wsum[wv] = incl;
    __syncthreads();
    uint32_t off = 0, total = 0;
    for (int i = 0; i < nw; ++i) {
        off += i < wv ? s_wsum[i] : 0u;
        total += s_wsum[i];
    }
    const uint32_t ex = off + incl - (n0 + n1);
    if (i0 < wn)
        s_span[i0].hi = ex;
    if (i1 < wn)
        s_span[i1].hi = ex + n0;
    __syncthreads();
    return total;
}

// The flattened span holding flat posting index g (last span whose first
// flat index is <= g).
__device__ __forceinline__ uint32_t flat_span(const Span* s_span, uint32_t wn, uint32_t g)
{
    uint32_t lo = 0, hi = wn;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_span[mid].hi <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// SPARSE: counting marks the touched counter words and the select visits
// only those (tile_select_sparse); for batches whose queries hold few
// postings per tile.  With a.qlist set, the grid covers the listed queries
// only (the others run the hash form).
template <int LB, bool SPARSE>
__global__ void __launch_bounds__(kScanThreads) k_scan_tile(ScanArgs a)
{
    __shared__ uint4 s_cnt4[kCntWords / 4 + kWave / 4];  // counters + per-lane scratch
    uint32_t* cnt = reinterpret_cast<uint32_t*>(s_cnt4);
    __shared__ Span s_span[kSpanCap];
    __shared__ uint32_t s_next;
    __shared__ SelShared s_sel;
    __shared__ uint32_t s_touch[SPARSE ? kTouchWords : 1];
    __shared__ uint32_t s_wsum[kScanWaves];

    const uint32_t nql = a.qlist ? a.n_qlist : a.n_queries;
    const uint32_t t = blockIdx.x / nql;
    const uint32_t q = a.qlist ? a.qlist[blockIdx.x % nql] : blockIdx.x % nql;
    const uint32_t obj0 = t * a.tile_objs;
    const uint32_t dt0 = t * a.dt_per_tile;
    const uint32_t dt1 = min(dt0 + a.dt_per_tile, a.n_dtiles);
    const int tid = threadIdx.x, lane = lane_id();

    for (int i = tid; i < kCntWords / 4; i += kScanThreads)
        s_cnt4[i] = make_uint4(0, 0, 0, 0);
    if constexpr (SPARSE)
        for (int i = tid; i < kTouchWords; i += kScanThreads)
            s_touch[i] = 0;

    // ---- count ------------------------------------------------------------
    // Windows of up to kSpanCap (chunk, key) pairs, resolved into spans, then
    // counted: by the waves taking spans dynamically, or (SPARSE) flattened.
    const uint32_t c0 = a.qc_ptr[q], c1 = a.qc_ptr[q + 1];
    const uint32_t pbase = a.qk_ptr[q], npairs = a.qk_ptr[q + 1] - pbase;
    for (uint32_t w0 = 0; w0 < npairs; w0 += kSpanCap) {
        const uint32_t wn = min((uint32_t)kSpanCap, npairs - w0);
        if (tid == 0)
            s_next = 0;
        resolve_spans(a, c0, c1, pbase, w0, wn, dt0, dt1, s_span);
        __syncthreads();
        if constexpr (SPARSE) {
            // Few postings per span: every thread takes postings g = tid,
            // tid + 512, ... of the flattened window, four loads at a time.
            const uint32_t total = flat_prefix(s_span, wn, s_wsum);
            for (uint32_t g0 = 0; g0 < total; g0 += 4 * kScanThreads) {
                uint32_t id[4], w[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t g = g0 + u * kScanThreads + tid;
                    id[u] = obj0 + a.tile_objs;  // outside the tile: counts nothing
                    w[u] = 0;
                    if (g < total) {
                        const Span sp = s_span[flat_span(s_span, wn, g)];
                        id[u] = a.post[sp.lo + (g - sp.hi)];
                        w[u] = sp.wf >> 1;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    count_one<LB, true, true>(id[u], obj0, a.tile_objs, w[u], cnt, s_touch);
            }
        } else {
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
        }
        __syncthreads();
    }
    if (npairs == 0)
        __syncthreads();  // the zeroing above must finish before the select passes

    plain_select<LB, SPARSE>(a, cnt, s_touch, q, obj0, s_sel);
}

// Compressed form of k_scan_tile: pairs resolve to ranges of compressed
// segments overlapping the tile (segment directory for long keys), the waves
// take single segments, stage and decode them (codecs_dev.h) and count the
// decoded ids that fall inside the tile.
template <int LB, bool SPARSE>
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
    __shared__ uint32_t s_touch[SPARSE ? kTouchWords : 1];

    const uint32_t nql = a.qlist ? a.n_qlist : a.n_queries;
    const uint32_t t = blockIdx.x / nql;
    const uint32_t q = a.qlist ? a.qlist[blockIdx.x % nql] : blockIdx.x % nql;
    const uint32_t obj0 = t * a.tile_objs;
    const uint32_t obj1 = obj0 + a.tile_objs;
    const uint32_t dt0 = t * a.dt_per_tile;
    const uint32_t dt1 = min(dt0 + a.dt_per_tile, a.n_dtiles);
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;

    for (int i = tid; i < kCntWords / 4; i += kScanThreads)
        s_cnt4[i] = make_uint4(0, 0, 0, 0);
    if constexpr (SPARSE)
        for (int i = tid; i < kTouchWords; i += kScanThreads)
            s_touch[i] = 0;

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
                if (doff != kNoDir)
                    sdir_clip(a.sdir, doff, ks, ke, dt0, dt1, obj1, a.seg_first, lo, hi);
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
            const uint32_t n = wave_decode(a.codec, D, nw, v, a.seg_bph ? a.seg_bph[sg] : uint4{}, a.seg_bph != nullptr);
            const uint32_t w = sp.wf >> 1;
#pragma unroll
            for (int j = 0; j < kPerLane; ++j) {
                const uint32_t i = lane * kPerLane + j;
                count_one<LB, true, SPARSE>(i < n ? v[j] : obj1, obj0, a.tile_objs, w, cnt, s_touch);
            }
            wave_sync();  // the staging buffer is reused by the next segment
        }
        __syncthreads();
    }
    if (npairs == 0)
        __syncthreads();
    plain_select<LB, SPARSE>(a, cnt, s_touch, q, obj0, s_sel);
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
    // the run's two rerun flags (candidate overflow, a hash table filled up)
    // folded into one word for an asynchronous download; every kernel that
    // sets them has finished when the merge starts
    const uint32_t* flag_a;
    const uint32_t* flag_b;
    uint32_t* flag_any;
};

// Queries with k <= kMergeSmallK and at most kMergeSmallM candidates are
// merged by the first wave of their workgroup alone, with no workgroup
// barrier (merge_wave).  A wave alone serializes its histogram atomics, so
// longer candidate lists (the headline batch's ~450 a query) merge faster on
// the whole workgroup.
constexpr uint32_t kMergeSmallK = 128;
constexpr uint32_t kMergeSmallM = 256;  // 4 keys per lane

__device__ __forceinline__ bool merge_small(uint32_t k, uint32_t m)
{
    return k <= kMergeSmallK && m <= kMergeSmallM;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m)
{
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, kWave);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, kWave);
    return ((uint64_t)hi << 32) | lo;
}

// One compare-exchange step of a descending bitonic sort: element idx keeps
// the larger of itself and its partner idx ^ stride when it is the lower of
// the two in a descending run (idx & size == 0), or the higher in an
// ascending one.
__device__ __forceinline__ uint64_t bitonic_step(uint64_t x, uint64_t y, uint32_t idx, uint32_t size, uint32_t stride)
{
    const bool desc = (idx & size) == 0u;
    const bool keep_max = ((idx & stride) == 0u) == desc;
    return keep_max ? (x > y ? x : y) : (x < y ? x : y);
}

// Sorts up to 128 keys descending across one wave (element l in e0, l + 64 in
// e1; zeros pad) by a bitonic network of shuffles, and writes the query's k
// results, keys of 0 and the slots past the 128th as (0, 0).
__device__ __forceinline__ void wave_sort_out(const MergeArgs& a, uint32_t q, uint32_t k, uint64_t e0, uint64_t e1,
                                              uint32_t n)
{
    const uint32_t l = (uint32_t)lane_id();
    if (n > (uint32_t)kWave) {  // 128 keys
        for (uint32_t size = 2; size <= 2u * kWave; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                if (stride == (uint32_t)kWave) {  // within the lane (size 128: one descending run)
                    const uint64_t hi = e0 > e1 ? e0 : e1, lo = e0 > e1 ? e1 : e0;
                    e0 = hi;
                    e1 = lo;
                } else {
                    const uint64_t p0 = shfl_xor_u64(e0, (int)stride), p1 = shfl_xor_u64(e1, (int)stride);
                    e0 = bitonic_step(e0, p0, l, size, stride);
                    e1 = bitonic_step(e1, p1, l + kWave, size, stride);
                }
            }
    } else {  // at most 64 keys, one a lane
        for (uint32_t size = 2; size <= (uint32_t)kWave; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1)
                e0 = bitonic_step(e0, shfl_xor_u64(e0, (int)stride), l, size, stride);
    }
    const uint64_t o = a.out_off[q];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const uint32_t i = l + (uint32_t)r * kWave;
        if (i < k) {
            const uint64_t key = r == 0 ? e0 : e1;
            a.out_ids[o + i] = key ? (int32_t)(~(uint32_t)key) : 0;
            a.out_counts[o + i] = key ? (int32_t)(key >> 32) : 0;
        }
    }
    for (uint32_t i = 2u * kWave + l; i < k; i += kWave) {
        a.out_ids[o + i] = 0;
        a.out_counts[o + i] = 0;
    }
}

// k_merge of one query by one wave (k <= 128, at most 256 candidates): the
// candidates in registers (4 per lane), the radix select of k_merge with the
// histogram in `hist` (256 words), the selected keys compacted into `sel`
// (kMergeSmallK keys) and sorted descending across the lanes by a bitonic
// network of shuffles (64 keys, one a lane, or 128, two), with no workgroup
// barrier.
__device__ __forceinline__ void merge_wave(const MergeArgs& a, uint32_t q, uint32_t k, uint32_t M, uint32_t* hist,
                                           uint64_t* sel)
{
    constexpr int PL = (int)(kMergeSmallM / kWave);
    const int lane = lane_id();
    const uint64_t* keys = a.cand + a.cand_base[q];
    const uint32_t rows = (M + kWave - 1) / kWave;  // uniform
    uint64_t v[PL];
#pragma unroll
    for (int j = 0; j < PL; ++j) {
        const uint32_t i = (uint32_t)lane + (uint32_t)j * kWave;
        v[j] = ((uint32_t)j < rows && i < M) ? keys[i] : 0ull;
    }
    uint64_t kth = 0;  // every key >= kth is in the result
    if (M > k) {
        uint64_t prefix = 0, mask = 0;
        uint32_t remaining = k;
        for (int shift = ((int)((a.key_bits + 7) / 8) - 1) * 8; shift >= 0; shift -= 8) {
#pragma unroll
            for (int b = 0; b < 4; ++b)
                hist[lane + b * kWave] = 0;
            wave_sync();
#pragma unroll
            for (int j = 0; j < PL; ++j) {
                const uint32_t i = (uint32_t)lane + (uint32_t)j * kWave;
                if ((uint32_t)j < rows && i < M && (v[j] & mask) == prefix)
                    atomicAdd(&hist[(uint32_t)(v[j] >> shift) & 255u], 1u);
            }
            wave_sync();
            // lane l holds bins 255 - 4l .. 252 - 4l, counted from the top
            const uint32_t top = 255u - 4u * (uint32_t)lane;
            uint32_t h[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                h[j] = hist[top - j];
            const uint32_t sum = h[0] + h[1] + h[2] + h[3];
            const uint32_t above = wave_incl_scan(sum) - sum;  // keys in higher bins
            const bool here = above < remaining && above + sum >= remaining;
            uint32_t bin = 0, rem = 0, inbin = 0;
            if (here) {
                uint32_t ab = above;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (ab + h[j] >= remaining) {
                        bin = top - j;
                        rem = remaining - ab;
                        inbin = h[j];
                        break;
                    }
                    ab += h[j];
                }
            }
            const int src = __builtin_ctzll(ballot(here));  // exactly one lane
            bin = (uint32_t)__shfl((int)bin, src, kWave);
            rem = (uint32_t)__shfl((int)rem, src, kWave);
            inbin = (uint32_t)__shfl((int)inbin, src, kWave);
            prefix |= (uint64_t)bin << shift;
            mask |= (uint64_t)255 << shift;
            remaining = rem;
            wave_sync();  // the histogram is read before the next pass clears it
            if (inbin == rem)
                break;  // the whole bin is needed: prefix with zero low bytes bounds it
        }
        kth = prefix;
    }
    // the selected keys (min(M, k) of them), compacted; all of them up to
    // kMergeSmallK are sorted, so that a select that took more (keys with
    // bits above key_bits) still yields the k largest
    uint32_t base = 0;
#pragma unroll
    for (int j = 0; j < PL; ++j) {
        const uint32_t i = (uint32_t)lane + (uint32_t)j * kWave;
        const bool take = (uint32_t)j < rows && i < M && v[j] >= kth;
        const uint64_t bm = ballot(take);
        const uint32_t slot = base + lane_prefix(bm);
        if (take && slot < kMergeSmallK)
            sel[slot] = v[j];
        base += (uint32_t)__popcll(bm);
    }
    wave_sync();
    const uint32_t nsel = min(base, kMergeSmallK);
    uint64_t e0 = (uint32_t)lane < nsel ? sel[lane] : 0ull;
    uint64_t e1 = (uint32_t)lane + kWave < nsel ? sel[lane + kWave] : 0ull;
    wave_sort_out(a, q, k, e0, e1, nsel);
}

// Per query: select the k largest candidate keys (radix select over the key
// bytes, most significant first), sort them descending, pad with (0, 0).
// Candidates are staged in LDS when they fit; the bin holding the k-th key
// is found with a wave-parallel suffix scan; the select stops as soon as a
// bin holds exactly the keys still needed.
__global__ void __launch_bounds__(kMergeThreads) k_merge(MergeArgs a)
{
    constexpr int PER = kMergeMaxK / kMergeThreads;  // staged keys per thread
    static_assert(kMergeThreads == 256, "one histogram bin per thread");
    __shared__ uint64_t s_keys[kMergeMaxK];
    __shared__ uint32_t s_hist[256];
    __shared__ uint32_t s_res[4];
    const uint32_t q = blockIdx.x;
    const int tid = threadIdx.x;
    if (q == 0 && tid == 0)
        *a.flag_any = *a.flag_a | *a.flag_b;
    const uint32_t k = a.kq[q];
    const uint32_t M = min(a.cand_cnt[q], a.cand_cap[q]);
    if (merge_small(k, M)) {  // one wave does it (the whole workgroup decides alike, before any barrier)
        if (tid < kWave)
            merge_wave(a, q, k, M, s_hist, s_keys);
        return;
    }
    const uint64_t* keys = a.cand + a.cand_base[q];
    const bool staged = M <= (uint32_t)kMergeMaxK;
    if (staged)
        for (uint32_t i = tid; i < M; i += kMergeThreads)
            s_keys[i] = keys[i];
    __syncthreads();

    uint64_t kth = 0;  // every key >= kth is in the result
    if (M > k) {
        uint64_t prefix = 0, mask = 0;
        uint32_t remaining = k;
        for (int shift = ((int)((a.key_bits + 7) / 8) - 1) * 8; shift >= 0; shift -= 8) {
            s_hist[tid] = 0;  // kMergeThreads == 256 bins
            __syncthreads();
            for (uint32_t i = tid; i < M; i += kMergeThreads) {
                const uint64_t key = staged ? s_keys[i] : keys[i];
                if ((key & mask) == prefix)
                    atomicAdd(&s_hist[(uint32_t)(key >> shift) & 255u], 1u);
            }
            __syncthreads();
            if (tid < kWave) {
                // lane l holds bins 255 - 4l .. 252 - 4l, counted from the top
                const uint32_t top = 255u - 4u * (uint32_t)tid;
                uint32_t h[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    h[j] = s_hist[top - j];
                const uint32_t sum = h[0] + h[1] + h[2] + h[3];
                uint32_t above = wave_incl_scan(sum) - sum;  // keys in higher bins
                if (above < remaining && above + sum >= remaining) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (above + h[j] >= remaining) {
                            s_res[0] = top - j;
                            s_res[1] = remaining - above;
                            s_res[2] = h[j];
                            break;
                        }
                        above += h[j];
                    }
                }
            }
            __syncthreads();
            const uint32_t bin = s_res[0], rem = s_res[1], inbin = s_res[2];
            prefix |= (uint64_t)bin << shift;
            mask |= (uint64_t)255 << shift;
            remaining = rem;
            __syncthreads();
            if (inbin == rem)
                break;  // the whole bin is needed: prefix with zero low bytes bounds it
        }
        kth = prefix;
    }
    // collect the selected keys (read first, then overwrite the stage)
    const uint32_t nsel = min(M, k);
    uint32_t n2 = 1;
    while (n2 < nsel)
        n2 <<= 1;
    uint64_t mine[PER];
    if (staged) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t i = tid + j * kMergeThreads;
            mine[j] = i < M ? s_keys[i] : 0;
        }
    }
    if (tid == 0)
        s_res[3] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n2; i += kMergeThreads)
        s_keys[i] = 0;
    __syncthreads();
    auto put = [&](uint64_t key, bool take) {
        const uint64_t bm = ballot(take);
        if (!bm)
            return;
        uint32_t wb = 0;
        if (lane_id() == __builtin_ctzll(bm))
            wb = atomicAdd(&s_res[3], (uint32_t)__popcll(bm));
        wb = (uint32_t)__shfl((int)wb, __builtin_ctzll(bm), kWave);
        if (take) {
            const uint32_t slot = wb + lane_prefix(bm);
            if (slot < n2)
                s_keys[slot] = key;
        }
    };
    if (staged) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t i = tid + j * kMergeThreads;
            put(mine[j], i < M && mine[j] >= kth);
        }
    } else {
        for (uint32_t i0 = 0; i0 < M; i0 += kMergeThreads) {
            const uint32_t i = i0 + tid;
            const uint64_t key = i < M ? keys[i] : 0;
            put(key, i < M && key >= kth);
        }
    }
    __syncthreads();
    if (n2 <= 2u * kWave) {  // at most 128 selected keys: the first wave sorts and writes them
        if (tid < kWave)
            wave_sort_out(a, q, k, tid < (int)n2 ? s_keys[tid] : 0ull,
                          tid + kWave < (int)n2 ? s_keys[tid + kWave] : 0ull, n2);  // the stage's first n2 keys
        return;
    }
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
