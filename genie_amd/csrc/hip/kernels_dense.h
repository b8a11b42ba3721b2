// Dense form of the batched search: bit-sliced counting over a forward code
// matrix derived from the inverted lists.
//
// For a single-value table every object holds exactly one key per dim, so
// the inverted lists of one dim are a permutation of all objects.  At index
// creation the lists are turned around into codes[d][o] = key of object o in
// dim d, relative to the dim's first key (one byte, four dims per word).  A
// batch whose items cover whole keys with unit weight, and never the same key
// twice for one query, then has the exact per-(query, key) weights 0/1.  For
// 32 queries at a time those weights are a 32-bit mask per key (`tab`,
// d * 256 + code), and
//
//     count(q, o) = sum over d of bit q of tab[d][codes[d][o]]
//
// is the same number the posting-list scan counts
// (src/genie/matching/match.cu:161-220 adds one per posting of every item).
// One LDS read serves 32 queries; the 32 sums are kept bit-sliced in eight
// registers (bit j of plane b = bit b of query j's count) and added with
// carry-save adders (v_bitop3: XOR3 and majority), 16 dims per step.
//
// The top-k is exact: a sample pass computes, per query, the k-th largest
// count among evenly spread sample tiles (a lower bound of the final k-th
// count, since the sample is a subset of all objects); the main pass emits
// every object whose count reaches that bound into the query's candidate
// list, and the merge kernel picks the top k under (count desc, id asc).
#pragma once
#include "kernels.h"

namespace gh {

constexpr int kDenseThreads = 1024;
constexpr uint32_t kDenseTile = 4096;     // objects per work tile
constexpr uint32_t kDenseMaxDims = 128;   // 128 dims x 256 keys x 4 B = 128 KiB of LDS
constexpr uint32_t kDenseKeysPerDim = 256;
constexpr uint32_t kDenseBins = 256;      // histogram bins per query (counts <= 128)

struct DenseArgs {
    const uint32_t* codes;  // [D_pad / 4][n_pad] words, byte d % 4 = code of dim d
    uint32_t n_pad;
    const uint32_t* tab;    // [G][D_pad * 256] masks, bit j = query 32 g + j
    uint32_t D_pad, G, n_queries, n_objects, n_tiles;
    const uint32_t* list;   // tiles to visit (sample pass), null = every tile
    uint32_t n_list;
    uint32_t* gthr;         // [n_queries] emission threshold (>= 1)
    uint32_t* hist;         // [n_queries][kDenseBins] sample histogram
    uint32_t* cand_cnt;
    const uint32_t* cand_cap;
    const uint64_t* cand_base;
    uint64_t* cand;
    uint32_t* overflow;
};

// ---- index creation --------------------------------------------------------

// One workgroup per key: write the key's code into the byte of every object
// in its list, and mark (object, dim) in a bitmap; a bit found already set
// means an object listed twice in one dim (flag |= 1).
__global__ void k_dense_codes(const uint32_t* __restrict__ post, const uint32_t* __restrict__ key_start,
                              const uint32_t* __restrict__ key_dim, const uint32_t* __restrict__ key_slot,
                              uint32_t n_pad, uint32_t bm_words, uint8_t* __restrict__ codes8,
                              uint32_t* __restrict__ bm, uint32_t* __restrict__ flag)
{
    const uint32_t k = blockIdx.x;
    const uint32_t d = key_dim[k];
    const uint8_t code = (uint8_t)(key_slot[k] & 255u);
    const uint32_t s = key_start[k], e = key_start[k + 1];
    for (uint32_t p = s + threadIdx.x; p < e; p += blockDim.x) {
        const uint32_t o = post[p];
        codes8[((uint64_t)(d >> 2) * n_pad + o) * 4 + (d & 3)] = code;
        const uint32_t bit = 1u << (d & 31);
        const uint32_t old = atomicOr(&bm[(uint64_t)o * bm_words + (d >> 5)], bit);
        if (old & bit)
            atomicOr(flag, 1u);
    }
}

// Every object must appear in every dim (flag |= 2 otherwise).
__global__ void k_dense_holes(const uint32_t* __restrict__ bm, uint32_t n_objects, uint32_t n_dims,
                              uint32_t bm_words, uint32_t* __restrict__ flag)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_objects)
        return;
    bool ok = true;
    for (uint32_t w = 0; w < bm_words; ++w) {
        const uint32_t lo = w * 32;
        const uint32_t nb = n_dims - lo >= 32 ? 32u : n_dims - lo;
        const uint32_t want = nb == 32 ? 0xFFFFFFFFu : ((1u << nb) - 1u);
        ok &= (bm[(uint64_t)o * bm_words + w] & want) == want;
    }
    if (!ok)
        atomicOr(flag, 2u);
}

// ---- per run ---------------------------------------------------------------

// One thread per item: set the query's bit in the mask of every key the item
// covers (the batch was checked at upload: whole keys, unit weight, no key
// twice per query).
__global__ void k_dense_masks(const genie_dim_t* __restrict__ items, uint32_t n_items,
                              const uint32_t* __restrict__ key_start, uint32_t n_keys,
                              const uint32_t* __restrict__ key_slot, uint32_t tab_words, uint32_t* __restrict__ tab)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items)
        return;
    const genie_dim_t it = items[i];
    if (it.end_pos <= it.start_pos || it.weight == 0.0f)
        return;
    // key containing position `pos`: last k with key_start[k] <= pos
    auto key_of = [&](uint32_t pos) {
        uint32_t lo = 0, hi = n_keys;  // key_start[lo] <= pos < key_start[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (key_start[mid] <= pos)
                lo = mid;
            else
                hi = mid;
        }
        return lo;
    };
    const uint32_t ka = key_of((uint32_t)it.start_pos), kb = key_of((uint32_t)it.end_pos - 1);
    const uint32_t q = (uint32_t)it.query;
    uint32_t* t = tab + (uint64_t)(q >> 5) * tab_words;
    const uint32_t bit = 1u << (q & 31);
    for (uint32_t k = ka; k <= kb; ++k)
        atomicOr(&t[key_slot[k]], bit);
}

// Threshold of every query from its sample histogram: the largest count c
// with at least k sample objects counting >= c, and 1 when fewer than k
// sample objects count at all.
__global__ void k_dense_thresh(const uint32_t* __restrict__ hist, const uint32_t* __restrict__ kq,
                               uint32_t n_queries, uint32_t* __restrict__ gthr)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_queries)
        return;
    const uint32_t k = kq[q];
    const uint32_t* h = hist + (uint64_t)q * kDenseBins;
    uint32_t cum = 0, thr = 1;
    for (uint32_t c = kDenseBins - 1; c >= 1; --c) {
        cum += h[c];
        if (cum >= k) {
            thr = c;
            break;
        }
    }
    gthr[q] = thr;
}

__device__ __forceinline__ void csa(uint32_t& hi, uint32_t& lo, uint32_t a, uint32_t b, uint32_t c)
{
    hi = __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);  // majority
    lo = __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);  // a ^ b ^ c
}

// Bit-sliced counts of object o over NBLK blocks of 16 dims: c[b] bit j =
// bit b of query j's count.  Harley-Seal: ones/twos/fours/eights absorb 16
// inputs per step, the weight-16 carry ripples into the upper planes.
template <int NBLK>
__device__ __forceinline__ void dense_accumulate(const uint32_t* __restrict__ codes, uint32_t n_pad, uint32_t o,
                                                 const uint32_t* s_tab, uint32_t (&c)[8])
{
    uint32_t w[4 * NBLK];
#pragma unroll
    for (int i = 0; i < 4 * NBLK; ++i)
        w[i] = __builtin_nontemporal_load(codes + (uint64_t)i * n_pad + o);
    uint32_t ones = 0, twos = 0, fours = 0, eights = 0, h16 = 0, h32 = 0, h64 = 0, h128 = 0;
#pragma unroll
    for (int blk = 0; blk < NBLK; ++blk) {
        uint32_t m[16];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                m[4 * i + j] = s_tab[(blk * 16 + 4 * i + j) * kDenseKeysPerDim + ((w[4 * blk + i] >> (8 * j)) & 255u)];
        uint32_t tA, tB, fA, fB, eA, eB, sx;
        csa(tA, ones, ones, m[0], m[1]);
        csa(tB, ones, ones, m[2], m[3]);
        csa(fA, twos, twos, tA, tB);
        csa(tA, ones, ones, m[4], m[5]);
        csa(tB, ones, ones, m[6], m[7]);
        csa(fB, twos, twos, tA, tB);
        csa(eA, fours, fours, fA, fB);
        csa(tA, ones, ones, m[8], m[9]);
        csa(tB, ones, ones, m[10], m[11]);
        csa(fA, twos, twos, tA, tB);
        csa(tA, ones, ones, m[12], m[13]);
        csa(tB, ones, ones, m[14], m[15]);
        csa(fB, twos, twos, tA, tB);
        csa(eB, fours, fours, fA, fB);
        csa(sx, eights, eights, eA, eB);
        uint32_t t = h16 & sx;
        h16 ^= sx;
        sx = t;
        t = h32 & sx;
        h32 ^= sx;
        sx = t;
        t = h64 & sx;
        h64 ^= sx;
        h128 ^= t;  // counts stay below 256 (at most 128 dims)
    }
    c[0] = ones;
    c[1] = twos;
    c[2] = fours;
    c[3] = eights;
    c[4] = h16;
    c[5] = h32;
    c[6] = h64;
    c[7] = h128;
}

__device__ __forceinline__ uint32_t dense_lane_count(const uint32_t (&c)[8], uint32_t j)
{
    uint32_t v = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b)
        v |= ((c[b] >> j) & 1u) << b;
    return v;
}

// Workgroup -> (query group, tiles) so that the groups of one tile run on one
// XCD (workgroup ids are dealt to the 8 XCDs round robin) and each
// workgroup loads one group's mask table once.
template <int NBLK, bool SAMPLE>
__global__ void __launch_bounds__(kDenseThreads) k_dense_count(DenseArgs a)
{
    extern __shared__ uint32_t s_lds[];
    uint32_t* s_tab = s_lds;                          // D_pad * 256
    uint32_t* s_hist = s_lds + NBLK * 16 * kDenseKeysPerDim;  // SAMPLE: 32 x (D_pad + 1)
    __shared__ uint32_t s_thr[8];
    constexpr uint32_t HB = NBLK * 16 + 1;
    const uint32_t tid = threadIdx.x;
    const uint32_t P = gridDim.x / 8, x = blockIdx.x % 8, s = blockIdx.x / 8;
    uint32_t g_first, g_step, t_first, t_step;
    if (a.G <= P) {
        const uint32_t R = P / a.G;
        if (s >= a.G * R)
            return;  // the whole workgroup leaves before any barrier
        g_first = s % a.G;
        g_step = a.G;
        t_first = x + 8 * (s / a.G);
        t_step = 8 * R;
    } else {
        g_first = s;
        g_step = P;
        t_first = x;
        t_step = 8;
    }
    const uint32_t n_list = a.list ? a.n_list : a.n_tiles;
    const uint32_t tab_words = NBLK * 16 * kDenseKeysPerDim;
    for (uint32_t g = g_first; g < a.G; g += g_step) {
        __syncthreads();
        {
            const uint4* src = reinterpret_cast<const uint4*>(a.tab + (uint64_t)g * tab_words);
            uint4* dst = reinterpret_cast<uint4*>(s_tab);
            for (uint32_t i = tid; i < tab_words / 4; i += kDenseThreads)
                dst[i] = src[i];
            if constexpr (SAMPLE)
                for (uint32_t i = tid; i < 32 * HB; i += kDenseThreads)
                    s_hist[i] = 0;
        }
        uint32_t tp[8] = {};
        for (uint32_t li = t_first; li < n_list; li += t_step) {
            const uint32_t t = a.list ? a.list[li] : li;
            if constexpr (!SAMPLE) {
                // threshold planes of the group (bit j of plane b = bit b of
                // query j's threshold; 255 for lanes past the last query)
                __syncthreads();
                if (tid < kWave) {
                    const uint32_t q = g * 32 + tid;
                    const uint32_t th = (tid < 32 && q < a.n_queries)
                                            ? __hip_atomic_load(&a.gthr[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                            : 255u;
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        const uint32_t pl = (uint32_t)ballot((th >> b) & 1u);
                        if (tid == 0)
                            s_thr[b] = pl;
                    }
                }
            }
            __syncthreads();
            if constexpr (!SAMPLE) {
#pragma unroll
                for (int b = 0; b < 8; ++b)
                    tp[b] = __builtin_amdgcn_readfirstlane(s_thr[b]);
            }
            const uint32_t o0 = t * kDenseTile;
            const uint32_t o1 = min(o0 + kDenseTile, a.n_objects);
            for (uint32_t o = o0 + tid; o < o1; o += kDenseThreads) {
                uint32_t c[8];
                dense_accumulate<NBLK>(a.codes, a.n_pad, o, s_tab, c);
                if constexpr (SAMPLE) {
                    for (uint32_t j = 0; j < 32; ++j) {
                        const uint32_t v = dense_lane_count(c, j);
                        if (v)
                            atomicAdd(&s_hist[j * HB + v], 1u);
                    }
                } else {
                    // count >= threshold, per lane, most significant plane first
                    uint32_t gt = 0, eq = 0xFFFFFFFFu;
#pragma unroll
                    for (int b = 7; b >= 0; --b) {
                        gt |= eq & c[b] & ~tp[b];
                        eq &= ~(c[b] ^ tp[b]);
                    }
                    uint32_t ge = gt | eq;
                    while (ge) {
                        const uint32_t j = (uint32_t)__builtin_ctz(ge);
                        ge &= ge - 1;
                        const uint32_t q = g * 32 + j;
                        const uint64_t key = ((uint64_t)dense_lane_count(c, j) << 32) | (uint32_t)~o;
                        const uint32_t slot = atomicAdd(&a.cand_cnt[q], 1u);
                        if (slot < a.cand_cap[q])
                            a.cand[a.cand_base[q] + slot] = key;
                        else
                            atomicOr(a.overflow, 1u);
                    }
                }
            }
        }
        if constexpr (SAMPLE) {
            __syncthreads();
            for (uint32_t i = tid; i < 32 * HB; i += kDenseThreads) {
                const uint32_t j = i / HB, bin = i % HB;
                const uint32_t q = g * 32 + j;
                const uint32_t v = s_hist[i];
                if (v && q < a.n_queries)
                    atomicAdd(&a.hist[(uint64_t)q * kDenseBins + bin], v);
            }
        }
    }
}

}  // namespace gh
