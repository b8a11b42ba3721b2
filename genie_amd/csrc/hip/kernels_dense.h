// Dense form of the batched search: bit-sliced counting over a forward code
// matrix derived from the inverted lists.
//
// For a single-value table every object holds exactly one key per dim, so
// the inverted lists of one dim are a permutation of all objects.  At index
// creation the lists are turned around into codes[d][o] = key of object o in
// dim d, as its slot within the dim (packed three per word, see below).  A
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
// Integer weights up to 15 take WP = 2-4 bits a query (the bits of the
// batch's largest weight): a word holds 32 / WP queries, bit WP j + p being
// bit p of query j's weight on the key, and the plane sums are weighted by
// 2^p afterwards (dense_weigh).
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

// Packed codes: three 10-bit fields per word, at bits 0, 11 and 22, each
// holding code * 4 -- the byte offset of the key's mask within its dim's
// 1 KiB table row -- so every LDS address takes one VALU op.  Rows of dims
// 0..63 start below 64 KiB and are reached through the 16-bit offset of
// ds_read; rows of dims 64..127 lie 64 KiB higher, which an op can add only
// for the outer fields ((x & 0x3FC) | 0x10000 and alignbit(0x40, x, 22)), so
// those dims occupy outer fields only and the middle fields hold low dims.
__host__ __device__ constexpr int dense_words(int dpad)
{
    const int L = dpad < 64 ? dpad : 64, U = dpad - L;
    return U / 2 + (L - U / 2 + 2) / 3;
}

// Dim stored in field f (0, 1, 2) of word w, or -1 for an unused field.
__host__ __device__ constexpr int dense_field_dim(int dpad, int w, int f)
{
    const int L = dpad < 64 ? dpad : 64, U = dpad - L;
    if (w < U / 2)
        return f == 0 ? L + 2 * w : (f == 2 ? L + 2 * w + 1 : w);
    const int i = U / 2 + 3 * (w - U / 2) + f;
    return i < L ? i : -1;
}

__host__ __device__ constexpr int dense_field_shift(int f) { return f == 0 ? 0 : (f == 1 ? 11 : 22); }

// The fields of a layout in load order, for the unrolled accumulation.
template <int DPAD>
struct DenseLayout {
    int w[DPAD], f[DPAD], d[DPAD];
    constexpr DenseLayout() : w(), f(), d()
    {
        int n = 0;
        for (int i = 0; i < dense_words(DPAD); ++i)
            for (int ff = 0; ff < 3; ++ff) {
                const int dd = dense_field_dim(DPAD, i, ff);
                if (dd >= 0) {
                    w[n] = i;
                    f[n] = ff;
                    d[n] = dd;
                    ++n;
                }
            }
    }
};

struct DenseArgs {
    const uint32_t* codes;  // [dense_words(D_pad)][n_pad] packed fields
    uint32_t n_pad;
    uint32_t* tab;          // [G][D_pad * 256] masks, bit WP j + p = bit p of query (32 / WP) g + j's weight (sample pass)
    const genie_dim_t* items;     // the batch, sorted by query
    const uint32_t* item_qptr;    // [n_queries + 1]
    const uint32_t* dkeys;        // [n_items][2]: first key | query's bit position << 24, end key | weight << 16
    const uint32_t* key_slot;     // [n_keys] dim * 256 + slot
    uint32_t D_pad, G, n_queries, n_objects, n_tiles;
    const uint32_t* list;   // tiles to visit (sample pass), null = every tile
    uint32_t n_list;
    uint32_t* gthr;         // [n_queries] emission threshold (>= 1)
    const uint32_t* redo;   // [G] groups to count again (null = every group)
    uint32_t* hist;         // [n_queries][kDenseBins] sample histogram
    uint32_t* cand_cnt;
    const uint32_t* cand_cap;
    const uint64_t* cand_base;
    uint64_t* cand;
    uint32_t* overflow;
    // set form (set-valued tables): forward lists and, per query group, its
    // keys with the queries' weights as bit planes
    const uint32_t* obj_perm;    // [n_slots] object of each slot (kSetEmpty: none)
    const uint32_t* blk_off;     // [n_slots / 64 + 1] first word of each 64-slot block
    const uint32_t* blk_key;     // key + 1 of occurrence i of slot 64 b + j at blk_off[b] + 64 i + j (0: none)
    uint32_t n_slots;            // objects rounded up to 64
    const uint32_t* set_keys;    // [entries] sorted keys of each group
    const uint4* set_mask;       // [entries][2] bit j of word b = bit b of query j's weight (b < 6)
    const uint32_t* set_gptr;    // [G + 1] entries of each group
    const uint32_t* hshift;      // [n_queries] sample histogram bin = count >> hshift (set form)
    // tile chunks per group (dense_chunks): the work units of a pass are
    // (group, chunk) pairs
    uint32_t chunks;
};

// Tile chunks per group for G groups over P workgroups per XCD.  Chunk c of
// a group takes the XCD's tiles c, c + C, c + 2 C, ...  With `balance` (the
// dense form's main pass), the C of 1 .. max(8, P / G) whose busiest
// workgroup does the least work, ceil(G C / P) / C group passes, the
// smallest on ties; else P / G when the groups are fewer than P (the sample
// pass and the set form, whose group passes are short or balanced), or 1.
__host__ __device__ inline uint32_t dense_chunks(uint32_t G, uint32_t P, bool balance)
{
    if (G == 0 || P == 0)
        return 1;
    if (!balance)
        return G <= P ? P / G : 1u;
    uint32_t best = 1, best_units = (G + P - 1) / P;  // ceil(G C / P) at C = best
    for (uint32_t c = 2; c <= (P / G > 8 ? P / G : 8u); ++c) {
        const uint32_t units = (uint32_t)(((uint64_t)G * c + P - 1) / P);
        if ((uint64_t)units * best < (uint64_t)best_units * c) {
            best = c;
            best_units = units;
        }
    }
    return best;
}

// Set form: LDS hash table of a group's keys (at most kSetEntries of them):
// kSetSlots slots of (key, entry), at most a quarter full, and the entries'
// weight planes apart from the slots (entry kSetEntries: zero weights, the
// entry of every empty slot).  Weights up to 63 (six planes); counts up to
// 2046 in eleven planes, 2047 being the threshold of "no query".
constexpr uint32_t kSetSlots = 8192;
constexpr uint32_t kSetEntries = 2048;
// LDS words of the table: slots, then the entries' planes 0-3 (uint4), then
// planes 4-5 (uint2), rounded up to a multiple of 8 words
constexpr uint32_t kSetMloWord = 2 * kSetSlots;
constexpr uint32_t kSetMhiWord = kSetMloWord + 4 * (kSetEntries + 1);
constexpr uint32_t kSetTabWords = (kSetMhiWord + 2 * (kSetEntries + 1) + 7) & ~7u;
constexpr uint32_t kSetEmpty = 0xFFFFFFFFu;
constexpr int kSetWeightPlanes = 6;
constexpr int kSetCountPlanes = 11;
constexpr uint32_t kSetMaxWeight = (1u << kSetWeightPlanes) - 1u;
constexpr uint32_t kSetNoQuery = (1u << kSetCountPlanes) - 1u;
constexpr uint32_t kDenseNoQuery = 255u;
constexpr int kSetPrefetch = 8;  // forward-list columns in flight per wave
constexpr uint32_t kSetWindowTiles = 8;  // tiles per XCD counted together (main pass)

// Home slot of a key, always even: bits 20..31 of key x 0x9E3779 (a
// full-rate 24-bit multiply; bits of a key above 24 do not enter the hash,
// which only spreads such keys less evenly) select one of 4,096 slot pairs.
// Inserts probe slots linearly from there, so the second slot of a pair is
// taken only after the first, and a lookup reads a pair at a time, stopping
// at a pair whose second slot is empty.
static_assert(kSetSlots == 1u << 13, "4,096 pairs of slots");
__device__ __forceinline__ uint32_t set_hash(uint32_t key) { return (((key & 0xFFFFFFu) * 0x9E3779u) >> 20) << 1; }

// ---- index creation --------------------------------------------------------

// One workgroup per key: or the key's field (code * 4 at the dim's word and
// shift) into the packed codes of every object in its list, and mark
// (object, dim) in a bitmap; a bit found already set means an object listed
// twice in one dim (flag |= 1).
__global__ void k_dense_codes(const uint32_t* __restrict__ post, const uint32_t* __restrict__ key_start,
                              const uint32_t* __restrict__ key_dim, const uint32_t* __restrict__ key_slot,
                              const uint32_t* __restrict__ dim_word, const uint32_t* __restrict__ dim_shift,
                              uint32_t n_pad, uint32_t bm_words, uint32_t* __restrict__ codes,
                              uint32_t* __restrict__ bm, uint32_t* __restrict__ flag)
{
    const uint32_t k = blockIdx.x;
    const uint32_t d = key_dim[k];
    const uint32_t field = (key_slot[k] & 255u) * 4u << dim_shift[d];
    uint32_t* row = codes + (uint64_t)dim_word[d] * n_pad;
    const uint32_t s = key_start[k], e = key_start[k + 1];
    for (uint32_t p = s + threadIdx.x; p < e; p += blockDim.x) {
        const uint32_t o = post[p];
        atomicOr(&row[o], field);
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

// Thresholds of every query from its sample histogram (S(c) = sample
// objects counting >= c):
//   safe: the largest c with S(c) >= k -- the k-th largest sample count, a
//         lower bound of the final k-th count, or 1;
//   spec: the largest c with S(c) >= j, j = min(k, m + 4 sqrt(m) + 4) with
//         m = k n_sample / N, the sample objects expected among the top k:
//         the final k-th count falls below spec only when the sample holds
//         4 standard deviations (of a Poisson count) more of the top k than
//         expected, or fewer than 4 objects in all.
// The main pass emits from spec; a query left with fewer than k candidates
// is counted again from safe (k_dense_check).  One wave per query, four bins
// per lane.
// Candidate space: the objects of the whole index expected at or above a
// threshold are S(c) x N / n_sample; a query asks for 3/2 of that at spec
// (want) and at safe (want_safe, for a recount), plus 4 k + 1024, times
// `mult`, at most N (the slack is address space only: a query writes what
// it emits).  Ties at the threshold are in S(c), so a query whose
// counts crowd at its threshold asks for what it will emit.
// ratio16 = ceil(N * 2^16 / n_sample), from the host: no division here.
__device__ __forceinline__ uint32_t dense_want(uint32_t s_at, uint32_t k, uint32_t ratio16, uint32_t n_objects,
                                               uint32_t mult)
{
    const uint64_t e = ((uint64_t)s_at * ratio16 + 0xFFFFu) >> 16;
    const uint64_t w = (e + e / 2 + 4ull * k + 1024) * mult;
    return (uint32_t)min<uint64_t>(w, n_objects);
}

// With `hshift` (set form), bin b of query q holds the sample counts
// [b << s, (b + 1) << s), s = hshift[q], and the thresholds are bin edges:
// S at an edge is exact, so both keep their meaning.
__global__ void __launch_bounds__(256) k_dense_thresh(const uint32_t* __restrict__ hist,
                                                      const uint32_t* __restrict__ kq, uint32_t n_queries,
                                                      uint32_t n_sample, uint32_t n_objects, uint32_t ratio16,
                                                      uint32_t mult, const uint32_t* __restrict__ hshift,
                                                      uint32_t* __restrict__ gthr, uint32_t* __restrict__ gsafe,
                                                      uint32_t* __restrict__ want, uint32_t* __restrict__ want_safe)
{
    static_assert(kDenseBins == 4 * kWave, "four bins per lane");
    const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= n_queries)
        return;  // whole waves leave together
    const uint32_t lane = lane_id();
    const uint4 h = reinterpret_cast<const uint4*>(hist + (uint64_t)q * kDenseBins)[lane];
    const uint32_t mine = h.x + h.y + h.z + h.w;
    // objects counting above this lane's bins: sum over higher lanes
    const uint32_t incl = wave_incl_scan(mine);
    const uint32_t total = (uint32_t)__shfl((int)incl, kWave - 1, kWave);
    const uint32_t above = total - incl;
    const uint32_t k = kq[q];
    const float m = (float)k * (float)n_sample / (float)n_objects;
    const uint32_t j = min(k, (uint32_t)ceilf(m + 4.0f * sqrtf(m) + 4.0f));
    const uint32_t s3 = above + h.w, s2 = s3 + h.z, s1 = s2 + h.y, s0 = s1 + h.x;
    const uint32_t b = 4 * lane;
    auto largest = [&](uint32_t need) {
        return s3 >= need ? b + 3 : (s2 >= need ? b + 2 : (s1 >= need ? b + 1 : (s0 >= need ? b : 0u)));
    };
    const uint32_t safe = max(wave_max(largest(k)), 1u);
    const uint32_t spec = max(wave_max(largest(j)), safe);
    // S(spec) and S(safe): from the lane holding that bin
    // (selects on the bits of c - b: a compare chain here becomes a table
    // lookup in private memory)
    auto s_of = [&](uint32_t c) {
        const uint32_t o = c & 3u;  // c - b, for c among this lane's bins
        const uint32_t lo = (o & 1u) ? s1 : s0, hi = (o & 1u) ? s3 : s2;
        return (o & 2u) ? hi : lo;
    };
    const uint32_t at_spec = (uint32_t)__shfl((int)(spec >= b && spec < b + 4 ? s_of(spec) : 0u), (int)(spec / 4), kWave);
    const uint32_t at_safe = (uint32_t)__shfl((int)(safe >= b && safe < b + 4 ? s_of(safe) : 0u), (int)(safe / 4), kWave);
    if (lane == 0) {
        const uint32_t sh = hshift ? hshift[q] : 0u;
        gthr[q] = max(spec << sh, 1u);
        gsafe[q] = max(safe << sh, 1u);
        want[q] = dense_want(at_spec, k, ratio16, n_objects, mult);
        want_safe[q] = dense_want(at_safe, k, ratio16, n_objects, mult);
    }
}

// Per-query candidate space from the wants: bases are the exclusive prefix of
// the wants (with `reserve_safe`, after a run that overflowed: the larger of
// the two wants, so that a recount needs no new space); a query that does not
// fit in `total` gets what is left (it will overflow, and the host grows the
// total).  `used` receives the end of the assigned space (k_dense_check
// places recounts after it).  One workgroup.
__global__ void __launch_bounds__(1024) k_cap_prefix(const uint32_t* __restrict__ want,
                                                    const uint32_t* __restrict__ want_safe, uint32_t reserve_safe,
                                                    uint32_t n_queries, uint64_t total, uint32_t* __restrict__ cap,
                                                    uint64_t* __restrict__ base, unsigned long long* __restrict__ used)
{
    __shared__ uint64_t s_w[16];
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
    uint64_t carry = 0;
    for (uint32_t q0 = 0; q0 < n_queries; q0 += 1024) {
        const uint32_t q = q0 + tid;
        const uint32_t w = q < n_queries ? (reserve_safe ? max(want[q], want_safe[q]) : want[q]) : 0u;
        uint64_t incl = w;  // 64-bit inclusive scan over the wave
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint64_t o = ((uint64_t)(uint32_t)__shfl_up((int)(uint32_t)(incl >> 32), off, kWave) << 32) |
                               (uint32_t)__shfl_up((int)(uint32_t)incl, off, kWave);
            if (lane >= off)
                incl += o;
        }
        if (lane == kWave - 1)
            s_w[wv] = incl;
        __syncthreads();
        uint64_t off = carry, all = carry;
        for (int i = 0; i < 16; ++i) {
            off += i < wv ? s_w[i] : 0;
            all += s_w[i];
        }
        if (q < n_queries) {
            const uint64_t b = off + incl - w;
            base[q] = b;
            cap[q] = b + w <= total ? w : (b < total ? (uint32_t)(total - b) : 0u);
        }
        carry = all;
        __syncthreads();
    }
    if (tid == 0)
        *used = carry;
}

// After the main pass: a query whose speculative threshold left it fewer than
// k candidates may miss objects counting between safe and spec.  Its group is
// marked for a second count; the query restarts its list from the safe
// threshold, the other queries of the group take `none` (a threshold no
// count reaches).  One wave per group.
// A restarted query takes new candidate space sized for its safe threshold
// after the assigned space (`used`, bumped atomically), when it fits in
// `total`; otherwise it keeps its space and overflows, and the host grows the
// total.
__global__ void __launch_bounds__(64) k_dense_check(const uint32_t* __restrict__ kq, uint32_t n_queries,
                                                    uint32_t* __restrict__ gthr, const uint32_t* __restrict__ gsafe,
                                                    uint32_t* __restrict__ cand_cnt, uint32_t* __restrict__ redo,
                                                    const uint32_t* __restrict__ want_safe,
                                                    uint32_t* __restrict__ cap, uint64_t* __restrict__ base,
                                                    unsigned long long* __restrict__ used, uint64_t total,
                                                    uint32_t none, uint32_t qpg)
{
    const uint32_t g = blockIdx.x, lane = threadIdx.x, q = g * qpg + lane;  // qpg queries a group
    const bool live = lane < qpg && q < n_queries;
    const bool flagged = live && gthr[q] > gsafe[q] && cand_cnt[q] < kq[q];
    const bool any = ballot(flagged) != 0;
    if (lane == 0)
        redo[g] = any ? 1u : 0u;
    if (any && live) {
        if (flagged) {
            cand_cnt[q] = 0;
            gthr[q] = gsafe[q];
            const uint32_t w = want_safe[q];
            if (w > cap[q]) {
                const unsigned long long b = atomicAdd(used, (unsigned long long)w);
                if (b + w <= total) {
                    base[q] = b;
                    cap[q] = w;
                }
            }
        } else {
            gthr[q] = none;
        }
    }
}

__device__ __forceinline__ void csa(uint32_t& hi, uint32_t& lo, uint32_t a, uint32_t b, uint32_t c)
{
    hi = __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);  // majority
    lo = __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);  // a ^ b ^ c
}

// (v & mask) | bits in one VALU op: the mask from a VGPR and the bits from an
// SGPR, since a gfx9 VOP3 encoding holds no literal.  Pure register op.
__device__ __forceinline__ uint32_t and_or_vs(uint32_t v, uint32_t mask_v, uint32_t bits_s)
{
    uint32_t r;
    asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(mask_v), "s"(bits_s));
    return r;
}

// Buffer descriptor of the code matrix: raw 32-bit loads, row offsets in
// scalar registers, the lane's object offset in one vector register.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t codes_rsrc(const uint32_t* codes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(codes), (short)0, 0x7FFFFFFF, 0x00020000);
}

// Bit-sliced counts of object o: c[b] bit j = bit b of query j's count.
// Harley-Seal: ones/twos/fours/eights absorb 16 inputs per step, the
// weight-16 carry ripples into the upper planes.  s_lds holds the group's
// table at LDS address 0 (dim d's row at d KiB).
template <int DPAD>
__device__ __forceinline__ void dense_accumulate(__amdgpu_buffer_rsrc_t rs, uint32_t row_bytes, uint32_t o,
                                                 const uint32_t* s_lds, uint32_t (&c)[8])
{
    constexpr int NW = dense_words(DPAD);
    constexpr DenseLayout<DPAD> L{};
    const char* tab = reinterpret_cast<const char*>(s_lds);
    uint32_t x[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i)
        x[i] = __builtin_amdgcn_raw_buffer_load_b32(rs, o * 4u, (uint32_t)i * row_bytes, 0);
    uint32_t ones = 0, twos = 0, fours = 0, eights = 0, h16 = 0, h32 = 0, h64 = 0, h128 = 0;
#pragma unroll
    for (int blk = 0; blk < DPAD / 16; ++blk) {
        uint32_t m[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int e = blk * 16 + i;
            const uint32_t v = x[L.w[e]];
            const bool up = L.d[e] >= 64;
            uint32_t a;
            if (L.f[e] == 0)
                a = up ? and_or_vs(v, 0x3FCu, 0x10000u) : (v & 0x3FCu);
            else if (L.f[e] == 1)
                a = (v >> 11) & 0x3FFu;
            else
                a = up ? __builtin_amdgcn_alignbit(0x40u, v, 22u) : (v >> 22);
            const uint32_t off = (uint32_t)(up ? L.d[e] - 64 : L.d[e]) * 1024u;
            m[i] = *reinterpret_cast<const uint32_t*>(tab + a + off);
        }
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

// All 32 counts of the lane at once: byte column q of the eight planes forms
// an 8 x 8 bit matrix (row b = plane b); its transpose (three delta swaps)
// holds in byte j the count of query 8 q + j.  out[q] = counts of queries
// 8 q .. 8 q + 7, one per byte.
__device__ __forceinline__ void dense_all_counts(const uint32_t (&c)[8], uint64_t (&out)[4])
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint64_t x = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b)
            x |= (uint64_t)((c[b] >> (8 * q)) & 255u) << (8 * b);
        uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
        x = x ^ t ^ (t << 7);
        t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
        x = x ^ t ^ (t << 14);
        t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
        x = x ^ t ^ (t << 28);
        out[q] = x;
    }
}

template <int NP>
__device__ __forceinline__ uint32_t dense_lane_count(const uint32_t (&c)[NP], uint32_t j)
{
    uint32_t v = 0;
#pragma unroll
    for (int b = 0; b < NP; ++b)
        v |= ((c[b] >> j) & 1u) << b;
    return v;
}

// LDS bytes of the dense kernels.
// Main pass: a tile's candidates are staged in LDS (up to kDenseStage of
// them), so that each query reserves its candidates of the tile with one
// global atomic.
constexpr uint32_t kDenseStage = 2048;
// Main pass: tiles whose candidates share the stage between two flushes
// (a tile of 4,096 objects stages about a hundred on the bench batches; a
// hit that finds the stage full goes straight to its query's list).
constexpr uint32_t kDenseFlushTiles = 8;
constexpr uint32_t kDenseStageWords = 2 + 32 + 32 + kDenseStage * 3;  // count, pad, per-query counts / bases, meta, keys

// wp: weight planes per query in a mask word (1: unit weights, 32 queries a
// group; 2, 3 or 4: weighted, 16, 10 or 8 queries, sample histograms of 256
// bins).
__host__ __device__ constexpr uint32_t dense_lds_bytes(int nblk, bool sample, int wp = 1)
{
    return (uint32_t)(nblk * 16 * kDenseKeysPerDim + 16 +
                      (sample ? (wp > 1 ? (32 / wp) * 256 : 32 * (nblk * 16 + 1)) : kDenseStageWords)) * 4u;
}

// Bits of the queries of a mask word with WP weight planes per query: bit
// WP q for q < 32 / WP (10 queries of three planes leave bits 30-31 unused).
__host__ __device__ constexpr uint32_t dense_query_bits(int wp)
{
    return wp == 4 ? 0x11111111u : (wp == 3 ? 0x09249249u : (wp == 2 ? 0x55555555u : 0xFFFFFFFFu));
}

// Weighted masks: bit WP q + p of a word is bit p of query q's weight on the
// key, so the eight bit-sliced count planes hold, at bit WP q + p, the number
// of dims whose weight has bit p set.  The query's count, sum over p of that
// number times 2^p, is assembled at bit WP q of 11 planes (the weights' bit
// groups aligned, then ripple adds of the shifted groups).
template <int WP>
__device__ __forceinline__ void dense_weigh(const uint32_t (&c)[8], uint32_t (&w)[kSetCountPlanes])
{
    constexpr uint32_t M = dense_query_bits(WP);
#pragma unroll
    for (int b = 0; b < kSetCountPlanes; ++b)
        w[b] = b < 8 ? c[b] & M : 0u;
#pragma unroll
    for (int p = 1; p < WP; ++p) {
        uint32_t carry = 0;
#pragma unroll
        for (int b = p; b < kSetCountPlanes; ++b) {
            const uint32_t x = b - p < 8 ? (c[b - p] >> p) & M : 0u;
            const uint32_t t = __builtin_amdgcn_bitop3_b32(w[b], x, carry, 0xE8);  // majority
            w[b] = __builtin_amdgcn_bitop3_b32(w[b], x, carry, 0x96);              // xor3
            carry = t;
        }
    }
}

// Set form: hash keys, weight planes 0-3 (uint4 per slot) and 4-5 (uint2 per
// slot), then the same threshold planes, sample histogram (256 bins per
// query) or candidate stage.
__host__ __device__ constexpr uint32_t set_lds_bytes(bool sample)
{
    return (kSetTabWords + 16 + (sample ? 32 * 256 : kDenseStageWords)) * 4u;
}

// Bit-sliced sum of two N-plane numbers (ripple of full adders, two VALU ops
// a plane): N + 1 planes.
template <int N>
__device__ __forceinline__ void bs_add(const uint32_t (&x)[N], const uint32_t (&y)[N], uint32_t (&z)[N + 1])
{
    uint32_t carry = x[0] & y[0];
    z[0] = x[0] ^ y[0];
#pragma unroll
    for (int b = 1; b < N; ++b) {
        z[b] = __builtin_amdgcn_bitop3_b32(x[b], y[b], carry, 0x96);              // xor3
        carry = __builtin_amdgcn_bitop3_b32(x[b], y[b], carry, 0xE8);           // majority
    }
    z[N] = carry;
}

// c += x for an N-plane x into the NP-plane count (N <= NP; counts stay
// below 2^NP).
template <int NP, int N>
__device__ __forceinline__ void bs_acc(uint32_t (&c)[NP], const uint32_t (&x)[N])
{
    uint32_t carry = c[0] & x[0];
    c[0] ^= x[0];
#pragma unroll
    for (int b = 1; b < N; ++b) {
        const uint32_t t = __builtin_amdgcn_bitop3_b32(c[b], x[b], carry, 0xE8);
        c[b] = __builtin_amdgcn_bitop3_b32(c[b], x[b], carry, 0x96);
        carry = t;
    }
#pragma unroll
    for (int b = N; b < NP; ++b) {
        const uint32_t t = c[b] & carry;
        c[b] ^= carry;
        carry = t;
    }
}

// Bit-sliced counts of the object in slot `sl` over its forward list: every
// occurrence whose key the group's queries hold adds each query's weight
// (counts stay below kSetNoQuery: the host admits a batch only when the
// longest list times each query's largest weight per key does).  The slots of
// a wave form one block, whose lists it reads a column (64 words) per step.
// WIDE: some weight of the group reaches 16 (planes 4-5 are read), chosen once
// per group.  A step looks up S = 8 columns (S = 4 when at most 4 are left,
// so that a list ends with fewer padding lookups): the S weights are summed
// by a tree of bit-sliced adders (4 + 2 + 1 adds of 4 to 6 planes for 8),
// then added to the count planes once: the carry ripples through the upper
// count planes once per step instead of once per occurrence.  No branch per
// occurrence but the rare extra probes.
template <int NP, bool WIDE, int S>
__device__ __forceinline__ void set_step(__amdgpu_buffer_rsrc_t rk, uint32_t lane_off, uint32_t col,
                                         const uint2* s_slot, const uint4* s_mlo, const uint2* s_mhi,
                                         uint32_t (&c)[NP])
{
    static_assert(S == 8 || S == 4, "an adder tree of four or eight weights");
    constexpr int P = WIDE ? kSetWeightPlanes : 4;  // weight planes read (weights < 16 unless WIDE)
    // S columns loaded at once, then looked up one by one
    uint32_t kk[S];
#pragma unroll
    for (int u = 0; u < S; ++u)
        kk[u] = __builtin_amdgcn_raw_buffer_load_b32(rk, lane_off + (col + (uint32_t)u) * 256u, 0, 0);
    // probe to the key or an empty slot (the table is at most a quarter
    // full, so every probe sequence ends there, mostly at the first pair);
    // an empty slot's entry has zero weights (a key the group does not hold,
    // or the padding key 0, which no table holds)
    uint32_t x[S][P];
#pragma unroll
    for (int u = 0; u < S; ++u) {
        const uint32_t key = kk[u];
        // a pair of (key, entry) slots in one 128-bit read: a lookup that
        // ends in its home pair takes two dependent LDS reads, the pair and
        // the weights.  A key is in the first pair that holds it or has an
        // empty second slot; a miss takes an empty slot's entry (the second
        // slot's, empty whenever the first is)
        const uint4* s_pair = reinterpret_cast<const uint4*>(s_slot);
        uint32_t pr = set_hash(key) >> 1;
        // the entries are pinned in registers where the pair is read (an
        // empty asm), so that the compiler reads the pair in one instruction
        // instead of the keys first and an entry afterwards
        uint4 pv = s_pair[pr];
        __asm__ volatile("" : "+v"(pv.y), "+v"(pv.w));
        for (uint32_t probe = 1;
             ((pv.x != key) & (pv.z != key) & (pv.z != kSetEmpty)) && probe < kSetSlots / 2; ++probe) {
            pr = (pr + 1u) & (kSetSlots / 2 - 1u);
            pv = s_pair[pr];
            __asm__ volatile("" : "+v"(pv.y), "+v"(pv.w));
        }
        const uint32_t e = pv.x == key ? pv.y : pv.w;
        const uint4 lo = s_mlo[e];
        x[u][0] = lo.x;
        x[u][1] = lo.y;
        x[u][2] = lo.z;
        x[u][3] = lo.w;
        if constexpr (WIDE) {
            const uint2 hi = s_mhi[e];
            x[u][4] = hi.x;
            x[u][5] = hi.y;
        }
    }
    uint32_t y[S / 2][P + 1], z[2][P + 2];
#pragma unroll
    for (int i = 0; i < S / 2; ++i)
        bs_add<P>(x[2 * i], x[2 * i + 1], y[i]);
    if constexpr (S == 8) {
        uint32_t t[P + 3];
        bs_add<P + 1>(y[0], y[1], z[0]);
        bs_add<P + 1>(y[2], y[3], z[1]);
        bs_add<P + 2>(z[0], z[1], t);
        bs_acc<NP, P + 3>(c, t);
    } else {
        bs_add<P + 1>(y[0], y[1], z[0]);
        bs_acc<NP, P + 2>(c, z[0]);
    }
}

template <int NP, bool WIDE>
__device__ __forceinline__ void set_accumulate(const DenseArgs& a, uint32_t sl, const uint2* s_slot,
                                               const uint4* s_mlo, const uint2* s_mhi, uint32_t (&c)[NP])
{
    constexpr uint32_t S = kSetPrefetch, S2 = kSetPrefetch / 2;
#pragma unroll
    for (int b = 0; b < NP; ++b)
        c[b] = 0;
    // the wave's 64 slots are one block (tiles and workgroups are multiples
    // of 64 slots): its columns through a buffer that ends at the block, so
    // that the loads past its last column read 0, the padding key
    const uint32_t blk = __builtin_amdgcn_readfirstlane(sl >> 6);
    const uint32_t w0 = __builtin_amdgcn_readfirstlane(a.blk_off[blk]);
    const uint32_t ncol = __builtin_amdgcn_readfirstlane((a.blk_off[blk + 1] - w0) >> 6);
    const __amdgpu_buffer_rsrc_t rk = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(a.blk_key) + w0,
                                                                        (short)0, (int)(ncol * 256u), 0x00020000);
    const uint32_t lane_off = (sl & 63u) * 4u;
    for (uint32_t col = 0; col < ncol;) {  // uniform
        if (ncol - col > S2) {
            set_step<NP, WIDE, S>(rk, lane_off, col, s_slot, s_mlo, s_mhi, c);
            col += S;
        } else {
            set_step<NP, WIDE, S2>(rk, lane_off, col, s_slot, s_mlo, s_mhi, c);
            col += S2;
        }
    }
}

// Sample-pass bins of 11-plane counts as 8 bit-sliced planes: bit b of
// plane k is bit k + s of the count at bit b, s its query's shift (the
// positions of the queries with shift s are set in by_shift[s]), all ones
// when the count reaches 256 << s.
template <int NP>
__device__ __forceinline__ void dense_bin_planes(const uint32_t (&c)[NP], const uint32_t (&by_shift)[4],
                                                 uint32_t (&bp)[8])
{
    static_assert(NP >= 11, "8-bit bins of 11-plane counts");
    uint32_t sat = 0;
#pragma unroll
    for (int sh = 0; sh < 4; ++sh)
#pragma unroll
        for (int b = 8 + sh; b < NP; ++b)
            sat |= c[b] & by_shift[sh];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint32_t v = sat;
#pragma unroll
        for (int sh = 0; sh < 4; ++sh)
            v |= c[k + sh] & by_shift[sh];
        bp[k] = v;
    }
}

// Workgroup -> (query group, tiles) so that the groups of one tile run on one
// XCD (workgroup ids are dealt to the 8 XCDs round robin) and each
// workgroup loads one group's mask table once.
// SETS (NBLK = 0): the set form -- each workgroup builds its group's key
// table from the host's entries and counts objects over their forward lists.
// NT: threads per workgroup, kDenseThreads (16 waves over the CU's table)
// or half of it for runtimes whose occupancy rule places no 1,024-thread
// workgroup of more than 64 VGPRs.  Both are compiled for 1,024 threads, so
// both stay within 128 VGPRs (two such waves per SIMD fit any such rule).
template <int NBLK, bool SAMPLE, bool SETS = false, int NT = kDenseThreads, int WP = 1>
__global__ void __launch_bounds__(kDenseThreads) k_dense_count(DenseArgs a)
{
    static_assert(WP >= 1 && WP <= 4, "weight planes per query");
    // the kernel's only LDS, static so that the table sits at address 0 and
    // every table offset folds into the ds_read instruction
    constexpr int DPAD = NBLK * 16;
    constexpr bool WEIGHTED = !SETS && WP > 1;               // weighted masks (dense_weigh)
    constexpr uint32_t QPG = SETS ? 32u : 32u / WP;          // queries of a group
    constexpr uint32_t QMASK = SETS ? 0xFFFFFFFFu : dense_query_bits(WP);  // their bits
    constexpr int NP = (SETS || WEIGHTED) ? kSetCountPlanes : 8;  // count planes
    constexpr uint32_t NONE = (SETS || WEIGHTED) ? kSetNoQuery : kDenseNoQuery;
    __shared__ uint32_t s_lds[(SETS ? set_lds_bytes(SAMPLE) : dense_lds_bytes(NBLK, SAMPLE, WP)) / 4];
    uint32_t* s_tab = s_lds;                             // DPAD * 256 masks
    uint2* s_slot = reinterpret_cast<uint2*>(s_lds);    // SETS: (key, entry) slots, then the entries' planes
    uint4* s_mlo = reinterpret_cast<uint4*>(s_lds + kSetMloWord);
    uint2* s_mhi = reinterpret_cast<uint2*>(s_lds + kSetMhiWord);
    uint32_t* s_thr = s_lds + (SETS ? kSetTabWords : DPAD * kDenseKeysPerDim);  // NP threshold planes
    uint32_t* s_hist = s_thr + 16;                       // SAMPLE: 32 x (DPAD + 1), SETS 32 x 256
    // main pass: the tile's staged candidates
    uint32_t* s_nhit = s_thr + 16;                       // staged so far
    uint32_t* s_qcnt = s_nhit + 2;                       // per query of the group
    uint32_t* s_qbase = s_qcnt + 32;                     // reserved base per query
    uint32_t* s_meta = s_qbase + 32;                     // query << 16 | rank within the query
    uint64_t* s_key = reinterpret_cast<uint64_t*>(s_meta + kDenseStage);
    constexpr uint32_t HB = (SETS || WEIGHTED) ? 256 : NBLK * 16 + 1;
    const uint32_t tid = threadIdx.x;
    // work units u = c G + g (chunk c of group g): workgroup s of XCD x takes
    // units s, s + P, ...; chunk-major, so that the workgroups of an XCD pass
    // over the same tiles together and read them from its L2
    const uint32_t P = gridDim.x / 8, x = blockIdx.x % 8, s = blockIdx.x / 8;
    const uint32_t n_units = a.G * a.chunks, t_step = 8 * a.chunks;
    if (s >= n_units)
        return;  // the whole workgroup leaves before any barrier
    const uint32_t n_list = a.list ? a.n_list : a.n_tiles;
    const uint32_t tab_words = NBLK * 16 * kDenseKeysPerDim;
    const __amdgpu_buffer_rsrc_t rs = codes_rsrc(a.codes);
    const uint32_t row_bytes = a.n_pad * 4u;
    // set form, main pass: tiles in windows of kSetWindowTiles per XCD, each
    // window counted for all of the workgroup's groups before the next, so
    // that the 32 workgroups of an XCD read one window's lists (about 3 MB)
    // from its L2 (a group's table is built again per window); elsewhere one
    // window holds every tile
    const uint32_t win = (SETS && !SAMPLE) ? t_step * kSetWindowTiles : n_list;
    for (uint32_t w0 = 0; w0 < n_list; w0 += win)
    for (uint32_t u = s; u < n_units; u += P) {
        const uint32_t g = u % a.G, t_first = x + 8 * (u / a.G);
        // sample pass: a unit without a sample tile builds no table, unless
        // it is the one that writes its group's table out (chunk 0 on XCD 0)
        const bool writer = !SETS && x == 0 && u < a.G;
        if ((SAMPLE && t_first >= n_list && !writer) || (a.redo && !a.redo[g]))
            continue;  // uniform across the workgroup
        __syncthreads();
        bool wide = false;
        if constexpr (SETS) {
            // the group's key table, built from its entries (unique keys);
            // empty slots point at the zero entry
            for (uint32_t i = tid; i < kSetSlots; i += NT)
                s_slot[i] = make_uint2(kSetEmpty, kSetEntries);
            if (tid == 0) {
                s_mlo[kSetEntries] = make_uint4(0, 0, 0, 0);
                s_mhi[kSetEntries] = make_uint2(0, 0);
            }
            if constexpr (SAMPLE)
                for (uint32_t i = tid; i < QPG * HB; i += NT)
                    s_hist[i] = 0;
            __syncthreads();
            int hi_any = 0;
            const uint32_t e0 = a.set_gptr[g];
            for (uint32_t e = e0 + tid; e < a.set_gptr[g + 1]; e += NT) {
                const uint32_t key = a.set_keys[e] + 1u;  // as the forward lists hold it
                const uint4 lo = a.set_mask[2 * e], hi = a.set_mask[2 * e + 1];
                hi_any |= (hi.x | hi.y) != 0u;
                s_mlo[e - e0] = lo;
                s_mhi[e - e0] = make_uint2(hi.x, hi.y);
                uint32_t h = set_hash(key);
                for (uint32_t probe = 0; probe < kSetSlots; ++probe) {
                    if (atomicCAS(&s_slot[h].x, kSetEmpty, key) == kSetEmpty) {
                        s_slot[h].y = e - e0;
                        break;
                    }
                    h = (h + 1u) & (kSetSlots - 1u);
                }
            }
            wide = __syncthreads_or(hi_any) != 0;
        } else if constexpr (SAMPLE) {
            // build the group's mask table from its items: one bit per
            // (query, covered key), at the key's slot
            uint4* dst = reinterpret_cast<uint4*>(s_tab);
            for (uint32_t i = tid; i < tab_words / 4; i += NT)
                dst[i] = make_uint4(0, 0, 0, 0);
            for (uint32_t i = tid; i < QPG * HB; i += NT)
                s_hist[i] = 0;
            __syncthreads();
            const uint32_t i0 = a.item_qptr[g * QPG], i1 = a.item_qptr[min(g * QPG + QPG, a.n_queries)];
            for (uint32_t i = i0 + tid; i < i1; i += NT) {
                // first key | query's bit position << 24, end key (| weight << 16
                // in a weighted batch)
                const uint32_t k0 = a.dkeys[2 * i], k1 = a.dkeys[2 * i + 1] & 0xFFFFu;
                const uint32_t bit = (WEIGHTED ? a.dkeys[2 * i + 1] >> 16 : 1u) << (k0 >> 24);
                // up to 16 slot loads in flight per item, then the LDS ors
                for (uint32_t kb = k0 & 0xFFFFFFu; kb < k1; kb += 16) {
                    uint32_t sl[16];
#pragma unroll
                    for (int j = 0; j < 16; ++j)
                        sl[j] = kb + j < k1 ? a.key_slot[kb + j] : 0u;
#pragma unroll
                    for (int j = 0; j < 16; ++j)
                        if (kb + j < k1)
                            atomicOr(&s_tab[sl[j]], bit);
                }
            }
            __syncthreads();
            if (writer) {  // one workgroup per group keeps it for the main pass
                uint4* out = reinterpret_cast<uint4*>(a.tab + (uint64_t)g * tab_words);
                for (uint32_t i = tid; i < tab_words / 4; i += NT)
                    out[i] = dst[i];
            }
        } else {
            const uint4* src = reinterpret_cast<const uint4*>(a.tab + (uint64_t)g * tab_words);
            uint4* dst = reinterpret_cast<uint4*>(s_tab);
            for (uint32_t i = tid; i < tab_words / 4; i += NT)
                dst[i] = src[i];
        }
        // SETS and weighted sample passes: the bin shift of each query of the
        // group, as four masks of the queries' bit positions by shift
        uint32_t by_shift[4] = {0u, 0u, 0u, 0u};
        if constexpr ((SETS || WEIGHTED) && SAMPLE) {
            for (uint32_t j = 0; j < QPG && g * QPG + j < a.n_queries; ++j)
                by_shift[min(a.hshift[g * QPG + j], 3u)] |= 1u << (WEIGHTED ? WP * j : j);
#pragma unroll
            for (int sh = 0; sh < 4; ++sh)  // uniform: scalar registers
                by_shift[sh] = __builtin_amdgcn_readfirstlane(by_shift[sh]);
        }
        uint32_t tp[NP] = {};
        if constexpr (!SAMPLE) {
            // threshold planes of the group (bit j of plane b = bit b of
            // query j's threshold; NONE for lanes past the last query): no
            // kernel changes a threshold while a pass runs, so once per group;
            // the stage counters start at zero
            __syncthreads();
            if (tid < 32)
                s_qcnt[tid] = 0;
            if (tid == 0)
                *s_nhit = 0;
            if (tid < kWave) {
                // bit WP q of the planes: query q of the group
                const uint32_t q = g * QPG + tid / WP;
                const uint32_t th = (tid < QPG * WP && tid % WP == 0 && q < a.n_queries) ? a.gthr[q] : NONE;
#pragma unroll
                for (int b = 0; b < NP; ++b) {
                    const uint32_t pl = (uint32_t)ballot((th >> b) & 1u);
                    if (tid == 0)
                        s_thr[b] = pl;
                }
            }
            __syncthreads();
#pragma unroll
            for (int b = 0; b < NP; ++b)
                tp[b] = __builtin_amdgcn_readfirstlane(s_thr[b]);
        }
        uint32_t staged_tiles = 0;
        const uint32_t li_end = min(n_list, w0 + win);
        for (uint32_t li = t_first + w0; li < li_end; li += t_step) {
            const uint32_t t = a.list ? a.list[li] : li;
            if constexpr (SAMPLE)
                __syncthreads();
            const uint32_t o0 = t * kDenseTile;
            const uint32_t o1 = min(o0 + kDenseTile, SETS ? a.n_slots : a.n_objects);
            for (uint32_t ol = o0 + tid; ol < o1; ol += NT) {
                uint32_t c[NP];
                uint32_t o = ol;
                if constexpr (SETS) {
                    if (wide)  // uniform across the workgroup
                        set_accumulate<NP, true>(a, ol, s_slot, s_mlo, s_mhi, c);
                    else
                        set_accumulate<NP, false>(a, ol, s_slot, s_mlo, s_mhi, c);
                    o = a.obj_perm[ol];  // kSetEmpty past the last object: counts 0, nothing emitted
                } else if constexpr (WEIGHTED) {
                    uint32_t c8[8];
                    dense_accumulate<DPAD>(rs, row_bytes, o, s_tab, c8);
                    dense_weigh<WP>(c8, c);
                } else {
                    dense_accumulate<DPAD>(rs, row_bytes, o, s_tab, c);
                }
                if constexpr (SAMPLE && SETS) {
                    // the bins as 8 planes, filed as the unit form files its
                    // counts (queries rotated per lane, zero counts too)
                    static_assert(HB == 256, "8-bit bins");
                    uint32_t bp[8];
                    dense_bin_planes<NP>(c, by_shift, bp);
                    const uint32_t r = tid & 31u;
                    uint32_t cr[8];
#pragma unroll
                    for (int b = 0; b < 8; ++b)
                        cr[b] = __builtin_amdgcn_alignbit(bp[b], bp[b], r);  // bit j: query (j + r) mod 32
                    uint64_t cnt[4];
                    dense_all_counts(cr, cnt);
#pragma unroll
                    for (int q = 0; q < 4; ++q)
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const uint32_t v = (uint32_t)(cnt[q] >> (8 * j)) & 255u;
                            atomicAdd(&s_hist[((8u * q + j + r) & 31u) * HB + v], 1u);
                        }
                } else if constexpr (SAMPLE && WEIGHTED) {
                    // the bins as 8 planes at the queries' bit positions WP j,
                    // transposed once; query j's bin is then byte WP j (zero
                    // counts are filed too)
                    static_assert(HB == 256, "8-bit bins");
                    uint32_t bp[8];
                    dense_bin_planes<NP>(c, by_shift, bp);
                    uint64_t cnt[4];
                    dense_all_counts(bp, cnt);
#pragma unroll
                    for (uint32_t jq = 0; jq < QPG; ++jq) {
                        const uint32_t p = WP * jq;
                        const uint32_t v = (uint32_t)(cnt[p >> 3] >> (8 * (p & 7))) & 255u;
                        atomicAdd(&s_hist[jq * HB + v], 1u);
                    }
                } else if constexpr (SAMPLE) {
                    // lane l files query (i + l) mod 32 at step i, so that one
                    // update of the 64 lanes spreads over 32 queries' rows (one
                    // query's row from every lane piles onto the few bins its
                    // counts share).  Zero counts are filed too: the thresholds
                    // read bins >= 1 only.
                    const uint32_t r = tid & 31u;
                    uint32_t cr[8];
#pragma unroll
                    for (int b = 0; b < 8; ++b)
                        cr[b] = __builtin_amdgcn_alignbit(c[b], c[b], r);  // bit j: query (j + r) mod 32
                    uint64_t cnt[4];
                    dense_all_counts(cr, cnt);
#pragma unroll
                    for (int q = 0; q < 4; ++q)
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const uint32_t v = (uint32_t)(cnt[q] >> (8 * j)) & 255u;
                            atomicAdd(&s_hist[((8u * q + j + r) & 31u) * HB + v], 1u);
                        }
                } else {
                    // count >= threshold, per lane, most significant plane first
                    uint32_t gt = 0, eq = 0xFFFFFFFFu;
#pragma unroll
                    for (int b = NP - 1; b >= 0; --b) {
                        gt |= eq & c[b] & ~tp[b];
                        eq &= ~(c[b] ^ tp[b]);
                    }
                    uint32_t ge = (gt | eq) & QMASK;
                    while (ge) {
                        const uint32_t j = (uint32_t)__builtin_ctz(ge);
                        ge &= ge - 1;
                        const uint32_t jq = j / WP;  // the group's query
                        const uint64_t key = ((uint64_t)dense_lane_count<NP>(c, j) << 32) | (uint32_t)~o;
                        // staged when the stage has room (the rank is taken only
                        // then, so a query's reserved slots have no holes)
                        const uint32_t i = atomicAdd(s_nhit, 1u);
                        if (i < kDenseStage) {
                            s_meta[i] = jq << 16 | atomicAdd(&s_qcnt[jq], 1u);
                            s_key[i] = key;
                        } else {
                            const uint32_t q = g * QPG + jq;
                            const uint32_t slot = atomicAdd(&a.cand_cnt[q], 1u);
                            if (slot < a.cand_cap[q])
                                a.cand[a.cand_base[q] + slot] = key;
                            else
                                atomicOr(a.overflow, 1u);
                        }
                    }
                }
            }
            // every kDenseFlushTiles tiles and after the group's last one
            // (uniform across the workgroup): one reservation per query of
            // the group, then the staged candidates to their slots; the stage
            // counters are reset once every thread has read them
            if constexpr (!SAMPLE)
            if (++staged_tiles == kDenseFlushTiles || li + t_step >= li_end) {
                staged_tiles = 0;
                __syncthreads();
                const uint32_t nh = min(*s_nhit, kDenseStage);
                if (tid < QPG && s_qcnt[tid])
                    s_qbase[tid] = atomicAdd(&a.cand_cnt[g * QPG + tid], s_qcnt[tid]);
                __syncthreads();
                for (uint32_t i = tid; i < nh; i += NT) {
                    const uint32_t m = s_meta[i], j = m >> 16;
                    const uint32_t q = g * QPG + j;
                    const uint32_t slot = s_qbase[j] + (m & 0xFFFFu);
                    if (slot < a.cand_cap[q])
                        a.cand[a.cand_base[q] + slot] = s_key[i];
                    else
                        atomicOr(a.overflow, 1u);
                }
                if (tid < 32)
                    s_qcnt[tid] = 0;
                if (tid == 0)
                    *s_nhit = 0;
                __syncthreads();  // the stage is free before the next tile's hits
            }
        }
        if constexpr (SAMPLE) {
            __syncthreads();
            for (uint32_t i = tid; i < QPG * HB; i += NT) {
                const uint32_t j = i / HB, bin = i % HB;
                const uint32_t q = g * QPG + j;
                const uint32_t v = s_hist[i];
                if (v && q < a.n_queries)
                    atomicAdd(&a.hist[(uint64_t)q * kDenseBins + bin], v);
            }
        }
    }
}

}  // namespace gh
