// Batched form: one persistent kernel over a device work queue of
// (query group, object tile) items.  The queries of a group share the
// workgroup: every object's counters of all of them sit in one u64 word
// (64 / LB lanes of LB bits), and a posting span that several queries of the
// group hold is read once and counted for all of them with one 64-bit LDS
// atomic whose increment carries each query's weight in its lane.  Then every
// lane goes through the same exact per-tile select as the plain form.
//
// The reference batches whole tables, one match_AT launch each, admitted by
// memory (knn.cu:185-283, match.cu:502-693); the queue here is per tile and
// per query group inside one launch.
#pragma once
#include "kernels.h"

namespace gh {

// One merged work pair of a group: postings [lo, hi) of `key` (positions;
// segment indices for a compressed index) and the lanes' summed increments.
struct GPair {
    uint32_t key, lo, hi, pad;
    uint64_t inc;
};

struct BSpan {
    uint32_t lo, hi, filter, pad;
    uint64_t inc;
};
constexpr int kBSpanCap = 512;

struct BatchArgs {
    ScanArgs s;            // index, per-query k / threshold / candidates
    const GPair* pairs;
    const uint32_t* gp_ptr;  // [n_groups + 1]
    const int32_t* gq;       // [n_groups * lanes] query of every lane, -1 if unused
    uint32_t n_groups, lanes, n_work;
    uint32_t* queue;         // next work item
};

__device__ __forceinline__ void count_one64(uint32_t id, uint32_t obj0, uint32_t tile_objs, uint64_t inc,
                                            uint64_t* cnt)
{
    const uint32_t loc = id - obj0;
    const bool ok = loc < tile_objs;
    const uint32_t word = ok ? loc : tile_objs + lane_id();  // per-lane scratch word
    atomicAdd((unsigned long long*)&cnt[word], (unsigned long long)(ok ? inc : 0ull));
}

__device__ __forceinline__ void count_span64(const uint32_t* __restrict__ post, uint32_t a, uint32_t b,
                                             uint32_t obj0, uint32_t tile_objs, uint64_t inc, uint64_t* cnt)
{
    const int lane = lane_id();
    const uint32_t a4 = (a + 3u) & ~3u, b4 = b & ~3u;
    if (a4 >= b4) {
        for (uint32_t p = a + lane; p < b; p += kWave)
            count_one64(post[p], obj0, tile_objs, inc, cnt);
        return;
    }
    if (lane < (int)(a4 - a))
        count_one64(post[a + lane], obj0, tile_objs, inc, cnt);
    if (lane < (int)(b - b4))
        count_one64(post[b4 + lane], obj0, tile_objs, inc, cnt);
    const uint4* body = reinterpret_cast<const uint4*>(post + a4);
    const uint32_t ng = (b4 - a4) >> 2;
    const uint32_t past = obj0 + tile_objs;
    for (uint32_t g0 = 0; g0 < ng; g0 += 4 * kWave) {
        uint4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            v[u] = body[min(g0 + u * kWave + lane, ng - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool in = g0 + u * kWave + lane < ng;
            count_one64(in ? v[u].x : past, obj0, tile_objs, inc, cnt);
            count_one64(in ? v[u].y : past, obj0, tile_objs, inc, cnt);
            count_one64(in ? v[u].z : past, obj0, tile_objs, inc, cnt);
            count_one64(in ? v[u].w : past, obj0, tile_objs, inc, cnt);
        }
    }
}

template <int LB, bool CMP, int NT>
__global__ void __launch_bounds__(NT) k_scan_batched(BatchArgs b)
{
    constexpr int NW = NT / kWave;
    extern __shared__ uint64_t s_cnt[];  // tile_objs counters + kWave scratch words
    __shared__ BSpan s_span[kBSpanCap];
    __shared__ uint32_t s_work, s_next;
    __shared__ uint32_t s_wsum[NW];
    __shared__ SelShared s_sel;
    const ScanArgs& a = b.s;
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
    const uint32_t T = a.tile_objs;

    for (;;) {
        if (tid == 0)
            s_work = atomicAdd(b.queue, 1u);
        __syncthreads();
        const uint32_t w = s_work;
        if (w >= b.n_work)
            break;  // every wave of every workgroup reaches this once the queue is empty
        const uint32_t t = w / b.n_groups, g = w % b.n_groups;
        const uint32_t obj0 = t * T, obj1 = obj0 + T;
        const uint32_t dt0 = t * a.dt_per_tile;
        const uint32_t dt1 = min(dt0 + a.dt_per_tile, a.n_dtiles);
        uint4* c4 = reinterpret_cast<uint4*>(s_cnt);
        for (uint32_t i = tid; i < (T + kWave) / 2; i += NT)
            c4[i] = make_uint4(0, 0, 0, 0);

        const uint32_t p0 = b.gp_ptr[g], np = b.gp_ptr[g + 1] - p0;
        for (uint32_t w0 = 0; w0 < np; w0 += kBSpanCap) {
            const uint32_t wn = min((uint32_t)kBSpanCap, np - w0);
            if (tid == 0)
                s_next = 0;
            for (uint32_t i = tid; i < wn; i += NT) {
                const GPair pr = b.pairs[p0 + w0 + i];
                BSpan sp;
                sp.inc = pr.inc;
                sp.pad = 0;
                uint32_t lo = pr.lo, hi = pr.hi;
                const uint32_t doff = CMP ? a.sdir_off[pr.key] : a.dir_off[pr.key];
                if constexpr (CMP) {
                    const uint32_t ks = a.key_seg[pr.key], ke = a.key_seg[pr.key + 1];
                    if (doff != kNoDir)
                        sdir_clip(a.sdir, doff, ks, ke, dt0, dt1, obj1, a.seg_first, lo, hi);
                    sp.filter = 1;
                } else {
                    sp.filter = 1;
                    if (doff != kNoDir)
                        sp.filter = dir_clip(a.dir, doff, a.key_start[pr.key], dt0, dt1, a.n_dtiles, lo, hi) ? 0 : 1;
                }
                sp.lo = lo;
                sp.hi = max(lo, hi);
                s_span[i] = sp;
            }
            __syncthreads();
            if constexpr (!CMP) {
                for (;;) {
                    uint32_t si = 0;
                    if (lane == 0)
                        si = atomicAdd(&s_next, 1u);
                    si = bcast_first(__shfl((int)si, 0, kWave));
                    if (si >= wn)
                        break;
                    const BSpan sp = s_span[si];
                    if (sp.lo < sp.hi)
                        count_span64(a.post, sp.lo, sp.hi, obj0, T, sp.inc, s_cnt);
                }
            } else {
                // waves take single segments: prefix of segment counts, then decode
                __shared__ uint32_t s_pre[kBSpanCap + 1];
                __shared__ DecodeScratch s_dec[NW];
                {
                    const uint32_t per = (kBSpanCap + NT - 1) / NT;  // spans per thread
                    uint32_t loc = 0;
                    for (uint32_t j = 0; j < per; ++j) {
                        const uint32_t i = tid * per + j;
                        loc += i < wn ? s_span[i].hi - s_span[i].lo : 0u;
                    }
                    const uint32_t incl = wave_incl_scan(loc);
                    if (lane == kWave - 1)
                        s_wsum[wv] = incl;
                    __syncthreads();
                    uint32_t off = 0;
                    for (int i = 0; i < wv; ++i)
                        off += s_wsum[i];
                    uint32_t run = off + incl - loc;
                    for (uint32_t j = 0; j < per; ++j) {
                        const uint32_t i = tid * per + j;
                        if (i <= wn)
                            s_pre[i] = run;
                        run += i < wn ? s_span[i].hi - s_span[i].lo : 0u;
                        if (i + 1 == wn)
                            s_pre[wn] = run;
                    }
                }
                __syncthreads();
                const uint32_t nsegs = s_pre[wn];
                DecodeScratch& D = s_dec[wv];
                for (;;) {
                    uint32_t gi = 0;
                    if (lane == 0)
                        gi = atomicAdd(&s_next, 1u);
                    gi = bcast_first(__shfl((int)gi, 0, kWave));
                    if (gi >= nsegs)
                        break;
                    uint32_t lo = 0, hi = wn;
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (s_pre[mid] <= gi)
                            lo = mid;
                        else
                            hi = mid;
                    }
                    const BSpan sp = s_span[lo];
                    const uint32_t sg = sp.lo + (gi - s_pre[lo]);
                    if (a.seg_last[sg] < obj0 || a.seg_first[sg] >= obj1)
                        continue;
                    const uint32_t cw0 = a.seg_cpos[sg], nw = a.seg_cpos[sg + 1] - cw0;
                    stage_words(D, a.words, cw0, nw);
                    uint32_t v[kPerLane];
                    const uint32_t n = wave_decode(a.codec, D, nw, v, a.seg_bph ? a.seg_bph[sg] : uint4{},
                                                     a.seg_bph != nullptr);
#pragma unroll
                    for (int j = 0; j < kPerLane; ++j) {
                        const uint32_t i = lane * kPerLane + j;
                        count_one64(i < n ? v[j] : obj1, obj0, T, sp.inc, s_cnt);
                    }
                    wave_sync();
                }
            }
            __syncthreads();
        }
        if (np == 0)
            __syncthreads();
        for (uint32_t l = 0; l < b.lanes; ++l) {
            const int32_t q = b.gq[g * b.lanes + l];
            if (q < 0)
                continue;
            tile_select(a, LaneAcc<LB>{s_cnt, T, LB * l}, (uint32_t)q, obj0, s_sel);
            __syncthreads();
        }
        __syncthreads();
    }
}

}  // namespace gh
