// Wavefront (wave64) primitives for CDNA4 used by the match-count kernels:
// ballots, lane prefix counts, wave reductions and small block scans.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gh {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Number of set lanes strictly below this lane in mask.
__device__ __forceinline__ uint32_t lane_prefix(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, (uint32_t)__shfl_xor((int)v, off, kWave));
    return v;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = min(v, (uint32_t)__shfl_xor((int)v, off, kWave));
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += (uint32_t)__shfl_xor((int)v, off, kWave);
    return v;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, off, kWave);
        uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), off, kWave);
        uint64_t o = ((uint64_t)hi << 32) | lo;
        v = o < v ? o : v;
    }
    return v;
}

// Inclusive prefix sum across the wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        uint32_t o = (uint32_t)__shfl_up((int)v, off, kWave);
        if (lane >= off)
            v += o;
    }
    return v;
}

__device__ __forceinline__ uint32_t bcast_first(uint32_t v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

// Add 1 to hist[bin] for every active lane with `take`, one LDS atomic per
// distinct bin value in the wave (ballot aggregation).
__device__ __forceinline__ void hist_add_aggregated(uint32_t* hist, uint32_t bin, bool take)
{
    uint64_t pending = ballot(take);
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const uint32_t lb = (uint32_t)__shfl((int)bin, leader, kWave);
        const uint64_t same = ballot(take && bin == lb) & pending;
        if (lane_id() == leader)
            atomicAdd(&hist[lb], (uint32_t)__popcll(same));
        pending &= ~same;
    }
}

// Add 1 to hist[bin] for every active lane with `take`: the lanes holding
// the first taking lane's bin add with one LDS atomic, every other lane with
// its own (cheap when one bin dominates, and when the bins are spread).
__device__ __forceinline__ void hist_add_leader(uint32_t* hist, uint32_t bin, bool take)
{
    const uint64_t act = ballot(take);
    if (!act)
        return;
    const int leader = __builtin_ctzll(act);
    const uint32_t lb = (uint32_t)__shfl((int)bin, leader, kWave);
    const bool same = take && bin == lb;
    const uint64_t sm = ballot(same);
    if (lane_id() == leader)
        atomicAdd(&hist[lb], (uint32_t)__popcll(sm));
    else if (take && !same)
        atomicAdd(&hist[bin], 1u);
}

} // namespace gh
