// Hash-table form of the plain scan, for selective queries (a few postings
// per object of the index): one workgroup per (query, object tile), where a
// query's tiles are sized to hold about kHashTarget of its postings each.  The
// tile's counts live in an open-addressing table of twice that many slots in
// LDS instead of one counter per object, so a workgroup covers hundreds of
// thousands of objects; an exact radix select then takes the tile's top k
// under (count desc, id asc).  A table that fills up flags its query and
// emits nothing for it; the host runs such queries again on the
// direct-mapped form.  The reference keeps a per-query c-PQ hash table in
// global memory for the same purpose (match_device_utils.h:161-419), with
// approximate eviction; these tables are exact and hold every touched id.
#pragma once
#include "kernels.h"

namespace gh {

constexpr int kHashLogMax = 13;         // 8,192 slots = 32 KiB of LDS
constexpr uint32_t kHashTarget = 4096;  // postings per (query, tile) aimed at

struct HashArgs {
    ScanArgs s;
    const uint32_t* work;  // [n_work][2]: query, tile
    const uint4* hq;       // [n_queries]: directory tiles per tile, key bits, log2 slots, tiles
    uint32_t* hov;         // [n_queries] set when a table of the query filled up
    uint32_t* hov_any;
    // [n_queries] largest k-th key (count << 32 | ~id) of the query's tiles
    // that emitted k candidates; 0 before any
    unsigned long long* gkey;
    // decoded spans of a compressed index (null otherwise): query q's spans
    // are spans[sq_ptr[q] .. sq_ptr[q + 1]), positions in s.post (the batch's
    // decoded postings), filter flag set
    const Span* spans;
    const uint32_t* sq_ptr;
};

// Decodes segments of a compressed index into a scratch posting array: job =
// (segment, destination offset); one wave per job.  Two uses: the segments
// of a batch's hash-form queries, once per query item, packed one after the
// other; or every segment a batch names, once, at its own posting positions
// (a scratch of the uncompressed layout).  Segments were validated at index
// creation.
template <int CODEC>
__global__ void __launch_bounds__(kDecodeWaves * kWave) __attribute__((amdgpu_waves_per_eu(5)))
    k_decode_jobs(const uint32_t* __restrict__ words, const uint32_t* __restrict__ cpos,
                  const uint4* __restrict__ bph, const uint2* __restrict__ jobs, uint32_t n_jobs,
                  uint32_t* __restrict__ out)
{
    __shared__ DecodeScratch S[kDecodeWaves];
    static_assert(sizeof(DecodeScratch) >= kWave * (kPerLane + 1) * 4, "transpose fits the scratch");
    const int wv = threadIdx.x / kWave, lane = lane_id();
    uint32_t* t = reinterpret_cast<uint32_t*>(&S[wv]);  // the transpose, once the decode is done
    const uint32_t stride = gridDim.x * kDecodeWaves;
    uint32_t j = blockIdx.x * kDecodeWaves + wv;
    // the next job's descriptor and word range are loaded while the current
    // job decodes
    constexpr bool HDR = CodecC<CODEC>::composite || CodecC<CODEC>::leaf == kBp32;
    uint2 jb = make_uint2(0, 0);
    uint32_t w0 = 0, w1 = 0;
    uint4 hd = make_uint4(0, 0, 0, 0);
    if (j < n_jobs) {
        jb = jobs[j];
        w0 = cpos[jb.x];
        w1 = cpos[jb.x + 1];
        if constexpr (HDR)
            hd = bph[jb.x];
    }
    for (; j < n_jobs; j += stride) {
        const uint32_t nw = __builtin_amdgcn_readfirstlane(w1 - w0);
        const uint32_t out0 = __builtin_amdgcn_readfirstlane(jb.y);
        stage_words_rs(S[wv], words, __builtin_amdgcn_readfirstlane(w0), nw);
        const uint4 hdr = hd;
        if (j + stride < n_jobs) {
            jb = jobs[j + stride];
            w0 = cpos[jb.x];
            w1 = cpos[jb.x + 1];
            if constexpr (HDR)
                hd = bph[jb.x];
        }
        uint32_t v[kPerLane];
        const uint32_t n = __builtin_amdgcn_readfirstlane(
            min(wave_decode_shape(CodecC<CODEC>{}, S[wv], nw, v, hdr, HDR), (uint32_t)kSegMax));
        wave_sync();  // every lane is done with the staged words
        // lane l holds values 16 l .. 16 l + 15: through LDS (rows of 17
        // words, so the writes of one k fall in distinct banks) to stores of
        // 64 consecutive values per instruction, through a buffer resource
        // that ends at the segment's n-th value (the stores past it are
        // dropped)
#pragma unroll
        for (int k = 0; k < kPerLane; ++k)
            t[lane * (kPerLane + 1) + k] = v[k];
        wave_sync();
        const __amdgpu_buffer_rsrc_t ro =
            __builtin_amdgcn_make_buffer_rsrc(out + out0, (short)0, (int)(n * 4u), 0x00020000);
#pragma unroll
        for (int k = 0; k < kPerLane; ++k) {
            const uint32_t i = (uint32_t)k * kWave + lane;
            if ((uint32_t)k * kWave < n)  // uniform: rows past the segment are not issued
                __builtin_amdgcn_raw_buffer_store_b32(t[(i / kPerLane) * (kPerLane + 1) + i % kPerLane], ro, i * 4u,
                                                      0, 0);
        }
        wave_sync();  // the scratch is reused by the next job
    }
}

// Adds w to the count of tile-local id `loc`.  A slot word holds key = loc + 1
// in its high kb bits and the count in the low 32 - kb bits (the host sizes
// kb and checks that every count fits); 0 is an empty slot.  Linear probing
// from a multiplicative hash; false when the table is full.
__device__ __forceinline__ bool hash_add(uint32_t* tab, uint32_t lhs, uint32_t kb, uint32_t loc, uint32_t w)
{
    const uint32_t key = (loc + 1u) << (32u - kb);
    const uint32_t kmask = ~0u << (32u - kb);
    const uint32_t smask = (1u << lhs) - 1u;
    uint32_t h = (loc * 0x9E3779B1u) >> (32u - lhs);
    for (uint32_t probe = 0; probe <= smask; ++probe) {
        // claim an empty slot outright (most ids of a selective query arrive
        // once); a taken slot returns its word
        const uint32_t v = atomicCAS(&tab[h], 0u, key | w);
        if (v == 0)
            return true;
        if ((v & kmask) == key) {
            atomicAdd(&tab[h], w);
            return true;
        }
        h = (h + 1u) & smask;
    }
    return false;
}

// Exact top-k of one (query, tile) from its table: the occupied slots with a
// count at or above the query's adaptive threshold are ordered by
// rank = count << kb | (2^kb - 1 - key), i.e. count descending, then id
// ascending (ranks are unique and nonzero).  Every thread loads its slots
// (tid + 512 j) once as ranks into registers (0 = not eligible); a radix
// select over 11-bit digits of the ranks, starting at the highest bit in
// which the eligible ranks differ, finds the k-th, its histograms in the
// table's LDS (free once the ranks are in registers).  Appends the selected
// keys to the query's candidates and raises the threshold to the tile's
// k-th count, and gkey to the tile's k-th key.
// Eligible: count >= the threshold, and ranked above gkey: a tile that emitted
// k candidates of keys >= gkey leaves no room in the top k for a key below
// it, so when gkey's id lies below this tile's ids (the work runs tile-major,
// lower tiles first), a count equal to gkey's is not eligible either.
constexpr int kHashPer = (1 << kHashLogMax) / kScanThreads;  // slots per thread at most
constexpr uint32_t kSelDigit = 11;                             // bits per radix pass
constexpr uint32_t kSelBins = 1u << kSelDigit;

__device__ __forceinline__ void hash_select(const ScanArgs& a, unsigned long long* gkey, uint32_t* tab, uint32_t ns,
                                            uint32_t kb, uint32_t q, uint32_t obj0, uint32_t* s_res, uint32_t* s_red)
{
    static_assert(kSelBins <= (1u << kHashLogMax) && kSelBins == 32 * kWave, "histogram in the table, 32 bins a lane");
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6, nw = blockDim.x / kWave;
    const uint32_t k = a.kq[q];
    const uint32_t cb = 32u - kb;
    const uint32_t cmask = (1u << cb) - 1u;
    const uint32_t kfull = (1u << kb) - 1u;
    const uint32_t per = ns / blockDim.x;  // uniform, 1..kHashPer
    if (tid == 0) {
        uint32_t lo = __hip_atomic_load(&a.gthr[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t g = __hip_atomic_load(&gkey[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (g) {
            const uint32_t gc = (uint32_t)(g >> 32), gid = ~(uint32_t)g;
            lo = max(lo, gc + (gid < obj0 ? 1u : 0u));
        }
        s_res[0] = lo;
    }
    __syncthreads();
    const uint32_t lower = max(1u, s_res[0]);
    uint32_t rk[kHashPer];
    uint32_t e = 0, rmax = 0, rmin = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < kHashPer; ++j) {
        rk[j] = 0;
        if ((uint32_t)j < per) {
            const uint32_t v = tab[tid + j * blockDim.x];
            const uint32_t c = v & cmask;
            if (c >= lower) {
                rk[j] = (c << kb) | (kfull - (v >> cb));
                ++e;
                rmax = max(rmax, rk[j]);
                rmin = min(rmin, rk[j]);
            }
        }
    }
    // eligible count, largest and smallest rank of the tile
    e = wave_sum(e);
    rmax = wave_max(rmax);
    rmin = wave_min(rmin);
    if (lane == 0) {
        s_red[wv] = e;
        s_red[kMaxWaves + wv] = rmax;
        s_red[2 * kMaxWaves + wv] = rmin;
    }
    __syncthreads();  // every slot is in registers: the table is free from here
    uint32_t E = 0, Rmax = 0, Rmin = 0xFFFFFFFFu;
    for (int i = 0; i < nw; ++i) {
        E += s_red[i];
        Rmax = max(Rmax, s_red[kMaxWaves + i]);
        Rmin = min(Rmin, s_red[2 * kMaxWaves + i]);
    }
    if (E == 0)
        return;

    uint32_t kth = 0;  // every nonzero rank >= kth is selected
    if (E > k) {
        // bits above the highest one in which Rmax and Rmin differ are common
        // to every eligible rank; digits are taken from there down
        uint32_t* hist = tab;
        uint32_t lo_bit = 32u - (uint32_t)__builtin_clz(Rmax ^ Rmin);  // Rmax != Rmin since E > k >= 1
        uint32_t mask = lo_bit >= 32u ? 0u : ~0u << lo_bit;
        uint32_t prefix = Rmax & mask, remaining = k;
#pragma unroll 1
        while (lo_bit > 0) {
            const uint32_t width = min(kSelDigit, lo_bit), shift = lo_bit - width;
            const uint32_t dmask = (1u << width) - 1u;
            for (uint32_t i = tid; i < kSelBins; i += blockDim.x)
                hist[i] = 0;
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kHashPer; ++j) {
                const uint32_t r = rk[j];
                if ((uint32_t)j < per && r != 0 && (r & mask) == prefix)
                    atomicAdd(&hist[(r >> shift) & dmask], 1u);
            }
            __syncthreads();
            if (tid < kWave) {
                // lane l holds bins 2047 - 32 l .. 2016 - 32 l, counted from the top
                const uint32_t top = kSelBins - 1u - 32u * (uint32_t)tid;
                uint32_t sum = 0;
#pragma unroll 8
                for (int j = 0; j < 32; ++j)
                    sum += hist[top - j];
                uint32_t above = wave_incl_scan(sum) - sum;
                if (above < remaining && above + sum >= remaining) {
                    for (int j = 0; j < 32; ++j) {
                        const uint32_t h = hist[top - j];
                        if (above + h >= remaining) {
                            s_res[1] = top - j;
                            s_res[2] = remaining - above;
                            s_res[3] = h;
                            break;
                        }
                        above += h;
                    }
                }
            }
            __syncthreads();
            const uint32_t bin = s_res[1], rem = s_res[2], inbin = s_res[3];
            prefix |= bin << shift;
            mask |= dmask << shift;
            remaining = rem;
            lo_bit = shift;
            __syncthreads();  // s_res and the histogram are read before the next pass
            if (inbin == rem)
                break;  // the whole bin is selected: the prefix bounds it
        }
        kth = prefix;
    }
    const uint32_t nemit = min(E, k);
    if (tid == 0) {
        uint32_t base = atomicAdd(&a.cand_cnt[q], nemit);
        if ((uint64_t)base + nemit > a.cand_cap[q]) {
            atomicOr(a.overflow, 1u);
            base = kEmitFail;
        }
        s_res[4] = base;
        s_res[5] = 0;            // slot counter
        s_res[6] = 0xFFFFFFFFu;  // smallest selected rank
    }
    __syncthreads();
    const uint32_t base = s_res[4];
    if (base == kEmitFail)
        return;
    uint64_t* out = a.cand + a.cand_base[q] + base;
    uint32_t rmin_sel = 0xFFFFFFFFu;  // smallest selected rank
#pragma unroll
    for (int j = 0; j < kHashPer; ++j) {
        const uint32_t r = (uint32_t)j < per ? rk[j] : 0u;
        const bool take = r != 0 && r >= kth;
        const uint64_t bm = ballot(take);
        if (bm) {
            uint32_t wb = 0;
            if (lane == __builtin_ctzll(bm))
                wb = atomicAdd(&s_res[5], (uint32_t)__popcll(bm));
            wb = (uint32_t)__shfl((int)wb, __builtin_ctzll(bm), kWave);
            if (take) {
                out[wb + lane_prefix(bm)] = make_key(r >> kb, obj0 + (kfull - (r & kfull)) - 1u);
                rmin_sel = min(rmin_sel, r);
            }
        }
    }
    rmin_sel = wave_min(rmin_sel);
    if (lane == 0 && rmin_sel != 0xFFFFFFFFu)
        atomicMin(&s_res[6], rmin_sel);
    __syncthreads();
    if (tid == 0 && E >= k) {
        const uint32_t r = s_res[6];
        atomicMax(&a.gthr[q], r >> kb);
        atomicMax(&gkey[q], (unsigned long long)make_key(r >> kb, obj0 + (kfull - (r & kfull)) - 1u));
    }
}

// Compressed spans: the segment range of each (chunk, key) pair that can hold
// ids of [obj0, obj1) (segment directory for long keys).
__device__ __forceinline__ void resolve_seg_spans(const ScanArgs& a, uint32_t c0, uint32_t c1, uint32_t pbase,
                                                  uint32_t w0, uint32_t wn, uint32_t dt0, uint32_t dt1,
                                                  uint32_t obj1, Span* s_span)
{
    for (uint32_t c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
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
}

// C: compressed index (the waves take single segments, stage and decode
// them, codecs_dev.h); otherwise the window's postings are flattened over
// the workgroup, four loads per thread at a time.
template <bool C>
__global__ void __launch_bounds__(kScanThreads) k_scan_hash(HashArgs h)
{
    __shared__ uint32_t s_tab[1 << kHashLogMax];
    __shared__ Span s_span[kSpanCap];
    __shared__ uint32_t s_wsum[kScanWaves];
    __shared__ uint32_t s_red[3 * kMaxWaves];
    __shared__ uint32_t s_res[8];
    __shared__ uint32_t s_next;

    const ScanArgs& a = h.s;
    const uint32_t q = h.work[2 * blockIdx.x], t = h.work[2 * blockIdx.x + 1];
    const uint4 prm = h.hq[q];
    const uint32_t kb = prm.y, lhs = prm.z, ns = 1u << lhs;
    const uint32_t tobj = prm.x * kDirTile;
    const uint32_t obj0 = t * tobj;
    const uint32_t nobj = min(tobj, a.n_objects - obj0);  // objects of this tile
    const uint32_t dt0 = t * prm.x, dt1 = min(dt0 + prm.x, a.n_dtiles);
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;

    for (uint32_t i = tid; i < ns; i += kScanThreads)
        s_tab[i] = 0;
    if (tid == 0)
        s_res[7] = 0;  // a table insert failed
    __syncthreads();

    bool ok = true;
    const uint32_t c0 = a.qc_ptr[q], c1 = a.qc_ptr[q + 1];
    const uint32_t* pptr = h.spans ? h.sq_ptr : a.qk_ptr;  // decoded spans, or (chunk, key) pairs
    const uint32_t pbase = pptr[q], npairs = pptr[q + 1] - pbase;
    for (uint32_t w0 = 0; w0 < npairs; w0 += kSpanCap) {
        const uint32_t wn = min((uint32_t)kSpanCap, npairs - w0);
        if constexpr (C) {
            if (tid == 0)
                s_next = 0;
            resolve_seg_spans(a, c0, c1, pbase, w0, wn, dt0, dt1, obj0 + nobj, s_span);
            __syncthreads();
            const uint32_t nsegs = flat_prefix(s_span, wn, s_wsum);
            __shared__ DecodeScratch s_dec[kScanWaves];  // the compressed variant's LDS only
            DecodeScratch& D = s_dec[wv];
            for (;;) {
                uint32_t g = 0;
                if (lane == 0)
                    g = atomicAdd(&s_next, 1u);
                g = bcast_first(__shfl((int)g, 0, kWave));
                if (g >= nsegs)
                    break;
                const Span sp = s_span[flat_span(s_span, wn, g)];
                const uint32_t sg = sp.lo + (g - sp.hi);
                if (a.seg_last[sg] < obj0 || a.seg_first[sg] >= obj0 + nobj)
                    continue;  // segment entirely outside the tile
                const uint32_t cw0 = a.seg_cpos[sg], nw = a.seg_cpos[sg + 1] - cw0;
                stage_words(D, a.words, cw0, nw);
                uint32_t v[kPerLane];
                const uint32_t n = wave_decode(a.codec, D, nw, v, a.seg_bph ? a.seg_bph[sg] : uint4{},
                                                     a.seg_bph != nullptr);
                const uint32_t w = sp.wf >> 1;
#pragma unroll
                for (int j = 0; j < kPerLane; ++j) {
                    const uint32_t loc = v[j] - obj0;
                    if (lane * kPerLane + j < n && loc < nobj && w)
                        ok &= hash_add(s_tab, lhs, kb, loc, w);
                }
                wave_sync();  // the staging buffer is reused by the next segment
            }
        } else {
            if (h.spans) {
                for (uint32_t i = tid; i < wn; i += kScanThreads)
                    s_span[i] = h.spans[pbase + w0 + i];
            } else {
                resolve_spans(a, c0, c1, pbase, w0, wn, dt0, dt1, s_span);
            }
            __syncthreads();
            const uint32_t total = flat_prefix(s_span, wn, s_wsum);
            for (uint32_t g0 = 0; g0 < total; g0 += 4 * kScanThreads) {
                uint32_t id[4], w[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t g = g0 + u * kScanThreads + tid;
                    id[u] = obj0 + nobj;  // outside the tile: counts nothing
                    w[u] = 0;
                    if (g < total) {
                        const Span sp = s_span[flat_span(s_span, wn, g)];
                        id[u] = a.post[sp.lo + (g - sp.hi)];
                        w[u] = sp.wf >> 1;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t loc = id[u] - obj0;
                    if (loc < nobj && w[u])
                        ok &= hash_add(s_tab, lhs, kb, loc, w[u]);
                }
            }
        }
        __syncthreads();  // the next window rewrites the spans
    }
    if (!ok)
        s_res[7] = 1;
    __syncthreads();
    if (s_res[7]) {
        if (tid == 0) {
            h.hov[q] = 1u;
            *h.hov_any = 1u;
        }
        return;
    }
    hash_select(a, h.gkey, s_tab, ns, kb, q, obj0, s_res, s_red);
}

}  // namespace gh
