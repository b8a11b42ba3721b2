// Wavefront decoders of the posting-list codecs (formats in
// include/genie/compression/codecs.h; COMPRESSION_TYPE ids of
// src/genie/compression/DeviceCodecFactory.h:38-53).
//
// One wave decodes one compressed segment of at most kSegMax values from
// kSegMax words staged in LDS.  Values land in registers: lane l holds values
// [16 l, 16 l + 16) of the segment.  The reference decodes with a 256-thread
// block per segment in shared memory (DeviceBitPackingCodec.cu:143-249,
// DeviceVarintCodec.cu:163-280, DeviceCodecs.cu:122-139); here a single
// wave does it with ballots and lane scans, so eight segments decode at once
// per workgroup.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wave.h"

namespace gh {

constexpr uint32_t kSegMax = 1024;       // values and words per segment
constexpr int kPerLane = kSegMax / kWave; // 16

// Per-wave LDS scratch of the decoders.
struct alignas(16) DecodeScratch {
    uint32_t words[kSegMax + 4];   // staged compressed segment (+ guard words)
    uint32_t mb[kSegMax / 32];     // BP32 miniblocks: start word << 8 | width
    uint16_t stop[kSegMax + 1];    // varint: byte position + 1 of each value's last byte
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

enum : int {
    kCopy = 0, kDelta = 1, kBp32 = 2, kVarint = 3
};

struct CodecShape {
    bool delta_first;  // Serial(Delta, inner)
    bool composite;    // Composite(BP32, leaf)
    int leaf;
};

__host__ __device__ constexpr bool codec_shape(int codec, CodecShape* s)
{
    switch (codec) {
    case 1: *s = {false, false, kCopy}; return true;    // COPY
    case 2: *s = {false, false, kDelta}; return true;   // DELTA
    case 3: *s = {false, false, kBp32}; return true;    // BP32
    case 4: *s = {false, false, kVarint}; return true;  // VARINT
    case 5: *s = {false, true, kCopy}; return true;     // COMP_BP32_COPY
    case 6: *s = {false, true, kVarint}; return true;   // COMP_BP32_VARINT
    case 7: *s = {false, false, kCopy}; return true;    // SERIAL_COPY_COPY
    case 8: *s = {true, false, kCopy}; return true;     // SERIAL_DELTA_COPY
    case 9: *s = {true, false, kDelta}; return true;    // SERIAL_DELTA_DELTA
    case 10: *s = {true, false, kVarint}; return true;  // SERIAL_DELTA_VARINT
    case 11: *s = {true, false, kBp32}; return true;    // SERIAL_DELTA_BP32
    case 12: *s = {true, true, kCopy}; return true;     // SERIAL_DELTA_COMP_BP32_COPY
    case 13: *s = {true, true, kVarint}; return true;   // SERIAL_DELTA_COMP_BP32_VARINT
    }
    return false;
}

// Inclusive prefix sum (mod 2^32) over the wave's 1,024 register values in
// index order: lane-local serial scan, then an exclusive scan of lane totals.
__device__ __forceinline__ void wave_prefix16(uint32_t (&v)[kPerLane])
{
#pragma unroll
    for (int j = 1; j < kPerLane; ++j)
        v[j] += v[j - 1];
    const uint32_t tot = v[kPerLane - 1];
    const uint32_t excl = wave_incl_scan(tot) - tot;
#pragma unroll
    for (int j = 0; j < kPerLane; ++j)
        v[j] += excl;
}

// BP32 at words [b0, ...) of the staged segment: lane 0 walks the block
// headers (at most 8 for 1,024 values) and records every miniblock's start
// word and width.  Returns the value count (the length word), or 0xFFFFFFFF
// when the headers run past `limit` words.
__device__ __forceinline__ uint32_t bp32_prepare(DecodeScratch& S, uint32_t b0, uint32_t limit)
{
    uint32_t n = S.words[b0];
    uint32_t bad = 0;
    if (lane_id() == 0) {
        if (n > kSegMax) {
            bad = 1;
        } else {
            uint32_t p = b0 + 1;
            const uint32_t nb = (n + 127) / 128;
            for (uint32_t j = 0; j < nb && !bad; ++j) {
                if (p >= limit) {
                    bad = 1;
                    break;
                }
                const uint32_t h = S.words[p++];
                for (int m = 0; m < 4; ++m) {
                    const uint32_t b = (h >> (24 - 8 * m)) & 255u;
                    if (b > 32) {
                        bad = 1;
                        break;
                    }
                    S.mb[j * 4 + m] = p << 8 | b;
                    p += b;
                }
            }
            if (p > limit)
                bad = 1;
        }
    }
    wave_sync();
    bad = (uint32_t)__shfl((int)bad, 0, kWave);
    return bad ? 0xFFFFFFFFu : n;
}

// The same for a segment of a created index, whose block header positions
// were recorded at creation (hdr: eight u16 word offsets within the segment,
// two per component) and whose headers were validated then: lane 4 j + m
// records miniblock m of block j, all in parallel.
__device__ __forceinline__ uint32_t bp32_prepare_hdr(DecodeScratch& S, uint32_t b0, uint4 hdr)
{
    const uint32_t n = S.words[b0];
    const uint32_t lane = lane_id(), j = lane >> 2, m = lane & 3u;
    if (lane < 32 && j < (n + 127) / 128) {
        const uint32_t c = j < 4 ? (j < 2 ? hdr.x : hdr.y) : (j < 6 ? hdr.z : hdr.w);
        const uint32_t p = (c >> ((j & 1u) * 16u)) & 0xFFFFu;
        const uint32_t h = S.words[p];
        const uint32_t w0 = h >> 24, w1 = (h >> 16) & 255u, w2 = (h >> 8) & 255u;
        const uint32_t before = (m > 0 ? w0 : 0u) + (m > 1 ? w1 : 0u) + (m > 2 ? w2 : 0u);
        S.mb[lane] = (p + 1u + before) << 8 | ((h >> (24u - 8u * m)) & 255u);
    }
    wave_sync();
    return n;
}

__device__ __forceinline__ uint32_t bp32_value(const DecodeScratch& S, uint32_t i)
{
    const uint32_t d = S.mb[i >> 5];
    const uint32_t b = d & 255u, start = d >> 8;
    if (b == 0)
        return 0;
    const uint32_t bit = (i & 31u) * b;
    const uint32_t w = start + (bit >> 5), sh = bit & 31u;
    uint64_t x = S.words[w];
    if (sh + b > 32)
        x |= (uint64_t)S.words[w + 1] << 32;
    x >>= sh;
    return b == 32 ? (uint32_t)x : (uint32_t)x & ((1u << b) - 1u);
}

// The 16 values [first, first + 16) of the BP32 part (first a multiple of
// 16, so they lie in one 32-value miniblock): one miniblock lookup, then per
// value two word reads and a funnel shift.  Word reads stay within the
// miniblock's words (+ one guard word).
__device__ __forceinline__ void bp32_lane16(const DecodeScratch& S, uint32_t first, uint32_t (&v)[kPerLane])
{
    const uint32_t d = S.mb[first >> 5];
    const uint32_t b = d & 255u, start = d >> 8;
    const uint32_t mask = b >= 32u ? 0xFFFFFFFFu : (1u << b) - 1u;
    uint32_t bit = (first & 31u) * b;
#pragma unroll
    for (int j = 0; j < kPerLane; ++j) {
        const uint32_t w = start + (bit >> 5);
        v[j] = __builtin_amdgcn_alignbit(S.words[w + 1], S.words[w], bit & 31u) & mask;
        bit += b;
    }
}

// Varint over words [v0, v0 + nw): records the end of every value and
// returns the value count (0xFFFFFFFF if more than kSegMax).
__device__ __forceinline__ uint32_t varint_prepare(DecodeScratch& S, uint32_t v0, uint32_t nw)
{
    const int lane = lane_id();
    const uint32_t per = (nw + kWave - 1) / kWave;  // words per lane, contiguous
    const uint32_t wa = min(nw, lane * per), wb = min(nw, wa + per);
    uint32_t c = 0;
    for (uint32_t w = wa; w < wb; ++w)
        c += __popc(S.words[v0 + w] & 0x80808080u);
    const uint32_t incl = wave_incl_scan(c);
    const uint32_t total = (uint32_t)__shfl((int)incl, kWave - 1, kWave);
    if (total > kSegMax)
        return 0xFFFFFFFFu;
    uint32_t idx = incl - c;
    for (uint32_t w = wa; w < wb; ++w) {
        uint32_t m = S.words[v0 + w] & 0x80808080u;
        while (m) {
            const uint32_t byte = (uint32_t)__builtin_ctz(m) >> 3;
            S.stop[idx++] = (uint16_t)(w * 4 + byte + 1);
            m &= m - 1;
        }
    }
    wave_sync();
    return total;
}

// Value i of the varint part at words [v0, ...): its bytes [b, e) (at most
// the first five count, 7 bits each, least significant first), gathered
// from the three words holding bytes b .. b + 7 with funnel shifts.  Reads
// stay within the staged words + two guard words.
__device__ __forceinline__ uint32_t varint_value(const DecodeScratch& S, uint32_t v0, uint32_t i)
{
    const uint32_t e = S.stop[i];                 // one past the last byte
    const uint32_t b = i ? S.stop[i - 1] : 0u;    // first byte
    const uint32_t wb = v0 + (b >> 2), sh = (b & 3u) * 8u;
    const uint32_t w0 = S.words[wb], w1 = S.words[wb + 1], w2 = S.words[wb + 2];
    const uint32_t lo = __builtin_amdgcn_alignbit(w1, w0, sh);  // bytes b .. b + 3
    const uint32_t hi = __builtin_amdgcn_alignbit(w2, w1, sh);  // bytes b + 4 .. b + 7
    const uint32_t v = (lo & 0x7Fu) | ((lo >> 1) & 0x3F80u) | ((lo >> 2) & 0x1FC000u) | ((lo >> 3) & 0xFE00000u) |
                       (hi << 28);
    const uint32_t len = e - b;
    return len >= 5u ? v : v & ((1u << (7u * len)) - 1u);
}

// Codecs with BP32 blocks (their segments' block header positions are
// recorded at index creation).
__host__ __device__ constexpr bool codec_has_bp32(int codec)
{
    CodecShape s{false, false, kCopy};
    return codec_shape(codec, &s) && (s.composite || s.leaf == kBp32);
}

// A codec's shape fixed at compile time (a kernel instantiated per codec
// keeps only that codec's steps and registers).
__host__ __device__ constexpr CodecShape codec_shape_of(int codec)
{
    CodecShape s{false, false, kCopy};
    codec_shape(codec, &s);
    return s;
}
template <int CODEC>
struct CodecC {
    static constexpr bool delta_first = codec_shape_of(CODEC).delta_first;
    static constexpr bool composite = codec_shape_of(CODEC).composite;
    static constexpr int leaf = codec_shape_of(CODEC).leaf;
};

// Decode the staged segment of `nw` words, shape `sh` (a CodecShape, or a
// CodecC with the same members fixed at compile time).  Returns the value
// count, or 0xFFFFFFFF on a malformed segment; vals[j] of lane l is value
// 16 l + j (undefined past the count).  With `use_hdr`, `hdr` holds the
// segment's BP32 block header positions from index creation (otherwise the
// decoder walks the headers).
template <class SH>
__device__ __forceinline__ uint32_t wave_decode_shape(const SH& sh, DecodeScratch& S, uint32_t nw,
                                                      uint32_t (&vals)[kPerLane], uint4 hdr = {}, bool use_hdr = false)
{
    const int lane = lane_id();
    const uint32_t base = lane * kPerLane;
    uint32_t n1 = 0, b1 = 0;       // composite: values of the BP32 part, its word count
    uint32_t l0 = 0, lnw = nw;     // leaf words
    if (sh.composite) {
        if (nw < 1)
            return 0xFFFFFFFFu;
        b1 = S.words[0];
        if (1 + b1 > nw)
            return 0xFFFFFFFFu;
        if (b1) {
            n1 = use_hdr ? bp32_prepare_hdr(S, 1, hdr) : bp32_prepare(S, 1, 1 + b1);
            if (n1 == 0xFFFFFFFFu)
                return n1;
        }
        l0 = 1 + b1;
        lnw = nw - l0;
    }
    uint32_t n2 = 0;
    switch (sh.leaf) {
    case kCopy:
    case kDelta: n2 = lnw; break;
    case kBp32:
        if (lnw < 1)
            return 0xFFFFFFFFu;
        n2 = use_hdr ? bp32_prepare_hdr(S, l0, hdr) : bp32_prepare(S, l0, l0 + lnw);
        break;
    case kVarint: n2 = varint_prepare(S, l0, lnw); break;
    }
    if (n2 == 0xFFFFFFFFu || n1 + n2 > kSegMax)
        return 0xFFFFFFFFu;
    const uint32_t n = n1 + n2;
    // values [0, nbp) come from BP32 miniblocks (the whole segment for a
    // BP32 leaf, the first part of a composite), the rest from the leaf
    const uint32_t nbp = sh.leaf == kBp32 ? n : n1;
#pragma unroll
    for (int j = 0; j < kPerLane; ++j)
        vals[j] = 0;
    if (base < nbp)
        bp32_lane16(S, base, vals);
    if (sh.leaf != kBp32 && base + kPerLane > n1 && base < n) {
#pragma unroll
        for (int j = 0; j < kPerLane; ++j) {
            const uint32_t i = base + j;
            if (i >= n1 && i < n) {
                const uint32_t k = i - n1;
                vals[j] = sh.leaf == kVarint ? varint_value(S, l0, k) : S.words[l0 + k];
            }
        }
    }
    if (sh.leaf == kDelta)  // Delta leaf: inclusive prefix sum of its part
        wave_prefix16(vals);
    if (sh.delta_first)     // Serial(Delta, X): prefix sum of the decoded values
        wave_prefix16(vals);
    return n;
}

// The same with the codec chosen at run time.
__device__ __forceinline__ uint32_t wave_decode(int codec, DecodeScratch& S, uint32_t nw,
                                                uint32_t (&vals)[kPerLane], uint4 hdr = {}, bool use_hdr = false)
{
    CodecShape sh{false, false, kCopy};
    if (!codec_shape(codec, &sh))
        return 0xFFFFFFFFu;
    return wave_decode_shape(sh, S, nw, vals, hdr, use_hdr);
}

// Stage words [w0, w0 + nw) of global memory into the wave's scratch
// (nw <= kSegMax, checked when the index was created).
__device__ __forceinline__ void stage_words(DecodeScratch& S, const uint32_t* __restrict__ g, uint32_t w0,
                                            uint32_t nw)
{
    for (uint32_t i = lane_id(); i < nw; i += kWave)
        S.words[i] = g[w0 + i];
    if (lane_id() < 4)
        S.words[nw + lane_id()] = 0;
    wave_sync();
}

// The same through a buffer resource that ends at the segment's last word:
// every lane issues all of its kPerLane loads unconditionally (the ones past
// the segment read zero), then stores them.  w0 and nw are wave-uniform.
__device__ __forceinline__ void stage_words_rs(DecodeScratch& S, const uint32_t* __restrict__ g, uint32_t w0,
                                               uint32_t nw)
{
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(g) + w0, (short)0, (int)(nw * 4u), 0x00020000);
    const uint32_t lane = lane_id();
    // rows of 64 words past the segment are not loaded (a short segment
    // issues a few loads, not 16); they are staged as zeros
    const uint32_t rows = __builtin_amdgcn_readfirstlane((nw + kWave - 1) / kWave);
    uint32_t t[kPerLane];
#pragma unroll
    for (int r = 0; r < kPerLane; ++r)
        t[r] = (uint32_t)r < rows ? __builtin_amdgcn_raw_buffer_load_b32(rs, (lane + (uint32_t)r * kWave) * 4u, 0, 0)
                                  : 0u;
#pragma unroll
    for (int r = 0; r < kPerLane; ++r)
        S.words[lane + (uint32_t)r * kWave] = t[r];
    if (lane < 4)
        S.words[kSegMax + lane] = 0;
    wave_sync();
}

}  // namespace gh
