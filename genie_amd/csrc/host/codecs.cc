// Host encoders/decoders of the posting-list codecs (formats documented in
// include/genie/compression/codecs.h).
#include <genie/compression/codecs.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <stdexcept>

namespace genie {
namespace compression {

const COMPRESSION_TYPE DEFAULT_COMPRESSION_TYPE = NO_COMPRESSION;
const COMPRESSION_TYPE LIGHTWEIGHT_COMPRESSION_TYPE = BP32;
const COMPRESSION_TYPE MIDDLEWEIGHT_COMPRESSION_TYPE = SERIAL_DELTA_BP32;
const COMPRESSION_TYPE HEAVYWEIGHT_COMPRESSION_TYPE = SERIAL_DELTA_COMP_BP32_VARINT;

namespace {

enum Leaf { kCopy, kDelta, kBp32, kVarint };

struct Shape {
    bool delta_first;  // Serial(Delta, inner)
    bool composite;    // Composite(BP32, leaf)
    Leaf leaf;
};

Shape shape_of(COMPRESSION_TYPE t)
{
    switch (t) {
    case COPY: return {false, false, kCopy};
    case DELTA: return {false, false, kDelta};
    case BP32: return {false, false, kBp32};
    case VARINT: return {false, false, kVarint};
    case COMP_BP32_COPY: return {false, true, kCopy};
    case COMP_BP32_VARINT: return {false, true, kVarint};
    case SERIAL_COPY_COPY: return {false, false, kCopy};
    case SERIAL_DELTA_COPY: return {true, false, kCopy};
    case SERIAL_DELTA_DELTA: return {true, false, kDelta};
    case SERIAL_DELTA_VARINT: return {true, false, kVarint};
    case SERIAL_DELTA_BP32: return {true, false, kBp32};
    case SERIAL_DELTA_COMP_BP32_COPY: return {true, true, kCopy};
    case SERIAL_DELTA_COMP_BP32_VARINT: return {true, true, kVarint};
    default: throw std::invalid_argument("unknown compression type");
    }
}

inline uint32_t width(uint32_t v) { return v ? 32u - (uint32_t)__builtin_clz(v) : 0u; }

void put_bits(std::vector<uint32_t>& out, size_t base, uint64_t bitpos, uint32_t v, uint32_t b)
{
    size_t w = base + (size_t)(bitpos >> 5);
    uint32_t sh = (uint32_t)(bitpos & 31);
    uint64_t x = (uint64_t)v << sh;
    out[w] |= (uint32_t)x;
    if (sh + b > 32)
        out[w + 1] |= (uint32_t)(x >> 32);
}

uint32_t get_bits(const uint32_t* in, uint64_t bitpos, uint32_t b)
{
    if (b == 0)
        return 0;
    const uint32_t* p = in + (bitpos >> 5);
    uint32_t sh = (uint32_t)(bitpos & 31);
    uint64_t x = p[0];
    if (sh + b > 32)
        x |= (uint64_t)p[1] << 32;
    x >>= sh;
    return b == 32 ? (uint32_t)x : (uint32_t)(x & ((1ull << b) - 1));
}

void bp32_enc(const uint32_t* in, size_t n, std::vector<uint32_t>& out)
{
    out.push_back((uint32_t)n);
    for (size_t i = 0; i < n; i += 128) {
        const size_t take = std::min<size_t>(128, n - i);
        const uint32_t mb = (uint32_t)((take + 31) / 32);
        uint32_t b[4] = {0, 0, 0, 0};
        for (uint32_t m = 0; m < mb; ++m) {
            uint32_t acc = 0;
            for (size_t j = m * 32; j < std::min<size_t>(take, (m + 1) * 32); ++j)
                acc |= in[i + j];
            b[m] = width(acc);
        }
        out.push_back(b[0] << 24 | b[1] << 16 | b[2] << 8 | b[3]);
        for (uint32_t m = 0; m < mb; ++m) {
            size_t base = out.size();
            out.resize(base + b[m], 0u);
            for (size_t j = 0; j < 32; ++j) {
                size_t src = m * 32 + j;
                uint32_t v = src < take ? in[i + src] : 0u;
                if (b[m])
                    put_bits(out, base, (uint64_t)j * b[m], v, b[m]);
            }
        }
    }
}

// Returns words consumed.
size_t bp32_dec(const uint32_t* in, size_t nw, std::vector<uint32_t>& out)
{
    if (nw < 1)
        throw std::runtime_error("BP32: empty input");
    const size_t n = in[0];
    size_t p = 1;
    const size_t o0 = out.size();
    out.resize(o0 + n);
    for (size_t i = 0; i < n; i += 128) {
        if (p >= nw)
            throw std::runtime_error("BP32: truncated header");
        const size_t take = std::min<size_t>(128, n - i);
        const uint32_t mb = (uint32_t)((take + 31) / 32);
        const uint32_t h = in[p++];
        const uint32_t b[4] = {h >> 24, (h >> 16) & 255u, (h >> 8) & 255u, h & 255u};
        for (uint32_t m = 0; m < mb; ++m) {
            if (b[m] > 32 || p + b[m] > nw)
                throw std::runtime_error("BP32: truncated block");
            for (size_t j = 0; j < 32 && m * 32 + j < take; ++j)
                out[o0 + i + m * 32 + j] = get_bits(in + p, (uint64_t)j * b[m], b[m]);
            p += b[m];
        }
    }
    return p;
}

void varint_enc(const uint32_t* in, size_t n, std::vector<uint32_t>& out)
{
    std::vector<uint8_t> bytes;
    bytes.reserve(n * 2 + 4);
    for (size_t i = 0; i < n; ++i) {
        uint32_t v = in[i];
        while (v >= 128u) {
            bytes.push_back((uint8_t)(v & 127u));
            v >>= 7;
        }
        bytes.push_back((uint8_t)(v | 128u));  // the stop bit sits on the last byte
    }
    while (bytes.size() & 3)
        bytes.push_back(0);
    size_t base = out.size();
    out.resize(base + bytes.size() / 4);
    std::memcpy(out.data() + base, bytes.data(), bytes.size());
}

void varint_dec(const uint32_t* in, size_t nw, std::vector<uint32_t>& out)
{
    const uint8_t* b = reinterpret_cast<const uint8_t*>(in);
    uint32_t v = 0, sh = 0;
    for (size_t i = 0; i < nw * 4; ++i) {
        v |= (uint32_t)(b[i] & 127u) << sh;
        if (b[i] & 128u) {
            out.push_back(v);
            v = 0;
            sh = 0;
        } else {
            sh += 7;
        }
    }
}

void leaf_enc(Leaf l, const uint32_t* in, size_t n, std::vector<uint32_t>& out)
{
    switch (l) {
    case kCopy: out.insert(out.end(), in, in + n); return;
    case kDelta: {
        uint32_t prev = 0;
        for (size_t i = 0; i < n; ++i) {
            out.push_back(in[i] - prev);
            prev = in[i];
        }
        return;
    }
    case kBp32: bp32_enc(in, n, out); return;
    case kVarint: varint_enc(in, n, out); return;
    }
}

void leaf_dec(Leaf l, const uint32_t* in, size_t nw, std::vector<uint32_t>& out)
{
    switch (l) {
    case kCopy: out.insert(out.end(), in, in + nw); return;
    case kDelta: {
        uint32_t acc = 0;
        for (size_t i = 0; i < nw; ++i) {
            acc += in[i];
            out.push_back(acc);
        }
        return;
    }
    case kBp32: {
        size_t used = bp32_dec(in, nw, out);
        if (used != nw)
            throw std::runtime_error("BP32: trailing words");
        return;
    }
    case kVarint: varint_dec(in, nw, out); return;
    }
}

}  // namespace

size_t encode(COMPRESSION_TYPE type, const uint32_t* in, size_t n, std::vector<uint32_t>& out)
{
    const Shape s = shape_of(type);
    std::vector<uint32_t> tmp;
    const uint32_t* src = in;
    if (s.delta_first) {
        tmp.resize(n);
        uint32_t prev = 0;
        for (size_t i = 0; i < n; ++i) {
            tmp[i] = in[i] - prev;
            prev = in[i];
        }
        src = tmp.data();
    }
    out.clear();
    if (!s.composite) {
        leaf_enc(s.leaf, src, n, out);
        return out.size();
    }
    const size_t n1 = n / 8 * 8;  // BP32's minimum effective length is 8
    out.push_back(0);
    if (n1) {
        bp32_enc(src, n1, out);
        out[0] = (uint32_t)(out.size() - 1);
    }
    if (n > n1)
        leaf_enc(s.leaf, src + n1, n - n1, out);
    return out.size();
}

size_t decode(COMPRESSION_TYPE type, const uint32_t* in, size_t n_words, std::vector<uint32_t>& out)
{
    const Shape s = shape_of(type);
    out.clear();
    if (!s.composite) {
        leaf_dec(s.leaf, in, n_words, out);
    } else {
        if (n_words < 1)
            throw std::runtime_error("composite: empty input");
        const size_t w1 = in[0];
        if (1 + w1 > n_words)
            throw std::runtime_error("composite: bad first length");
        if (w1 && bp32_dec(in + 1, w1, out) != w1)
            throw std::runtime_error("composite: BP32 part length mismatch");
        if (n_words > 1 + w1)
            leaf_dec(s.leaf, in + 1 + w1, n_words - 1 - w1, out);
    }
    if (s.delta_first)
        for (size_t i = 1; i < out.size(); ++i)
            out[i] += out[i - 1];
    return out.size();
}

std::string DeviceCodecFactory::getCompressionName(COMPRESSION_TYPE type)
{
    // Names as the reference codecs report them (name(), DeviceCodecs.h,
    // DeviceBitPackingCodec.h:114, DeviceVarintCodec.h:46, Composite/Serial).
    switch (type) {
    case NO_COMPRESSION: return "no";
    case COPY: return "Copy";
    case DELTA: return "Delta";
    case BP32: return "BitPacking32";
    case VARINT: return "Varint";
    case COMP_BP32_COPY: return "Composite(BitPacking32-Copy)";
    case COMP_BP32_VARINT: return "Composite(BitPacking32-Varint)";
    case SERIAL_COPY_COPY: return "Serial(Copy-Copy)";
    case SERIAL_DELTA_COPY: return "Serial(Delta-Copy)";
    case SERIAL_DELTA_DELTA: return "Serial(Delta-Delta)";
    case SERIAL_DELTA_VARINT: return "Serial(Delta-Varint)";
    case SERIAL_DELTA_BP32: return "Serial(Delta-BitPacking32)";
    case SERIAL_DELTA_COMP_BP32_COPY: return "Serial(Delta-Composite(BitPacking32-Copy))";
    case SERIAL_DELTA_COMP_BP32_VARINT: return "Serial(Delta-Composite(BitPacking32-Varint))";
    }
    return "Unknown-Compression-Type";
}

COMPRESSION_TYPE DeviceCodecFactory::getCompressionType(const std::string& name)
{
    for (COMPRESSION_TYPE t : allCompressionTypes())
        if (getCompressionName(t) == name)
            return t;
    return NO_COMPRESSION;
}

const std::vector<COMPRESSION_TYPE>& DeviceCodecFactory::allCompressionTypes()
{
    static const std::vector<COMPRESSION_TYPE> all = {
        NO_COMPRESSION, COPY, DELTA, BP32, VARINT, COMP_BP32_COPY, COMP_BP32_VARINT,
        SERIAL_COPY_COPY, SERIAL_DELTA_COPY, SERIAL_DELTA_DELTA, SERIAL_DELTA_VARINT,
        SERIAL_DELTA_BP32, SERIAL_DELTA_COMP_BP32_COPY, SERIAL_DELTA_COMP_BP32_VARINT};
    return all;
}

}  // namespace compression
}  // namespace genie
