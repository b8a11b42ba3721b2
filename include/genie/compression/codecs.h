// Posting-list codecs of GENIE's compressed tables: names, ids and host
// encode/decode.  The ids are COMPRESSION_TYPE
// (src/genie/compression/DeviceCodecFactory.h:38-53) and the word formats are
// those of the reference codecs, so a compressed table written by either side
// decodes on the other:
//   Copy                       DeviceCodecs.h:96-100
//   Delta (x[i]-x[i-1], u32)   DeviceCodecs.cu:90-104, DeviceDeltaHelper.h:18-50
//   BP32  (w0=length; per 128 values a b0..b3 header byte word, then 4 x 32
//          values packed LSB-first in b_i words; tail ceil(rem/32) miniblocks)
//                              DeviceBitPackingCodec.cu:16-99
//   Varint (7-bit groups, high bit on the LAST byte, zero-padded to a word)
//                              DeviceVarintCodec.cu:14-152
//   Composite(BP32, X): w0 = BP32 words for floor(n/8)*8 values, then X
//                              DeviceCompositeCodec.cu:28-97
//   Serial(Delta, X): delta first, X on its output DeviceSerialCodec.cu:47-91
#ifndef GENIE_COMPRESSION_CODECS_H_
#define GENIE_COMPRESSION_CODECS_H_

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace genie {
namespace compression {

enum COMPRESSION_TYPE {
    NO_COMPRESSION = 0,
    COPY,
    DELTA,
    BP32,
    VARINT,
    COMP_BP32_COPY,
    COMP_BP32_VARINT,
    SERIAL_COPY_COPY,
    SERIAL_DELTA_COPY,
    SERIAL_DELTA_DELTA,
    SERIAL_DELTA_VARINT,
    SERIAL_DELTA_BP32,
    SERIAL_DELTA_COMP_BP32_COPY,
    SERIAL_DELTA_COMP_BP32_VARINT,
};

extern const COMPRESSION_TYPE DEFAULT_COMPRESSION_TYPE;      // NO_COMPRESSION
extern const COMPRESSION_TYPE LIGHTWEIGHT_COMPRESSION_TYPE;  // BP32
extern const COMPRESSION_TYPE MIDDLEWEIGHT_COMPRESSION_TYPE; // SERIAL_DELTA_BP32
extern const COMPRESSION_TYPE HEAVYWEIGHT_COMPRESSION_TYPE;  // SERIAL_DELTA_COMP_BP32_VARINT

// Largest uncompressed segment any codec here decodes in one piece
// (decodeArrayParallel_lengthPerBlock, DeviceBitPackingCodec.h:118).
constexpr size_t kMaxSegmentLength = 1024;

class DeviceCodecFactory {
public:
    static std::string getCompressionName(COMPRESSION_TYPE type);
    static COMPRESSION_TYPE getCompressionType(const std::string& name);
    static const std::vector<COMPRESSION_TYPE>& allCompressionTypes();
};

// Encode n values into out (resized); returns the number of words.
size_t encode(COMPRESSION_TYPE type, const uint32_t* in, size_t n, std::vector<uint32_t>& out);
// Decode n_words words; returns the number of values written to out (resized).
size_t decode(COMPRESSION_TYPE type, const uint32_t* in, size_t n_words, std::vector<uint32_t>& out);

}  // namespace compression
}  // namespace genie

#endif
