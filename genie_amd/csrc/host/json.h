// A small JSON reader for the serving precedent's documents: objects, arrays,
// strings (with escapes), numbers, true / false / null.  Numbers keep the
// distinction the reference's validation relies on (rapidjson IsInt /
// IsFloat, src/dist_genie/parser.cu:16-76): an integer literal that fits 32
// bits is an int; a literal with a fraction or an exponent is a double.
#ifndef GENIE_AMD_HOST_JSON_H_
#define GENIE_AMD_HOST_JSON_H_

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace genie_amd {

struct Json {
    enum Kind { Null, Bool, Int, BigInt, Double, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    int64_t i = 0;   // Int and BigInt (an integer literal outside int32)
    double d = 0;    // every number
    std::string s;
    std::vector<Json> a;
    std::vector<std::pair<std::string, Json> > o;

    bool is_int() const { return kind == Int; }
    // rapidjson IsFloat: a double literal within float range
    bool is_float() const { return kind == Double && d >= -3.4e38 && d <= 3.4e38; }
    const Json* get(const std::string& key) const
    {
        for (const auto& kv : o)
            if (kv.first == key)
                return &kv.second;
        return nullptr;
    }
};

// Parses the whole document; false with a message (position included) on
// malformed input or trailing characters.
bool parse_json(const std::string& text, Json& out, std::string* error);

}  // namespace genie_amd

#endif
