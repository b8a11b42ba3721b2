// Compressed inverted index (src/genie/table/inv_compr_table.cu:31-119): the
// uncompressed build, then every _inv_pos segment encoded independently with
// one codec.  inv() / inv_pos() return the compressed words and word offsets
// (inv_compr_table.cu:174-190), which is what compressed queries refer to.
#ifndef GENIE_TABLE_INV_COMPR_TABLE_H_
#define GENIE_TABLE_INV_COMPR_TABLE_H_

#include <genie/compression/codecs.h>
#include <genie/table/inv_table.h>

namespace genie {
namespace table {

class inv_compr_table : public inv_table {
public:
    inv_compr_table() = default;
    ~inv_compr_table() override = default;

    genie::compression::COMPRESSION_TYPE getCompression() const { return m_compression; }
    void setCompression(genie::compression::COMPRESSION_TYPE c);
    size_t getUncompressedPostingListMaxLength() const { return m_max_length; }
    void setUncompressedPostingListMaxLength(size_t n) { m_max_length = n; }
    double getCompressionRatio();

    std::vector<int>* inv() override { return reinterpret_cast<std::vector<int>*>(&m_comprInv); }
    std::vector<int>* inv_pos() override { return &m_comprInvPos; }
    std::vector<uint32_t>* compressedInv() { return &m_comprInv; }
    std::vector<int>* compressedInvPos() { return &m_comprInvPos; }
    std::vector<int>* uncompressedInv() { return inv_table::inv(); }
    std::vector<int>* uncompressedInvPos() { return inv_table::inv_pos(); }

    void save_fields(std::string& out) const override;
    void load_fields(const std::string& in, size_t& at) override;

protected:
    int upload(genie_index** out) override;
    void after_build(size_t max_length) override;

private:
    genie::compression::COMPRESSION_TYPE m_compression = genie::compression::HEAVYWEIGHT_COMPRESSION_TYPE;
    size_t m_max_length = genie::compression::kMaxSegmentLength;
    std::vector<uint32_t> m_comprInv;
    std::vector<int> m_comprInvPos;
};

}  // namespace table
}  // namespace genie

#endif
