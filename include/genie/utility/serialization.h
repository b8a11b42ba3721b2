// Index persistence: the names and signatures of the reference
// (src/genie/utility/serialization.h:14-34).  The file format is this build's
// own little-endian binary format; it holds the reference archive's fields in
// the archive's order (serialization.cc:68-146) and is specified in
// genie_amd/csrc/host/serialization.cc.
#ifndef GENIE_UTILITY_SERIALIZATION_H_
#define GENIE_UTILITY_SERIALIZATION_H_

#include <memory>
#include <string>

namespace genie {
namespace table {
class inv_table;
}
namespace utility {

// Saves a built table (inv_table or inv_compr_table).  Throws
// cpu_runtime_error for a table of a multi-table set (table index != 0 or
// more than one table), as the reference does, or when the file cannot be
// written.
void SaveTable(const std::string& filename, const std::shared_ptr<const genie::table::inv_table>& table);

// Loads a table saved by SaveTable; the dynamic type (inv_table or
// inv_compr_table) is restored.  Throws cpu_runtime_error when the file is
// missing, truncated, of another format or fails its checksum.
std::shared_ptr<genie::table::inv_table> LoadTable(const std::string& filename);

}  // namespace utility
}  // namespace genie

#endif
