// Table and query IO of the 2nd-level API (src/genie/interface/io.h:14-30).
#ifndef GENIE_INTERFACE_IO_H_
#define GENIE_INTERFACE_IO_H_

#include <memory>
#include <string>

#include <genie/interface/execution_policy.h>
#include <genie/interface/types.h>

namespace genie {

TableData LoadTableDataFromCsv(const std::string& filename);
QueryData LoadQueryDataFromCsv(const std::string& filename, const std::shared_ptr<const ExecutionPolicy>& policy);
// A table saved by SaveTableToBinary (format: genie/utility/serialization.h).
std::shared_ptr<genie::table::inv_table> LoadTableFromBinary(const std::string& filename);
void SaveTableToBinary(const std::string& filename, const std::shared_ptr<const genie::table::inv_table>& table);

}  // namespace genie

#endif
