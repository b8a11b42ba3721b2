// Data types of the GENIE 2nd-level API (src/genie/interface/types.h:12-20).
#ifndef GENIE_INTERFACE_TYPES_H_
#define GENIE_INTERFACE_TYPES_H_

#include <utility>
#include <vector>

namespace genie {

typedef std::vector<std::vector<int> > TableData;
typedef std::vector<std::vector<int> > QueryData;
// (top-k ids, top-k counts), query-major.
typedef std::pair<std::vector<int>, std::vector<int> > SearchResult;

}  // namespace genie

#endif
