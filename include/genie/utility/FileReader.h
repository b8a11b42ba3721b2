// CSV loaders of the GENIE API (src/genie/utility/FileReader.cc:54-130).
#ifndef GENIE_UTILITY_FILEREADER_H_
#define GENIE_UTILITY_FILEREADER_H_

#include <vector>

namespace genie {
namespace utility {

// One multirange query line "qid dim value selectivity weight" (FileReader.h:29-36).
typedef struct _GPUGenie_Query_Data {
    int qid;
    int dim;
    int value;
    float sel;
    float weight;
} attr_t;

const unsigned int GPUGENIE_QUERY_QID_INDEX = 0u;
const unsigned int GPUGENIE_QUERY_DIM_INDEX = 1u;
const unsigned int GPUGENIE_QUERY_VALUE_INDEX = 2u;
const unsigned int GPUGENIE_QUERY_SELECTIVITY_INDEX = 3u;
const unsigned int GPUGENIE_QUERY_WEIGHT_INDEX = 4u;
const unsigned int GPUGENIE_QUERY_NUM_OF_FIELDS = 5u;

// Rows of ints separated by ',' or ' ' (at most num rows; num < 0 = all).
void read_file(std::vector<std::vector<int> >& dest, const char* fname, int num);
// Multirange query file, at most num lines (num < 0 = all).
void read_query(std::vector<attr_t>& data, const char* file_name, int num);

}  // namespace utility
}  // namespace genie

#endif
