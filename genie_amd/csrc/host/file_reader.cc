// CSV readers (src/genie/utility/FileReader.cc:54-130): fields split on ','
// and ' ', converted with atoi / atof.
#include <genie/utility/FileReader.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

namespace genie {
namespace utility {

static std::vector<std::string> tokens(const std::string& line)
{
    std::vector<std::string> out;
    std::string cur;
    for (char c : line) {
        if (c == ',' || c == ' ' || c == '\r' || c == '\t') {
            if (!cur.empty())
                out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(c);
        }
    }
    if (!cur.empty())
        out.push_back(cur);
    return out;
}

void read_file(std::vector<std::vector<int> >& dest, const char* fname, int num)
{
    dest.clear();
    std::ifstream in(fname);
    std::string line;
    int count = 0;
    while ((num < 0 || count < num) && std::getline(in, line)) {
        std::vector<int> row;
        for (const std::string& t : tokens(line))
            row.push_back(std::atoi(t.c_str()));
        dest.push_back(row);
        ++count;
    }
}

void read_query(std::vector<attr_t>& data, const char* file_name, int num)
{
    data.clear();
    std::ifstream in(file_name);
    std::string line;
    int left = num;
    while (left != 0 && std::getline(in, line)) {
        std::vector<std::string> f = tokens(line);
        if (f.size() != GPUGENIE_QUERY_NUM_OF_FIELDS)
            continue;
        --left;
        attr_t a;
        a.qid = std::atoi(f[GPUGENIE_QUERY_QID_INDEX].c_str());
        a.dim = std::atoi(f[GPUGENIE_QUERY_DIM_INDEX].c_str());
        a.value = std::atoi(f[GPUGENIE_QUERY_VALUE_INDEX].c_str());
        a.sel = (float)std::atof(f[GPUGENIE_QUERY_SELECTIVITY_INDEX].c_str());
        a.weight = (float)std::atof(f[GPUGENIE_QUERY_WEIGHT_INDEX].c_str());
        data.push_back(a);
    }
}

}  // namespace utility
}  // namespace genie
