// User configuration of GENIE (src/genie/interface/config.h:20-95,
// config.cc:10-106).  K and the number of queries are required.
#ifndef GENIE_INTERFACE_CONFIG_H_
#define GENIE_INTERFACE_CONFIG_H_

#include <cstdint>
#include <string>

namespace genie {

class Config {
public:
    uint32_t GetK() const { return k_; }
    uint32_t GetQueryRange() const { return query_range_; }
    uint32_t GetNumOfQueries() const { return num_of_queries_; }
    bool GetSaveToGpu() const { return save_to_gpu_; }
    uint8_t GetGpuId() const { return gpu_id_; }
    Config& SetK(const uint32_t k);
    Config& SetQueryRange(const uint32_t query_range);
    Config& SetNumOfQueries(const uint32_t num_of_queries);
    Config& SetSaveToGpu(const bool save_to_gpu);
    Config& SetGpuId(const uint8_t gpu_id);
    // Not implemented in GENIE either (config.cc:68-71).
    Config& LoadFromFile(const std::string& filename);
    bool IsKSet() const { return k_initialized_; }
    bool IsQueryRangeSet() const { return query_range_initialized_; }
    bool IsNumOfQueriesSet() const { return num_of_queries_initialized_; }
    void Validate() const;
    void DisplayConfiguration();

private:
    uint32_t k_ = 0;
    uint32_t query_range_ = 0;
    uint32_t num_of_queries_ = 0;
    bool save_to_gpu_ = false;
    uint8_t gpu_id_ = 0;
    bool k_initialized_ = false;
    bool query_range_initialized_ = false;
    bool num_of_queries_initialized_ = false;
};

}  // namespace genie

#endif
