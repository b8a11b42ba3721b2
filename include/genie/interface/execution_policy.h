// Execution policies (src/genie/interface/execution_policy.h:21-71,
// execution_policy.cc:14-66; execution_policy/single_value.cc,
// single_range.cc, validation.cc).
#ifndef GENIE_INTERFACE_EXECUTION_POLICY_H_
#define GENIE_INTERFACE_EXECUTION_POLICY_H_

#include <memory>
#include <vector>

#include <genie/interface/config.h>
#include <genie/interface/types.h>
#include <genie/query/query.h>
#include <genie/table/inv_table.h>

namespace genie {

class ExecutionPolicy {
public:
    virtual ~ExecutionPolicy() = default;
    virtual std::shared_ptr<genie::table::inv_table> BuildTable(const genie::TableData& table_data) = 0;
    virtual std::vector<genie::query::Query> BuildQuery(const std::shared_ptr<const genie::table::inv_table>& table,
                                                        const genie::QueryData& query_data) = 0;
    virtual genie::SearchResult Match(const std::shared_ptr<const genie::table::inv_table>& table,
                                      const std::vector<genie::query::Query>& queries) = 0;
    // Several tables, each with its own queries; results concatenated in
    // table order.
    virtual genie::SearchResult Match(const std::vector<std::shared_ptr<const genie::table::inv_table> >& tables,
                                      const std::vector<std::vector<genie::query::Query> >& queries) = 0;
    virtual void Validate();
    void SetK(const uint32_t k) { k_ = k; }
    void SetNumOfQueries(const uint32_t n) { num_of_queries_ = n; }
    uint32_t GetNumOfQueries() const { return num_of_queries_; }

protected:
    uint32_t k_ = 0;
    uint32_t dim_ = 0;
    uint32_t num_of_queries_ = 0;
};

std::shared_ptr<ExecutionPolicy> MakePolicy(const Config& config);

namespace execution_policy {

class SingleValueExecutionPolicy : public genie::ExecutionPolicy {
    friend std::shared_ptr<genie::ExecutionPolicy> genie::MakePolicy(const genie::Config& config);

protected:
    SingleValueExecutionPolicy() = default;

public:
    std::shared_ptr<genie::table::inv_table> BuildTable(const genie::TableData& table_data) override;
    std::vector<genie::query::Query> BuildQuery(const std::shared_ptr<const genie::table::inv_table>& table,
                                                const genie::QueryData& query_data) override;
    genie::SearchResult Match(const std::shared_ptr<const genie::table::inv_table>& table,
                              const std::vector<genie::query::Query>& queries) override;
    genie::SearchResult Match(const std::vector<std::shared_ptr<const genie::table::inv_table> >& tables,
                              const std::vector<std::vector<genie::query::Query> >& queries) override;
};

class SingleRangeExecutionPolicy : public SingleValueExecutionPolicy {
    friend std::shared_ptr<genie::ExecutionPolicy> genie::MakePolicy(const genie::Config& config);
    SingleRangeExecutionPolicy() = default;
    uint32_t query_range_ = 0;

public:
    std::vector<genie::query::Query> BuildQuery(const std::shared_ptr<const genie::table::inv_table>& table,
                                                const genie::QueryData& query_data) override;
    void Validate() override;
    void SetQueryRange(const uint32_t r) { query_range_ = r; }
};

namespace validation {
void ValidateK(uint32_t k);
void ValidateNumOfQueries(uint32_t num_of_queries);
void ValidateQueryRange(uint32_t query_range);
}  // namespace validation

}  // namespace execution_policy
}  // namespace genie

#endif
