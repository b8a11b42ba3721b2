// The reference's own API tests, restated against this build's C++ API
// (test/test_csv_input_sift_data.cc, test_refactoring_new_interface.cc,
// test_csv_input_tweets.cc, test_multiple_search.cc, test_hashtable_overflow.cc,
// test_csv_input_multiload.cc, test_csv_multiload_tweets.cc,
// test_serialization.cc, example/example_sift.cc).  Counts are the exact ones
// (the reference's 16s on tweets_20 are wrapped 4-bit counters).  Where the
// reference asserts nothing (test_hashtable_overflow.cc:37, example_sift.cc),
// every result slot is compared with the CPU oracle (oracle/oracle.h, linked
// into this test binary as the checker).  Usage: test_api <golden dir> [test]
#include "oracle.h"

#include <genie/genie.h>
#include <genie/original/interface.h>

#include <genie/interface/wire.h>
#include <genie/table/inv_compr_table.h>
#include <genie/utility/serialization.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <map>
#include <string>
#include <vector>

using namespace genie::original;
using genie::table::inv_table;
using std::string;
using std::vector;

static string G;
static int failures = 0;

#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::printf("  CHECK FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);   \
            ++failures;                                                          \
        }                                                                        \
    } while (0)

static GPUGenie_Config base_config(vector<vector<int> >& data, vector<vector<int> >& queries, int dim, int k)
{
    GPUGenie_Config config;
    config.dim = dim;
    config.count_threshold = 14;
    config.num_of_topk = k;
    config.hashtable_size = 14 * k * 1.5f;
    config.query_radius = 0;
    config.use_device = 0;
    config.use_adaptive_range = false;
    config.selectivity = 0.0f;
    config.query_points = &queries;
    config.data_points = &data;
    config.use_load_balance = false;
    config.posting_list_max_length = 6400;
    config.multiplier = 1.5f;
    config.use_multirange = false;
    config.data_type = 0;
    config.search_type = 0;
    config.max_data_size = 0;
    config.num_of_queries = 3;
    return config;
}

static void test_sift_legacy()
{
    vector<vector<int> > data, queries;
    GPUGenie_Config config = base_config(data, queries, 5, 5);
    genie::utility::read_file(data, (G + "/sift_20.csv").c_str(), -1);
    genie::utility::read_file(queries, (G + "/sift_20.csv").c_str(), config.num_of_queries);
    init_genie(config);
    inv_table* table = nullptr;
    preprocess_for_knn_csv(config, table);
    vector<int>& inv = *table[0].inv();
    const int want[6] = {8, 9, 7, 0, 2, 4};
    for (int i = 0; i < 6; ++i)
        CHECK(inv[i] == want[i]);
    vector<int> result, count;
    knn_search_after_preprocess(config, table, result, count);
    CHECK(result[0] == 0 && count[0] == 5);
    CHECK(result[1] == 4 && count[1] == 2);
    CHECK(result[5] == 1 && count[5] == 5);
    CHECK(result[10] == 2 && count[10] == 5);
    delete[] table;
}

static void test_sift_new_interface()
{
    genie::Config config = genie::Config().SetK(5).SetNumOfQueries(5);
    std::shared_ptr<genie::ExecutionPolicy> policy = genie::MakePolicy(config);
    genie::SearchResult r = genie::Search(policy, G + "/sift_20.csv", G + "/sift_20.csv");
    CHECK(r.first.size() == 25);
    CHECK(r.first[0] == 0 && r.second[0] == 5);
    CHECK(r.first[1] == 4 && r.second[1] == 2);
    CHECK(r.first[5] == 1 && r.second[5] == 5);
    CHECK(r.first[10] == 2 && r.second[10] == 5);
    // 2nd-level interface with a range policy
    genie::Config c2 = genie::Config().SetK(5).SetNumOfQueries(3).SetQueryRange(1);
    std::shared_ptr<genie::ExecutionPolicy> p2 = genie::MakePolicy(c2);
    genie::TableData td = genie::LoadTableDataFromCsv(G + "/sift_20.csv");
    genie::QueryData qd = genie::LoadQueryDataFromCsv(G + "/sift_20.csv", p2);
    auto table = genie::BuildTable(p2, td);
    auto qs = genie::BuildQuery(p2, table, qd);
    genie::SearchResult r2 = genie::Match(p2, table, qs);
    CHECK(r2.first.size() == 15);
    CHECK(r2.second[0] == 5 && r2.first[0] == 0);  // a row always matches itself in every dim
}

static void test_invalid_config()
{
    bool threw = false;
    try {
        genie::MakePolicy(genie::Config().SetNumOfQueries(3));
    } catch (genie::exception::InvalidConfigurationException&) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        genie::MakePolicy(genie::Config().SetK(0).SetNumOfQueries(3));
    } catch (genie::exception::InvalidConfigurationException&) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        genie::Config().LoadFromFile("x");
    } catch (genie::exception::NotImplementedException&) {
        threw = true;
    }
    CHECK(threw);
}

static void test_tweets()
{
    vector<vector<int> > data, queries;
    GPUGenie_Config config = base_config(data, queries, 14, 5);
    config.hashtable_size = 100 * 5 * 1.5f;
    config.search_type = 1;
    genie::utility::read_file(data, (G + "/tweets_20.csv").c_str(), -1);
    genie::utility::read_file(queries, (G + "/tweets_20.csv").c_str(), config.num_of_queries);
    init_genie(config);
    inv_table* table = nullptr;
    preprocess_for_knn_csv(config, table);
    vector<int>& inv = *table[0].inv();
    const int want[6] = {0, 11, 0, 11, 0, 11};
    for (int i = 0; i < 6; ++i)
        CHECK(inv[i] == want[i]);
    vector<int> result, count;
    knn_search_after_preprocess(config, table, result, count);
    CHECK((result[0] == 0 && result[1] == 11) || (result[0] == 11 && result[1] == 0));
    CHECK(count[0] == 17 && count[1] == 17);
    CHECK(result[5] == 1 && count[5] == 14);
    CHECK(result[10] == 2 && count[10] == 18);
    delete[] table;
}

static void test_multiple_search()
{
    vector<vector<int> > data, queries;
    GPUGenie_Config config = base_config(data, queries, 5, 5);
    genie::utility::read_file(data, (G + "/sift_20.csv").c_str(), -1);
    genie::utility::read_file(queries, (G + "/sift_20.csv").c_str(), config.num_of_queries);
    config.save_to_gpu = true;
    init_genie(config);
    inv_table* table = nullptr;
    preprocess_for_knn_csv(config, table);
    CHECK(table[0].is_stored_in_gpu);
    vector<int> r1, c1, r2, c2;
    knn_search_after_preprocess(config, table, r1, c1);
    knn_search_after_preprocess(config, table, r2, c2);
    CHECK(r1 == r2 && c1 == c2);
    CHECK(r1[0] == 0 && c1[0] == 5);
    delete[] table;
}

// The oracle's top-k of one query from its final ranges (Query::dump): the
// exhaustive closed-form count over the raw rows (orc_count_exhaustive_single
// / _biject), then the exact top-k under (count desc, id asc).
static void oracle_topk_of(genie::query::Query& q, const vector<vector<int> >& rows, bool biject, int k,
                           vector<int>& ids, vector<int>& counts)
{
    vector<genie::query::Query::range> rs;
    q.dump(rs);
    vector<orc_range> ors;
    for (const auto& r : rs)
        ors.push_back(orc_range{0, r.dim, r.low, r.up, r.weight, r.order});
    const int32_t n = (int32_t)rows.size();
    vector<uint32_t> c((size_t)n, 0);
    if (biject) {
        vector<int64_t> ptr(1, 0);
        vector<int32_t> vals;
        for (const auto& row : rows) {
            vals.insert(vals.end(), row.begin(), row.end());
            ptr.push_back((int64_t)vals.size());
        }
        orc_count_exhaustive_biject(ptr.data(), vals.data(), n, ors.data(), (int64_t)ors.size(), c.data());
    } else {
        const int32_t d = (int32_t)rows[0].size();
        vector<int32_t> flat;
        for (const auto& row : rows)
            flat.insert(flat.end(), row.begin(), row.end());
        orc_count_exhaustive_single(flat.data(), n, d, ors.data(), (int64_t)ors.size(), c.data());
    }
    ids.assign((size_t)k, 0);
    counts.assign((size_t)k, 0);
    orc_topk(c.data(), n, k, ids.data(), counts.data());
}

// Every slot of a knn result against the oracle, query by query (queries in
// the order the search numbered them).
static void check_against_oracle(vector<genie::query::Query>& qs, const vector<vector<int> >& rows, bool biject,
                                 int k, const vector<int>& result, const vector<int>& count)
{
    CHECK(result.size() == qs.size() * (size_t)k && count.size() == result.size());
    int bad = 0;
    for (size_t q = 0; q < qs.size() && result.size() == qs.size() * (size_t)k; ++q) {
        vector<int> ri, rc;
        oracle_topk_of(qs[q], rows, biject, k, ri, rc);
        for (int j = 0; j < k; ++j)
            if (result[q * k + j] != ri[(size_t)j] || count[q * k + j] != rc[(size_t)j])
                ++bad;
    }
    if (bad)
        std::printf("  %d of %zu slots differ from the oracle\n", bad, result.size());
    CHECK(bad == 0);
}

static void test_hashtable_overflow()
{
    vector<vector<int> > data, queries;
    GPUGenie_Config config = base_config(data, queries, 14, 10);
    config.hashtable_size = 3;  // the reference's overflow-retry trigger
    config.search_type = 1;
    config.num_of_queries = 20;
    genie::utility::read_file(data, (G + "/tweets_20.csv").c_str(), -1);
    genie::utility::read_file(queries, (G + "/tweets_20.csv").c_str(), config.num_of_queries);
    init_genie(config);
    inv_table* table = nullptr;
    preprocess_for_knn_csv(config, table);
    vector<int> result, count;
    knn_search_after_preprocess(config, table, result, count);
    CHECK(result.size() == 200);
    // every row matches itself with its full length
    for (int q = 0; q < 20; ++q)
        CHECK(count[(size_t)q * 10] >= (int)data[(size_t)q].size());
    // and all 200 slots are the oracle's
    vector<genie::query::Query> qs;
    load_query(table[0], qs, config);
    check_against_oracle(qs, data, true, 10, result, count);
    delete[] table;
}

// example/example_sift.cc: sift_1k.csv (1,000 x 128) and the multirange query
// file sift_1k_query.csv (qid, dim, value, selectivity 0.04, weight), adaptive
// query range (query.cc:196-261) from radius 1, load balance with posting
// lists of 256, k = 10.  The search's results against the oracle's counts of
// the adapted ranges.
static void test_sift_1k_multirange()
{
    vector<vector<int> > data, queries;
    vector<genie::utility::attr_t> mq;
    GPUGenie_Config config = base_config(data, queries, 128, 10);
    config.num_of_queries = 500;
    config.count_threshold = -128;
    config.hashtable_size = 128 * 10 * 1.5f;
    config.query_radius = 1;
    config.use_adaptive_range = true;
    config.selectivity = 0.2f;
    config.multiplier = 1.5f;
    config.posting_list_max_length = 256;
    config.use_load_balance = true;
    config.use_multirange = true;
    genie::utility::read_file(data, (G + "/sift_1k.csv").c_str(), -1);
    genie::utility::read_query(mq, (G + "/sift_1k_query.csv").c_str(), -1);
    config.multirange_query_points = &mq;
    CHECK(data.size() == 1000 && mq.size() == 12800);
    init_genie(config);
    inv_table* table = nullptr;
    preprocess_for_knn_csv(config, table);
    vector<int> result, count;
    knn_search_after_preprocess(config, table, result, count);
    vector<genie::query::Query> qs;
    load_query(table[0], qs, config);
    CHECK(qs.size() == 100);
    // the adaptive range widened the ranges: some query covers more than one value somewhere
    bool widened = false;
    for (auto& q : qs) {
        vector<genie::query::Query::range> rs;
        q.dump(rs);
        for (const auto& r : rs)
            widened = widened || r.up - r.low > 2;
    }
    CHECK(widened);
    check_against_oracle(qs, data, false, 10, result, count);
    delete[] table;
}

// tweets_4k.csv (4,096 token lists) as a bijectMap table: 64 of its rows as
// queries, k = 25, every slot against the oracle's multiset counts.
static void test_tweets_4k_bijectmap()
{
    vector<vector<int> > data, queries;
    GPUGenie_Config config = base_config(data, queries, 14, 25);
    config.search_type = 1;
    config.num_of_queries = 64;
    genie::utility::read_file(data, (G + "/tweets_4k.csv").c_str(), -1);
    genie::utility::read_file(queries, (G + "/tweets_4k.csv").c_str(), config.num_of_queries);
    CHECK(data.size() == 4096 && queries.size() == 64);
    init_genie(config);
    inv_table* table = nullptr;
    preprocess_for_knn_csv(config, table);
    vector<int> result, count;
    knn_search_after_preprocess(config, table, result, count);
    vector<genie::query::Query> qs;
    load_query(table[0], qs, config);
    check_against_oracle(qs, data, true, 25, result, count);
    // each query row finds itself with its full length at the top
    for (int q = 0; q < 64; ++q)
        CHECK(count[(size_t)q * 25] >= (int)queries[(size_t)q].size());
    delete[] table;
}

static void test_multiload_tweets()
{
    vector<vector<int> > data, queries;
    GPUGenie_Config config = base_config(data, queries, 14, 5);
    config.hashtable_size = 100 * 5 * 1.5f;
    config.search_type = 1;
    config.max_data_size = 10;
    genie::utility::read_file(data, (G + "/tweets_20.csv").c_str(), -1);
    genie::utility::read_file(queries, (G + "/tweets_20.csv").c_str(), config.num_of_queries);
    init_genie(config);
    inv_table* table = nullptr;
    preprocess_for_knn_csv(config, table);
    CHECK(table[0].get_total_num_of_table() == 2);
    vector<int>& inv = *table[0].inv();
    const int want[6] = {0, 0, 0, 0, 0, 7};
    for (int i = 0; i < 6; ++i)
        CHECK(inv[i] == want[i]);
    vector<int> result, count;
    knn_search_after_preprocess(config, table, result, count);
    CHECK((result[0] == 0 && result[1] == 11) || (result[0] == 11 && result[1] == 0));
    CHECK(count[0] == 17 && count[1] == 17);
    CHECK(result[5] == 1 && count[5] == 14);
    CHECK(result[10] == 2 && count[10] == 18);
    delete[] table;
}

static void test_multiload_sift()
{
    vector<vector<int> > data, queries;
    GPUGenie_Config config = base_config(data, queries, 5, 5);
    config.max_data_size = 7;  // 20 rows -> tables of 7, 6, 7
    genie::utility::read_file(data, (G + "/sift_20.csv").c_str(), -1);
    genie::utility::read_file(queries, (G + "/sift_20.csv").c_str(), config.num_of_queries);
    init_genie(config);
    inv_table* table = nullptr;
    preprocess_for_knn_csv(config, table);
    CHECK(table[0].get_total_num_of_table() == 3);
    vector<int> result, count;
    knn_search_after_preprocess(config, table, result, count);
    CHECK(result[0] == 0 && count[0] == 5);
    CHECK(result[1] == 4 && count[1] == 2);
    CHECK(result[5] == 1 && count[5] == 5);
    CHECK(result[10] == 2 && count[10] == 5);
    delete[] table;
}

static void test_compressed_sift()
{
    // compressed tables need load-balanced segments of <= 1024 ids
    for (int codec = 1; codec <= 13; ++codec) {
        vector<vector<int> > data, queries;
        GPUGenie_Config config = base_config(data, queries, 5, 5);
        config.use_load_balance = true;
        config.posting_list_max_length = 1024;
        config.compression = (genie::compression::COMPRESSION_TYPE)codec;
        genie::utility::read_file(data, (G + "/sift_20.csv").c_str(), -1);
        genie::utility::read_file(queries, (G + "/sift_20.csv").c_str(), config.num_of_queries);
        init_genie(config);
        inv_table* table = nullptr;
        preprocess_for_knn_csv(config, table);
        vector<int> result, count;
        knn_search_after_preprocess(config, table, result, count);
        CHECK(result[0] == 0 && count[0] == 5);
        CHECK(result[1] == 4 && count[1] == 2);
        CHECK(result[5] == 1 && count[5] == 5);
        CHECK(result[10] == 2 && count[10] == 5);
        delete[] table;
    }
}

// test_serialization.cc:25-81: save, load, compare every field the
// reference compares; then the loaded table answers the same queries.
static void test_serialization()
{
    for (int codec : {0, 11}) {
        vector<vector<int> > data, queries;
        GPUGenie_Config config = base_config(data, queries, 5, 5);
        if (codec) {
            config.use_load_balance = true;
            config.posting_list_max_length = 1024;
            config.compression = (genie::compression::COMPRESSION_TYPE)codec;
        }
        genie::utility::read_file(data, (G + "/sift_20.csv").c_str(), -1);
        genie::utility::read_file(queries, (G + "/sift_20.csv").c_str(), config.num_of_queries);
        init_genie(config);
        inv_table* raw = nullptr;
        preprocess_for_knn_csv(config, raw);
        CHECK(raw != nullptr && raw->build_status() == inv_table::builded);
        std::shared_ptr<inv_table> table(raw, std::default_delete<inv_table[]>());
        const string path = "/tmp/genie_test_api_" + std::to_string(getpid()) + "_" + std::to_string(codec) + ".gix";
        genie::SaveTableToBinary(path, table);
        std::shared_ptr<inv_table> loaded = genie::LoadTableFromBinary(path);
        std::remove(path.c_str());
        CHECK(loaded->d_inv_p == nullptr);
        CHECK(loaded->is_stored_in_gpu == false);
        CHECK(loaded->get_table_index() == 0);
        CHECK(loaded->get_total_num_of_table() == 1);
        CHECK(table->shift_bits_sequence == loaded->shift_bits_sequence);
        CHECK(table->m_size() == loaded->m_size());
        CHECK(table->i_size() == loaded->i_size());
        CHECK(table->shifter() == loaded->shifter());
        CHECK(table->_shift_bits_subsequence() == loaded->_shift_bits_subsequence());
        CHECK(table->build_status() == loaded->build_status());
        CHECK(*table->inv() == *loaded->inv());
        CHECK(*table->inv_pos() == *loaded->inv_pos());
        CHECK(*table->inv_index_map() == *loaded->inv_index_map());
        CHECK(table->get_min_value_sequence() == loaded->get_min_value_sequence());
        CHECK(table->get_max_value_sequence() == loaded->get_max_value_sequence());
        CHECK(table->get_gram_length_sequence() == loaded->get_gram_length_sequence());
        if (codec) {
            auto* c = dynamic_cast<genie::table::inv_compr_table*>(table.get());
            auto* cl = dynamic_cast<genie::table::inv_compr_table*>(loaded.get());
            CHECK(c && cl);
            if (c && cl) {
                CHECK(*c->compressedInv() == *cl->compressedInv());
                CHECK(*c->uncompressedInv() == *cl->uncompressedInv());
                CHECK(*c->compressedInvPos() == *cl->compressedInvPos());
                CHECK(*c->uncompressedInvPos() == *cl->uncompressedInvPos());
                CHECK(c->getUncompressedPostingListMaxLength() == cl->getUncompressedPostingListMaxLength());
                CHECK(c->getCompression() == cl->getCompression());
            }
        }
        // the loaded table answers like the original (sift_20 goldens)
        vector<int> r1, c1, r2, c2;
        inv_table* tp = table.get();
        inv_table* lp = loaded.get();
        knn_search_after_preprocess(config, tp, r1, c1);
        knn_search_after_preprocess(config, lp, r2, c2);
        CHECK(r1 == r2 && c1 == c2);
        CHECK(r2[0] == 0 && c2[0] == 5 && r2[1] == 4 && c2[1] == 2);
    }
    // a table of a multi-table set is refused, as by the reference
    vector<vector<int> > data, queries;
    GPUGenie_Config config = base_config(data, queries, 5, 5);
    genie::utility::read_file(data, (G + "/sift_20.csv").c_str(), -1);
    inv_table* raw = nullptr;
    preprocess_for_knn_csv(config, raw);
    raw->set_table_index(1);
    raw->set_total_num_of_table(2);
    std::shared_ptr<inv_table> t(raw, std::default_delete<inv_table[]>());
    bool threw = false;
    try {
        genie::utility::SaveTable("/tmp/genie_test_api_refused.gix", t);
    } catch (genie::exception::cpu_runtime_error&) {
        threw = true;
    }
    CHECK(threw);
}

// The serving precedent's path end to end: cluster files -> ReadData ->
// query document -> ExecuteMultitableQuery -> GenerateOutput; the csv and
// the table-file (data_format 1) forms of the same clusters answer alike.
static void test_wire()
{
    vector<vector<int> > rows;
    genie::utility::read_file(rows, (G + "/sift_20.csv").c_str(), -1);
    const string base = "/tmp/genie_test_wire_" + std::to_string(getpid());
    // clusters: rows 0..9 and 10..19, five dims each
    for (int c = 0; c < 2; ++c) {
        FILE* f = std::fopen((base + "_" + std::to_string(c) + "_0.csv").c_str(), "w");
        for (int r = 10 * c; r < 10 * c + 10; ++r)
            std::fprintf(f, "%d,%d,%d,%d,%d\n", rows[r][0], rows[r][1], rows[r][2], rows[r][3], rows[r][4]);
        std::fclose(f);
    }
    const string cfg = "{\"data_file\": \"" + base + "\", \"dim\": 5, \"count_threshold\": 5, \"data_type\": 0, "
                       "\"search_type\": 0, \"max_data_size\": 0, \"num_of_cluster\": 2, \"data_format\": 0, "
                       "\"radius\": 0, \"selectivity\": 0.0}";
    char q[512];
    std::snprintf(q, sizeof(q),
                  "{\"topk\": 3, \"queries\": [{\"content\": [%d, %d, %d, %d, %d], \"clusters\": [0, 1]}, "
                  "{\"content\": [%d, %d, %d, %d, %d], \"clusters\": [1]}]}",
                  rows[4][0], rows[4][1], rows[4][2], rows[4][3], rows[4][4], rows[12][0], rows[12][1], rows[12][2],
                  rows[12][3], rows[12][4]);
    string outputs[2];
    for (int fmt = 0; fmt < 2; ++fmt) {
        GPUGenie_Config config;
        genie::wire::DistConfig extra;
        genie::wire::QueryBatch batch;
        string err;
        CHECK(genie::wire::ParseConfiguration(cfg, config, extra, &err));
        extra.data_format = fmt;
        CHECK(genie::wire::ParseQuery(q, config, batch, &err));
        auto tables = genie::wire::ReadData(config, extra);
        CHECK(tables.size() == 2);
        if (fmt == 0)  // keep the csv-built tables as table files for the second pass
            for (int c = 0; c < 2; ++c)
                genie::utility::SaveTable(base + "_" + std::to_string(c) + "_0.dat", tables[c]);
        auto res = genie::wire::ExecuteMultitableQuery(config, tables, batch);
        outputs[fmt] = genie::wire::GenerateOutput(res, config.num_of_topk);
    }
    CHECK(outputs[0] == outputs[1]);
    // query 0 is row 4 of cluster 0: itself with count 5 first; query 1 is
    // row 12, i.e. row 2 of cluster 1, global id 10 + 2 = 12
    CHECK(outputs[0].rfind("4,5\n", 0) == 0);
    CHECK(outputs[0].find("\n12,5\n") != string::npos);
    for (int c = 0; c < 2; ++c) {
        std::remove((base + "_" + std::to_string(c) + "_0.csv").c_str());
        std::remove((base + "_" + std::to_string(c) + "_0.dat").c_str());
    }
}

// Adaptive query range (query.cc:196-261): ranges widened one value each
// side until the selectivity's posting count is reached.  Known answers
// worked out by hand from the reference's loop on a table whose dim 0 holds
// values 0..9 with list sizes {5,0,3,2,7,1,0,4,6,2} and dim 1 values 3..9
// with sizes {4,5,6,5,4,3,3}.  Host only: runs without a GPU.
static void test_adaptive_range()
{
    const int size0[10] = {5, 0, 3, 2, 7, 1, 0, 4, 6, 2};
    const int size1[7] = {4, 5, 6, 5, 4, 3, 3};
    vector<int> col0, col1;
    for (int v = 0; v < 10; ++v)
        col0.insert(col0.end(), size0[v], v);
    for (int v = 0; v < 7; ++v)
        col1.insert(col1.end(), size1[v], v + 3);
    CHECK(col0.size() == 30 && col1.size() == 30);
    genie::table::inv_table table;
    genie::table::inv_list l0, l1;
    l0.invert(col0);
    l1.invert(col1);
    table.append(l0);
    table.append(l1);
    table.build(0, false);
    struct Case {
        int dim, value, threshold, low, up;
    };
    const Case cases[] = {
        {0, 5, 12, 3, 7},   // both sides grow in step: 1 -> 8 -> 14 postings
        {0, 1, 20, 0, 7},   // the low side stops at 0, the high side goes on
        {0, 9, 10, 7, 9},   // nothing above the dim's max: only the low side grows
        {1, 4, 25, 0, 8},   // below the dim's min the high side waits for low to reach 0
        {0, 5, 40, 0, 9},   // unreachable: the whole dim
    };
    for (const Case& c : cases) {
        genie::query::Query q(table, 0);
        // selectivity s with ceil(s * 30) == threshold
        q.attr(c.dim, c.value, 1.0f, (c.threshold - 0.5f) / 30.0f, 0);
        q.apply_adaptive_query_range();
        vector<genie::query::Query::range> rs;
        q.dump(rs);
        CHECK(rs.size() == 1);
        if (rs.size() == 1) {
            if (rs[0].low != c.low || rs[0].up != c.up)
                std::printf("  dim %d value %d threshold %d: [%d, %d], expected [%d, %d]\n", c.dim, c.value,
                            c.threshold, rs[0].low, rs[0].up, c.low, c.up);
            CHECK(rs[0].low == c.low && rs[0].up == c.up);
        }
    }
}

// Multirange query file "qid dim value selectivity weight"
// (FileReader.cc:85-130, interface.cc:427-472): weighted exact matches on
// sift_20, checked against a brute-force count over the rows.
static void test_multirange()
{
    vector<vector<int> > data, queries;
    GPUGenie_Config config = base_config(data, queries, 5, 4);
    genie::utility::read_file(data, (G + "/sift_20.csv").c_str(), -1);
    const string path = "/tmp/genie_test_multirange_" + std::to_string(getpid()) + ".txt";
    FILE* f = std::fopen(path.c_str(), "w");
    // query 0: row 4 with dim 0 weighted 2; query 1: dims 1 and 3 of row 12
    std::fprintf(f, "0 0 %d 0.0 2\n0 1 %d 0.0 1\n0 2 %d 0.0 1\n0 3 %d 0.0 1\n0 4 %d 0.0 1\n", data[4][0],
                 data[4][1], data[4][2], data[4][3], data[4][4]);
    std::fprintf(f, "1 1 %d 0.0 3\n1 3 %d 0.0 1\n", data[12][1], data[12][3]);
    std::fclose(f);
    vector<genie::utility::attr_t> attrs;
    genie::utility::read_query(attrs, path.c_str(), -1);
    std::remove(path.c_str());
    CHECK(attrs.size() == 7);
    config.use_multirange = true;
    config.multirange_query_points = &attrs;
    config.num_of_queries = 2;
    init_genie(config);
    inv_table* table = nullptr;
    preprocess_for_knn_csv(config, table);
    vector<int> result, count;
    knn_search_after_preprocess(config, table, result, count);
    delete[] table;
    // brute force: per query the weighted number of matching dims
    for (int q = 0; q < 2; ++q) {
        vector<std::pair<int, int> > score;  // (-count, id)
        for (int r = 0; r < (int)data.size(); ++r) {
            int c = 0;
            for (const auto& a : attrs)
                if (a.qid == q && data[r][a.dim] == a.value)
                    c += (int)a.weight;
            if (c > 0)
                score.push_back({-c, r});
        }
        std::sort(score.begin(), score.end());
        for (int i = 0; i < 4; ++i) {
            const int want_id = i < (int)score.size() ? score[i].second : 0;
            const int want_c = i < (int)score.size() ? -score[i].first : 0;
            CHECK(result[q * 4 + i] == want_id && count[q * 4 + i] == want_c);
        }
    }
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::printf("usage: %s <golden dir> [test]\n", argv[0]);
        return 2;
    }
    G = argv[1];
    std::map<string, std::function<void()> > tests = {
        {"sift_legacy", test_sift_legacy},
        {"sift_new_interface", test_sift_new_interface},
        {"invalid_config", test_invalid_config},
        {"tweets", test_tweets},
        {"multiple_search", test_multiple_search},
        {"hashtable_overflow", test_hashtable_overflow},
        {"multiload_tweets", test_multiload_tweets},
        {"multiload_sift", test_multiload_sift},
        {"compressed_sift", test_compressed_sift},
        {"serialization", test_serialization},
        {"wire", test_wire},
        {"adaptive_range", test_adaptive_range},
        {"multirange", test_multirange},
        {"sift_1k_multirange", test_sift_1k_multirange},
        {"tweets_4k_bijectmap", test_tweets_4k_bijectmap},
    };
    for (auto& kv : tests) {
        if (argc > 2 && kv.first != argv[2])
            continue;
        std::printf("[ RUN ] %s\n", kv.first.c_str());
        std::fflush(stdout);
        const int before = failures;
        try {
            kv.second();
        } catch (std::exception& e) {
            std::printf("  EXCEPTION: %s\n", e.what());
            ++failures;
        }
        std::printf("[ %s ] %s\n", failures == before ? " OK " : "FAIL", kv.first.c_str());
        std::fflush(stdout);
    }
    return failures ? 1 : 0;
}
