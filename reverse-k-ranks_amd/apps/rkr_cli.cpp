// rkr_cli -- command-line drivers of the US path over librkr.so, with the
// reference's binaries' names, options and files:
//   rkr_cli fsr   --dataset_dir D --dataset_name N --index_dir X --n_sample S
//       FindSampleRank with method US (src/app/build-index/FindSampleRank.cpp:71-167):
//       writes X/qrs_to_sample_index/US-N-n_sample_S.index (header + ranks).
//   rkr_cli bsibc --dataset_dir D --dataset_name N --index_dir X --n_sample S
//       BuildSampleIndexByComputation (src/app/build-index/BuildSampleIndexByComputation.cpp:66-214):
//       builds the sampled score table on the GPU and writes
//       X/memory_index/US-N-n_sample_S.index.
//   rkr_cli qdibc --dataset_dir D --dataset_name N --index_dir X [--n_sample_item Q] [--sample_topk K]
//       QueryDistributionIndexByComputation (src/app/build-index/QueryDistributionIndexByComputation.cpp,
//       QueryRankByComputation.hpp:244-297): samples Q items, ranks each for every user on the GPU and
//       writes X/query_distribution/N-n_sample_item_Q-sample_topk_K/ (FileIO.hpp:56-137).
//   --method_name QS (fsr, bsibc, rri; with --n_sample_query Q --sample_topk K): fsr trains the ranks
//       from qdibc's files with FindOptimalSampleFast into
//       X/qrs_to_sample_index/QS-N-n_sample_S-n_sample_query_Q-sample_topk_K.index; bsibc builds the
//       table on them into X/memory_index/ under the same name; rri answers from that table.
//   rkr_cli bst   --dataset_dir D --dataset_name N --index_dir X [--eval_size_gb G]
//       BuildScoreTableBatch (src/app/BuildScoreTableBatch.cpp:56-144, 159-169): writes every user's
//       sorted scores (ReadAll's layout, float[n_user][n_item]) to X/N.index; with eval_size_gb below
//       the table's size, only the first eval_size_gb GiB of rows.
//   rkr_cli rri   --dataset_dir D --dataset_name N --index_dir X --n_sample S --topk K
//                 [--form plain|screened] [--output exact|reference] [--result_dir R]
//       ReverseRankIndex with method US (src/app/ReverseRankIndex.cpp:83-219):
//       answers every query item and writes R/rank/N-US-top<K>-n_sample_<S>-{userID,rank,IP}.csv
//       and R/single_query_performance/...csv (FileIO.hpp:84-206 layouts).
//   rkr_cli rriu  --dataset_dir D --dataset_name N --index_dir X --n_sample S --topk K
//                 --update_type user|data_item --update_operator insert|delete --updateID 0..4
//       ReverseRankIndexUpdate (src/app/ReverseRankIndexUpdate.cpp:81-227): loads the table bsibc
//       built on the base sets, inserts the first updateID quarters of N_{user,data_item}_update.fvecs
//       or deletes as many users/items from the front, answers every query over the updated sets
//       and writes the result CSVs with -updateID_<i>-<type>-<operator> in their names.  Methods
//       QSUpdate and QSRPNormalLPUpdate read the QS table, US the US table.
// Datasets are <D>/<N>/<N>_{data_item,user,query_item}.fvecs
// (src/include/util/VectorIO.hpp:40-68); users are normalised on load.
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "rkr.h"
#include "rkr_us.hpp"

namespace {

using Args = std::map<std::string, std::string>;

Args parse(int argc, char **argv, int first) {
    Args a;
    for (int i = first; i < argc; ++i) {
        std::string k = argv[i];
        if (k.rfind("--", 0) != 0) continue;
        k = k.substr(2);
        std::string v = (i + 1 < argc && std::strncmp(argv[i + 1], "--", 2) != 0) ? argv[++i] : "1";
        a[k] = v;
    }
    return a;
}

std::string get(const Args &a, const char *k, const char *def) {
    auto it = a.find(k);
    return it == a.end() ? def : it->second;
}

void mkdirs(const std::string &p) {
    std::string cur;
    for (size_t i = 0; i < p.size(); ++i) {
        cur += p[i];
        if (p[i] == '/' || i + 1 == p.size()) mkdir(cur.c_str(), 0755);
    }
}

// loadVector (VectorIO.hpp:14-37)
bool load_fvecs(const std::string &path, rkr::VectorMatrix &m) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    int32_t d = 0;
    if (std::fread(&d, 4, 1, f) != 1 || d <= 0) { std::fclose(f); return false; }
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    const long n = size / (4L * (d + 1));
    m.n_vector_ = (int)n;
    m.vec_dim_ = d;
    m.data.resize((size_t)n * d);
    for (long i = 0; i < n; ++i) {
        int32_t dd;
        if (std::fread(&dd, 4, 1, f) != 1 || dd != d ||
            std::fread(m.data.data() + (size_t)i * d, 4, (size_t)d, f) != (size_t)d) {
            std::fclose(f);
            return false;
        }
    }
    std::fclose(f);
    return true;
}

bool read_data(const Args &a, rkr::VectorMatrix &user, rkr::VectorMatrix &item, rkr::VectorMatrix &query) {
    const std::string base = get(a, "dataset_dir", ".") + "/" + get(a, "dataset_name", "fake-normal") + "/" +
                             get(a, "dataset_name", "fake-normal");
    if (!load_fvecs(base + "_data_item.fvecs", item) || !load_fvecs(base + "_user.fvecs", user) ||
        !load_fvecs(base + "_query_item.fvecs", query)) {
        std::fprintf(stderr, "cannot read dataset %s_*.fvecs\n", base.c_str());
        return false;
    }
    rkr_normalize_users(user.data.data(), user.n_vector_, user.vec_dim_);  // readData, VectorIO.hpp:61
    std::printf("n_data_item %d, n_query_item %d, n_user %d, vec_dim %d\n", item.n_vector_, query.n_vector_,
                user.n_vector_, user.vec_dim_);
    return true;
}

// QSRPNormalLP and the update methods run on the QS table (SampleSearchIndexName,
// NameTranslation.hpp:36-48)
bool is_qsrp(const Args &a) { return get(a, "method_name", "US") == "QSRPNormalLP"; }
bool is_qs(const Args &a) {
    const std::string m = get(a, "method_name", "US");
    return m == "QS" || m == "QSUpdate" || m == "QSRPNormalLPUpdate" || is_qsrp(a);
}

// SampleSearch::SaveIndex / LoadIndex names (SampleSearch.hpp:158-179): QS adds the training parameters
std::string index_file(const Args &a, const char *sub, int S) {
    std::string name = get(a, "index_dir", "index") + "/" + sub + "/" + (is_qs(a) ? "QS-" : "US-") +
                       get(a, "dataset_name", "fake-normal") + "-n_sample_" + std::to_string(S);
    if (is_qs(a))
        name += "-n_sample_query_" + get(a, "n_sample_query", "150") + "-sample_topk_" + get(a, "sample_topk", "50");
    return name + ".index";
}

std::string query_distribution_dir(const Args &a, const std::string &n_sample_item, const std::string &topk) {
    return get(a, "index_dir", "index") + "/query_distribution/" + get(a, "dataset_name", "fake-normal") +
           "-n_sample_item_" + n_sample_item + "-sample_topk_" + topk;
}

bool write_file(const std::string &path, const void *p, size_t bytes) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}

bool read_file(const std::string &path, void *p, size_t bytes) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    const bool ok = std::fread(p, 1, bytes, f) == bytes && std::fgetc(f) == EOF;
    std::fclose(f);
    return ok;
}

int cmd_qdibc(const Args &a) {
    rkr::VectorMatrix user, item, query;
    if (!read_data(a, user, item, query)) return 1;
    const std::string qs = get(a, "n_sample_item", "150"), ks = get(a, "sample_topk", "50");
    const int Q = std::stoi(qs), K = std::stoi(ks);
    const int64_t I = item.n_vector_, W = I + 1;
    std::vector<int32_t> ids((size_t)std::max(Q, 1)), kth((size_t)std::max(Q, 1)), order((size_t)std::max(Q, 1));
    if (rkr_sample_items(I, Q, ids.data()) != RKR_OK) {
        std::fprintf(stderr, "n_sample_item must be in [1, n_item]\n");
        return 1;
    }
    try {
        rkr::Index idx(std::stoi(get(a, "device", "0")));
        idx.SetData(user, item, false);
        std::vector<int32_t> accu((size_t)(Q * W)), sorted((size_t)(Q * W));
        double ms = 0;
        const auto t0 = std::chrono::steady_clock::now();
        if (int rc = rkr_query_rank_distribution(idx.ctx(), ids.data(), Q, accu.data(), &ms))
            throw rkr::Error(rc, rkr_last_error(idx.ctx()));
        if (rkr_sort_kth_rank(accu.data(), I, Q, K, kth.data(), order.data()) != RKR_OK)
            throw rkr::Error(RKR_ERR_INVALID, "bad sample_topk");
        const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int s = 0; s < Q; ++s)
            std::copy(accu.begin() + (int64_t)order[s] * W, accu.begin() + (int64_t)(order[s] + 1) * W,
                      sorted.begin() + (int64_t)s * W);
        // WriteDistributionBelowTopk (FileIO.hpp:56-137): samples in k-th-rank order
        const std::string dir = query_distribution_dir(a, qs, ks);
        mkdirs(dir);
        std::ofstream fi(dir + "/sample_itemID_l.txt");
        for (int s = 0; s < Q; ++s) fi << ids[order[s]] << "\n";
        fi.close();
        if (!fi || !write_file(dir + "/sort_kth_rank_l.index", kth.data(), sizeof(int32_t) * Q) ||
            !write_file(dir + "/accu_n_user_rank_l.index", sorted.data(), sizeof(int32_t) * sorted.size())) {
            std::fprintf(stderr, "cannot write %s\n", dir.c_str());
            return 1;
        }
        std::printf("report build_index_time %.3fs (device rank distribution %.3fs), wrote %s\n", total, ms * 1e-3,
                    dir.c_str());
    } catch (const rkr::Error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

// FindSampleRank with method QS (FindSampleRank.cpp:71-167, FindOptimalSampleFast) from qdibc's files
int train_qs_ranks(const Args &a, int64_t n_user, int64_t n_item, int S, std::vector<int32_t> &r) {
    const std::string qs = get(a, "n_sample_query", "150"), ks = get(a, "sample_topk", "50");
    const int Q = std::stoi(qs);
    const int64_t W = n_item + 1;
    const std::string dir = query_distribution_dir(a, qs, ks);
    std::vector<int32_t> kth((size_t)std::max(Q, 1)), accu((size_t)(std::max(Q, 1) * W));
    if (Q <= 0 || !read_file(dir + "/sort_kth_rank_l.index", kth.data(), sizeof(int32_t) * Q) ||
        !read_file(dir + "/accu_n_user_rank_l.index", accu.data(), sizeof(int32_t) * accu.size())) {
        std::fprintf(stderr, "cannot read the query distribution in %s (run qdibc)\n", dir.c_str());
        return 1;
    }
    try {
        rkr::Index idx(std::stoi(get(a, "device", "0")));
        int64_t cost = 0;
        r.assign((size_t)S, 0);
        if (int rc = rkr_find_optimal_sample(idx.ctx(), n_user, n_item, Q, kth.data(), accu.data(),
                                             RKR_SAMPLE_OPTIMAL_PART, S, r.data(), &cost))
            throw rkr::Error(rc, rkr_last_error(idx.ctx()));
        std::printf("train io cost: %lld\n", (long long)cost);
    } catch (const rkr::Error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

int cmd_fsr(const Args &a) {
    rkr::VectorMatrix user, item, query;
    if (!read_data(a, user, item, query)) return 1;
    const int S = std::stoi(get(a, "n_sample", "-1"));
    std::vector<int32_t> r((size_t)std::max(S, 1));
    if (S <= 0 || S >= item.n_vector_) {
        std::fprintf(stderr, "n_sample too small or too large\n");
        return 1;
    }
    if (is_qs(a)) {
        if (train_qs_ranks(a, user.n_vector_, item.n_vector_, S, r)) return 1;
    } else if (rkr_uniform_sample_ranks(item.n_vector_, S, r.data()) != RKR_OK) {
        std::fprintf(stderr, "n_sample too small or too large\n");
        return 1;
    }
    mkdirs(get(a, "index_dir", "index") + "/qrs_to_sample_index");
    const std::string path = index_file(a, "qrs_to_sample_index", S);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) { std::fprintf(stderr, "cannot write %s\n", path.c_str()); return 1; }
    const size_t hdr[3] = {(size_t)S, (size_t)item.n_vector_, (size_t)user.n_vector_};
    std::fwrite(hdr, sizeof(size_t), 3, f);
    std::fwrite(r.data(), sizeof(int32_t), (size_t)S, f);
    std::fclose(f);
    std::printf("wrote %s\n", path.c_str());
    return 0;
}

int cmd_bsibc(const Args &a) {
    rkr::VectorMatrix user, item, query;
    if (!read_data(a, user, item, query)) return 1;
    const int S = std::stoi(get(a, "n_sample", "-1"));
    // ranks from fsr's file when present (BuildSampleIndexByComputation.cpp:77-79), else uniform
    std::vector<int32_t> r((size_t)std::max(S, 1));
    bool have = false;
    if (FILE *f = std::fopen(index_file(a, "qrs_to_sample_index", S).c_str(), "rb")) {
        size_t hdr[3];
        have = std::fread(hdr, sizeof(size_t), 3, f) == 3 && hdr[0] == (size_t)S &&
               std::fread(r.data(), sizeof(int32_t), (size_t)S, f) == (size_t)S;
        std::fclose(f);
    }
    if (!have && is_qs(a)) {
        std::fprintf(stderr, "no trained ranks at %s (run fsr --method_name QS)\n",
                     index_file(a, "qrs_to_sample_index", S).c_str());
        return 1;
    }
    if (!have && rkr_uniform_sample_ranks(item.n_vector_, S, r.data()) != RKR_OK) {
        std::fprintf(stderr, "n_sample too small or too large\n");
        return 1;
    }
    try {
        rkr::Index idx(std::stoi(get(a, "device", "0")));
        const auto t0 = std::chrono::steady_clock::now();
        idx.SetData(user, item, false);
        const double build_ms = idx.BuildTable(r);
        mkdirs(get(a, "index_dir", "index") + "/memory_index");
        const std::string path = index_file(a, "memory_index", S);
        idx.SaveTable(path);
        const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("total_index_construction_time: %.3fs (device table build %.3fs), wrote %s\n", total,
                    build_ms * 1e-3, path.c_str());
    } catch (const rkr::Error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

int cmd_bst(const Args &a) {
    rkr::VectorMatrix user, item, query;
    if (!read_data(a, user, item, query)) return 1;
    // ComputeEvalUserSize (BuildScoreTableBatch.cpp:56-64): the rows that fit eval_size_gb GiB
    const uint64_t eval_gb = std::stoull(get(a, "eval_size_gb", "500"));
    const uint64_t row_bytes = (uint64_t)item.n_vector_ * sizeof(float);
    const uint64_t total_gb = (uint64_t)user.n_vector_ * row_bytes / 1024 / 1024 / 1024;
    const int64_t n_eval = eval_gb > total_gb ? user.n_vector_
                                              : (int64_t)std::min<uint64_t>(eval_gb * 1024 * 1024 * 1024 / row_bytes,
                                                                            (uint64_t)user.n_vector_);
    try {
        rkr::Index idx(std::stoi(get(a, "device", "0")));
        idx.SetData(user, item, false);
        mkdirs(get(a, "index_dir", "index"));
        const std::string path = get(a, "index_dir", "index") + "/" + get(a, "dataset_name", "fake-normal") + ".index";
        const auto t = idx.SaveScoreTable(path, n_eval);
        std::printf("build index time: total %.3fs, compute sort time %.3f s, save index time %.3f s, "
                    "%lld of %d users, wrote %s\n", (t.first + t.second) * 1e-3, t.first * 1e-3, t.second * 1e-3,
                    (long long)n_eval, user.n_vector_, path.c_str());
    } catch (const rkr::Error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

int cmd_rri(const Args &a) {
    rkr::VectorMatrix user, item, query;
    if (!read_data(a, user, item, query)) return 1;
    const int S = std::stoi(get(a, "n_sample", "20"));
    const int topk = std::stoi(get(a, "topk", "10"));
    const std::string form = get(a, "form", "plain"), output = get(a, "output", "exact");
    const std::string res_dir = get(a, "result_dir", "result");
    const std::string ds = get(a, "dataset_name", "fake-normal");
    try {
        auto t0 = std::chrono::steady_clock::now();
        auto idx = is_qs(a) ? rkr::BuildIndexQS(item, user, ds.c_str(), S, std::stoi(get(a, "n_sample_query", "150")),
                                                std::stoi(get(a, "sample_topk", "50")), -1,
                                                get(a, "index_dir", "index").c_str(), std::stoi(get(a, "device", "0")))
                            : rkr::BuildIndex(item, user, ds.c_str(), get(a, "index_dir", "index").c_str(), S, -1,
                                   std::stoi(get(a, "device", "0")));
        idx->SetForm(form == "screened" ? RKR_FORM_SCREENED : RKR_FORM_PLAIN);
        idx->SetOutput(output == "reference" ? RKR_OUTPUT_REFERENCE : RKR_OUTPUT_EXACT);
        if (is_qsrp(a)) idx->BuildQSRP(std::stof(get(a, "SIGMA", "0.7")));  // BuildSVDIndex's --SIGMA
        const double build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::vector<rkr::SingleQueryPerformance> perf;
        auto res = idx->Retrieval(query, topk, query.n_vector_, perf);
        std::printf("%s\nbuild index time: total %.3fs\n", idx->PerformanceStatistics(topk).c_str(), build_s);
        // WriteRankResult (FileIO.hpp:84-148) and WriteQueryPerformance (FileIO.hpp:175-206)
        char param[256];
        if (is_qs(a))  // ReverseRankIndex.cpp:122-123
            std::snprintf(param, sizeof param, "top%d-n_sample_%d-n_sample_query_%s-sample_topk_%s-n_thread_%d", topk,
                          S, get(a, "n_sample_query", "150").c_str(), get(a, "sample_topk", "50").c_str(), -1);
        else
            std::snprintf(param, sizeof param, "top%d-n_sample_%d-n_thread_%d", topk, S, -1);
        const std::string method = is_qsrp(a) ? "-QSRPNormalLP-" : is_qs(a) ? "-QS-" : "-US-";
        mkdirs(res_dir + "/rank");
        mkdirs(res_dir + "/single_query_performance");
        const std::string stem = res_dir + "/rank/" + ds + method + param;
        std::ofstream fu(stem + "-userID.csv"), fr(stem + "-rank.csv"), fi(stem + "-IP.csv");
        for (const auto &row : res) {
            for (size_t j = 0; j < row.size(); ++j) {
                const char *sep = j + 1 < row.size() ? "," : "\n";
                fu << row[j].userID_ << sep;
                fr << row[j].rank_ << sep;
                fi << row[j].queryIP_ << sep;
            }
        }
        std::ofstream fp(res_dir + "/single_query_performance/" + ds + method + param + "-single-query-performance.csv");
        fp << "queryID,n_prune_user,n_result_user,n_refine_user,ip_cost,io_cost,total_time,memory_index_time,io_time\n";
        char buf[512];
        for (const auto &p : perf) {
            std::snprintf(buf, sizeof buf, "%10d,%10d,%10d,%10d,%10ld,%10ld,%10.2f,%10.2f,%10.2f", p.queryID_,
                          p.n_prune_user_, p.n_result_user_, p.n_refine_user_, (long)p.ip_cost_, (long)p.io_cost_,
                          p.total_time_, p.memory_index_time_, p.io_time_);
            fp << buf << "\n";
        }
        std::printf("wrote %s-{userID,rank,IP}.csv\n", stem.c_str());
    } catch (const rkr::Error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

// The update run of ReverseRankIndexUpdate.cpp:81-227 on the table of the base sets.
int cmd_rriu(const Args &a) {
    rkr::VectorMatrix user, item, query, upd;
    if (!read_data(a, user, item, query)) return 1;
    const std::string type = get(a, "update_type", "data_item"), op = get(a, "update_operator", "insert");
    const int update_id = std::stoi(get(a, "updateID", "0"));
    if (update_id < 0 || update_id > 4) {  // :153-156
        std::fprintf(stderr, "updateID should be in [0, 4]\n");
        return 2;
    }
    if (type != "user" && type != "data_item") {
        std::fprintf(stderr, "no such update type %s\n", type.c_str());
        return 2;
    }
    if (op != "insert" && op != "delete") {
        std::fprintf(stderr, "no such update operator %s\n", op.c_str());
        return 2;
    }
    // readDataUpdate (VectorIOUpdate.hpp:43-98): N_user_update.fvecs (normalised) or
    // N_data_item_update.fvecs
    const std::string base = get(a, "dataset_dir", ".") + "/" + get(a, "dataset_name", "fake-normal") + "/" +
                             get(a, "dataset_name", "fake-normal");
    if (!load_fvecs(base + (type == "user" ? "_user_update.fvecs" : "_data_item_update.fvecs"), upd)) {
        std::fprintf(stderr, "cannot read the %s update vectors\n", type.c_str());
        return 1;
    }
    if (type == "user") rkr_normalize_users(upd.data.data(), upd.n_vector_, upd.vec_dim_);
    // the first updateID quarters of the update set (:158-171)
    const int n_vecs = upd.n_vector_ / 4 * update_id;
    const int S = std::stoi(get(a, "n_sample", "20"));
    const int topk = std::stoi(get(a, "topk", "10"));
    const std::string ds = get(a, "dataset_name", "fake-normal");
    const std::string res_dir = get(a, "result_dir", "result");
    try {
        auto idx = is_qs(a) ? rkr::BuildIndexQS(item, user, ds.c_str(), S, std::stoi(get(a, "n_sample_query", "150")),
                                                std::stoi(get(a, "sample_topk", "50")), -1,
                                                get(a, "index_dir", "index").c_str(), std::stoi(get(a, "device", "0")))
                            : rkr::BuildIndex(item, user, ds.c_str(), get(a, "index_dir", "index").c_str(), S, -1,
                                              std::stoi(get(a, "device", "0")));
        rkr_ctx *c = idx->ctx();
        auto check = [&](int rc) {
            if (rc != RKR_OK) throw rkr::Error(rc, rkr_last_error(c));
        };
        int n_user = user.n_vector_;
        if (update_id != 0) {  // :174-200
            if (op == "insert") {
                if (type == "user") {
                    check(rkr_insert_users(c, upd.data.data(), n_vecs));
                    n_user += n_vecs;
                } else {
                    check(rkr_insert_items(c, upd.data.data(), n_vecs));
                }
            } else {
                std::vector<int64_t> ids((size_t)n_vecs);
                for (int i = 0; i < n_vecs; ++i) ids[i] = i;
                if (type == "user") {
                    check(rkr_delete_users(c, ids.data(), n_vecs));
                    n_user -= n_vecs;
                } else {
                    check(rkr_delete_items(c, ids.data(), n_vecs));
                }
            }
        }
        // Retrieval over the updated sets (:202-205)
        const int nq = query.n_vector_;
        std::vector<int32_t> ks((size_t)nq, topk), uo((size_t)nq * topk), ro((size_t)nq * topk);
        std::vector<float> qo((size_t)nq * topk);
        if (topk > n_user) throw rkr::Error(RKR_ERR_INVALID, "top-k is too large");
        check(rkr_query(c, query.data.data(), nq, ks.data(), topk, uo.data(), ro.data(), qo.data(), nullptr, nullptr));
        int64_t cu = 0, ci = 0, ni = 0, nd = 0;
        check(rkr_get_counts(c, &cu, &ci, &ni, &nd));
        std::printf("UpdateID: %d, n_data_item %lld, n_user %lld\n", update_id, (long long)ci, (long long)cu);
        char param[512];
        std::snprintf(param, sizeof param, "top%d-n_sample_%d-n_thread_%d-updateID_%d-%s-%s", topk, S, -1, update_id,
                      type.c_str(), op.c_str());
        mkdirs(res_dir + "/rank");
        const std::string stem = res_dir + "/rank/" + ds + "-" + get(a, "method_name", "US") + "-" + param;
        std::ofstream fu(stem + "-userID.csv"), fr(stem + "-rank.csv"), fi(stem + "-IP.csv");
        for (int q = 0; q < nq; ++q)
            for (int j = 0; j < topk; ++j) {
                const char *sep = j + 1 < topk ? "," : "\n";
                const size_t o = (size_t)q * topk + j;
                fu << uo[o] << sep;
                fr << ro[o] << sep;
                fi << qo[o] << sep;
            }
        std::printf("wrote %s-{userID,rank,IP}.csv\n", stem.c_str());
    } catch (const rkr::Error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: rkr_cli qdibc|fsr|bsibc|bst|rri|rriu --dataset_dir D --dataset_name N --index_dir X "
                             "--n_sample S [--topk K] [--form plain|screened] [--output exact|reference]\n");
        return 2;
    }
    const std::string cmd = argv[1];
    const Args a = parse(argc, argv, 2);
    if (cmd == "qdibc") return cmd_qdibc(a);
    if (cmd == "rriu") return cmd_rriu(a);
    if (get(a, "method_name", "US") != "US" && !is_qs(a)) {
        std::fprintf(stderr, "method_name US, QS and QSRPNormalLP are implemented\n");
        return 2;
    }
    if (cmd == "fsr") return cmd_fsr(a);
    if (cmd == "bsibc") return cmd_bsibc(a);
    if (cmd == "bst") return cmd_bst(a);
    if (cmd == "rri") return cmd_rri(a);
    std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
    return 2;
}
