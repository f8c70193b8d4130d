// rkr_cli -- command-line drivers of the US path over librkr.so, with the
// reference's binaries' names, options and files:
//   rkr_cli fsr   --dataset_dir D --dataset_name N --index_dir X --n_sample S
//       FindSampleRank with method US (src/app/build-index/FindSampleRank.cpp:71-167):
//       writes X/qrs_to_sample_index/US-N-n_sample_S.index (header + ranks).
//   rkr_cli bsibc --dataset_dir D --dataset_name N --index_dir X --n_sample S
//       BuildSampleIndexByComputation (src/app/build-index/BuildSampleIndexByComputation.cpp:66-214):
//       builds the sampled score table on the GPU and writes
//       X/memory_index/US-N-n_sample_S.index.
//   rkr_cli bst   --dataset_dir D --dataset_name N --index_dir X [--eval_size_gb G]
//       BuildScoreTableBatch (src/app/BuildScoreTableBatch.cpp:56-144, 159-169): writes every user's
//       sorted scores (ReadAll's layout, float[n_user][n_item]) to X/N.index; with eval_size_gb below
//       the table's size, only the first eval_size_gb GiB of rows.
//   rkr_cli rri   --dataset_dir D --dataset_name N --index_dir X --n_sample S --topk K
//                 [--form plain|screened] [--output exact|reference] [--result_dir R]
//       ReverseRankIndex with method US (src/app/ReverseRankIndex.cpp:83-219):
//       answers every query item and writes R/rank/N-US-top<K>-n_sample_<S>-{userID,rank,IP}.csv
//       and R/single_query_performance/...csv (FileIO.hpp:84-206 layouts).
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

std::string index_file(const Args &a, const char *sub, int S) {
    return get(a, "index_dir", "index") + "/" + sub + "/US-" + get(a, "dataset_name", "fake-normal") + "-n_sample_" +
           std::to_string(S) + ".index";
}

int cmd_fsr(const Args &a) {
    rkr::VectorMatrix user, item, query;
    if (!read_data(a, user, item, query)) return 1;
    const int S = std::stoi(get(a, "n_sample", "-1"));
    std::vector<int32_t> r((size_t)std::max(S, 1));
    if (rkr_uniform_sample_ranks(item.n_vector_, S, r.data()) != RKR_OK) {
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
        auto idx = rkr::BuildIndex(item, user, ds.c_str(), get(a, "index_dir", "index").c_str(), S, -1,
                                   std::stoi(get(a, "device", "0")));
        idx->SetForm(form == "screened" ? RKR_FORM_SCREENED : RKR_FORM_PLAIN);
        idx->SetOutput(output == "reference" ? RKR_OUTPUT_REFERENCE : RKR_OUTPUT_EXACT);
        const double build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::vector<rkr::SingleQueryPerformance> perf;
        auto res = idx->Retrieval(query, topk, query.n_vector_, perf);
        std::printf("%s\nbuild index time: total %.3fs\n", idx->PerformanceStatistics(topk).c_str(), build_s);
        // WriteRankResult (FileIO.hpp:84-148) and WriteQueryPerformance (FileIO.hpp:175-206)
        char param[256];
        std::snprintf(param, sizeof param, "top%d-n_sample_%d-n_thread_%d", topk, S, -1);
        mkdirs(res_dir + "/rank");
        mkdirs(res_dir + "/single_query_performance");
        const std::string stem = res_dir + "/rank/" + ds + "-US-" + param;
        std::ofstream fu(stem + "-userID.csv"), fr(stem + "-rank.csv"), fi(stem + "-IP.csv");
        for (const auto &row : res) {
            for (size_t j = 0; j < row.size(); ++j) {
                const char *sep = j + 1 < row.size() ? "," : "\n";
                fu << row[j].userID_ << sep;
                fr << row[j].rank_ << sep;
                fi << row[j].queryIP_ << sep;
            }
        }
        std::ofstream fp(res_dir + "/single_query_performance/" + ds + "-US-" + param + "-single-query-performance.csv");
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

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: rkr_cli fsr|bsibc|bst|rri --dataset_dir D --dataset_name N --index_dir X "
                             "--n_sample S [--topk K] [--form plain|screened] [--output exact|reference]\n");
        return 2;
    }
    const std::string cmd = argv[1];
    const Args a = parse(argc, argv, 2);
    if (get(a, "method_name", "US") != "US") {
        std::fprintf(stderr, "only method_name US is implemented\n");
        return 2;
    }
    if (cmd == "fsr") return cmd_fsr(a);
    if (cmd == "bsibc") return cmd_bsibc(a);
    if (cmd == "bst") return cmd_bst(a);
    if (cmd == "rri") return cmd_rri(a);
    std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
    return 2;
}
