// rkr_us.hpp -- C++ mirror of the reference's US index interface over the
// C ABI (rkr.h).  Header-only; link with librkr.so.
//
// The reference declares the query interface as BaseIndex::Retrieval
// (src/include/struct/MethodBase.hpp:14-30), implements it in
// US::Index::Retrieval (src/impl/US.hpp:96-213) and builds the index with
// US::BuildIndex (src/impl/US.hpp:246-261).  This header keeps those names,
// argument meanings and result shape (one vector of topk UserRankElement per
// query, plus SingleQueryPerformance records), with two deliberate changes:
// errors throw rkr::Error instead of calling exit(), and every query runs
// (the reference answers only the first 100 when n_sample <= 1000, US.hpp:112).
#pragma once

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rkr.h"

namespace rkr {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

// Row-major FP32 matrix view (the role of VectorMatrix,
// src/include/struct/VectorMatrix.hpp:12-69).
struct VectorMatrix {
    std::vector<float> data;
    int n_vector_ = 0, vec_dim_ = 0;
    const float *getVector(int i) const { return data.data() + (size_t)i * vec_dim_; }
    const float *getRawData() const { return data.data(); }
};

// src/include/struct/UserRankElement.hpp:11-92: ordered by (rank, userID).
struct UserRankElement {
    int userID_ = -1, rank_ = -1;
    float queryIP_ = -1.0f;
    bool operator<(const UserRankElement &o) const {
        return rank_ != o.rank_ ? rank_ < o.rank_ : userID_ < o.userID_;
    }
};

// src/include/util/FileIO.hpp:150-173.
struct SingleQueryPerformance {
    int queryID_ = 0, n_prune_user_ = 0, n_result_user_ = 0, n_refine_user_ = 0;
    size_t io_cost_ = 0, ip_cost_ = 0;
    double total_time_ = 0, memory_index_time_ = 0, io_time_ = 0;
};

class Index {
  public:
    Index(int device = 0, size_t mem_budget = 0) {
        rkr_ctx *c = nullptr;
        check(rkr_create(device, mem_budget, &c));
        ctx_.reset(c);
    }

    void SetData(const VectorMatrix &user, const VectorMatrix &data_item, bool normalize_users = true) {
        check(rkr_set_data(ctx_.get(), user.getRawData(), user.n_vector_, data_item.getRawData(),
                           data_item.n_vector_, user.vec_dim_, normalize_users ? 1 : 0));
        n_user_ = user.n_vector_;
    }
    void SetForm(int form) { check(rkr_set_option(ctx_.get(), RKR_OPT_FORM, form)); }
    void SetOutput(int output) { check(rkr_set_option(ctx_.get(), RKR_OPT_OUTPUT, output)); }
    double BuildTable(const std::vector<int32_t> &ranks) {
        double ms = 0;
        check(rkr_build_table(ctx_.get(), ranks.data(), (int32_t)ranks.size(), &ms));
        return ms;
    }
    void LoadTable(const std::string &path, float eps = 0.0f) { check(rkr_load_table(ctx_.get(), path.c_str(), eps)); }
    // QSRPNormalLP's screen (SVD.hpp:62-123 + NormalLinearRegressionLP.hpp:149-283) fitted to the
    // loaded table and users, then used by Retrieval; SIGMA as BuildSVDIndex's --SIGMA (default 0.7).
    double BuildQSRP(float sigma = 0.7f) {
        double ms = 0;
        check(rkr_qsrp_build(ctx_.get(), sigma, &ms));
        check(rkr_set_option(ctx_.get(), RKR_OPT_QSRP, 1));
        return ms;
    }
    void SaveTable(const std::string &path) { check(rkr_save_table(ctx_.get(), path.c_str())); }
    // ComputeScoreTableBatch::ComputeSortItemsBatch (ComputeScoreTableBatch.hpp:79-92): batch_n_user rows of
    // n_item scores, each sorted in descending order, into distance_l.
    void ComputeSortItemsBatch(int start_userID, int batch_n_user, float *distance_l) {
        check(rkr_sorted_scores(ctx_.get(), start_userID, batch_n_user, distance_l));
    }
    // ReadAll's index file for the first n_user users (all when negative); returns the wall time in ms
    // of computing the rows and of writing them.
    std::pair<double, double> SaveScoreTable(const std::string &path, int64_t n_user = -1) {
        double t[4] = {0, 0, 0, 0};
        check(rkr_save_score_table(ctx_.get(), path.c_str(), n_user, t));
        return {t[2], t[3]};
    }

    // BaseIndex::Retrieval (MethodBase.hpp:16-18) / US::Index::Retrieval (US.hpp:96-213).
    std::vector<std::vector<UserRankElement>> Retrieval(const VectorMatrix &query_item, const int &topk,
                                                        const int &n_execute_query,
                                                        std::vector<SingleQueryPerformance> &query_performance_l) {
        if (n_execute_query > query_item.n_vector_)
            throw Error(RKR_ERR_INVALID, "n_execute_query larger than n_query_item");
        if (topk > n_user_) throw Error(RKR_ERR_INVALID, "top-k is too large");
        const int nq = n_execute_query;
        std::vector<int32_t> ks((size_t)nq, topk), users((size_t)nq * topk), ranks((size_t)nq * topk);
        std::vector<float> qips((size_t)nq * topk);
        std::vector<rkr_query_stats> st((size_t)nq);
        rkr_batch_timing tm{};
        check(rkr_query(ctx_.get(), query_item.getRawData(), nq, ks.data(), topk, users.data(), ranks.data(),
                        qips.data(), st.data(), &tm));
        last_timing_ = tm;
        std::vector<std::vector<UserRankElement>> out((size_t)nq, std::vector<UserRankElement>((size_t)topk));
        if ((int)query_performance_l.size() < nq) query_performance_l.resize((size_t)nq);
        for (int q = 0; q < nq; ++q) {
            for (int j = 0; j < topk; ++j) {
                const size_t o = (size_t)q * topk + j;
                out[q][j].userID_ = users[o];
                out[q][j].rank_ = ranks[o];
                out[q][j].queryIP_ = qips[o];
            }
            SingleQueryPerformance &p = query_performance_l[q];
            p.queryID_ = q;
            p.n_prune_user_ = st[q].n_prune_user;
            p.n_result_user_ = st[q].n_result_user;
            p.n_refine_user_ = st[q].n_refine_user;
            p.ip_cost_ = (size_t)st[q].ip_cost;
            p.total_time_ = tm.total_ms * 1e-3 / (nq ? nq : 1);
        }
        return out;
    }

    // US::Index::PerformanceStatistics (US.hpp:216-233), same field order.
    std::string PerformanceStatistics(const int &topk) const {
        char buf[512];
        std::snprintf(buf, sizeof buf,
                      "top%d retrieval time: total %.3fs\n\tinner product %.3fs, rank bound %.3fs, prune user %.3fs, "
                      "refine user %.3fs\n\ttotal refine user %lld, rechecked %lld of %lld screened",
                      topk, last_timing_.total_ms * 1e-3, last_timing_.ip_ms * 1e-3, last_timing_.bound_ms * 1e-3,
                      last_timing_.prune_ms * 1e-3, (last_timing_.refine_ms + last_timing_.recheck_ms) * 1e-3,
                      (long long)last_timing_.n_refined, (long long)last_timing_.n_rechecked,
                      (long long)last_timing_.n_screened);
        return buf;
    }

    uint64_t IndexSizeByte() const { return (uint64_t)rkr_index_size_bytes(ctx_.get()); }
    void FinishCompute() { check(rkr_sync(ctx_.get())); }
    const rkr_batch_timing &LastTiming() const { return last_timing_; }
    rkr_ctx *ctx() const { return ctx_.get(); }

  private:
    struct Del {
        void operator()(rkr_ctx *c) const { rkr_destroy(c); }
    };
    void check(int rc) const {
        if (rc != RKR_OK) throw Error(rc, ctx_ ? rkr_last_error(ctx_.get()) : "rkr_create failed");
    }
    std::unique_ptr<rkr_ctx, Del> ctx_;
    int n_user_ = 0;
    rkr_batch_timing last_timing_{};
};

// US::BuildIndex (US.hpp:246-261): loads <basic_index_path>/memory_index/
// US-<dataset_name>-n_sample_<n_sample>.index written by bsibc.  The
// reference's n_thread has no meaning on the device and is accepted only for
// call compatibility.
inline std::unique_ptr<Index> BuildIndex(VectorMatrix &data_item, VectorMatrix &user, const char *dataset_name,
                                         const char *basic_index_path, const int &n_sample, const int &n_thread,
                                         int device = 0) {
    (void)n_thread;
    auto idx = std::make_unique<Index>(device);
    idx->SetData(user, data_item, /*normalize_users=*/false);  // readData already normalised them
    char path[1024];
    std::snprintf(path, sizeof path, "%s/memory_index/US-%s-n_sample_%d.index", basic_index_path, dataset_name,
                  n_sample);
    idx->LoadTable(path);
    return idx;
}

// QS::BuildIndex (src/impl/QS.hpp, called at ReverseRankIndex.cpp:113-123): loads
// <index_dir>/memory_index/QS-<dataset_name>-n_sample_<n_sample>-n_sample_query_<q>-sample_topk_<k>.index,
// the table bsibc --method_name QS builds on the trained ranks.  Answers come from the same exact query
// path as US.
inline std::unique_ptr<Index> BuildIndexQS(VectorMatrix &data_item, VectorMatrix &user, const char *dataset_name,
                                           const int &n_sample, const int &n_sample_query, const int &sample_topk,
                                           const int &n_thread, const char *index_dir, int device = 0) {
    (void)n_thread;
    auto idx = std::make_unique<Index>(device);
    idx->SetData(user, data_item, /*normalize_users=*/false);
    char path[1024];
    std::snprintf(path, sizeof path, "%s/memory_index/QS-%s-n_sample_%d-n_sample_query_%d-sample_topk_%d.index",
                  index_dir, dataset_name, n_sample, n_sample_query, sample_topk);
    idx->LoadTable(path);
    return idx;
}

// QSRPNormalLP::BuildIndex (src/impl/QSRPNormalLP.hpp:258-322, called at ReverseRankIndex.cpp:125-135):
// the QS table, then the SVD rotation and the per-user regression, which the reference loads from the
// files bsvdi and bri write and which are fitted here when the index is built.
inline std::unique_ptr<Index> BuildIndexQSRP(VectorMatrix &data_item, VectorMatrix &user, const char *dataset_name,
                                             const int &n_sample, const int &n_sample_query, const int &sample_topk,
                                             const int &n_thread, const char *index_dir, float sigma = 0.7f,
                                             int device = 0) {
    auto idx = BuildIndexQS(data_item, user, dataset_name, n_sample, n_sample_query, sample_topk, n_thread, index_dir,
                            device);
    idx->BuildQSRP(sigma);
    return idx;
}

}  // namespace rkr
