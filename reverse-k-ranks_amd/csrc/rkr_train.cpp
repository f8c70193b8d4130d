// rkr_train.cpp -- QS sample-rank training: the host half.
//
// The reference trains the sampled ranks of its QS index in two binaries:
//   qdibc (QueryDistributionIndexByComputation) samples items, computes for
//     every sample item the distribution of its rank over the users
//     (rkr_query_rank_distribution, rkr_api.cpp) and sorts the samples by the
//     rank of their k-th best user (rkr_sort_kth_rank);
//   fsr --method_name QS (FindSampleRank with FindOptimalSampleFast) picks the
//     n_sample ranks that minimise the summed number of users whose rank falls
//     in the sample bucket holding a query's k-th rank (rkr_find_optimal_sample).
// The DP is integer work with a strict-improvement rule; the device runs its
// layers (k_train.hip) and ties resolve as in the reference's loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "rkr_internal.hpp"

using namespace rkr;

struct rkr_ctx : public rkr::Ctx {};

namespace {

// Fills rank_set up to n_sample ranks the way FindOptimalSampleFast does when
// the DP cannot supply n_sample distinct ranks
// (OptimalSampleRankFast.hpp:121-141 and 245-266): first every `delta` ranks
// below last_sample (stepping by one past ranks already taken), then the
// smallest ranks not yet taken.
void fill_ranks(std::set<int> &rank_set, int n_sample, int last_sample, int n_item) {
    const int n_remain = n_sample - (int)rank_set.size();
    if (n_remain <= 0) return;
    const double delta = last_sample * 1.0 / n_remain;
    double insert_rank = 0;
    while ((int)rank_set.size() < n_sample && (int)insert_rank < last_sample) {
        if (rank_set.find((int)insert_rank) == rank_set.end()) {
            rank_set.insert((int)insert_rank);
            insert_rank += delta;
        } else {
            insert_rank += 1;
        }
    }
    for (int r = 0; (int)rank_set.size() < n_sample && r < n_item; ++r) rank_set.insert(r);
}

template <class T>
struct DevArr {
    Ctx *c;
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t count, const char *what) {
        n = count;
        return dev_alloc_t(c, &p, std::max<size_t>(count, 1), what);
    }
    ~DevArr() { dev_free(c, p, std::max<size_t>(n, 1) * sizeof(T)); }
};

}  // namespace

extern "C" {

int rkr_sample_items(int64_t n_item, int32_t n_sample_item, int32_t *out_item_ids) {
    // SampleItem (src/app/build-index/QueryRankBySample.hpp:20-37): the first
    // n_sample_item entries of 0..n_item-1 after std::shuffle with std::mt19937(0)
    if (!out_item_ids || n_item <= 0 || n_item > INT_MAX || n_sample_item <= 0 || n_sample_item > n_item)
        return RKR_ERR_INVALID;
    std::vector<int> idx((size_t)n_item);
    std::iota(idx.begin(), idx.end(), 0);
    std::mt19937 eng(0);
    std::shuffle(idx.begin(), idx.end(), eng);
    std::memcpy(out_item_ids, idx.data(), sizeof(int32_t) * (size_t)n_sample_item);
    return RKR_OK;
}

int rkr_sort_kth_rank(const int32_t *accu, int64_t n_item, int32_t n_sample_item, int32_t sample_topk,
                      int32_t *out_sort_kth_rank, int32_t *out_sort_sample_ids) {
    // ComputeSortKthRank (src/app/build-index/QueryRankBySample.hpp:103-134)
    if (!accu || !out_sort_kth_rank || !out_sort_sample_ids || n_item <= 0 || n_item >= INT_MAX ||
        n_sample_item <= 0 || sample_topk < 0)
        return RKR_ERR_INVALID;
    const int64_t w = n_item + 1;
    std::vector<int> kth((size_t)n_sample_item);
    for (int32_t s = 0; s < n_sample_item; ++s) {
        const int32_t *row = accu + (int64_t)s * w;
        kth[s] = (int)(std::lower_bound(row, row + w, sample_topk) - row);
    }
    // std::sort, not a stable sort: samples with equal k-th ranks keep the
    // order libstdc++'s introsort gives the reference
    std::vector<int> ids((size_t)n_sample_item);
    std::iota(ids.begin(), ids.end(), 0);
    std::sort(ids.begin(), ids.end(), [&kth](int i1, int i2) { return kth[i1] < kth[i2]; });
    for (int32_t s = 0; s < n_sample_item; ++s) {
        out_sort_sample_ids[s] = ids[s];
        out_sort_kth_rank[s] = kth[ids[s]];
    }
    return RKR_OK;
}

int rkr_find_optimal_sample(rkr_ctx *c, int64_t n_user, int64_t n_item, int32_t n_sample_query,
                            const int32_t *sort_kth_rank, const int32_t *accu, int32_t method, int32_t n_sample,
                            int32_t *out_ranks, int64_t *train_io_cost) {
    // FindOptimalSampleFast (OptimalSampleRankFast.hpp:102-305) over the costs
    // of QueryDistribution (OptimalSampleRank.hpp:17-330)
    if (!c) return RKR_ERR_INVALID;
    if (!sort_kth_rank || !accu || !out_ranks || n_user <= 0 || n_item <= 1 || n_item >= INT_MAX - 1 ||
        n_sample_query <= 0 || n_sample <= 0 || n_sample >= n_item ||
        (method != RKR_SAMPLE_OPTIMAL_PART && method != RKR_SAMPLE_OPTIMAL_UNIFORM))
        return set_err(c, RKR_ERR_INVALID, "find_optimal_sample: bad arguments");
    const int64_t Q = n_sample_query, W = n_item + 1;
    for (int64_t q = 0; q < Q; ++q)
        if (sort_kth_rank[q] < 0 || sort_kth_rank[q] > n_item || (q && sort_kth_rank[q] < sort_kth_rank[q - 1]))
            return set_err(c, RKR_ERR_INVALID, "find_optimal_sample: sort_kth_rank must be ascending in [0, n_item]");

    // candidate ranks (QueryDistribution constructor, OptimalSampleRank.hpp:48-91)
    std::vector<int> rank;
    if (method == RKR_SAMPLE_OPTIMAL_PART) {
        std::set<int> rs;
        for (int64_t q = 0; q < Q; ++q) {
            rs.insert(sort_kth_rank[q]);
            if (sort_kth_rank[q] != 0) rs.insert(sort_kth_rank[q] - 1);
        }
        rank.assign(rs.begin(), rs.end());
    } else {
        const double delta = (double)(n_item - 1) / (double)Q;
        rank.resize((size_t)Q);
        for (int64_t s = 0; s < Q; ++s) rank[s] = (int)std::floor(s * delta);
        std::sort(rank.begin(), rank.end());
    }
    const int64_t R = (int64_t)rank.size();
    if (rank.back() >= n_item)
        return set_err(c, RKR_ERR_INVALID, "find_optimal_sample: a k-th rank equals n_item (sample_topk above the user count?)");
    if (train_io_cost) *train_io_cost = 0;

    if (R <= n_sample) {
        // take every candidate, then fill (OptimalSampleRankFast.hpp:114-154)
        std::set<int> rs(rank.begin(), rank.end());
        fill_ranks(rs, n_sample, n_sample_query < 500 ? (int)n_item : rank.back(), (int)n_item);
        if ((int64_t)rs.size() != n_sample) return set_err(c, RKR_ERR_INVALID, "find_optimal_sample: too few ranks");
        std::copy(rs.begin(), rs.end(), out_ranks);
        return RKR_OK;
    }

    // pos[j]: first query whose k-th rank exceeds rank[j] (BuildRankIDMap,
    // OptimalSampleRank.hpp:206-231: "this last" and "next first" coincide for
    // integer ranks).  C[j][q]: prefix sums of accu[.][rank[j]] over queries.
    std::vector<int32_t> pos((size_t)R);
    for (int64_t j = 0; j < R; ++j)
        pos[j] = (int32_t)(std::upper_bound(sort_kth_rank, sort_kth_rank + Q, rank[j]) - sort_kth_rank);
    std::vector<int64_t> C((size_t)(R * (Q + 1)));
    for (int64_t j = 0; j < R; ++j) C[(size_t)(j * (Q + 1))] = 0;
    for (int64_t q = 0; q < Q; ++q) {
        const int32_t *row = accu + q * W;
        for (int64_t j = 0; j < R; ++j) C[(size_t)(j * (Q + 1) + q + 1)] = C[(size_t)(j * (Q + 1) + q)] + row[rank[j]];
    }
    std::vector<int64_t> end_cost((size_t)R), dp0((size_t)R);
    std::vector<int32_t> pos0((size_t)R);
    for (int64_t j = 0; j < R; ++j) {
        const int64_t *Cj = &C[(size_t)(j * (Q + 1))];
        const int64_t p = pos[j];
        // End: Upsilon([rank+1, n_item]); Begin: Upsilon([0, rank]) (OptimalSampleRank.hpp:270-308)
        end_cost[j] = (Q - p) * n_user - (Cj[Q] - Cj[p]);
        dp0[j] = Cj[p] + end_cost[j];
        pos0[j] = (int32_t)j;
    }

    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    DevArr<int64_t> dC{c}, dEnd{c}, dT{c}, dDp{c};
    DevArr<int32_t> dPos{c}, dPath{c};
    int rc;
    if ((rc = dC.alloc(C.size(), "training prefix sums")) || (rc = dEnd.alloc((size_t)R, "training costs")) ||
        (rc = dT.alloc((size_t)(R * (R - 1) / 2), "transit costs")) || (rc = dDp.alloc((size_t)(2 * R), "DP row")) ||
        (rc = dPos.alloc((size_t)R, "query positions")) ||
        (rc = dPath.alloc((size_t)n_sample * R, "DP positions")))
        return rc;
    hipStream_t s = c->stream;
    RKR_HIP(c, hipMemcpyAsync(dC.p, C.data(), sizeof(int64_t) * C.size(), hipMemcpyHostToDevice, s));
    RKR_HIP(c, hipMemcpyAsync(dEnd.p, end_cost.data(), sizeof(int64_t) * R, hipMemcpyHostToDevice, s));
    RKR_HIP(c, hipMemcpyAsync(dPos.p, pos.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, s));
    RKR_HIP(c, hipMemcpyAsync(dDp.p, dp0.data(), sizeof(int64_t) * R, hipMemcpyHostToDevice, s));
    RKR_HIP(c, hipMemcpyAsync(dPath.p, pos0.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, s));
    RKR_HIP(c, launch_transit(s, dC.p, (int32_t)Q, dPos.p, dEnd.p, (int32_t)R, dT.p));
    for (int32_t l = 1; l < n_sample; ++l)
        RKR_HIP(c, launch_dp_layer(s, dT.p, dDp.p + ((l - 1) & 1) * R, dPath.p + (int64_t)(l - 1) * R, (int32_t)R,
                                   dDp.p + (l & 1) * R, dPath.p + (int64_t)l * R));
    std::vector<int64_t> last((size_t)R);
    std::vector<int32_t> path((size_t)n_sample * R);
    RKR_HIP(c, hipMemcpyAsync(last.data(), dDp.p + ((n_sample - 1) & 1) * R, sizeof(int64_t) * R,
                              hipMemcpyDeviceToHost, s));
    RKR_HIP(c, hipMemcpyAsync(path.data(), dPath.p, sizeof(int32_t) * path.size(), hipMemcpyDeviceToHost, s));
    RKR_HIP(c, hipStreamSynchronize(s));

    // smallest final cost, first index on ties (unsigned compare as the reference's)
    uint64_t min_cost = UINT64_MAX;
    int64_t idx = -1;
    for (int64_t j = 0; j < R; ++j)
        if ((uint64_t)last[j] < min_cost) {
            min_cost = (uint64_t)last[j];
            idx = j;
        }
    if (idx < 0) return set_err(c, RKR_ERR_INVALID, "find_optimal_sample: no finite cost");
    if (train_io_cost) *train_io_cost = (int64_t)min_cost;
    for (int32_t l = n_sample - 1; l >= 0; --l) {
        out_ranks[l] = rank[idx];
        idx = path[(size_t)l * R + idx];
    }
    // repeated ranks (the DP kept a rank for several samples): fill as the reference does
    std::set<int> rs(out_ranks, out_ranks + n_sample);
    if ((int64_t)rs.size() != n_sample) {
        fill_ranks(rs, n_sample, rank.back(), (int)n_item);
        if ((int64_t)rs.size() != n_sample) return set_err(c, RKR_ERR_INVALID, "find_optimal_sample: too few ranks");
        std::copy(rs.begin(), rs.end(), out_ranks);
    }
    return RKR_OK;
}

}  // extern "C"
