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
#include <thread>
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
    // integer ranks).  The costs read the prefix sums
    // S_j(x) = sum_{q < x} accu[q][rank[j]] only at x = 0, x = Q and x = pos[j'],
    // so they are kept at those P breakpoints cols[0..P) only, one row of R per
    // breakpoint: C[c * R + j] = S_j(cols[c]), and col[j] is pos[j]'s index in cols.
    std::vector<int32_t> pos((size_t)R), col((size_t)R), cols;
    for (int64_t j = 0; j < R; ++j)
        pos[j] = (int32_t)(std::upper_bound(sort_kth_rank, sort_kth_rank + Q, rank[j]) - sort_kth_rank);
    cols.reserve((size_t)R + 2);
    cols.push_back(0);
    for (int64_t j = 0; j < R; ++j)  // pos ascends with the ranks
        if (pos[j] != cols.back()) cols.push_back(pos[j]);
    if (cols.back() != (int32_t)Q) cols.push_back((int32_t)Q);
    const int64_t P = (int64_t)cols.size();
    for (int64_t j = 0; j < R; ++j) col[j] = (int32_t)(std::lower_bound(cols.begin(), cols.end(), pos[j]) - cols.begin());
    std::vector<int64_t> C((size_t)(P * R)), run((size_t)R, 0);
    // one pass over the queries in order per range of ranks; ranges are disjoint
    // columns of C, so host threads take them without sharing anything they write
    auto sums = [&](int64_t j0, int64_t j1) {
        if (j0 >= j1) return;
        int64_t ci = 0;
        for (int64_t q = 0;; ++q) {
            for (; ci < P && cols[ci] == q; ++ci) std::copy(&run[j0], &run[j0] + (j1 - j0), &C[(size_t)(ci * R + j0)]);
            if (q == Q) break;
            const int32_t *row = accu + q * W;
            for (int64_t j = j0; j < j1; ++j) run[j] += row[rank[j]];
        }
    };
    const int64_t nt = std::min<int64_t>({(int64_t)std::max(1u, std::thread::hardware_concurrency()), 16,
                                          std::max<int64_t>(1, Q * R >> 14), R});
    std::vector<std::thread> th;
    int64_t launched = 1;
    try {
        th.reserve((size_t)nt);
        for (; launched < nt; ++launched) th.emplace_back(sums, R * launched / nt, R * (launched + 1) / nt);
    } catch (...) {  // a range without a thread runs on this one
    }
    sums(0, R / nt);
    for (int64_t t = launched; t < nt; ++t) sums(R * t / nt, R * (t + 1) / nt);
    for (auto &x : th) x.join();
    std::vector<int64_t> end_cost((size_t)R), dp0((size_t)R);
    std::vector<int32_t> pos0((size_t)R);
    const int64_t *CQ = &C[(size_t)((P - 1) * R)];  // S_j(Q)
    for (int64_t j = 0; j < R; ++j) {
        const int64_t p = pos[j], Sp = C[(size_t)((int64_t)col[j] * R + j)];
        // End: Upsilon([rank+1, n_item]); Begin: Upsilon([0, rank]) (OptimalSampleRank.hpp:270-308)
        end_cost[j] = (Q - p) * n_user - (CQ[j] - Sp);
        dp0[j] = Sp + end_cost[j];
        pos0[j] = (int32_t)j;
    }

    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    DevArr<int64_t> dC{c}, dEnd{c}, dT{c}, dDp{c};
    DevArr<int32_t> dPos{c}, dPath{c};
    int rc;
    if ((rc = dC.alloc(C.size(), "training prefix sums")) || (rc = dEnd.alloc((size_t)R, "training costs")) ||
        (rc = dT.alloc((size_t)(R * (R - 1) / 2), "transit costs")) || (rc = dDp.alloc((size_t)(2 * R), "DP row")) ||
        (rc = dPos.alloc((size_t)(2 * R), "query positions")) ||
        (rc = dPath.alloc((size_t)n_sample * R, "DP positions")))
        return rc;
    hipStream_t s = c->stream;
    RKR_HIP(c, hipMemcpyAsync(dC.p, C.data(), sizeof(int64_t) * C.size(), hipMemcpyHostToDevice, s));
    RKR_HIP(c, hipMemcpyAsync(dEnd.p, end_cost.data(), sizeof(int64_t) * R, hipMemcpyHostToDevice, s));
    RKR_HIP(c, hipMemcpyAsync(dPos.p, pos.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, s));
    RKR_HIP(c, hipMemcpyAsync(dPos.p + R, col.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, s));
    RKR_HIP(c, hipMemcpyAsync(dDp.p, dp0.data(), sizeof(int64_t) * R, hipMemcpyHostToDevice, s));
    RKR_HIP(c, hipMemcpyAsync(dPath.p, pos0.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, s));
    RKR_HIP(c, launch_transit(s, dC.p, (int32_t)Q, dPos.p, dPos.p + R, dEnd.p, (int32_t)R, dT.p));
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
