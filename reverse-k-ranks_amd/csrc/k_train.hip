// k_train.hip -- the device half of QS sample-rank training.
//
// 1. Query-rank distribution (ComputeQueryRankByComputation,
//    src/app/build-index/QueryRankByComputation.hpp:122-242, with the rank of
//    QueryRankGPU::ComputeRankOfQuery, src/app/build-index/QueryRankGPU.hpp:63-92,225-240):
//    for every user u and sample item s, rank(u, s) = #{items whose score
//    beats <u, s>} -- the lower_bound with greater<float> in u's descending
//    row.  k_rank_hist_lds (row staged in LDS) or k_rank_hist counts
//    hist[s][rank] over the users of a batch.
// 2. The sample-rank DP of FindOptimalSampleFast
//    (src/include/alg/RankBoundRefinement/FindSampleRank/OptimalSampleRankFast.hpp:160-205):
//    k_transit fills the transit-cost triangle of QueryDistribution
//    (OptimalSampleRank.hpp:233-318) from per-rank prefix sums over the
//    queries, in O(1) per pair; k_dp_layer runs one layer of the DP.
#include "rkr_internal.hpp"

#include <climits>

namespace rkr {
namespace {

// One workgroup per user row; threads take the samples.
__global__ void __launch_bounds__(256) k_rank_hist(const float *__restrict__ sorted, int64_t ld, int32_t n_item,
                                                   const float *__restrict__ q, int64_t ldq, int32_t ns,
                                                   int32_t *__restrict__ hist) {
    const int64_t u = blockIdx.x;
    const float *row = sorted + u * ld;
    for (int s = threadIdx.x; s < ns; s += blockDim.x) {
        const float x = q[u * ldq + s];
        int lo = 0, hi = n_item;  // first position whose value is not > x
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (row[mid] > x) lo = mid + 1;
            else hi = mid;
        }
        atomicAdd(hist + (int64_t)s * (n_item + 1) + lo, 1);
    }
}

// The same count with the row staged in LDS first (n_item floats of dynamic
// LDS): each binary search then reads LDS instead of issuing its ~15 dependent
// requests to the caches.  Used when the row fits the workgroup's LDS.
__global__ void __launch_bounds__(512) k_rank_hist_lds(const float *__restrict__ sorted, int64_t ld, int32_t n_item,
                                                       const float *__restrict__ q, int64_t ldq, int32_t ns,
                                                       int32_t *__restrict__ hist) {
    extern __shared__ float srow[];
    const int64_t u = blockIdx.x;
    const float *row = sorted + u * ld;
    for (int i = threadIdx.x; i < n_item; i += blockDim.x) srow[i] = row[i];
    __syncthreads();
    for (int s = threadIdx.x; s < ns; s += blockDim.x) {
        const float x = q[u * ldq + s];
        int lo = 0, hi = n_item;  // first position whose value is not > x
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (srow[mid] > x) lo = mid + 1;
            else hi = mid;
        }
        atomicAdd(hist + (int64_t)s * (n_item + 1) + lo, 1);
    }
}

__device__ inline int64_t tri_off(int64_t lb) { return lb * (lb - 1) / 2; }

// T[tri_off(lb) + ub] = -End(ub) + Between(ub, lb) + End(lb) for ub < lb, where
// Between(ub, lb) = sum over queries q in [pos[ub], pos[lb]) of
// accu[q][rank_lb] - accu[q][rank_ub], read as differences of the prefix sums
// S_j(x) = sum_{q' < x} accu[q'][rank_j], which C keeps at the breakpoints:
// C[col[j'] * R + j] = S_j(pos[j']).  For one lb the threads read C's rows
// col[lb] and col[ub] at consecutive ub.
__global__ void __launch_bounds__(256) k_transit(const int64_t *__restrict__ C, int32_t Q,
                                                 const int32_t *__restrict__ pos, const int32_t *__restrict__ col,
                                                 const int64_t *__restrict__ end_cost, int32_t R,
                                                 int64_t *__restrict__ T) {
    const int64_t lb = blockIdx.x;
    if (lb < 1 || lb >= R) return;
    const int32_t pe = pos[lb];
    const int64_t *Ce = C + (int64_t)col[lb] * R;
    const int64_t el = end_cost[lb];
    int64_t *Tl = T + tri_off(lb);
    for (int64_t ub = threadIdx.x; ub < lb; ub += blockDim.x) {
        const int32_t pb = pos[ub];
        int64_t between = 0;
        if (pb < Q && pb < pe) {
            const int64_t *Cb = C + (int64_t)col[ub] * R;
            between = (Ce[lb] - Cb[lb]) - (Ce[ub] - Cb[ub]);
        }
        Tl[ub] = -end_cost[ub] + between + el;
    }
}

// One wavefront per rank r: the carried value dp_prev[r] wins ties, then the
// smallest p, as the reference's sequential loop with a strict comparison does.
__global__ void __launch_bounds__(256) k_dp_layer(const int64_t *__restrict__ T, const int64_t *__restrict__ dp_prev,
                                                  const int32_t *__restrict__ pos_prev, int32_t R,
                                                  int64_t *__restrict__ dp_cur, int32_t *__restrict__ pos_cur) {
    const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    if (r >= R) return;  // whole wavefronts leave together
    const int64_t *Tr = T + tri_off(r);
    long long best = LLONG_MAX;
    int bp = INT_MAX;
    for (int p = lane; p < r; p += 64) {
        const long long c = dp_prev[p] + Tr[p];
        if (c < best) {
            best = c;
            bp = p;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const long long oc = __shfl_xor(best, off, 64);
        const int op = __shfl_xor(bp, off, 64);
        if (oc < best || (oc == best && op < bp)) {
            best = oc;
            bp = op;
        }
    }
    if (lane == 0) {
        if (bp != INT_MAX && best < dp_prev[r]) {
            dp_cur[r] = best;
            pos_cur[r] = bp;
        } else {
            dp_cur[r] = dp_prev[r];
            pos_cur[r] = pos_prev[r];
        }
    }
}

}  // namespace

hipError_t launch_rank_hist(hipStream_t s, const float *sorted, int64_t ld, int64_t nb, int64_t n_item, const float *q,
                            int64_t ldq, int32_t ns, int32_t *hist) {
    if (nb <= 0 || ns <= 0) return hipSuccess;
    if (n_item <= 0 || n_item >= INT_MAX || nb > INT_MAX) return hipErrorInvalidValue;
    int dev = 0, lds_max = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    const size_t lds = sizeof(float) * (size_t)n_item;
    bool use_lds = lds <= (size_t)lds_max;
    if (use_lds && hipFuncSetAttribute((const void *)k_rank_hist_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds) != hipSuccess) {
        (void)hipGetLastError();  // a refused attribute: the row stays in global memory
        use_lds = false;
    }
    if (use_lds)
        hipLaunchKernelGGL(k_rank_hist_lds, dim3((unsigned)nb), dim3(512), lds, s, sorted, ld, (int32_t)n_item, q, ldq,
                           ns, hist);
    else
        hipLaunchKernelGGL(k_rank_hist, dim3((unsigned)nb), dim3(256), 0, s, sorted, ld, (int32_t)n_item, q, ldq, ns,
                           hist);
    return hipGetLastError();
}

hipError_t launch_transit(hipStream_t s, const int64_t *C, int32_t Q, const int32_t *pos, const int32_t *col,
                          const int64_t *end_cost, int32_t R, int64_t *T) {
    if (R < 2) return hipSuccess;
    hipLaunchKernelGGL(k_transit, dim3((unsigned)R), dim3(256), 0, s, C, Q, pos, col, end_cost, R, T);
    return hipGetLastError();
}

hipError_t launch_dp_layer(hipStream_t s, const int64_t *T, const int64_t *dp_prev, const int32_t *pos_prev, int32_t R,
                           int64_t *dp_cur, int32_t *pos_cur) {
    if (R <= 0) return hipSuccess;
    const int64_t waves = R;
    hipLaunchKernelGGL(k_dp_layer, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, T, dp_prev, pos_prev, R, dp_cur,
                       pos_cur);
    return hipGetLastError();
}

}  // namespace rkr
