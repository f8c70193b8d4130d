// k_sort.hip -- every user's scores against every item, sorted in descending order.
//
// The full score table (ReadAll's index file, src/include/alg/DiskIndex/ReadAll.hpp:80-87, written by
// src/app/BuildScoreTableBatch.cpp:69-144) holds each user's n_item scores sorted with std::greater
// (src/include/score_computation/ComputeScoreTableBatch.hpp:83-89).  Only the values are kept, so the
// order of equal values is irrelevant and the sort need not be stable.
//
// One 1024-thread workgroup per row:
//   1. tiles of TILE values are sorted in LDS by a bitonic network (a partial last tile is padded with
//      -inf, which sorts to its end and is not written back);
//   2. sorted runs are merged pairwise -- run length TILE, 2*TILE, ... -- between the output row and
//      a scratch row.  Each thread finds where its share of a merged run starts by a binary search
//      along the merge path, then merges that share sequentially.  The tiles go to whichever of the
//      two rows makes the last pass land in the output row.
#include "rkr_internal.hpp"

namespace rkr {
namespace {

constexpr int NT = 1024;                 // threads per row
constexpr int TILE = 8192;               // values sorted in LDS at a time (32 KB)
constexpr int PER = TILE / 2 / NT;       // compare-exchange pairs per thread and stage

// Sorts s[0..TILE) in descending order.  Stage (k, j) compares i with i + j for
// the i whose bit j is clear; blocks of k elements alternate direction so that
// each merge step sees bitonic sequences, and the final k = TILE is descending.
__device__ inline void bitonic_desc(float *s) {
    for (int k = 2; k <= TILE; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                const int p = threadIdx.x + e * NT;                   // pair index < TILE / 2
                const int i = ((p & ~(j - 1)) << 1) | (p & (j - 1));  // lower element of the pair
                const float a = s[i], b = s[i + j];
                const bool desc = (i & k) == 0;
                if (desc ? (a < b) : (a > b)) {
                    s[i] = b;
                    s[i + j] = a;
                }
            }
            __syncthreads();
        }
    }
}

// Number of elements of A among the first d outputs of the descending merge of
// A[0..na) and B[0..nb), where A goes first on ties.
__device__ inline int merge_split(const float *A, int na, const float *B, int nb, int d) {
    int lo = max(0, d - nb), hi = min(d, na);
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A[mid] >= B[d - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// src and out may be the same rows: every value of src is read before the
// output row is written where it lies.
__global__ void __launch_bounds__(NT) k_sort_rows(const float *src, int64_t ld_src, int32_t n, float *out,
                                                  float *tmp, int64_t ld) {
    __shared__ float s[TILE];
    const int64_t row = blockIdx.x;
    const float *x = src + row * ld_src;
    int passes = 0;
    for (int64_t w = TILE; w < n; w <<= 1) ++passes;
    float *a = (passes & 1) ? tmp + row * ld : out + row * ld;  // sorted tiles
    float *b = (passes & 1) ? out + row * ld : tmp + row * ld;

    for (int t0 = 0; t0 < n; t0 += TILE) {
        const int m = min(TILE, n - t0);
        for (int i = threadIdx.x; i < TILE; i += NT) s[i] = i < m ? x[t0 + i] : -INFINITY;
        __syncthreads();
        bitonic_desc(s);
        for (int i = threadIdx.x; i < m; i += NT) a[t0 + i] = s[i];
        __syncthreads();
    }

    for (int w = TILE; w < n; w <<= 1) {
        __syncthreads();  // the runs written by the previous pass are visible to the workgroup
        for (int s0 = 0; s0 < n; s0 += 2 * w) {
            const int na = min(w, n - s0);
            const int nb = min(w, n - s0 - na);
            const int no = na + nb;
            const int per = (no + NT - 1) / NT;
            const int d = (int)threadIdx.x * per;
            if (d >= no) continue;
            const int cnt = min(per, no - d);
            const float *A = a + s0, *B = a + s0 + na;
            float *D = b + s0 + d;
            int i = merge_split(A, na, B, nb, d), j = d - i;
            float av = i < na ? A[i] : 0.0f, bv = j < nb ? B[j] : 0.0f;
            for (int o = 0; o < cnt; ++o) {
                if (j >= nb || (i < na && av >= bv)) {
                    D[o] = av;
                    if (++i < na) av = A[i];
                } else {
                    D[o] = bv;
                    if (++j < nb) bv = B[j];
                }
            }
        }
        float *t = a;
        a = b;
        b = t;
    }
}

}  // namespace

hipError_t launch_sort_rows(hipStream_t s, const float *src, int64_t ld_src, int64_t nb, int64_t n_item, float *out,
                            float *tmp, int64_t ld) {
    if (nb <= 0 || n_item <= 0) return hipSuccess;
    if (n_item > (int64_t)1 << 30 || ld < n_item || ld_src < n_item) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sort_rows, dim3((unsigned)nb), dim3(NT), 0, s, src, ld_src, (int32_t)n_item, out, tmp, ld);
    return hipGetLastError();
}

}  // namespace rkr
