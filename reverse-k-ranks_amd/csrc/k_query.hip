// k_query.hip -- query-side kernels between the <u,q> scores and the
// exact-rank refinement: table binary search, per-query k-th bound,
// candidate compaction and top-k selection.
//
// Bucket convention (SampleSearch::CoarseBinarySearch,
// src/include/alg/RankBoundRefinement/SampleSearch.hpp:91-130): T[u] is
// non-increasing and bucket(q) = #{j : T[u][j] >= q}.  Rank bounds are
// functions of the bucket alone, so per-query k-th bounds come from a
// (S+1)-bin histogram of buckets instead of the reference's two max-heaps
// (PruneCandidateByBound.hpp:35-77).
//
// Exact output (RKR_OUTPUT_EXACT): with b_lo = #{T > q - eps} and
// b_hi = #{T > q + eps} (eps bounds |T - exact order statistic|, 0 for tables
// built with the score rule),
//   rank <= lbs = (b_lo == S ? I : r[b_lo])       (no item at or below sample
//                                                   position r[b_lo] beats q),
//   rank >= ubs = (b_hi == 0 ? 0 : r[b_hi - 1] + 1) (items up to r[b_hi-1] do),
// which hold under score ties too (q equal to a table entry: the reference's
// #{T >= q} bucket is also valid but looser).  beta = bucket of the k-th
// smallest lbs; a user can be in the answer only if ubs <= r[beta], i.e.
// b_hi <= beta.
//
// Reference output (RKR_OUTPUT_REFERENCE): lb/ub and the result / prune /
// refine classification of PruneCandidateByBound.hpp:83-150, and the
// US.hpp:173-186 answer layout.
#include "rkr_internal.hpp"

namespace rkr {
namespace {

constexpr int NT = 1024;
constexpr int NW = NT / 64;
constexpr int SMAXQ = kMaxSample + 1;

enum { MODE_EXACT = 0, MODE_REF_EQUAL = 1, MODE_REF_SPLIT = 2, MODE_REF_ALL = 3 };

__device__ __forceinline__ int32_t bucket_ge(const float *__restrict__ row, int32_t S, float q) {
    int32_t lo = 0, hi = S;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (row[mid] >= q) lo = mid + 1; else hi = mid;
    }
    return lo;
}
// #{T > q} when it is known to lie in [lo, hi] (T[0..lo) > q >= T[hi..S)).
__device__ __forceinline__ int32_t bucket_gt_in(const float *__restrict__ row, int32_t lo, int32_t hi, float q) {
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (row[mid] > q) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ int32_t bucket_gt(const float *__restrict__ row, int32_t S, float q) {
    int32_t lo = 0, hi = S;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (row[mid] > q) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Bucket pair of one (user, query).  Reference output keeps the reference's
// bucket #{T >= q} (SampleSearch.hpp:99-102) in both halves.
__device__ __forceinline__ void bucket_pair(const float *row, int32_t S, float q, float eps, int ref_mode,
                                            uint32_t &lo, uint32_t &hi) {
    if (ref_mode) {
        lo = hi = (uint32_t)bucket_ge(row, S, q);
    } else if (eps > 0.0f) {
        lo = (uint32_t)bucket_gt(row, S, __double2float_rd((double)q - (double)eps));
        hi = (uint32_t)bucket_gt(row, S, __double2float_ru((double)q + (double)eps));
    } else {
        lo = hi = (uint32_t)bucket_gt(row, S, q);
    }
}

// grid: (user blocks, queries).  bk[b * U + u] = b_lo | b_hi << 16.
__global__ void __launch_bounds__(256) k_bounds(const float *__restrict__ qip, const float *__restrict__ table,
                                                int64_t n_user, int32_t S, float eps, int ref_mode,
                                                uint32_t *__restrict__ bk) {
    const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t b = blockIdx.y;
    if (u >= n_user) return;
    const float q = qip[b * n_user + u];
    const float *row = table + u * (int64_t)S;
    uint32_t lo, hi;
    bucket_pair(row, S, q, eps, ref_mode, lo, hi);
    bk[b * n_user + u] = lo | (hi << 16);
}

// Batched form: a workgroup stages ut consecutive table rows in LDS and runs
// all nq searches of those users from there (each row is read from HBM once
// per batch instead of once per query).  Thread t serves user t % ut.  Rows
// sit at a stride of S + 1 floats, so the ut searches of one step, which
// probe the same index of different rows, fall in different LDS banks.
__global__ void __launch_bounds__(256) k_bounds_lds(const float *__restrict__ qip, const float *__restrict__ table,
                                                    int64_t n_user, int32_t S, int32_t nq, int32_t ut, float eps,
                                                    int ref_mode, uint32_t *__restrict__ bk) {
    extern __shared__ float rows[];  // ut x (S + 1)
    const int64_t u0 = (int64_t)blockIdx.x * ut;
    const int64_t nu = min((int64_t)ut, n_user - u0);
    stage_rows(rows, table + u0 * S, nu, S, S);
    __syncthreads();
    const int ul = threadIdx.x % ut;
    const int lanes_q = 256 / ut;
    if (ul >= nu) return;
    const int64_t u = u0 + ul;
    const float *row = rows + (int64_t)ul * (S + 1);
    for (int32_t b = threadIdx.x / ut; b < nq; b += lanes_q) {
        uint32_t lo, hi;
        bucket_pair(row, S, qip[(int64_t)b * n_user + u], eps, ref_mode, lo, hi);
        bk[(int64_t)b * n_user + u] = lo | (hi << 16);
    }
}

// Reference bounds of a bucket (SampleSearch.hpp:107-125).
__device__ __forceinline__ int32_t ref_lb(int32_t b, int32_t S, const int32_t *r, int64_t I) {
    return b == S ? (int32_t)I : r[b];
}
__device__ __forceinline__ int32_t ref_ub(int32_t b, int32_t S, const int32_t *r, int64_t I) {
    if (b == S) return r[b - 1];
    if (b == 0) return 0;
    if ((uint32_t)r[b] - (uint32_t)r[b - 1] <= 1u) return r[b];
    return r[b - 1];
}

__device__ uint32_t wg_excl_scan(uint32_t v, uint32_t *red, uint32_t *total) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) red[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int i = 0; i < NW; ++i) { const uint32_t t = red[i]; red[i] = s; s += t; }
        red[NW] = s;
    }
    __syncthreads();
    const uint32_t r = red[w] + x - v;
    *total = red[NW];
    __syncthreads();
    return r;
}

// grid (chunks, queries): bucket histograms of the chunk's users, added into
// the query's global histograms ghist[b] = [h_lo | h_hi], (S + 1) bins each.
// same != 0: b_lo == b_hi for every user (exact output, eps = 0), so only
// h_lo is built and the selection reads it for both.
__global__ void __launch_bounds__(NT) k_hist_chunks(const uint32_t *__restrict__ bk, int64_t n_user, int32_t S,
                                                    int same, int64_t cu, uint32_t *__restrict__ ghist) {
    __shared__ uint32_t h_lo[SMAXQ + 1];
    __shared__ uint32_t h_hi[SMAXQ + 1];
    const int64_t b = blockIdx.y;
    const int64_t u0 = (int64_t)blockIdx.x * cu, u1 = min(n_user, u0 + cu);
    for (int i = threadIdx.x; i <= S; i += NT) { h_lo[i] = 0; h_hi[i] = 0; }
    __syncthreads();
    const uint32_t *v = bk + b * n_user;
    if (same) {
        for (int64_t u = u0 + threadIdx.x; u < u1; u += NT) atomicAdd(&h_lo[v[u] & 0xFFFFu], 1u);
    } else {
        for (int64_t u = u0 + threadIdx.x; u < u1; u += NT) {
            const uint32_t x = v[u];
            atomicAdd(&h_lo[x & 0xFFFFu], 1u);
            atomicAdd(&h_hi[x >> 16], 1u);
        }
    }
    __syncthreads();
    uint32_t *g = ghist + b * 2 * (int64_t)(S + 1);
    for (int i = threadIdx.x; i <= S; i += NT) {
        if (h_lo[i]) atomicAdd(&g[i], h_lo[i]);
        if (!same && h_hi[i]) atomicAdd(&g[S + 1 + i], h_hi[i]);
    }
}

// One workgroup per query: k-th bound and class counts from the histograms.
__global__ void __launch_bounds__(NT) k_hist_kth(const uint32_t *__restrict__ ghist, int32_t S, int same,
                                                 const int32_t *__restrict__ ranks, int64_t n_item, int output_mode,
                                                 QueryInfo *__restrict__ qi) {
    __shared__ uint32_t h_lo[SMAXQ + 1];
    __shared__ uint32_t h_hi[SMAXQ + 1];
    const int64_t b = blockIdx.x;
    const uint32_t *g = ghist + b * 2 * (int64_t)(S + 1);
    for (int i = threadIdx.x; i <= S; i += NT) {
        const uint32_t lo = g[i];
        h_lo[i] = lo;
        h_hi[i] = same ? lo : g[S + 1 + i];
    }
    __syncthreads();
    if (output_mode == RKR_OUTPUT_EXACT) {
        // beta = first bin where the cumulative h_lo reaches k, and the
        // candidate count sum_{j <= beta} h_hi[j]: block scans over runs of
        // ceil((S + 1) / NT) bins per thread (counts fit in 32 bits)
        __shared__ uint32_t red[NW + 1];
        __shared__ int32_t s_beta;
        const uint32_t k = (uint32_t)qi[b].topk;
        const int32_t per = (S + 1 + NT - 1) / NT;
        const int32_t j0 = threadIdx.x * per, j1 = min(S + 1, j0 + per);
        uint32_t mine = 0;
        for (int32_t j = j0; j < j1; ++j) mine += h_lo[j];
        uint32_t tot;
        uint32_t before = wg_excl_scan(mine, red, &tot);
        if (threadIdx.x == 0) s_beta = S;
        __syncthreads();
        if (before < k && before + mine >= k) {
            uint32_t cum = before;
            for (int32_t j = j0; j < j1; ++j) {
                cum += h_lo[j];
                if (cum >= k) { s_beta = j; break; }
            }
        }
        __syncthreads();
        const int32_t beta = s_beta;
        uint32_t nh = 0;
        for (int32_t j = j0; j < min(j1, beta + 1); ++j) nh += h_hi[j];
        (void)wg_excl_scan(nh, red, &tot);
        if (threadIdx.x == 0) {
            QueryInfo q = qi[b];
            q.beta = beta;
            q.n_refine = 0;
            q.fill = 0;
            q.mode = MODE_EXACT;
            q.gamma = beta;
            q.n_cand = (int32_t)tot;
            q.n_settled = 0;
            qi[b] = q;
        }
        return;
    }
    if (threadIdx.x != 0) return;
    QueryInfo q = qi[b];
    const int64_t k = q.topk;
    int32_t beta = S;
    int64_t cum = 0;
    for (int32_t j = 0; j <= S; ++j) {
        cum += h_lo[j];
        if (cum >= k) { beta = j; break; }
    }
    q.beta = beta;
    q.n_refine = 0;
    q.fill = 0;
    {
        const int32_t Lk = ref_lb(beta, S, ranks, n_item), Uk = ref_ub(beta, S, ranks, n_item);
        int64_t n_res = 0, n_ref = 0;
        for (int32_t j = 0; j <= S; ++j) {
            const int64_t h = h_lo[j];
            if (!h) continue;
            const int32_t lb = ref_lb(j, S, ranks, n_item), ub = ref_ub(j, S, ranks, n_item);
            if (Lk == Uk) {
                if (ub <= Lk && lb <= Lk) n_res += h;
            } else if (lb <= Uk) n_res += h;
            else if (Lk <= ub) continue;
            else n_ref += h;
        }
        if (Lk == Uk) {
            q.mode = MODE_REF_EQUAL;  // PruneCandidateByBound.hpp:83-111
            q.n_settled = (int32_t)(n_res < k ? n_res : k);
            q.n_cand = 0;
        } else if (n_res + n_ref == k) {
            q.mode = MODE_REF_ALL;    // :139-150: every live user is a result
            q.n_settled = (int32_t)k;
            q.n_cand = 0;
        } else {
            q.mode = MODE_REF_SPLIT;  // :114-136
            q.n_settled = (int32_t)(n_res < k ? n_res : k);
            q.n_cand = (int32_t)n_ref;
        }
        q.gamma = Uk;
        q.pad = Lk;
    }
    qi[b] = q;
}

// Exclusive offsets of each query's [settled | candidates] slice.
__global__ void __launch_bounds__(NT) k_offsets(QueryInfo *__restrict__ qi, int32_t nq, int32_t *__restrict__ refine_total) {
    __shared__ uint32_t red[NW + 1];
    uint32_t base = 0;
    for (int32_t c0 = 0; c0 < nq; c0 += NT) {
        const int32_t i = c0 + threadIdx.x;
        const uint32_t v = i < nq ? (uint32_t)(qi[i].n_settled + qi[i].n_cand) : 0u;
        uint32_t tot;
        const uint32_t e = wg_excl_scan(v, red, &tot);
        if (i < nq) qi[i].cand_off = (int64_t)base + e;
        base += tot;
    }
    if (threadIdx.x == 0) *refine_total = 0;
}

// One workgroup per query, users in ascending order: writes the settled users
// (reference output, at most k, with rank = bound) and the candidates (with
// their rank when the bounds fix it, else -1 plus a refine row).
__global__ void __launch_bounds__(NT) k_compact(const uint32_t *__restrict__ bk, const float *__restrict__ qip,
                                                int64_t n_user, int32_t S, const int32_t *__restrict__ ranks,
                                                int64_t n_item, QueryInfo *__restrict__ qi,
                                                int32_t *__restrict__ cand_user, int32_t *__restrict__ cand_rank,
                                                int32_t *__restrict__ refine_total, int32_t *__restrict__ refine_cand,
                                                int32_t *__restrict__ refine_user, int32_t *__restrict__ refine_query,
                                                float *__restrict__ refine_thr) {
    __shared__ uint32_t red[NW + 1];
    __shared__ int32_t n_refine;
    const int64_t b = blockIdx.x;
    const QueryInfo q = qi[b];
    const uint32_t *v = bk + b * n_user;
    const float *qb = qip + b * n_user;
    uint32_t n_set = 0, n_can = 0;  // running counts
    int32_t my_refine = 0;
    if (threadIdx.x == 0) n_refine = 0;
    const int64_t set_base = q.cand_off;
    const int64_t can_base = q.cand_off + q.n_settled;
    const int32_t Lk = q.pad, Uk = q.gamma;
    for (int64_t u0 = 0; u0 < n_user; u0 += NT) {
        const int64_t u = u0 + threadIdx.x;
        bool settled = false, cand = false;
        int32_t rank = -1;
        if (u < n_user) {
            const uint32_t x = v[u];
            const int32_t blo = (int32_t)(x & 0xFFFFu), bhi = (int32_t)(x >> 16);
            if (q.mode == MODE_EXACT) {
                if (bhi <= q.beta) {
                    cand = true;
                    const int32_t lbs = blo == S ? (int32_t)n_item : ranks[blo];
                    const int32_t ubs = bhi == 0 ? 0 : ranks[bhi - 1] + 1;
                    if (lbs == ubs) rank = lbs;
                }
            } else {
                const int32_t lb = ref_lb(blo, S, ranks, n_item), ub = ref_ub(blo, S, ranks, n_item);
                if (q.mode == MODE_REF_EQUAL) {
                    settled = (ub <= Lk && lb <= Lk);
                } else if (q.mode == MODE_REF_ALL) {
                    settled = !(lb > Uk && Lk <= ub);
                } else {
                    if (lb <= Uk) settled = true;
                    else if (!(Lk <= ub)) { cand = true; if (lb == ub) rank = lb; }
                }
                if (settled) rank = lb;
            }
        }
        uint32_t tot;
        const uint32_t e = wg_excl_scan((settled ? 1u : 0u) | (cand ? 0x10000u : 0u), red, &tot);
        const uint32_t my_set = n_set + (e & 0xFFFFu), my_can = n_can + (e >> 16);
        if (settled && my_set < (uint32_t)q.n_settled) {
            cand_user[set_base + my_set] = (int32_t)u;
            cand_rank[set_base + my_set] = rank;
        }
        if (cand) {
            const int64_t slot = can_base + my_can;
            cand_user[slot] = (int32_t)u;
            cand_rank[slot] = rank;
            if (rank < 0) {
                ++my_refine;
                const int32_t r = atomicAdd(refine_total, 1);
                refine_cand[r] = (int32_t)slot;
                refine_user[r] = (int32_t)u;
                refine_query[r] = (int32_t)b;
                refine_thr[r] = qb[u];
            }
        }
        n_set += tot & 0xFFFFu;
        n_can += tot >> 16;
    }
    if (my_refine) atomicAdd(&n_refine, my_refine);
    __syncthreads();
    if (threadIdx.x == 0) qi[b].n_refine = n_refine;
}

// Exact output, grid (chunks, queries): the candidates b_hi <= beta of the
// chunk go to ranges reserved with one atomic per 1024 users on the query's
// cursor, so the slice is in no fixed order (the top-k selection orders by
// (rank, userID)); refine rows likewise.
__global__ void __launch_bounds__(NT) k_compact_chunks(const uint32_t *__restrict__ bk, const float *__restrict__ qip,
                                                       int64_t n_user, int32_t S, const int32_t *__restrict__ ranks,
                                                       int64_t n_item, int64_t cu, QueryInfo *__restrict__ qi,
                                                       int32_t *__restrict__ cand_user, int32_t *__restrict__ cand_rank,
                                                       int32_t *__restrict__ refine_total,
                                                       int32_t *__restrict__ refine_cand,
                                                       int32_t *__restrict__ refine_user,
                                                       int32_t *__restrict__ refine_query,
                                                       float *__restrict__ refine_thr) {
    __shared__ uint32_t red[NW + 1];
    __shared__ int32_t s_can, s_ref;
    const int64_t b = blockIdx.y;
    const int32_t beta = qi[b].beta;
    const int64_t off = qi[b].cand_off;
    const uint32_t *v = bk + b * n_user;
    const float *qb = qip + b * n_user;
    const int64_t u0 = (int64_t)blockIdx.x * cu, u1 = min(n_user, u0 + cu);
    for (int64_t ub0 = u0; ub0 < u1; ub0 += NT) {
        const int64_t u = ub0 + threadIdx.x;
        bool cand = false;
        int32_t rank = -1;
        if (u < u1) {
            const uint32_t x = v[u];
            const int32_t blo = (int32_t)(x & 0xFFFFu), bhi = (int32_t)(x >> 16);
            if (bhi <= beta) {
                cand = true;
                const int32_t lbs = blo == S ? (int32_t)n_item : ranks[blo];
                const int32_t ubs = bhi == 0 ? 0 : ranks[bhi - 1] + 1;
                if (lbs == ubs) rank = lbs;
            }
        }
        const bool refine = cand && rank < 0;
        uint32_t tot;
        const uint32_t e = wg_excl_scan((cand ? 1u : 0u) | (refine ? 0x10000u : 0u), red, &tot);
        if (tot == 0) continue;
        if (threadIdx.x == 0) {
            s_can = atomicAdd(&qi[b].fill, (int32_t)(tot & 0xFFFFu));
            if (tot >> 16) {
                s_ref = atomicAdd(refine_total, (int32_t)(tot >> 16));
                atomicAdd(&qi[b].n_refine, (int32_t)(tot >> 16));
            }
        }
        __syncthreads();
        if (cand) {
            const int64_t slot = off + s_can + (e & 0xFFFFu);
            cand_user[slot] = (int32_t)u;
            cand_rank[slot] = rank;
            if (refine) {
                const int32_t r = s_ref + (int32_t)(e >> 16);
                refine_cand[r] = (int32_t)slot;
                refine_user[r] = (int32_t)u;
                refine_query[r] = (int32_t)b;
                refine_thr[r] = qb[u];
            }
        }
    }
}

__global__ void k_scatter_counts(const int32_t *__restrict__ refine_total, const int32_t *__restrict__ refine_cand,
                                 const int32_t *__restrict__ cnt, int64_t max_rows, int32_t *__restrict__ cand_rank) {
    const int64_t n = min((int64_t)*refine_total, max_rows);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        cand_rank[refine_cand[i]] = cnt[i];
}

// ---- top-k -----------------------------------------------------------------
constexpr int SORT_CAP = kMaxTopk;  // keys sorted in LDS

__device__ void bitonic_sort_u64(uint64_t *keys, int n_pow2) {
    for (int size = 2; size <= n_pow2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n_pow2 / 2; i += NT) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const uint64_t a = keys[lo], c = keys[hi];
                if ((a > c) == up) { keys[lo] = c; keys[hi] = a; }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ uint64_t make_key(int32_t rank, int32_t user) {
    return ((uint64_t)(uint32_t)rank << 32) | (uint32_t)user;
}

// One workgroup per query.  Output row: the settled users (reference
// output) followed by the smallest (rank, userID) candidates.
__global__ void __launch_bounds__(NT) k_topk(const QueryInfo *__restrict__ qi, const int32_t *__restrict__ cand_user,
                                             const int32_t *__restrict__ cand_rank, const float *__restrict__ qip,
                                             int64_t n_user, int32_t max_k, int32_t *__restrict__ out_user,
                                             int32_t *__restrict__ out_rank, float *__restrict__ out_qip) {
    __shared__ uint64_t keys[SORT_CAP];
    __shared__ uint32_t hist[256];
    __shared__ uint64_t sel_prefix;
    __shared__ uint32_t sel_rem, n_sel;
    const int64_t b = blockIdx.x;
    const QueryInfo q = qi[b];
    const int32_t k = q.topk;
    int32_t *ou = out_user + b * (int64_t)max_k;
    int32_t *orank = out_rank + b * (int64_t)max_k;
    float *oq = out_qip + b * (int64_t)max_k;
    const float *qb = qip + b * n_user;

    for (int i = threadIdx.x; i < max_k; i += NT) { ou[i] = -1; orank[i] = -1; oq[i] = -1.0f; }
    __syncthreads();
    const int32_t ns = q.n_settled < k ? q.n_settled : k;
    for (int i = threadIdx.x; i < ns; i += NT) {
        const int32_t u = cand_user[q.cand_off + i];
        ou[i] = u;
        orank[i] = cand_rank[q.cand_off + i];
        oq[i] = qb[u];
    }
    const int32_t need = k - ns;
    const int64_t nc = q.n_cand;
    if (need <= 0 || nc <= 0) return;
    const int32_t *cu = cand_user + q.cand_off + q.n_settled;
    const int32_t *cr = cand_rank + q.cand_off + q.n_settled;
    const int32_t take = (int32_t)(nc < need ? nc : need);

    int n_keys;
    if (nc <= SORT_CAP) {
        for (int64_t i = threadIdx.x; i < nc; i += NT) keys[i] = make_key(cr[i], cu[i]);
        n_keys = (int)nc;
    } else {
        // MSD radix select of the take-th smallest key, 8 bits at a time;
        // keys are distinct (userIDs are), so exactly `take` keys are <= it.
        if (threadIdx.x == 0) { sel_prefix = 0; sel_rem = (uint32_t)take - 1; }
        __syncthreads();
        for (int shift = 56; shift >= 0; shift -= 8) {
            for (int i = threadIdx.x; i < 256; i += NT) hist[i] = 0;
            __syncthreads();
            const uint64_t pre = sel_prefix;
            const uint64_t mask_hi = (shift == 56) ? 0ull : (~0ull << (shift + 8));
            for (int64_t i = threadIdx.x; i < nc; i += NT) {
                const uint64_t key = make_key(cr[i], cu[i]);
                if ((key & mask_hi) == pre) atomicAdd(&hist[(key >> shift) & 255u], 1u);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t rem = sel_rem, d = 0;
                for (; d < 255; ++d) {
                    if (rem < hist[d]) break;
                    rem -= hist[d];
                }
                sel_prefix = pre | ((uint64_t)d << shift);
                sel_rem = rem;
            }
            __syncthreads();
        }
        const uint64_t kth = sel_prefix;
        if (threadIdx.x == 0) n_sel = 0;
        __syncthreads();
        for (int64_t i = threadIdx.x; i < nc; i += NT) {
            const uint64_t key = make_key(cr[i], cu[i]);
            if (key <= kth) keys[atomicAdd(&n_sel, 1u)] = key;
        }
        __syncthreads();
        n_keys = (int)n_sel;
    }
    int p2 = 1;
    while (p2 < n_keys) p2 <<= 1;
    for (int i = n_keys + threadIdx.x; i < p2; i += NT) keys[i] = ~0ull;
    __syncthreads();
    bitonic_sort_u64(keys, p2);
    for (int i = threadIdx.x; i < take; i += NT) {
        const uint64_t key = keys[i];
        const int32_t u = (int32_t)(uint32_t)(key & 0xFFFFFFFFu);
        ou[ns + i] = u;
        orank[ns + i] = (int32_t)(key >> 32);
        oq[ns + i] = qb[u];
    }
}


// ---- coarse bounds (screened form, exact output) ---------------------------
// k_ip_coarse (k_screen.hip) gives every (query, user) a byte c_lo | c_hi << 4:
// how many of the user's guard columns j_0 < ... < j_{G-1} of the table lie
// above the lowest / highest value the score can take.  A guard above the
// lowest value bounds b_lo = #{T > s - eps} <= j_{c_lo} (S when c_lo = G); a
// guard above the highest bounds b_hi >= j_{c_hi - 1} + 1 (0 when c_hi = 0).
// The k-th smallest bound j_{beta_c} over all users is then >= the exact
// beta, so the users with c_hi <= beta_c hold every exact candidate and the k
// users of smallest exact b_lo: scoring and searching only them yields the
// exact path's beta, candidates and ranks.

constexpr int CNT = 256;

// Workgroups are dispatched to the 8 XCDs in turn (block p on XCD p % 8), and
// each XCD has its own L2.  For a list of n_live entries in CNT-entry blocks,
// this gives block p the entries of logical block (p % 8) * per + p / 8, so
// each XCD takes one contiguous eighth of the list: entries close in the
// list (the same users, for the coarse list's chunk-major order) share one
// L2.  Returns -1 for a block past the list.  The grid must hold a multiple
// of 8 blocks covering the largest list (xcd_grid), so that every block p <
// 8 * per exists and the mapping is a bijection onto them.
__device__ __forceinline__ int64_t xcd_block(int64_t n_live) {
    const int64_t blocks = (n_live + CNT - 1) / CNT, per = (blocks + 7) / 8;
    const int64_t p = blockIdx.x;
    if (p >= per * 8) return -1;
    const int64_t lb = (p % 8) * per + p / 8;
    return lb < blocks ? lb : -1;
}
inline unsigned xcd_grid(int64_t n_max) { return (unsigned)(((n_max + CNT - 1) / CNT + 7) / 8 * 8); }

// The score rule: FP32 products summed in ascending dimension order.
__device__ __forceinline__ float rule_dot(const float *__restrict__ a, const float *__restrict__ b, int32_t d) {
    float s = 0.0f;
    int32_t k = 0;
    // 32 terms are loaded before they are summed: lanes of a wavefront read
    // different rows, and taking a whole 128-byte stretch per lane at once
    // uses each cache line fully before it can be evicted (the sum's order is
    // unchanged)
    for (; k + 32 <= d; k += 32) {
        float4 x[8], y[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            x[t] = *reinterpret_cast<const float4 *>(a + k + 4 * t);
            y[t] = *reinterpret_cast<const float4 *>(b + k + 4 * t);
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            s = s + x[t].x * y[t].x;
            s = s + x[t].y * y[t].y;
            s = s + x[t].z * y[t].z;
            s = s + x[t].w * y[t].w;
        }
    }
    for (; k + 4 <= d; k += 4) {
        const float4 x = *reinterpret_cast<const float4 *>(a + k);
        const float4 y = *reinterpret_cast<const float4 *>(b + k);
        s = s + x.x * y.x;
        s = s + x.y * y.y;
        s = s + x.z * y.z;
        s = s + x.w * y.w;
    }
    for (; k < d; ++k) s = s + a[k] * b[k];
    return s;
}

// Nibble counters of 8 bytes' c_lo (low 3 bits) or c_hi (bits 4..6): bin c
// at bits 4c..4c+3 (at most 8 per nibble).
template <int SH>
__device__ __forceinline__ uint32_t nibble_count8(uint32_t x0, uint32_t x1) {
    uint32_t n = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        n += 1u << (__builtin_amdgcn_ubfe(x0, 8 * j + SH, 3) << 2);
        n += 1u << (__builtin_amdgcn_ubfe(x1, 8 * j + SH, 3) << 2);
    }
    return n;
}

// Adds two nibble-counter words into byte counters: even bins (0, 2, 4, 6)
// in the bytes of e, odd bins in the bytes of o.
__device__ __forceinline__ void nibbles_to_bytes(uint32_t a, uint32_t b, uint32_t &e, uint32_t &o) {
    e += (a & 0x0F0F0F0Fu) + (b & 0x0F0F0F0Fu);
    o += ((a >> 4) & 0x0F0F0F0Fu) + ((b >> 4) & 0x0F0F0F0Fu);
}

// Adds byte counters (even bins in e, odd bins in o) into four words of
// 16-bit counters (bins 2w, 2w + 1 in word w).
__device__ __forceinline__ void widen_add(uint32_t e, uint32_t o, uint32_t (&w)[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] += ((e >> (8 * k)) & 0xFFu) | (((o >> (8 * k)) & 0xFFu) << 16);
}

// grid (chunks, queries): counts of c_lo and c_hi values 0..6 of the chunk's
// users, part[(b * n_chunk + chunk) * 16 + {0..6 | 8..14}] (b = qmap[y] when
// given).  16 users per thread and load, four loads in flight (rows of cb
// are ldc bytes, ldc % 16 == 0, chunks start at multiples of 1024).  Each
// 16-byte load is counted into nibble counters (8 bytes per word, so no
// nibble passes 8), added into byte counters (at most 16 per load), and those
// are widened to 16 bits every 15 loads, before they can carry; the workgroup
// sums them by wave reductions (cu < 2^18 keeps the sums in 16 bits).  Only
// a chunk's last load tests its bytes against the chunk's end.
__global__ void __launch_bounds__(CNT) k_coarse_hist(const uint8_t *__restrict__ cb, int64_t ldc, int64_t n_user,
                                                     int64_t cu, int32_t n_chunk, const int32_t *__restrict__ qmap,
                                                     uint32_t *__restrict__ part) {
    __shared__ uint32_t red[CNT / 64][8];
    const int64_t b = qmap ? qmap[blockIdx.y] : blockIdx.y;
    const uint8_t *v = cb + b * ldc;
    const int64_t u0 = (int64_t)blockIdx.x * cu, u1 = min(n_user, u0 + cu);
    uint32_t elo = 0, olo = 0, ehi = 0, ohi = 0;
    uint32_t wlo[4] = {0u, 0u, 0u, 0u}, whi[4] = {0u, 0u, 0u, 0u};
    int n = 0;
    auto count16 = [&](const uint4 w) {
        nibbles_to_bytes(nibble_count8<0>(w.x, w.y), nibble_count8<0>(w.z, w.w), elo, olo);
        nibbles_to_bytes(nibble_count8<4>(w.x, w.y), nibble_count8<4>(w.z, w.w), ehi, ohi);
        if (++n == 15) {
            widen_add(elo, olo, wlo);
            widen_add(ehi, ohi, whi);
            elo = olo = ehi = ohi = 0;
            n = 0;
        }
    };
    constexpr int64_t STEP = 16 * CNT;
    int64_t u = u0 + 16 * (int64_t)threadIdx.x;
    for (; u + 3 * STEP + 16 <= u1; u += 4 * STEP) {
        uint4 w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = *reinterpret_cast<const uint4 *>(v + u + k * STEP);
#pragma unroll
        for (int k = 0; k < 4; ++k) count16(w[k]);
    }
    for (; u + 16 <= u1; u += STEP) count16(*reinterpret_cast<const uint4 *>(v + u));
    if (u < u1) {
        // the chunk's last, partial load: bytes past the end count as bin 7
        // (not reported)
        uint4 w = *reinterpret_cast<const uint4 *>(v + u);
        const int rem = (int)(u1 - u);
        uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (j >= rem) ws[j >> 2] |= 0x77u << (8 * (j & 3));
        count16(make_uint4(ws[0], ws[1], ws[2], ws[3]));
    }
    widen_add(elo, olo, wlo);
    widen_add(ehi, ohi, whi);
    uint32_t t[8] = {wlo[0], wlo[1], wlo[2], wlo[3], whi[0], whi[1], whi[2], whi[3]};
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int i = 0; i < 8; ++i) t[i] += __shfl_xor(t[i], o, 64);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < 8; ++i) red[wave][i] = t[i];
    __syncthreads();
    if (threadIdx.x < 16) {
        // bin threadIdx.x: lo bins 0..7 then hi bins 0..7
        const int i = threadIdx.x >> 1, sh = (threadIdx.x & 1) * 16;
        uint32_t sum = 0;
#pragma unroll
        for (int w = 0; w < CNT / 64; ++w) sum += (red[w][i] >> sh) & 0xFFFFu;
        const int bin = threadIdx.x & 7;
        part[((int64_t)b * n_chunk + blockIdx.x) * 16 + threadIdx.x] = bin == 7 ? 0u : sum;
    }
}

// One workgroup per query: beta_c = smallest c < n_guard with
// #{c_lo <= c} >= k (-1 when there is none: the query takes the full path)
// and the number of users with c_hi <= beta_c.  cq[b] = {beta_c, count, -, 0}.
__global__ void __launch_bounds__(CNT) k_coarse_kth(const uint32_t *__restrict__ part, int32_t n_chunk,
                                                    int32_t n_guard, const int32_t *__restrict__ qmap,
                                                    const QueryInfo *__restrict__ qi, int32_t *__restrict__ cq,
                                                    int32_t *__restrict__ hcnt) {
    __shared__ uint32_t sh[CNT];
    const int64_t b = qmap ? qmap[blockIdx.x] : blockIdx.x;
    uint32_t acc = 0;
    for (int32_t ch = threadIdx.x >> 4; ch < n_chunk; ch += CNT / 16)
        acc += part[((int64_t)b * n_chunk + ch) * 16 + (threadIdx.x & 15)];
    sh[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < 16) {
        uint32_t t = 0;
        for (int j = 0; j < CNT / 16; ++j) t += sh[j * 16 + threadIdx.x];
        sh[threadIdx.x] = t;  // rows j >= 1 are no longer read
    }
    __syncthreads();
    if (threadIdx.x < 8) hcnt[b * 8 + threadIdx.x] = threadIdx.x < 7 ? (int32_t)sh[8 + threadIdx.x] : 0;
    if (threadIdx.x != 0) return;
    const int64_t k = qi[b].topk;
    int32_t beta = -1;
    int64_t cum = 0;
    for (int32_t c = 0; c < n_guard; ++c) {
        cum += sh[c];
        if (cum >= k) { beta = c; break; }
    }
    int64_t n = 0;
    for (int32_t c = 0; c <= beta; ++c) n += sh[8 + c];
    cq[b * 4 + 0] = beta;
    cq[b * 4 + 1] = (int32_t)min(n, (int64_t)INT32_MAX);
    cq[b * 4 + 3] = 0;
}

// grid = n_chunk workgroups: per chunk, the listed users of every query
// (c_hi <= beta_c, none when beta_c < 0) from the chunk histograms, their
// exclusive offsets within the chunk, loc[chunk * nq + b], and the chunk's
// total, ctot[chunk].
__global__ void __launch_bounds__(NT) k_coarse_chunk_offsets(const uint32_t *__restrict__ part, int32_t n_chunk,
                                                             int32_t nq, const int32_t *__restrict__ cq,
                                                             int32_t *__restrict__ loc, int32_t *__restrict__ ctot) {
    __shared__ uint32_t red[NW + 1];
    const int64_t ch = blockIdx.x;
    uint32_t base = 0;
    for (int32_t b0 = 0; b0 < nq; b0 += NT) {
        const int32_t b = b0 + (int32_t)threadIdx.x;
        uint32_t v = 0;
        if (b < nq) {
            const int32_t beta = cq[(int64_t)b * 4];
            const uint32_t *p = part + ((int64_t)b * n_chunk + ch) * 16 + 8;
            for (int32_t c = 0; c <= beta; ++c) v += p[c];
        }
        uint32_t tot;
        const uint32_t e = wg_excl_scan(v, red, &tot);
        if (b < nq) loc[ch * nq + b] = (int32_t)(base + e);
        base += tot;
    }
    if (threadIdx.x == 0) ctot[ch] = (int32_t)base;
}

// One workgroup: cbase[ch] = exclusive scan of ctot (chunk-major list).
__global__ void __launch_bounds__(NT) k_coarse_chunk_base(const int32_t *__restrict__ ctot, int32_t n_chunk,
                                                          int32_t *__restrict__ cbase, int32_t *__restrict__ total) {
    __shared__ uint32_t red[NW + 1];
    uint32_t base = 0;
    for (int32_t c0 = 0; c0 < n_chunk; c0 += NT) {
        const int32_t i = c0 + (int32_t)threadIdx.x;
        const uint32_t v = i < n_chunk ? (uint32_t)ctot[i] : 0u;
        uint32_t tot;
        const uint32_t e = wg_excl_scan(v, red, &tot);
        if (i < n_chunk) cbase[i] = (int32_t)(base + e);
        base += tot;
    }
    if (total && threadIdx.x == 0) *total = (int32_t)base;  // the list's length
}

// One workgroup: a query whose coarse record has no k-th bound, a list over
// cap, or a list longer than the bound the host put in field 2 (its first
// pass's length) is sent to the full path (beta = -1, nothing listed).
__global__ void __launch_bounds__(NT) k_coarse_mark(int32_t *__restrict__ cq, int32_t nq, int64_t cap) {
    for (int32_t b = threadIdx.x; b < nq; b += NT) {
        int32_t *e = cq + (int64_t)b * 4;
        if (e[0] < 0 || e[1] > cap || e[1] > e[2]) {
            e[0] = -1;
            e[1] = 0;
        }
        e[3] = 0;
    }
}

// grid (chunks, queries): lists the users with c_hi <= beta_c of every query
// that has one.  The list is chunk-major (all queries' users of chunk 0,
// then chunk 1, ...), so entries processed together share users' rows; the
// order within a (chunk, query) slice is arbitrary.  cl_b = b | byte << 16.
// 16 users per thread and load; a byte-parallel compare marks the listed ones.
__global__ void __launch_bounds__(CNT) k_coarse_compact(const uint8_t *__restrict__ cb, int64_t ldc, int64_t n_user,
                                                        int64_t cu, int32_t nq, const int32_t *__restrict__ cq,
                                                        const int32_t *__restrict__ loc,
                                                        const int32_t *__restrict__ cbase,
                                                        int32_t *__restrict__ cl_u, int32_t *__restrict__ cl_b) {
    __shared__ int32_t s_fill;
    const int64_t b = blockIdx.y;
    const int32_t beta = cq[b * 4 + 0];
    if (beta < 0) return;
    const int64_t off = (int64_t)cbase[blockIdx.x] + loc[(int64_t)blockIdx.x * nq + b];
    if (threadIdx.x == 0) s_fill = 0;
    __syncthreads();
    const uint8_t *v = cb + b * ldc;
    const int64_t u0 = (int64_t)blockIdx.x * cu, u1 = min(n_user, u0 + cu);
    const int lane = threadIdx.x & 63;
    const uint32_t bb = (uint32_t)(beta + 1) * 0x01010101u;
    for (int64_t ub = u0; ub < u1; ub += 16 * CNT) {
        const int64_t u = ub + 16 * (int64_t)threadIdx.x;
        uint32_t bits = 0;  // bit j: user u + j is listed
        uint32_t ws[4] = {0u, 0u, 0u, 0u};
        if (u < u1) {
            const uint4 w = *reinterpret_cast<const uint4 *>(v + u);
            ws[0] = w.x; ws[1] = w.y; ws[2] = w.z; ws[3] = w.w;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // bit 7 of a byte of t is set iff its c_hi >= beta + 1
                const uint32_t t = (((ws[i] >> 4) & 0x07070707u) | 0x80808080u) - bb;
                const uint32_t m = ~t & 0x80808080u;
                bits |= (((m >> 7) & 1u) | ((m >> 14) & 2u) | ((m >> 21) & 4u) | ((m >> 28) & 8u)) << (4 * i);
            }
            const int64_t rem = u1 - u;
            if (rem < 16) bits &= (1u << rem) - 1u;
        }
        const uint32_t cnt = (uint32_t)__popc(bits);
        if (!__ballot(cnt != 0)) continue;
        uint32_t x = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        int32_t base = 0;
        if (lane == 63) base = atomicAdd(&s_fill, (int32_t)x);
        base = __shfl(base, 63, 64);
        int64_t pos = off + base + (int64_t)(x - cnt);
        while (bits) {
            const int j = __ffs(bits) - 1;
            bits &= bits - 1;
            cl_u[pos] = (int32_t)(u + j);
            cl_b[pos] = (int32_t)b | (int32_t)(((ws[j >> 2] >> (8 * (j & 3))) & 0xFFu) << 16);
            ++pos;
        }
    }
}

// One thread per listed (query, user): the score rule's <u,q>, the exact
// bucket pair against the full table row, and the query's bucket histograms
// (ghist[b] = [h_lo | h_hi], S + 1 bins each; h_lo only when same != 0).
__global__ void __launch_bounds__(CNT) k_coarse_exact(const float *__restrict__ users,
                                                      const float *__restrict__ queries, int32_t ld, int32_t dim,
                                                      const float *__restrict__ table, int32_t S, float eps, int same,
                                                      const GuardCols *__restrict__ gsets,
                                                      const int32_t *__restrict__ gset, const int32_t *__restrict__ cl_u,
                                                      const int32_t *__restrict__ cl_b, int64_t n,
                                                      const int32_t *__restrict__ n_dev,
                                                      uint32_t *__restrict__ bk_c, float *__restrict__ s_c,
                                                      uint32_t *__restrict__ ghist) {
    const int64_t n_live = n_dev ? min(n, (int64_t)*n_dev) : n;
    const int64_t lb = xcd_block(n_live);
    const int64_t i = lb * CNT + threadIdx.x;
    if (lb < 0 || i >= n_live) return;
    const int64_t u = cl_u[i];
    const uint32_t packed = (uint32_t)cl_b[i];
    const int64_t b = packed & 0xFFFFu;
    const int32_t clo = (int32_t)((packed >> 16) & 7u), chi = (int32_t)((packed >> 20) & 7u);
    const GuardCols gc = gsets[gset[b]];
    const float sc = rule_dot(users + u * ld, queries + b * ld, dim);
    // both buckets lie in [j_{c_hi - 1} + 1, j_{c_lo}] (0 / S at the ends),
    // j the query's guard columns
    int32_t r0 = 0, r1 = S;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j == chi - 1) r0 = gc.col[j] + 1;
        if (j == clo && j < gc.n) r1 = gc.col[j];
    }
    const float *row = table + u * (int64_t)S;
    uint32_t lo, hi;
    if (eps > 0.0f) {
        lo = (uint32_t)bucket_gt_in(row, r0, r1, __double2float_rd((double)sc - (double)eps));
        hi = (uint32_t)bucket_gt_in(row, r0, r1, __double2float_ru((double)sc + (double)eps));
    } else {
        lo = hi = (uint32_t)bucket_gt_in(row, r0, r1, sc);
    }
    bk_c[i] = lo | (hi << 16);
    s_c[i] = sc;
    uint32_t *g = ghist + b * 2 * (int64_t)(S + 1);
    atomicAdd(&g[lo], 1u);
    if (!same) atomicAdd(&g[S + 1 + hi], 1u);
}

// Update path, one thread per listed (query, user): the rank bound values of
// the exact bounds with pending inserts and deletes (base bucket within the
// guard range, side rows searched at the score rule's <u,q>), as
// k_bounds_upd computes them, and <u,q> for the refinement and top-k.
__global__ void __launch_bounds__(CNT) k_coarse_exact_upd(
    const float *__restrict__ users, const float *__restrict__ queries, int32_t ld, int32_t dim,
    const float *__restrict__ table, int32_t S, const int32_t *__restrict__ ranks, int64_t n_item_base, float eps,
    const GuardCols *__restrict__ gsets, const int32_t *__restrict__ gset, const int32_t *__restrict__ cl_u,
    const int32_t *__restrict__ cl_b, int64_t n, const float *__restrict__ ins, int32_t n_ins, int32_t ld_ins,
    const float *__restrict__ del, int32_t n_del, int32_t ld_del, int64_t n_user, int32_t *__restrict__ lbs,
    int32_t *__restrict__ ubs, float *__restrict__ qip) {
    const int64_t lb = xcd_block(n);
    const int64_t i = lb * CNT + threadIdx.x;
    if (lb < 0 || i >= n) return;
    const int64_t u = cl_u[i];
    const uint32_t packed = (uint32_t)cl_b[i];
    const int64_t b = packed & 0xFFFFu;
    const int32_t clo = (int32_t)((packed >> 16) & 7u), chi = (int32_t)((packed >> 20) & 7u);
    const GuardCols gc = gsets[gset[b]];
    const float sc = rule_dot(users + u * ld, queries + b * ld, dim);
    int32_t r0 = 0, r1 = S;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j == chi - 1) r0 = gc.col[j] + 1;
        if (j == clo && j < gc.n) r1 = gc.col[j];
    }
    const float *row = table + u * (int64_t)S;
    int32_t blo, bhi;
    if (eps > 0.0f) {
        blo = bucket_gt_in(row, r0, r1, __double2float_rd((double)sc - (double)eps));
        bhi = bucket_gt_in(row, r0, r1, __double2float_ru((double)sc + (double)eps));
    } else {
        blo = bhi = bucket_gt_in(row, r0, r1, sc);
    }
    int32_t off = 0;
    if (n_ins) off += bucket_gt(ins + u * (int64_t)ld_ins, n_ins, sc);
    if (n_del) off -= bucket_gt(del + u * (int64_t)ld_del, n_del, sc);
    const int64_t o = b * n_user + u;
    lbs[o] = (blo == S ? (int32_t)n_item_base : ranks[blo]) + off;
    ubs[o] = max((bhi == 0 ? 0 : ranks[bhi - 1] + 1) + off, 0);
    qip[o] = sc;
}

// One thread per listed (query, user): the candidates (b_hi <= beta) with
// their rank when the bounds fix it, else -1 and a refine row; <u,q> of each
// candidate goes to qip for the top-k.
__global__ void __launch_bounds__(CNT) k_coarse_final(const int32_t *__restrict__ cl_u, const int32_t *__restrict__ cl_b,
                                                      const uint32_t *__restrict__ bk_c, const float *__restrict__ s_c,
                                                      int64_t n, const int32_t *__restrict__ n_dev, int32_t S,
                                                      const int32_t *__restrict__ ranks,
                                                      int64_t n_item, int64_t n_user, QueryInfo *__restrict__ qi,
                                                      int32_t *__restrict__ cand_user, int32_t *__restrict__ cand_rank,
                                                      int32_t *__restrict__ refine_total,
                                                      int32_t *__restrict__ refine_cand,
                                                      int32_t *__restrict__ refine_user,
                                                      int32_t *__restrict__ refine_query,
                                                      float *__restrict__ refine_thr, float *__restrict__ qip) {
    const int64_t i = (int64_t)blockIdx.x * CNT + threadIdx.x;
    if (i >= (n_dev ? min(n, (int64_t)*n_dev) : n)) return;
    const int32_t b = cl_b[i] & 0xFFFF;
    const uint32_t x = bk_c[i];
    const int32_t blo = (int32_t)(x & 0xFFFFu), bhi = (int32_t)(x >> 16);
    if (bhi > qi[b].beta) return;
    const int32_t u = cl_u[i];
    const int32_t lbs = blo == S ? (int32_t)n_item : ranks[blo];
    const int32_t ubs = bhi == 0 ? 0 : ranks[bhi - 1] + 1;
    const int32_t rank = lbs == ubs ? lbs : -1;
    const int64_t slot = qi[b].cand_off + atomicAdd(&qi[b].fill, 1);
    const float sc = s_c[i];
    cand_user[slot] = u;
    cand_rank[slot] = rank;
    qip[(int64_t)b * n_user + u] = sc;
    if (rank < 0) {
        const int32_t r = atomicAdd(refine_total, 1);
        atomicAdd(&qi[b].n_refine, 1);
        refine_cand[r] = (int32_t)slot;
        refine_user[r] = u;
        refine_query[r] = b;
        refine_thr[r] = sc;
    }
}

}  // namespace

hipError_t launch_bounds(hipStream_t s, const float *qip, const float *table, int64_t n_user, int32_t S, int32_t nq,
                         float eps, int ref_mode, uint32_t *bk) {
    if (n_user <= 0 || nq <= 0) return hipSuccess;
    const int ut = bounds_rows_per_wg(S);
    if (nq >= 4 && ut > 0) {
        const unsigned grid = (unsigned)((n_user + ut - 1) / ut);
        hipLaunchKernelGGL(k_bounds_lds, dim3(grid), dim3(256), (size_t)ut * (S + 1) * 4, s, qip, table, n_user, S, nq,
                           ut, eps, ref_mode, bk);
        return hipGetLastError();
    }
    dim3 grid((unsigned)((n_user + 255) / 256), (unsigned)nq);
    hipLaunchKernelGGL(k_bounds, grid, dim3(256), 0, s, qip, table, n_user, S, eps, ref_mode, bk);
    return hipGetLastError();
}

hipError_t launch_hist_kth(hipStream_t s, const uint32_t *bk, int64_t n_user, int32_t S, const int32_t *ranks,
                           int64_t n_item, int32_t nq, int output_mode, int same, int64_t cu, uint32_t *ghist,
                           QueryInfo *qi) {
    if (nq <= 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(ghist, 0, sizeof(uint32_t) * 2 * (size_t)(S + 1) * nq, s);
    if (e != hipSuccess) return e;
    const unsigned n_chunk = (unsigned)((n_user + cu - 1) / cu);
    if (n_user > 0)
        hipLaunchKernelGGL(k_hist_chunks, dim3(n_chunk, (unsigned)nq), dim3(NT), 0, s, bk, n_user, S, same, cu, ghist);
    hipLaunchKernelGGL(k_hist_kth, dim3((unsigned)nq), dim3(NT), 0, s, ghist, S, same, ranks, n_item, output_mode, qi);
    return hipGetLastError();
}

hipError_t launch_offsets(hipStream_t s, QueryInfo *qi, int32_t nq, int32_t *refine_total) {
    hipLaunchKernelGGL(k_offsets, dim3(1), dim3(NT), 0, s, qi, nq, refine_total);
    return hipGetLastError();
}

hipError_t launch_compact(hipStream_t s, const uint32_t *bk, const float *qip, int64_t n_user, int32_t S,
                          const int32_t *ranks, int64_t n_item, int32_t nq, int output_mode, int64_t cu,
                          QueryInfo *qi, int32_t *cand_user, int32_t *cand_rank, int32_t *refine_total,
                          int32_t *refine_cand, int32_t *refine_user, int32_t *refine_query, float *refine_thr) {
    if (nq <= 0 || n_user <= 0) return hipSuccess;
    if (output_mode == RKR_OUTPUT_EXACT) {
        const unsigned n_chunk = (unsigned)((n_user + cu - 1) / cu);
        hipLaunchKernelGGL(k_compact_chunks, dim3(n_chunk, (unsigned)nq), dim3(NT), 0, s, bk, qip, n_user, S, ranks,
                           n_item, cu, qi, cand_user, cand_rank, refine_total, refine_cand, refine_user, refine_query,
                           refine_thr);
    } else {
        // reference output keeps the users in ascending order (settled users
        // are taken by userID, US.hpp:173-186)
        hipLaunchKernelGGL(k_compact, dim3((unsigned)nq), dim3(NT), 0, s, bk, qip, n_user, S, ranks, n_item, qi,
                           cand_user, cand_rank, refine_total, refine_cand, refine_user, refine_query, refine_thr);
    }
    return hipGetLastError();
}

hipError_t launch_scatter_counts(hipStream_t s, const int32_t *refine_total, const int32_t *refine_cand,
                                 const int32_t *cnt, int64_t max_rows, int32_t *cand_rank) {
    if (max_rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_counts, dim3(1024), dim3(256), 0, s, refine_total, refine_cand, cnt, max_rows,
                       cand_rank);
    return hipGetLastError();
}

hipError_t launch_topk(hipStream_t s, const QueryInfo *qi, int32_t nq, const int32_t *cand_user,
                       const int32_t *cand_rank, const float *qip, int64_t n_user, int32_t max_k, int32_t *out_user,
                       int32_t *out_rank, float *out_qip, int64_t *scratch_keys) {
    (void)scratch_keys;
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_topk, dim3((unsigned)nq), dim3(NT), 0, s, qi, cand_user, cand_rank, qip, n_user, max_k,
                       out_user, out_rank, out_qip);
    return hipGetLastError();
}

hipError_t launch_coarse_kth(hipStream_t s, const uint8_t *cb, int64_t ldc, int64_t n_user, int32_t nq, int64_t cu,
                             int32_t n_guard, const int32_t *qmap, uint32_t *part, const QueryInfo *qi, int32_t *cq,
                             int32_t *hcnt) {
    if (nq <= 0 || n_user <= 0) return hipSuccess;
    if (ldc % 16 || cu % 1024 || cu >= (1 << 18)) return hipErrorInvalidValue;
    const int32_t n_chunk = (int32_t)((n_user + cu - 1) / cu);
    hipLaunchKernelGGL(k_coarse_hist, dim3((unsigned)n_chunk, (unsigned)nq), dim3(CNT), 0, s, cb, ldc, n_user, cu,
                       n_chunk, qmap, part);
    hipLaunchKernelGGL(k_coarse_kth, dim3((unsigned)nq), dim3(CNT), 0, s, part, n_chunk, n_guard, qmap, qi, cq,
                       hcnt);
    return hipGetLastError();
}

hipError_t launch_coarse_compact(hipStream_t s, const uint8_t *cb, int64_t ldc, int64_t n_user, int32_t nq,
                                 int64_t cu, const uint32_t *part, const int32_t *cq, int32_t *loc, int32_t *ctot,
                                 int32_t *cbase, int32_t *cl_u, int32_t *cl_b, int32_t *total) {
    if (nq <= 0 || n_user <= 0) return hipSuccess;
    if (ldc % 16 || cu % 1024 || nq > 65535) return hipErrorInvalidValue;
    const int32_t n_chunk = (int32_t)((n_user + cu - 1) / cu);
    hipLaunchKernelGGL(k_coarse_chunk_offsets, dim3((unsigned)n_chunk), dim3(NT), 0, s, part, n_chunk, nq, cq, loc,
                       ctot);
    hipLaunchKernelGGL(k_coarse_chunk_base, dim3(1), dim3(NT), 0, s, ctot, n_chunk, cbase, total);
    hipLaunchKernelGGL(k_coarse_compact, dim3((unsigned)n_chunk, (unsigned)nq), dim3(CNT), 0, s, cb, ldc, n_user, cu,
                       nq, cq, loc, cbase, cl_u, cl_b);
    return hipGetLastError();
}

hipError_t launch_coarse_exact(hipStream_t s, const float *users, const float *queries, int32_t ld, int32_t dim,
                               const float *table, int32_t S, const int32_t *ranks, int64_t n_item, float eps,
                               const GuardCols *gsets, const int32_t *gset, int32_t nq, const int32_t *cl_u,
                               const int32_t *cl_b, int64_t n, const int32_t *n_dev, uint32_t *bk_c, float *s_c,
                               uint32_t *ghist, QueryInfo *qi) {
    if (nq <= 0) return hipSuccess;
    const int same = eps == 0.0f ? 1 : 0;
    hipError_t e = hipMemsetAsync(ghist, 0, sizeof(uint32_t) * 2 * (size_t)(S + 1) * nq, s);
    if (e != hipSuccess) return e;
    if (n > 0)
        hipLaunchKernelGGL(k_coarse_exact, dim3(xcd_grid(n)), dim3(CNT), 0, s, users, queries, ld,
                           dim, table, S, eps, same, gsets, gset, cl_u, cl_b, n, n_dev, bk_c, s_c, ghist);
    hipLaunchKernelGGL(k_hist_kth, dim3((unsigned)nq), dim3(NT), 0, s, ghist, S, same, ranks, n_item,
                       (int)RKR_OUTPUT_EXACT, qi);
    return hipGetLastError();
}

hipError_t launch_coarse_exact_upd(hipStream_t s, const float *users, const float *queries, int32_t ld, int32_t dim,
                                   const float *table, int32_t S, const int32_t *ranks, int64_t n_item_base,
                                   float eps, const GuardCols *gsets, const int32_t *gset, const int32_t *cl_u,
                                   const int32_t *cl_b, int64_t n, const float *ins, int32_t n_ins, int32_t ld_ins,
                                   const float *del, int32_t n_del, int32_t ld_del, int64_t n_user, int32_t *lbs,
                                   int32_t *ubs, float *qip) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_coarse_exact_upd, dim3(xcd_grid(n)), dim3(CNT), 0, s, users, queries, ld,
                       dim, table, S, ranks, n_item_base, eps, gsets, gset, cl_u, cl_b, n, ins, n_ins, ld_ins, del,
                       n_del, ld_del, n_user, lbs, ubs, qip);
    return hipGetLastError();
}

hipError_t launch_coarse_mark(hipStream_t s, int32_t *cq, int32_t nq, int64_t cap) {
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_coarse_mark, dim3(1), dim3(NT), 0, s, cq, nq, cap);
    return hipGetLastError();
}

hipError_t launch_coarse_final(hipStream_t s, const int32_t *cl_u, const int32_t *cl_b, const uint32_t *bk_c,
                               const float *s_c, int64_t n, const int32_t *n_dev, int32_t S, const int32_t *ranks,
                               int64_t n_item,
                               int64_t n_user, QueryInfo *qi, int32_t *cand_user, int32_t *cand_rank,
                               int32_t *refine_total, int32_t *refine_cand, int32_t *refine_user,
                               int32_t *refine_query, float *refine_thr, float *qip) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_coarse_final, dim3((unsigned)((n + CNT - 1) / CNT)), dim3(CNT), 0, s, cl_u, cl_b, bk_c, s_c, n,
                       n_dev, S, ranks, n_item, n_user, qi, cand_user, cand_rank, refine_total, refine_cand, refine_user,
                       refine_query, refine_thr, qip);
    return hipGetLastError();
}

}  // namespace rkr
