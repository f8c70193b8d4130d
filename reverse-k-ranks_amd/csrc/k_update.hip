// k_update.hip -- kernels of the update path (inserted / deleted items and
// users).
//
// After items are inserted or deleted, the table T still holds the order
// statistics of the base item set it was built from.  As the reference does
// (SampleSearchUpdate::InsertItem / DeleteItem / SearchInsertDelete,
// src/include/alg/RankBoundRefinement/SampleSearchUpdate.hpp:111-166,212-243),
// each user keeps two sorted side arrays: its scores against the inserted
// items and against the deleted base items.  The exact rank over the current
// item set is
//   rank = rank_base + #{inserted : s > q} - #{deleted : s > q},
// so the base bounds shift by that exact offset (the reference uses
// #{s >= q}; the strict count keeps the bounds exact under ties).  Bounds then
// vary per user, so the k-th smallest upper bound is found by a two-level
// histogram over bound values instead of over buckets.
#include <algorithm>

#include "rkr_internal.hpp"

namespace rkr {
namespace {

constexpr int NT = 1024;
constexpr int NW = NT / 64;

// Sorts each row of `x` (rows of n <= 1024 floats, stride ld) descending in
// place.  One single-wavefront workgroup per row, bitonic sort in LDS.
__global__ void __launch_bounds__(64) k_sort_rows_desc(float *__restrict__ x, int64_t n_rows, int32_t n, int64_t ld) {
    __shared__ float b[1024];
    const int lane = threadIdx.x;
    const int64_t row = blockIdx.x;
    float *src = x + row * ld;
    int p2 = 1;
    while (p2 < n) p2 <<= 1;
    for (int i = lane; i < p2; i += 64) b[i] = i < n ? src[i] : -INFINITY;
    __syncthreads();
    for (int size = 2; size <= p2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = lane; i < p2 / 2; i += 64) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool desc = ((lo & size) == 0);
                const float a = b[lo], c = b[hi];
                if ((a < c) == desc) { b[lo] = c; b[hi] = a; }
            }
            __syncthreads();
        }
    }
    for (int i = lane; i < n; i += 64) src[i] = b[i];
}

// Same for rows of up to 16384 floats: one 1024-thread workgroup per row.
__global__ void __launch_bounds__(1024) k_sort_rows_desc_big(float *__restrict__ x, int64_t n_rows, int32_t n,
                                                             int64_t ld) {
    __shared__ float b[16384];
    const int64_t row = blockIdx.x;
    float *src = x + row * ld;
    int p2 = 1;
    while (p2 < n) p2 <<= 1;
    for (int i = threadIdx.x; i < p2; i += 1024) b[i] = i < n ? src[i] : -INFINITY;
    __syncthreads();
    for (int size = 2; size <= p2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < p2 / 2; i += 1024) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool desc = ((lo & size) == 0);
                const float a = b[lo], c = b[hi];
                if ((a < c) == desc) { b[lo] = c; b[hi] = a; }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += 1024) src[i] = b[i];
}

// #{j < n : a[j] > v} for a non-increasing row a.
__device__ __forceinline__ int32_t count_gt(const float *a, int32_t n, float v) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (a[mid] > v) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ int32_t count_ge(const float *a, int32_t n, float v) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (a[mid] >= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// out[u] = merge of two non-increasing rows a[u] (na) and b[u] (nb).  Each
// output position i finds its split by a co-rank search: element a[j] goes
// to j + #{b > a[j]} and b[j] to j + #{a >= b[j]}, so ties keep a first and
// every position is written once.
__global__ void k_merge_rows(const float *__restrict__ a, int64_t lda, int32_t na, const float *__restrict__ b,
                             int64_t ldb, int32_t nb, float *__restrict__ out, int64_t ldo, int64_t n_rows) {
    const int64_t row = blockIdx.x;
    if (row >= n_rows) return;
    const float *ar = a + row * lda;
    const float *br = b + row * ldb;
    float *o = out + row * ldo;
    for (int64_t i = threadIdx.x; i < (int64_t)na + nb; i += blockDim.x) {
        if (i < na) {
            const float v = ar[i];
            o[i + count_gt(br, nb, v)] = v;
        } else {
            const int64_t j = i - na;
            const float v = br[j];
            o[j + count_ge(ar, na, v)] = v;
        }
    }
}

// out[u] = a[u] minus the multiset b[u] (both non-increasing; b is a
// sub-multiset of a).  The t-th copy of a value in b removes the t-th copy of
// that value in a; kept entries are compacted in order by one workgroup per row.
__global__ void __launch_bounds__(256) k_subtract_rows(const float *__restrict__ a, int64_t lda, int32_t na,
                                                       const float *__restrict__ b, int64_t ldb, int32_t nb,
                                                       float *__restrict__ out, int64_t ldo, int64_t n_rows) {
    __shared__ uint32_t red[5];
    const int64_t row = blockIdx.x;
    if (row >= n_rows) return;
    const float *ar = a + row * lda;
    const float *br = b + row * ldb;
    float *o = out + row * ldo;
    uint32_t base = 0;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int32_t i0 = 0; i0 < na; i0 += 256) {
        const int32_t i = i0 + threadIdx.x;
        bool keep = false;
        float v = 0.f;
        if (i < na) {
            v = ar[i];
            // copies of v before i in a, and copies of v in b
            const int32_t first_a = count_gt(ar, na, v);
            const int32_t nb_v = count_ge(br, nb, v) - count_gt(br, nb, v);
            keep = (i - first_a) >= nb_v;
        }
        const uint64_t ball = __ballot(keep);
        const uint32_t before = __popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) red[w] = __popcll(ball);
        __syncthreads();
        uint32_t wbase = base;
        for (int k = 0; k < w; ++k) wbase += red[k];
        const uint32_t tot = red[0] + red[1] + red[2] + red[3];
        if (keep && wbase + before < (uint32_t)ldo) o[wbase + before] = v;  // ldo = na - nb kept entries
        __syncthreads();
        base += tot;
    }
}

// Bounds with side-array offsets (exact output).  lbs/ubs as values:
//   off = #{ins > q} - #{del > q};  lbs' = lbs + off;  ubs' = ubs + off.
__device__ __forceinline__ void bounds_values(const float *row, int32_t S, float q, float eps,
                                              const int32_t *__restrict__ ranks, int64_t n_item_base,
                                              const float *ins_row, int32_t n_ins, const float *del_row,
                                              int32_t n_del, int32_t &lb, int32_t &ub) {
    float qlo = q, qhi = q;
    if (eps > 0.0f) {
        qlo = __double2float_rd((double)q - (double)eps);
        qhi = __double2float_ru((double)q + (double)eps);
    }
    const int32_t blo = count_gt(row, S, qlo), bhi = eps > 0.0f ? count_gt(row, S, qhi) : blo;
    int32_t off = 0;
    if (n_ins) off += count_gt(ins_row, n_ins, q);
    if (n_del) off -= count_gt(del_row, n_del, q);
    lb = (blo == S ? (int32_t)n_item_base : ranks[blo]) + off;
    // rank >= ub + off holds for any off; 0 is the tighter bound when it is negative
    ub = max((bhi == 0 ? 0 : ranks[bhi - 1] + 1) + off, 0);
}

// One thread per user: #{inserted scores > T[u][gc.col[j]]} for each guard,
// maximised over the wave and then over the device (out zeroed before).
__global__ void __launch_bounds__(256) k_ins_guard_max(const float *__restrict__ table, int64_t n_user, int32_t S,
                                                       GuardCols gc, const float *__restrict__ ins, int32_t n_ins,
                                                       int32_t ld_ins, int32_t *__restrict__ out) {
    const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (int j = 0; j < gc.n; ++j) {
        int32_t v = 0;
        if (u < n_user) v = count_gt(ins + u * (int64_t)ld_ins, n_ins, table[u * (int64_t)S + gc.col[j]]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0 && v > 0) atomicMax(&out[j], v);
    }
}

// grid (user blocks, queries); side rows at strides ld_ins / ld_del
__global__ void __launch_bounds__(256) k_bounds_upd(const float *__restrict__ qip, const float *__restrict__ table,
                                                    int64_t n_user, int32_t S, const int32_t *__restrict__ ranks,
                                                    int64_t n_item_base, float eps, const float *__restrict__ ins,
                                                    int32_t n_ins, int32_t ld_ins, const float *__restrict__ del,
                                                    int32_t n_del, int32_t ld_del, int32_t *__restrict__ lbs,
                                                    int32_t *__restrict__ ubs) {
    const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t b = blockIdx.y;
    if (u >= n_user) return;
    int32_t lb, ub;
    bounds_values(table + u * (int64_t)S, S, qip[b * n_user + u], eps, ranks, n_item_base, ins + u * (int64_t)ld_ins,
                  n_ins, del + u * (int64_t)ld_del, n_del, lb, ub);
    lbs[b * n_user + u] = lb;
    ubs[b * n_user + u] = ub;
}

// Batched form: ut table rows staged in LDS serve all nq queries of those
// users (thread t serves user t % ut).  With stage_side, the users' side rows
// are staged too; otherwise they are searched in place.  LDS rows carry one
// pad float so the searches of one step spread over the banks.
__global__ void __launch_bounds__(256) k_bounds_upd_lds(const float *__restrict__ qip, const float *__restrict__ table,
                                                        int64_t n_user, int32_t S, int32_t nq, int32_t ut,
                                                        int stage_side, const int32_t *__restrict__ ranks,
                                                        int64_t n_item_base, float eps, const float *__restrict__ ins,
                                                        int32_t n_ins, int32_t ld_ins, const float *__restrict__ del,
                                                        int32_t n_del, int32_t ld_del, int32_t *__restrict__ lbs,
                                                        int32_t *__restrict__ ubs) {
    // ut x (S + 1), then (stage_side) ut x (n_ins + 1) if n_ins, ut x (n_del + 1) if n_del
    extern __shared__ float rows[];
    float *rows_ins = rows + (int64_t)ut * (S + 1);
    float *rows_del = rows_ins + (n_ins ? (int64_t)ut * (n_ins + 1) : 0);
    const int64_t u0 = (int64_t)blockIdx.x * ut;
    const int64_t nu = min((int64_t)ut, n_user - u0);
    stage_rows(rows, table + u0 * S, nu, S, S);
    if (stage_side) {
        if (n_ins) stage_rows(rows_ins, ins + u0 * ld_ins, nu, n_ins, ld_ins);
        if (n_del) stage_rows(rows_del, del + u0 * ld_del, nu, n_del, ld_del);
    }
    __syncthreads();
    const int ul = threadIdx.x % ut;
    const int lanes_q = 256 / ut;
    if (ul >= nu) return;
    const int64_t u = u0 + ul;
    const float *row = rows + (int64_t)ul * (S + 1);
    const float *ins_row = stage_side ? rows_ins + (int64_t)ul * (n_ins + 1) : ins + u * (int64_t)ld_ins;
    const float *del_row = stage_side ? rows_del + (int64_t)ul * (n_del + 1) : del + u * (int64_t)ld_del;
    for (int32_t b = threadIdx.x / ut; b < nq; b += lanes_q) {
        int32_t lb, ub;
        bounds_values(row, S, qip[(int64_t)b * n_user + u], eps, ranks, n_item_base, ins_row, n_ins, del_row, n_del,
                      lb, ub);
        lbs[(int64_t)b * n_user + u] = lb;
        ubs[(int64_t)b * n_user + u] = ub;
    }
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

// k-th smallest lbs' per query by a two-level histogram of the bound values
// (value >> 11, then value & 2047; the host keeps values below 2^22), and
// the candidate count #{ubs' <= L} from the same two levels of ubs'.
// ghist per query: [level-1 lbs | level-1 ubs | level-2 lbs | level-2 ubs],
// 2048 bins each.  Level kernels run on grid (chunks, queries), the
// selections on one workgroup per query.
__global__ void __launch_bounds__(NT) k_vhist1(const int32_t *__restrict__ lbs, const int32_t *__restrict__ ubs,
                                               int64_t n_user, int64_t cu, uint32_t *__restrict__ ghist) {
    __shared__ uint32_t hl[2048], hu[2048];
    const int64_t b = blockIdx.y;
    const int64_t u0 = (int64_t)blockIdx.x * cu, u1 = min(n_user, u0 + cu);
    for (int i = threadIdx.x; i < 2048; i += NT) { hl[i] = 0; hu[i] = 0; }
    __syncthreads();
    const int32_t *lv = lbs + b * n_user, *uv = ubs + b * n_user;
    for (int64_t u = u0 + threadIdx.x; u < u1; u += NT) {
        // pairs without bounds (kNoBound) are never the k-th nor candidates
        if (lv[u] == kNoBound) continue;
        atomicAdd(&hl[((uint32_t)lv[u] >> 11) & 2047u], 1u);
        atomicAdd(&hu[((uint32_t)uv[u] >> 11) & 2047u], 1u);
    }
    __syncthreads();
    uint32_t *g = ghist + b * 4 * 2048;
    for (int i = threadIdx.x; i < 2048; i += NT) {
        if (hl[i]) atomicAdd(&g[i], hl[i]);
        if (hu[i]) atomicAdd(&g[2048 + i], hu[i]);
    }
}

// Level-1 selection: bin of the k-th smallest lbs' (QueryInfo.beta), the rank
// left inside it (gamma) and #{ubs' in lower bins} (n_cand, partial).
__global__ void __launch_bounds__(NT) k_vsel1(const uint32_t *__restrict__ ghist, QueryInfo *__restrict__ qi) {
    __shared__ uint32_t hl[2048], hu[2048];
    const int64_t b = blockIdx.x;
    const uint32_t *g = ghist + b * 4 * 2048;
    for (int i = threadIdx.x; i < 2048; i += NT) { hl[i] = g[i]; hu[i] = g[2048 + i]; }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const uint32_t k = (uint32_t)qi[b].topk;
    uint32_t cum = 0, j = 0;
    for (; j < 2047; ++j) { if (cum + hl[j] >= k) break; cum += hl[j]; }
    uint32_t nc = 0;
    for (uint32_t i = 0; i < j; ++i) nc += hu[i];
    QueryInfo q = qi[b];
    q.beta = (int32_t)j;
    q.gamma = (int32_t)(k - cum);  // 1-based rank inside the bin
    q.n_cand = (int32_t)nc;
    qi[b] = q;
}

__global__ void __launch_bounds__(NT) k_vhist2(const int32_t *__restrict__ lbs, const int32_t *__restrict__ ubs,
                                               int64_t n_user, int64_t cu, const QueryInfo *__restrict__ qi,
                                               uint32_t *__restrict__ ghist) {
    __shared__ uint32_t hl[2048], hu[2048];
    const int64_t b = blockIdx.y;
    const uint32_t hi = (uint32_t)qi[b].beta;
    const int64_t u0 = (int64_t)blockIdx.x * cu, u1 = min(n_user, u0 + cu);
    for (int i = threadIdx.x; i < 2048; i += NT) { hl[i] = 0; hu[i] = 0; }
    __syncthreads();
    const int32_t *lv = lbs + b * n_user, *uv = ubs + b * n_user;
    for (int64_t u = u0 + threadIdx.x; u < u1; u += NT) {
        const uint32_t l = (uint32_t)lv[u], w = (uint32_t)uv[u];
        if (l == (uint32_t)kNoBound) continue;
        if (((l >> 11) & 2047u) == hi) atomicAdd(&hl[l & 2047u], 1u);
        if (((w >> 11) & 2047u) == hi) atomicAdd(&hu[w & 2047u], 1u);
    }
    __syncthreads();
    uint32_t *g = ghist + b * 4 * 2048 + 2 * 2048;
    for (int i = threadIdx.x; i < 2048; i += NT) {
        if (hl[i]) atomicAdd(&g[i], hl[i]);
        if (hu[i]) atomicAdd(&g[2048 + i], hu[i]);
    }
}

// Level-2 selection: L = k-th smallest lbs' (QueryInfo.gamma) and the
// candidate count #{ubs' <= L}.
__global__ void __launch_bounds__(NT) k_vsel2(const uint32_t *__restrict__ ghist, QueryInfo *__restrict__ qi) {
    __shared__ uint32_t hl[2048], hu[2048];
    const int64_t b = blockIdx.x;
    const uint32_t *g = ghist + b * 4 * 2048 + 2 * 2048;
    for (int i = threadIdx.x; i < 2048; i += NT) { hl[i] = g[i]; hu[i] = g[2048 + i]; }
    __syncthreads();
    if (threadIdx.x != 0) return;
    QueryInfo q = qi[b];
    const uint32_t rem = (uint32_t)q.gamma;
    uint32_t cum = 0, j = 0;
    for (; j < 2047; ++j) { if (cum + hl[j] >= rem) break; cum += hl[j]; }
    uint32_t nc = (uint32_t)q.n_cand;
    for (uint32_t i = 0; i <= j; ++i) nc += hu[i];
    q.mode = 0;
    q.gamma = (q.beta << 11) + (int32_t)j;
    q.beta = -1;
    q.n_cand = (int32_t)nc;
    q.n_settled = 0;
    q.n_refine = 0;
    q.fill = 0;
    qi[b] = q;
}

// Chunked compaction of the candidates ubs' <= L (rank known when
// lbs' == ubs'), into ranges reserved on the query's cursor; the top-k
// selection orders the slice by (rank, userID).
__global__ void __launch_bounds__(NT) k_compact_values(const int32_t *__restrict__ lbs, const int32_t *__restrict__ ubs,
                                                       const float *__restrict__ qip, int64_t n_user, int64_t cu,
                                                       QueryInfo *__restrict__ qi, int32_t *__restrict__ cand_user,
                                                       int32_t *__restrict__ cand_rank, int32_t *__restrict__ refine_total,
                                                       int32_t *__restrict__ refine_cand, int32_t *__restrict__ refine_user,
                                                       int32_t *__restrict__ refine_query, float *__restrict__ refine_thr) {
    __shared__ uint32_t red[NW + 1];
    __shared__ int32_t s_can, s_ref;
    const int64_t b = blockIdx.y;
    const int32_t L = qi[b].gamma;
    const int64_t off = qi[b].cand_off;
    const int32_t *lv = lbs + b * n_user, *uv = ubs + b * n_user;
    const float *qb = qip + b * n_user;
    const int64_t u0 = (int64_t)blockIdx.x * cu, u1 = min(n_user, u0 + cu);
    for (int64_t ub0 = u0; ub0 < u1; ub0 += NT) {
        const int64_t u = ub0 + threadIdx.x;
        bool cand = false;
        int32_t rank = -1;
        // a pair without bounds is never a candidate (a query whose bounds
        // are all kNoBound selects L = kNoBound and counts no candidates)
        if (u < u1 && uv[u] <= L && uv[u] != kNoBound) {
            cand = true;
            if (lv[u] == uv[u]) rank = lv[u];
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

__global__ void __launch_bounds__(256) k_gather_cols(const float *__restrict__ src, int64_t ld_src,
                                                     const int32_t *__restrict__ cols, int32_t n_cols, int64_t n_rows,
                                                     float *__restrict__ dst, int64_t ld_dst) {
    for (int64_t r = blockIdx.x; r < n_rows; r += gridDim.x)
        for (int32_t j = threadIdx.x; j < n_cols; j += blockDim.x) dst[r * ld_dst + j] = src[r * ld_src + cols[j]];
}

__global__ void k_gather_rows(const float *__restrict__ src, int64_t lds, const int64_t *__restrict__ idx,
                              int64_t n_rows, int64_t width, float *__restrict__ dst, int64_t ldd) {
    const int64_t r = blockIdx.y;
    if (r >= n_rows) return;
    const float *s = src + idx[r] * lds;
    float *d = dst + r * ldd;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < width; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

}  // namespace

hipError_t launch_sort_rows_desc(hipStream_t s, float *x, int64_t n_rows, int32_t n, int64_t ld) {
    if (n_rows <= 0 || n <= 1) return hipSuccess;
    if (n > 16384) return hipErrorInvalidValue;
    if (n > 1024)
        hipLaunchKernelGGL(k_sort_rows_desc_big, dim3((unsigned)n_rows), dim3(1024), 0, s, x, n_rows, n, ld);
    else
        hipLaunchKernelGGL(k_sort_rows_desc, dim3((unsigned)n_rows), dim3(64), 0, s, x, n_rows, n, ld);
    return hipGetLastError();
}

hipError_t launch_merge_rows(hipStream_t s, const float *a, int64_t lda, int32_t na, const float *b, int64_t ldb,
                             int32_t nb, float *out, int64_t ldo, int64_t n_rows) {
    if (n_rows <= 0 || na + nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_merge_rows, dim3((unsigned)n_rows), dim3(256), 0, s, a, lda, na, b, ldb, nb, out, ldo, n_rows);
    return hipGetLastError();
}

hipError_t launch_subtract_rows(hipStream_t s, const float *a, int64_t lda, int32_t na, const float *b, int64_t ldb,
                                int32_t nb, float *out, int64_t ldo, int64_t n_rows) {
    if (n_rows <= 0 || na <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_subtract_rows, dim3((unsigned)n_rows), dim3(256), 0, s, a, lda, na, b, ldb, nb, out, ldo,
                       n_rows);
    return hipGetLastError();
}

hipError_t launch_bounds_upd(hipStream_t s, const float *qip, const float *table, int64_t n_user, int32_t S,
                             const int32_t *ranks, int64_t n_item_base, float eps, const float *ins, int32_t n_ins,
                             int32_t ld_ins, const float *del, int32_t n_del, int32_t ld_del, int32_t nq,
                             int32_t *lbs, int32_t *ubs) {
    if (n_user <= 0 || nq <= 0) return hipSuccess;
    if (nq >= 4) {
        // users per workgroup: with their side rows when they fit 40 KB
        const int64_t per_user = (int64_t)(S + 1) + (n_ins ? n_ins + 1 : 0) + (n_del ? n_del + 1 : 0);
        int ut = 16;
        while (ut > 1 && (int64_t)ut * per_user * 4 > 40960) ut >>= 1;
        // staging every side row costs a full pass over them: worth it for
        // larger batches only
        int stage_side = nq >= 32 && (int64_t)ut * per_user * 4 <= 65536;
        size_t lds = (size_t)ut * per_user * 4;
        if (!stage_side) {
            ut = bounds_rows_per_wg(S);
            lds = (size_t)ut * (S + 1) * 4;
        }
        if (ut > 0) {
            const unsigned grid = (unsigned)((n_user + ut - 1) / ut);
            hipLaunchKernelGGL(k_bounds_upd_lds, dim3(grid), dim3(256), lds, s, qip, table, n_user, S, nq, ut,
                               stage_side, ranks, n_item_base, eps, ins, n_ins, ld_ins, del, n_del, ld_del, lbs, ubs);
            return hipGetLastError();
        }
    }
    dim3 grid((unsigned)((n_user + 255) / 256), (unsigned)nq);
    hipLaunchKernelGGL(k_bounds_upd, grid, dim3(256), 0, s, qip, table, n_user, S, ranks, n_item_base, eps, ins, n_ins,
                       ld_ins, del, n_del, ld_del, lbs, ubs);
    return hipGetLastError();
}

hipError_t launch_ins_guard_max(hipStream_t s, const float *table, int64_t n_user, int32_t S, const GuardCols &gc,
                                const float *ins, int32_t n_ins, int32_t ld_ins, int32_t *out) {
    hipError_t e = hipMemsetAsync(out, 0, sizeof(int32_t) * 8, s);
    if (e != hipSuccess || n_user <= 0 || n_ins <= 0) return e;
    hipLaunchKernelGGL(k_ins_guard_max, dim3((unsigned)((n_user + 255) / 256)), dim3(256), 0, s, table, n_user, S, gc,
                       ins, n_ins, ld_ins, out);
    return hipGetLastError();
}

hipError_t launch_kth_values(hipStream_t s, const int32_t *lbs, const int32_t *ubs, int64_t n_user, int32_t nq,
                             int64_t cu, uint32_t *ghist, QueryInfo *qi) {
    if (nq <= 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(ghist, 0, sizeof(uint32_t) * 4 * 2048 * (size_t)nq, s);
    if (e != hipSuccess) return e;
    const dim3 grid((unsigned)((n_user + cu - 1) / cu), (unsigned)nq);
    if (n_user > 0) hipLaunchKernelGGL(k_vhist1, grid, dim3(NT), 0, s, lbs, ubs, n_user, cu, ghist);
    hipLaunchKernelGGL(k_vsel1, dim3((unsigned)nq), dim3(NT), 0, s, ghist, qi);
    if (n_user > 0) hipLaunchKernelGGL(k_vhist2, grid, dim3(NT), 0, s, lbs, ubs, n_user, cu, qi, ghist);
    hipLaunchKernelGGL(k_vsel2, dim3((unsigned)nq), dim3(NT), 0, s, ghist, qi);
    return hipGetLastError();
}

hipError_t launch_compact_values(hipStream_t s, const int32_t *lbs, const int32_t *ubs, const float *qip,
                                 int64_t n_user, int32_t nq, int64_t cu, QueryInfo *qi, int32_t *cand_user,
                                 int32_t *cand_rank, int32_t *refine_total, int32_t *refine_cand,
                                 int32_t *refine_user, int32_t *refine_query, float *refine_thr) {
    if (nq <= 0 || n_user <= 0) return hipSuccess;
    const dim3 grid((unsigned)((n_user + cu - 1) / cu), (unsigned)nq);
    hipLaunchKernelGGL(k_compact_values, grid, dim3(NT), 0, s, lbs, ubs, qip, n_user, cu, qi, cand_user, cand_rank,
                       refine_total, refine_cand, refine_user, refine_query, refine_thr);
    return hipGetLastError();
}

hipError_t launch_gather_cols(hipStream_t s, const float *src, int64_t ld_src, const int32_t *cols, int32_t n_cols,
                              int64_t n_rows, float *dst, int64_t ld_dst) {
    if (n_rows <= 0 || n_cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_cols, dim3((unsigned)std::min<int64_t>(n_rows, 65536)), dim3(256), 0, s, src, ld_src,
                       cols, n_cols, n_rows, dst, ld_dst);
    return hipGetLastError();
}

hipError_t launch_gather_rows(hipStream_t s, const float *src, int64_t lds, const int64_t *idx, int64_t n_rows,
                              int64_t width, float *dst, int64_t ldd) {
    if (n_rows <= 0 || width <= 0) return hipSuccess;
    const unsigned gx = (unsigned)std::min<int64_t>((width + 255) / 256, 16);
    for (int64_t r0 = 0; r0 < n_rows; r0 += 65535) {
        const int64_t nr = std::min<int64_t>(65535, n_rows - r0);
        hipLaunchKernelGGL(k_gather_rows, dim3(gx, (unsigned)nr), dim3(256), 0, s, src, lds, idx + r0, nr, width,
                           dst + r0 * ldd, ldd);
    }
    return hipGetLastError();
}

}  // namespace rkr
