/*
 * rkr_oracle.c -- CPU restatement of the reverse k-ranks "US" (uniform sample)
 * path of DBGroup-SUSTech/reverse-k-ranks.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (librkr.so) never links, calls or falls back to it.
 *
 * Every function cites the reference file:line (paths under src/) it
 * restates.  Score arithmetic follows src/include/alg/SpaceInnerProduct.hpp:14-20
 * as ISO C evaluates it: FP32 products and sums, each rounded, summed in
 * ascending dimension order.  This file must be compiled with
 * -ffp-contract=off and without -ffast-math (see oracle/Makefile) so that no
 * product is fused into an FMA and no sum is reassociated.
 *
 * How each function is pinned (checked by tests/test_oracle_*.py):
 *   orc_ip_f32, orc_refine_rank, orc_prune, orc_score_rows,
 *   orc_normalize_users  -- bit-exact against the reference's own headers
 *                           compiled into oracle/_ref (RefineCPU.hpp,
 *                           PruneCandidateByBound.hpp, CPUScoreTable.hpp,
 *                           VectorMatrix.hpp, SpaceInnerProduct.hpp), and
 *                           against committed fixtures made from them.
 *   orc_uniform_sample   -- closed form floor(j*(I-1)/S) and the invariants
 *                           asserted at UniformSample.hpp:26-35.
 *   orc_coarse_bsearch   -- brute-force bucket count #{j : T[j] >= q} and the
 *                           case table of SampleSearch.hpp:107-125.
 *   orc_build_table      -- numpy sort + gather on small inputs.
 *   orc_us_retrieval     -- composition of the pinned pieces above; checked
 *                           against the brute-force exact answer wherever the
 *                           reference's bounds are exact.
 *   orc_exact_query      -- equals orc_bruteforce (rank every user) on every
 *                           input.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* a1: user normalisation, src/include/struct/VectorMatrix.hpp:51-66.   */
/* FP32 sequential sum of squares, FP32 sqrt, skip norms <= 0.001       */
/* (the float is promoted to double for that compare), FP32 division.   */
/* ------------------------------------------------------------------ */
void orc_normalize_users(float *u, int64_t n, int32_t d) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        float *row = u + i * (int64_t)d;
        float l2 = 0.0f;
        for (int32_t j = 0; j < d; j++) {
            float sq = row[j] * row[j];
            l2 = l2 + sq;
        }
        l2 = sqrtf(l2);
        if ((double)l2 <= 0.001) continue;
        for (int32_t j = 0; j < d; j++) row[j] = row[j] / l2;
    }
}

/* ------------------------------------------------------------------ */
/* Score rule, src/include/alg/SpaceInnerProduct.hpp:14-20.             */
/* ------------------------------------------------------------------ */
float orc_ip_f32(const float *a, const float *b, int32_t d) {
    float res = 0.0f;
    for (int32_t i = 0; i < d; i++) {
        float p = a[i] * b[i];
        res = res + p;
    }
    return res;
}

/* FP64 sequential sum of the (exact) FP32 products: used only to measure
 * the error of screened scores; not a ranking rule of the reference.      */
double orc_ip_f64(const float *a, const float *b, int32_t d) {
    double res = 0.0;
    for (int32_t i = 0; i < d; i++) res = res + (double)a[i] * (double)b[i];
    return res;
}

/* ------------------------------------------------------------------ */
/* a2: uniform sample ranks,                                           */
/* src/include/alg/RankBoundRefinement/FindSampleRank/UniformSample.hpp:17-25 */
/* r_j = floor(j * ((I-1)*1.0/S)), ascending.  Returns 0, or -1 when the */
/* SampleSearch constructor would exit (S <= 0 or S >= I,               */
/* SampleSearch.hpp:35-38).                                             */
/* ------------------------------------------------------------------ */
int orc_uniform_sample(int64_t n_item, int32_t n_sample, int32_t *ranks) {
    if (n_sample <= 0 || (int64_t)n_sample >= n_item) return -1;
    const int end_rank = (int)n_item - 1;
    const double delta = (end_rank - 0) * 1.0 / n_sample;
    for (int32_t j = 0; j < n_sample; j++) ranks[j] = (int32_t)floor(j * delta);
    return 0;
}

/* ------------------------------------------------------------------ */
/* a3: score rows, src/include/score_computation/CPUScoreTable.hpp:55-68 */
/* out[b*I + i] = InnerProduct(user[u0+b], item[i]).                    */
/* ------------------------------------------------------------------ */
void orc_score_rows(const float *users, int64_t u0, int64_t nu, const float *items,
                    int64_t n_item, int32_t d, float *out) {
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nu; b++) {
        const float *u = users + (u0 + b) * (int64_t)d;
        for (int64_t i = 0; i < n_item; i++) out[b * n_item + i] = orc_ip_f32(u, items + i * (int64_t)d, d);
    }
}

static int cmp_float_desc(const void *x, const void *y) {
    float a = *(const float *)x, b = *(const float *)y;
    return (a < b) - (a > b);
}

/* ------------------------------------------------------------------ */
/* a3+a4: sampled score table.  Restates the CPU branch of              */
/* ComputeScoreTableBatch::ComputeSortItemsBatch                        */
/* (src/include/score_computation/ComputeScoreTableBatch.hpp:79-92:     */
/* scores, then a descending sort of each row) and                      */
/* SampleSearch::LoopPreprocess (SampleSearch.hpp:80-85:                */
/* T[u][j] = sorted_u[r_j]), as driven by                               */
/* src/app/build-index/BuildSampleIndexByComputation.cpp:111-125.       */
/* ------------------------------------------------------------------ */
void orc_build_table(const float *users, int64_t n_user, const float *items, int64_t n_item,
                     int32_t d, const int32_t *ranks, int32_t n_sample, float *table) {
#pragma omp parallel
    {
        float *row = (float *)malloc(sizeof(float) * (size_t)n_item);
#pragma omp for schedule(dynamic, 16)
        for (int64_t u = 0; u < n_user; u++) {
            const float *uv = users + u * (int64_t)d;
            for (int64_t i = 0; i < n_item; i++) row[i] = orc_ip_f32(uv, items + i * (int64_t)d, d);
            qsort(row, (size_t)n_item, sizeof(float), cmp_float_desc);
            for (int32_t j = 0; j < n_sample; j++) table[u * (int64_t)n_sample + j] = row[ranks[j]];
        }
        free(row);
    }
}

/* ------------------------------------------------------------------ */
/* a6: coarse binary search, SampleSearch.hpp:91-130.                   */
/* bucket = length of the prefix of the (non-increasing) row with       */
/* T[j] >= q (the std::lower_bound with predicate arrIP >= queryIP at   */
/* :99-102); then the case table of :107-125.                           */
/* ------------------------------------------------------------------ */
static int32_t bucket_ge(const float *row, int32_t S, float q) {
    int32_t lo = 0, hi = S; /* partition point of (row[j] >= q) */
    while (lo < hi) {
        int32_t mid = lo + (hi - lo) / 2;
        if (row[mid] >= q) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static int32_t bucket_gt(const float *row, int32_t S, float q) {
    int32_t lo = 0, hi = S; /* partition point of (row[j] > q) */
    while (lo < hi) {
        int32_t mid = lo + (hi - lo) / 2;
        if (row[mid] > q) lo = mid + 1; else hi = mid;
    }
    return lo;
}

void orc_coarse_bsearch(const float *row, int32_t S, const int32_t *ranks, int64_t n_item,
                        float qip, int32_t *rank_lb, int32_t *rank_ub) {
    const uint32_t b = (uint32_t)bucket_ge(row, S, qip);
    const uint32_t tlb = b == (uint32_t)S ? (uint32_t)n_item : (uint32_t)ranks[b];
    const uint32_t tub = b == 0 ? 0u : (uint32_t)ranks[b - 1];
    if (b == (uint32_t)S) {
        *rank_lb = (int32_t)n_item;
        *rank_ub = (int32_t)tub;
    } else if (b == 0) {
        *rank_lb = (int32_t)tlb;
        *rank_ub = 0;
    } else if (tlb - tub <= 1) {
        *rank_lb = (int32_t)tlb;
        *rank_ub = (int32_t)tlb;
    } else {
        *rank_lb = (int32_t)tlb;
        *rank_ub = (int32_t)tub;
    }
}

/* ------------------------------------------------------------------ */
/* a7: PruneCandidateByBound,                                           */
/* src/include/alg/RankBoundRefinement/PruneCandidateByBound.hpp:19-179 */
/* Restated with a counting pass instead of the two max-heaps: the heap */
/* fronts at :76-77 are the heap_size-th smallest rank_lb and rank_ub   */
/* over live users, which is what kth_smallest computes.                */
/* ------------------------------------------------------------------ */
static int cmp_int_asc(const void *x, const void *y) {
    int a = *(const int *)x, b = *(const int *)y;
    return (a > b) - (a < b);
}

static int32_t kth_smallest_live(const int32_t *v, int64_t n, const char *prune, const char *result,
                                 int32_t k) {
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++)
        if (!prune[i] && !result[i]) tmp[m++] = v[i];
    /* The reference fills the heap with heap_size entries that start at 0
     * (std::vector<int>(heap_size), :37-38); with fewer live users than
     * heap_size the unfilled zeros stay in the heap. */
    int32_t kth;
    if (m < k) {
        kth = 0;
        for (int64_t i = 0; i < m; i++) if (tmp[i] > kth) kth = tmp[i];
    } else {
        qsort(tmp, (size_t)m, sizeof(int32_t), cmp_int_asc);
        kth = tmp[k - 1];
    }
    free(tmp);
    return kth;
}

void orc_prune(const int32_t *rank_lb, const int32_t *rank_ub, int64_t n_user, int32_t topk,
               int32_t *refine_user_size, int32_t *n_result_user, int32_t *n_prune_user,
               char *prune, char *result) {
    if (*n_result_user >= topk) return; /* :31-33 */
    const int32_t heap_size = topk - *n_result_user;
    const int32_t Lk = kth_smallest_live(rank_lb, n_user, prune, result, heap_size); /* :76 */
    const int32_t Uk = kth_smallest_live(rank_ub, n_user, prune, result, heap_size); /* :77 */

    if (Lk == Uk) { /* :83-111 */
        *n_result_user = topk;
        *refine_user_size = 0;
        for (int64_t u = 0; u < n_user; u++)
            if (rank_ub[u] <= Lk && rank_lb[u] <= Lk) result[u] = 1;
        return;
    }

    int32_t refine = 0, nres = *n_result_user, npr = *n_prune_user; /* :114-136 */
    for (int64_t u = 0; u < n_user; u++) {
        if (prune[u] || result[u]) continue;
        if (rank_lb[u] <= Uk) { result[u] = 1; nres++; }
        else if (Lk <= rank_ub[u]) { prune[u] = 1; npr++; }
        else refine++;
    }
    if (nres + refine == topk) { /* :139-150 */
        for (int64_t u = 0; u < n_user; u++)
            if (!prune[u] && !result[u]) result[u] = 1;
        refine = 0;
        nres = topk;
    }
    *refine_user_size = refine;
    *n_result_user = nres;
    *n_prune_user = npr;
}

/* ------------------------------------------------------------------ */
/* a8: exact rank of one user,                                          */
/* src/include/alg/DiskIndex/RefineByComputation/RefineCPU.hpp:33-55:   */
/* #{items : InnerProduct(user, item) > queryIP}.                       */
/* ------------------------------------------------------------------ */
int32_t orc_refine_rank(const float *user, const float *items, int64_t n_item, int32_t d, float qip) {
    int32_t rank = 0;
    for (int64_t i = 0; i < n_item; i++)
        if (orc_ip_f32(user, items + i * (int64_t)d, d) > qip) rank++;
    return rank;
}

/* orc_refine_rank for n (user, query) pairs, OpenMP over the pairs:       */
/* user_id[i] indexes users, q_id[i] indexes queries; out[i] receives the  */
/* strict count and qip_out[i] (may be NULL) the pair's InnerProduct.      */
void orc_refine_ranks(const float *users, const float *queries, const float *items, int64_t n_item, int32_t d,
                      const int64_t *user_id, const int64_t *q_id, int64_t n, int32_t *out, float *qip_out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < n; i++) {
        const float *u = users + user_id[i] * (int64_t)d;
        const float qip = orc_ip_f32(queries + q_id[i] * (int64_t)d, u, d);
        out[i] = orc_refine_rank(u, items, n_item, d, qip);
        if (qip_out) qip_out[i] = qip;
    }
}

typedef struct { int32_t user, rank; float qip; } orc_elem;

/* UserRankElement::operator<, src/include/struct/UserRankElement.hpp:59-64 */
static int cmp_elem(const void *x, const void *y) {
    const orc_elem *a = (const orc_elem *)x, *b = (const orc_elem *)y;
    if (a->rank != b->rank) return (a->rank > b->rank) - (a->rank < b->rank);
    return (a->user > b->user) - (a->user < b->user);
}

/* Per-query counters, mirroring SingleQueryPerformance
 * (src/include/util/FileIO.hpp:150-173). */
typedef struct {
    int32_t n_prune_user, n_result_user, n_refine_user, n_candidate;
    int64_t ip_cost;
} orc_qstat;

/* ------------------------------------------------------------------ */
/* a5-a9: one US query as src/impl/US.hpp:117-187 runs it:              */
/* GEMV (:124-132), RankBound (:139 -> SampleSearch.hpp:132-156),       */
/* PruneCandidateByBound (:148-151), ComputeAll::GetRank                */
/* (:163-167 -> ComputeAll.hpp:54-125) and result assembly (:173-186).  */
/* Output: topk (user, rank, queryIP); result-flagged users first in    */
/* userID order carrying rank_lb, then refined users by (rank, userID). */
/* ------------------------------------------------------------------ */
int orc_us_query(const float *users, int64_t n_user, const float *items, int64_t n_item, int32_t d,
                 const float *table, const int32_t *ranks, int32_t S, const float *q, int32_t topk,
                 int32_t *out_user, int32_t *out_rank, float *out_qip, orc_qstat *st) {
    if (topk <= 0 || topk > n_user) return -1; /* US.hpp:106-109 */
    float *qip = (float *)malloc(sizeof(float) * (size_t)n_user);
    int32_t *lb = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_user);
    int32_t *ub = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_user);
    char *prune = (char *)calloc((size_t)n_user, 1);
    char *result = (char *)calloc((size_t)n_user, 1);

#pragma omp parallel for schedule(static)
    for (int64_t u = 0; u < n_user; u++) qip[u] = orc_ip_f32(q, users + u * (int64_t)d, d);

#pragma omp parallel for schedule(dynamic, 1000)
    for (int64_t u = 0; u < n_user; u++)
        orc_coarse_bsearch(table + u * (int64_t)S, S, ranks, n_item, qip[u], &lb[u], &ub[u]);

    int32_t refine_size = (int32_t)n_user, n_result = 0, n_prune = 0;
    orc_prune(lb, ub, n_user, topk, &refine_size, &n_result, &n_prune, prune, result);

    /* ComputeAll::GetRank (ComputeAll.hpp:54-125) */
    const int32_t remain = topk - n_result > 0 ? topk - n_result : 0;
    orc_elem *cache = NULL;
    int64_t n_cand = 0, n_refined = 0;
    if (remain > 0) {
        int64_t *cand = (int64_t *)malloc(sizeof(int64_t) * (size_t)n_user);
        for (int64_t u = 0; u < n_user; u++)
            if (!prune[u] && !result[u]) cand[n_cand++] = u;
        cache = (orc_elem *)malloc(sizeof(orc_elem) * (size_t)(n_cand > 0 ? n_cand : 1));
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : n_refined)
        for (int64_t c = 0; c < n_cand; c++) {
            const int64_t u = cand[c];
            int32_t r;
            if (lb[u] == ub[u]) r = lb[u];
            else { r = orc_refine_rank(users + u * (int64_t)d, items, n_item, d, qip[u]); n_refined++; }
            cache[c].user = (int32_t)u; cache[c].rank = r; cache[c].qip = qip[u];
        }
        qsort(cache, (size_t)n_cand, sizeof(orc_elem), cmp_elem);
        free(cand);
    }

    /* result assembly, US.hpp:173-186 */
    int32_t nc = 0;
    for (int64_t u = 0; u < n_user && nc < topk; u++) {
        if (result[u]) { out_user[nc] = (int32_t)u; out_rank[nc] = lb[u]; out_qip[nc] = qip[u]; nc++; }
    }
    for (int32_t c = nc; c < topk; c++) {
        const int64_t src = c - nc;
        if (cache && src < n_cand) { out_user[c] = cache[src].user; out_rank[c] = cache[src].rank; out_qip[c] = cache[src].qip; }
        else { out_user[c] = -1; out_rank[c] = -1; out_qip[c] = -1.0f; }
    }
    if (st) {
        st->n_prune_user = n_prune; st->n_result_user = n_result; st->n_refine_user = (int32_t)n_refined;
        st->n_candidate = (int32_t)n_cand; st->ip_cost = n_user + n_refined * n_item;
    }
    free(cache); free(qip); free(lb); free(ub); free(prune); free(result);
    return 0;
}

/* All queries of a batch, sequential over queries as US.hpp:117 runs them
 * (every query runs; the reference's 100-query cap at US.hpp:112 is not
 * restated).  Threads: OpenMP inside each query, as the reference.        */
int orc_us_retrieval(const float *users, int64_t n_user, const float *items, int64_t n_item, int32_t d,
                     const float *table, const int32_t *ranks, int32_t S, const float *queries,
                     int32_t n_query, int32_t topk, int32_t *out_user, int32_t *out_rank, float *out_qip,
                     orc_qstat *st) {
    for (int32_t qi = 0; qi < n_query; qi++) {
        int rc = orc_us_query(users, n_user, items, n_item, d, table, ranks, S, queries + (int64_t)qi * d, topk,
                              out_user + (int64_t)qi * topk, out_rank + (int64_t)qi * topk,
                              out_qip + (int64_t)qi * topk, st ? st + qi : NULL);
        if (rc) return rc;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Exact reverse k-ranks through the sampled table (the pipeline the    */
/* GPU path implements).  Same score rule, table and tie rule as the    */
/* reference (strict count RefineCPU.hpp:42; order UserRankElement.hpp:59-64), */
/* with bounds that stay valid under score ties:                        */
/*   b  = #{j : T[j] > q - eps},   b2 = #{j : T[j] > q + eps}           */
/*   lbs = (b == S) ? I : r[b]     (items at or after sample position  */
/*                                   r[b] score <= T[b] <= q: none beats q) */
/*   ubs = (b2 == 0) ? 0 : r[b2-1] + 1  (items up to r[b2-1] beat q)     */
/* eps bounds |T - exact order statistic| (0 for tables built with this */
/* score rule).  L = topk-th smallest lbs; candidates are users with    */
/* ubs <= L; each candidate's rank is lbs when lbs == ubs, else counted. */
/* ------------------------------------------------------------------ */
void orc_safe_bounds(const float *row, int32_t S, const int32_t *ranks, int64_t n_item, float qip, float eps,
                     int32_t *lbs, int32_t *ubs) {
    float qlo = qip, qhi = qip;
    if (eps > 0.0f) {
        qlo = (float)((double)qip - (double)eps);
        if ((double)qlo > (double)qip - (double)eps) qlo = nextafterf(qlo, -INFINITY);
        qhi = (float)((double)qip + (double)eps);
        if ((double)qhi < (double)qip + (double)eps) qhi = nextafterf(qhi, INFINITY);
    }
    const int32_t b = bucket_gt(row, S, qlo);
    const int32_t b2 = bucket_gt(row, S, qhi);
    *lbs = b == S ? (int32_t)n_item : ranks[b];
    *ubs = b2 == 0 ? 0 : ranks[b2 - 1] + 1;
}

int orc_exact_query(const float *users, int64_t n_user, const float *items, int64_t n_item, int32_t d,
                    const float *table, const int32_t *ranks, int32_t S, float eps, const float *q, int32_t topk,
                    int32_t *out_user, int32_t *out_rank, float *out_qip, orc_qstat *st) {
    if (topk <= 0 || topk > n_user) return -1;
    float *qip = (float *)malloc(sizeof(float) * (size_t)n_user);
    int32_t *lbs = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_user);
    int32_t *ubs = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_user);
    int64_t *hist = (int64_t *)calloc((size_t)S + 1, sizeof(int64_t));

#pragma omp parallel for schedule(static)
    for (int64_t u = 0; u < n_user; u++) {
        qip[u] = orc_ip_f32(q, users + u * (int64_t)d, d);
        orc_safe_bounds(table + u * (int64_t)S, S, ranks, n_item, qip[u], eps, &lbs[u], &ubs[u]);
    }
    /* topk-th smallest lbs: lbs takes values r[0..S-1] or I, in bucket order */
    for (int64_t u = 0; u < n_user; u++) {
        int32_t lo = 0, hi = S; /* index of lbs among {r[0..S-1], I} */
        while (lo < hi) { int32_t m = (lo + hi) / 2; if (ranks[m] < lbs[u]) lo = m + 1; else hi = m; }
        hist[lo]++;
    }
    int64_t cum = 0; int32_t L = (int32_t)n_item;
    for (int32_t j = 0; j <= S; j++) { cum += hist[j]; if (cum >= topk) { L = j == S ? (int32_t)n_item : ranks[j]; break; } }

    orc_elem *cand = (orc_elem *)malloc(sizeof(orc_elem) * (size_t)n_user);
    int64_t n_cand = 0, n_refined = 0;
    for (int64_t u = 0; u < n_user; u++)
        if (ubs[u] <= L) { cand[n_cand].user = (int32_t)u; cand[n_cand].qip = qip[u]; n_cand++; }
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : n_refined)
    for (int64_t c = 0; c < n_cand; c++) {
        const int64_t u = cand[c].user;
        if (lbs[u] == ubs[u]) cand[c].rank = lbs[u];
        else { cand[c].rank = orc_refine_rank(users + u * (int64_t)d, items, n_item, d, qip[u]); n_refined++; }
    }
    qsort(cand, (size_t)n_cand, sizeof(orc_elem), cmp_elem);
    for (int32_t c = 0; c < topk; c++) { out_user[c] = cand[c].user; out_rank[c] = cand[c].rank; out_qip[c] = cand[c].qip; }
    if (st) {
        st->n_candidate = (int32_t)n_cand; st->n_refine_user = (int32_t)n_refined;
        st->n_result_user = 0; st->n_prune_user = (int32_t)(n_user - n_cand);
        st->ip_cost = n_user + n_refined * n_item;
    }
    free(cand); free(hist); free(qip); free(lbs); free(ubs);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Brute force: rank every user with the strict count, order by         */
/* (rank, userID), keep topk.  The definition the exact path must meet. */
/* ------------------------------------------------------------------ */
int orc_bruteforce(const float *users, int64_t n_user, const float *items, int64_t n_item, int32_t d,
                   const float *q, int32_t topk, int32_t *out_user, int32_t *out_rank, float *out_qip) {
    if (topk <= 0 || topk > n_user) return -1;
    orc_elem *all = (orc_elem *)malloc(sizeof(orc_elem) * (size_t)n_user);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t u = 0; u < n_user; u++) {
        const float *uv = users + u * (int64_t)d;
        const float qi = orc_ip_f32(q, uv, d);
        all[u].user = (int32_t)u; all[u].qip = qi;
        all[u].rank = orc_refine_rank(uv, items, n_item, d, qi);
    }
    qsort(all, (size_t)n_user, sizeof(orc_elem), cmp_elem);
    for (int32_t c = 0; c < topk; c++) { out_user[c] = all[c].user; out_rank[c] = all[c].rank; out_qip[c] = all[c].qip; }
    free(all);
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
