/*
 * rkr.h -- C ABI of the MI355X reverse k-ranks library (librkr.so).
 *
 * Reverse k-ranks over users U (n_user x d), items P (n_item x d) and query
 * items q: rank(u, q) = #{p in P : <u,p> > <u,q>} (strict count), and a query
 * returns the k users with the smallest (rank, userID).  Scores follow the
 * reference's InnerProduct (src/include/alg/SpaceInnerProduct.hpp:14-20):
 * FP32 products and sums, each rounded, summed in ascending dimension order.
 * Every rank this library returns is the exact strict count under that rule.
 *
 * The path is the reference's "US" (uniform sample) pipeline:
 *   build: score GEMM -> per-user order statistics at sample ranks r_j
 *          -> sampled score table T (n_user x n_sample), resident in HBM;
 *   query: <u,q> for every user -> binary search of <u,q> in T[u] -> rank
 *          bounds -> bound pruning against the k-th bound -> exact ranks of
 *          the surviving users -> top-k by (rank, userID).
 *
 * Conventions: every entry point returns RKR_OK (0) or a negative status and
 * never exits the process; rkr_last_error() describes the last failure on a
 * context.  Host buffers are owned by the caller and read only during the
 * call.  Device buffers are owned by the context and stay resident until
 * rkr_destroy().  A context is driven by one host thread at a time.
 */
#ifndef RKR_H_
#define RKR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RKR_OK 0
#define RKR_ERR_INVALID (-1)   /* bad argument (shape, k, n_sample, ...)      */
#define RKR_ERR_STATE (-2)     /* call out of order (no data / no table)      */
#define RKR_ERR_NOMEM (-3)     /* device or host allocation failed / budget   */
#define RKR_ERR_HIP (-4)       /* HIP runtime error                           */
#define RKR_ERR_IO (-5)        /* file missing, short or malformed            */
#define RKR_ERR_NODEVICE (-6)  /* no usable GPU                               */

typedef struct rkr_ctx rkr_ctx;

/* Refinement forms (rkr_set_option RKR_OPT_FORM). */
#define RKR_FORM_PLAIN 0     /* exact FP32 score for every (user, item)        */
#define RKR_FORM_SCREENED 1  /* BF16 MFMA screen + exact recheck near <u,q>    */

/* Output orders (rkr_set_option RKR_OPT_OUTPUT). */
#define RKR_OUTPUT_EXACT 0      /* k smallest (exact rank, userID), ascending   */
#define RKR_OUTPUT_REFERENCE 1  /* US.hpp:173-186 layout: users the bounds alone
                                   settle first (userID order, rank = bound),
                                   then refined users by (rank, userID)         */

#define RKR_OPT_FORM 1
#define RKR_OPT_OUTPUT 2
#define RKR_OPT_QUERY_BATCH 3   /* max queries per device batch (0 = auto)      */
#define RKR_OPT_BUILD_BATCH 4   /* max users per build batch (0 = auto)         */
#define RKR_OPT_SCREEN_CHUNK 5  /* items per screened work-unit chunk (0 = no limit) */
#define RKR_OPT_DEDUP 6         /* 1 (default): byte-equal queries of one rkr_query
                                   call are answered once, with their largest k   */
#define RKR_OPT_SCREEN_ROWS 7   /* refine rows per wavefront of the screen: 32 or 64
                                   (0 = 32 for d <= 160, 64 above)                */
#define RKR_OPT_DENSE_TOP 8     /* query table that exact output searches, kept
                                   by rkr_build_table beside the unchanged sampled
                                   table: 2 (default) every rank below the largest
                                   D that fits 4096 columns; 1: 0..255, every 8th
                                   to 1023, every 32nd to 4095; 0: none.  Answers
                                   are unchanged; refinement shrinks.            */
#define RKR_OPT_QUERY_SLOTS 10  /* 2 (default): rkr_query calls of >= 512 queries
                                   that span two or more device batches run in two
                                   parts on two streams, each driven by its own host
                                   thread, so one part's host work between kernels
                                   overlaps the other's kernels; 1: one stream.
                                   Answers are unchanged.                        */
#define RKR_OPT_COARSE 9        /* 1 (default): screened form, exact output: the
                                   batch's <u,q> runs on BF16 matrix cores against
                                   7 guard columns per user, and only the users
                                   this cannot rule out are scored exactly and
                                   searched; answers are unchanged.             */
#define RKR_OPT_QSRP 11         /* 0 (default); 1: exact queries without pending
                                   updates screen users with the QSRP bounds that
                                   rkr_qsrp_build made (partial-dimension inner
                                   product bound + per-user regression rank
                                   bound) before exact scores and the table
                                   search; it replaces the coarse bounds.
                                   Answers are unchanged.                        */

/* Per-query counters; mirrors SingleQueryPerformance
 * (src/include/util/FileIO.hpp:150-173) plus the screening figures. */
typedef struct rkr_query_stats {
    int32_t query_id;
    int32_t n_prune_user;    /* users the bounds exclude                        */
    int32_t n_result_user;   /* users the bounds place in the answer            */
    int32_t n_refine_user;   /* users whose rank needed counting                */
    int32_t n_candidate;     /* users kept after bound pruning                  */
    int32_t reserved;
    int64_t ip_cost;         /* inner products evaluated (GEMV + refinement)    */
    int64_t n_rechecked;     /* screened form: (user, item) scores recomputed   */
    int64_t n_screened;      /* screened form: (user, item) scores screened     */
} rkr_query_stats;

/* Batch-level timings (device time from HIP events, milliseconds). */
typedef struct rkr_batch_timing {
    double total_ms;        /* host wall time of the rkr_query call            */
    double ip_ms;           /* <u,q> GEMV/GEMM                                 */
    double bound_ms;        /* table binary search                             */
    double prune_ms;        /* k-th bound, candidate compaction                */
    double refine_ms;       /* exact-rank refinement (screening kernel when screened) */
    double recheck_ms;      /* screened form: exact recheck of undecided items */
    double topk_ms;         /* top-k selection                                 */
    double h2d_ms, d2h_ms;  /* copies of queries and results                   */
    int64_t n_batches;
    int64_t n_refined;      /* (query, user) pairs counted                     */
    int64_t n_rechecked;
    int64_t n_screened;
    double max_screen_err;  /* screened form: max |screen - FP64| observed     */
    int64_t n_listed;       /* coarse bounds: (query, user) pairs scored exactly */
    int64_t n_fallback;     /* coarse bounds: queries that took the full path  */
    int64_t n_coarse_refined; /* coarse bounds: queries given a second, finer pass */
} rkr_batch_timing;

/* ---- context ------------------------------------------------------------ */

/* Opens a context on HIP device `device`.  mem_budget_bytes caps the device
 * memory the context may hold (0 = 90% of free memory).  Replaces the device
 * set-up of GPUComputeSort (src/include/score_computation/GPUComputeSort.hpp:79-110). */
int rkr_create(int device, size_t mem_budget_bytes, rkr_ctx **out);
void rkr_destroy(rkr_ctx *ctx);
const char *rkr_last_error(const rkr_ctx *ctx);
int rkr_set_option(rkr_ctx *ctx, int option, int64_t value);
int rkr_get_option(const rkr_ctx *ctx, int option, int64_t *value);
/* Waits for all work of the context. */
int rkr_sync(rkr_ctx *ctx);
/* Library version string. */
const char *rkr_version(void);

/* ---- data ---------------------------------------------------------------- */

/* L2-normalises users in place exactly as VectorMatrix::vectorNormalize
 * (src/include/struct/VectorMatrix.hpp:51-66): FP32 sequential sum of
 * squares, sqrt, rows with norm <= 0.001 left unchanged.  Host-side helper. */
int rkr_normalize_users(float *users, int64_t n_user, int32_t dim);

/* Uploads users and items (row-major FP32, caller-owned, copied).  When
 * normalize_users != 0, users are normalised as rkr_normalize_users first
 * (readData does this, src/include/util/VectorIO.hpp:57-61).  Replaces the
 * H2D copies of GPUComputeSort (GPUComputeSort.hpp:90-97) and the data
 * hand-over of US::BuildIndex (src/impl/US.hpp:246-261). */
int rkr_set_data(rkr_ctx *ctx, const float *users, int64_t n_user, const float *items, int64_t n_item,
                 int32_t dim, int normalize_users);

/* ---- sampled score table --------------------------------------------------- */

/* r_j = floor(j * (n_item-1) / n_sample), j < n_sample
 * (FindUniformSample, src/include/alg/RankBoundRefinement/FindSampleRank/UniformSample.hpp:17-25).
 * Fails when n_sample <= 0 or n_sample >= n_item (SampleSearch.hpp:35-38). */
int rkr_uniform_sample_ranks(int64_t n_item, int32_t n_sample, int32_t *out_ranks);

/* Builds T[u][j] = (r_j-th largest score of user u) on the device for every
 * user: the fused product of ComputeScoreTableBatch::ComputeSortItemsBatch
 * (src/include/score_computation/ComputeScoreTableBatch.hpp:79-92) and
 * SampleSearch::LoopPreprocess (SampleSearch.hpp:80-85) as driven by
 * BuildSampleIndexByComputation.cpp:111-144.  ranks must be strictly
 * increasing in [0, n_item).  build_ms (optional) receives device time. */
int rkr_build_table(rkr_ctx *ctx, const int32_t *ranks, int32_t n_sample, double *build_ms);

/* Installs a table computed elsewhere (n_user x n_sample, row-major,
 * rows non-increasing).  eps bounds |T - exact order statistic| under this
 * library's score rule (0 for tables built by the rule; the query path
 * widens its binary search by eps so answers stay exact). */
int rkr_set_table(rkr_ctx *ctx, const int32_t *ranks, int32_t n_sample, const float *table, float eps);

/* Copies table rows [user_begin, user_begin + n_rows) to out (n_rows x n_sample). */
int rkr_get_table(rkr_ctx *ctx, int64_t user_begin, int64_t n_rows, float *out);
int rkr_get_ranks(const rkr_ctx *ctx, int32_t *out_ranks, int32_t *out_n_sample);

/* Table file in the layout of SampleSearch::SaveIndex / LoadIndex
 * (SampleSearch.hpp:186-194, 229-241): size_t n_sample, size_t n_item,
 * size_t n_user, int32 ranks[n_sample], float table[n_user][n_sample].
 * rkr_save_table returns RKR_ERR_STATE while item inserts or deletes are
 * pending (the rows describe the base item set): rebuild first. */
int rkr_save_table(rkr_ctx *ctx, const char *path);
int rkr_load_table(rkr_ctx *ctx, const char *path, float eps);

/* ---- full score table ------------------------------------------------------ */

/* Scores of users [user_begin, user_begin + n_rows) against every current item,
 * each row sorted in descending order, into out[n_rows][n_item] (host):
 * ComputeScoreTableBatch::ComputeSortItemsBatch
 * (src/include/score_computation/ComputeScoreTableBatch.hpp:79-92).  Scores
 * follow the same rule as the table build.  The rows reach out through two
 * pinned host staging buffers of up to 32 MB each, allocated for the call,
 * with up to 16 host threads moving data out of them. */
int rkr_sorted_scores(rkr_ctx *ctx, int64_t user_begin, int64_t n_rows, float *out);

/* rkr_sorted_scores without any allocation in the call.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  out_dev is
 * caller-owned DEVICE memory on the context's GPU, n_rows x n_item floats,
 * row-major; the rows are identical to rkr_sorted_scores'.  The call only
 * enqueues work on the context's stream (rkr_sync waits for it).  The row
 * sort's scratch is a fixed 32 MiB device array per GPU, held until the
 * process ends: n_item may be at most 8,388,608 (RKR_ERR_INVALID above), and
 * rows run in chunks of 8,388,608 / n_item.  Every call on one GPU shares that
 * array, so calls that overlap on one GPU (two contexts' streams at once)
 * race; calls on one context's stream follow each other and are safe. */
int rkr_sorted_scores_noalloc(rkr_ctx *ctx, int64_t user_begin, int64_t n_rows, float *out_dev);

/* Writes the sorted score rows of the first n_rows users (n_rows < 0: every
 * user) to path in the layout of ReadAll's index file
 * (src/include/alg/DiskIndex/ReadAll.hpp:80-87, as
 * src/app/BuildScoreTableBatch.cpp:69-144 writes it): float[n_rows][n_item],
 * row-major, no header.  timing_ms (optional, 4 doubles) receives the
 * reference's split of the build time (BuildScoreTableBatch.cpp:113-127):
 * [0] device time of the score GEMMs, [1] device time of the row sorts,
 * [2] wall time of computing and copying the rows to the host ("compute sort
 * time"), [3] wall time of writing the file ("save index time").  The file
 * is written from two pinned staging buffers of up to 32 MB each, one while
 * the next copies, so [2] and [3] overlap. */
int rkr_save_score_table(rkr_ctx *ctx, const char *path, int64_t n_rows, double *timing_ms);

/* ---- QS sample-rank training ---------------------------------------------- */

/* Sample-rank methods of rkr_find_optimal_sample (the sample names of
 * SampleMethodName, src/include/util/NameTranslation.hpp:12-34). */
#define RKR_SAMPLE_OPTIMAL_PART 0     /* QS: candidates are every k-th rank and the rank before it */
#define RKR_SAMPLE_OPTIMAL_UNIFORM 1  /* candidates floor(j * (n_item-1) / n_sample_query)          */

/* SampleItem (src/app/build-index/QueryRankBySample.hpp:20-37): the first
 * n_sample_item item IDs of 0..n_item-1 after std::shuffle with
 * std::mt19937 seeded 0.  Host only. */
int rkr_sample_items(int64_t n_item, int32_t n_sample_item, int32_t *out_item_ids);

/* ComputeQueryRankByComputation
 * (src/app/build-index/QueryRankByComputation.hpp:122-242): for each sample
 * item s (an ID among the current items) and every user u,
 * rank(u, s) = #{items whose score beats <u, s>} (QueryRankGPU.hpp:63-92);
 * accu[s][r] = #{u : rank(u, s) <= r}, r = 0..n_item, so accu is
 * n_sample_item x (n_item + 1) int32 (host).  device_ms (optional) receives
 * the device time.  The counts leave through two pinned 32 MB staging
 * buffers; up to 16 host threads take the prefix sums. */
int rkr_query_rank_distribution(rkr_ctx *ctx, const int32_t *item_ids, int32_t n_sample_item, int32_t *accu,
                                double *device_ms);

/* ComputeSortKthRank (src/app/build-index/QueryRankBySample.hpp:103-134):
 * kth[s] = first r with accu[s][r] >= sample_topk; the samples sorted by kth
 * ascending (std::sort) give out_sort_sample_ids and out_sort_kth_rank.
 * Host only. */
int rkr_sort_kth_rank(const int32_t *accu, int64_t n_item, int32_t n_sample_item, int32_t sample_topk,
                      int32_t *out_sort_kth_rank, int32_t *out_sort_sample_ids);

/* FindOptimalSampleFast
 * (src/include/alg/RankBoundRefinement/FindSampleRank/OptimalSampleRankFast.hpp:102-305)
 * with the costs of QueryDistribution (OptimalSampleRank.hpp:17-330): picks
 * n_sample strictly increasing ranks minimising the summed number of users
 * whose rank lies in the sample bucket holding each query's k-th rank.
 * sort_kth_rank (ascending) and accu (n_sample_query x (n_item + 1), rows in
 * the same order) are what qdibc writes.  The DP runs on the device.
 * train_io_cost (optional) receives the minimum cost (0 when every candidate
 * fits). */
int rkr_find_optimal_sample(rkr_ctx *ctx, int64_t n_user, int64_t n_item, int32_t n_sample_query,
                            const int32_t *sort_kth_rank, const int32_t *accu, int32_t method, int32_t n_sample,
                            int32_t *out_ranks, int64_t *train_io_cost);

/* ---- queries --------------------------------------------------------------- */

/* Answers n_query reverse k-ranks queries (queries: n_query x dim, row-major
 * FP32, caller-owned).  topk[i] is query i's k (1 <= k <= min(n_user, max_k)).
 * Outputs are n_query x max_k, row-major; entries past topk[i] are -1.
 * out_qip receives <u,q> of each returned user (may be NULL); stats
 * (n_query entries) and timing may be NULL.  With RKR_OPT_DEDUP on, byte-equal
 * query vectors are computed once, with the largest k asked for them, and share
 * that run's stats (counters of the largest k); under RKR_OUTPUT_REFERENCE only
 * queries equal in vector and k share a run.  Replaces US::Index::Retrieval
 * (src/impl/US.hpp:96-213), which BaseIndex::Retrieval declares
 * (src/include/struct/MethodBase.hpp:14-30). */
int rkr_query(rkr_ctx *ctx, const float *queries, int32_t n_query, const int32_t *topk, int32_t max_k,
              int32_t *out_user, int32_t *out_rank, float *out_qip, rkr_query_stats *stats,
              rkr_batch_timing *timing);

/* Bytes of device memory the index holds (table + ranks), as
 * SampleSearch::IndexSizeByte (SampleSearch.hpp:246-250). */
int64_t rkr_index_size_bytes(const rkr_ctx *ctx);

/* Columns of the query table exact output searches (RKR_OPT_DENSE_TOP), or 0
 * when the context has none and searches the sampled table.  No counterpart
 * in the reference, whose only table is the sampled one. */
int32_t rkr_query_table_cols(const rkr_ctx *ctx);

/* ---- QSRP screen ------------------------------------------------------------
 * The reference's QSRPNormalLP method (src/impl/QSRPNormalLP.hpp:137-253):
 * users rotated into the SVD basis of the user matrix (SVD::PerformSVD,
 * src/include/alg/SVD.hpp:62-123), an inner-product interval from the first
 * check_dim dimensions plus a Cauchy-Schwarz bound of the rest
 * (PartDimPartNorm::IPBound, src/include/alg/QueryIPBound/PartDimPartNorm.hpp:
 * 52-74), and a per-user line from the normal CDF of a score to its sample
 * position, fitted on the host as the minimax (LP) fit of
 * NormalLinearRegressionLP (src/include/alg/RegressionPruning/
 * NormalLinearRegressionLP.hpp:149-283) and evaluated as ComputeRankBound
 * (:326-398).  Here every bound carries its rounding margins, so the screen
 * only ever removes users that cannot be in the answer; answers stay exact.
 *
 * rkr_qsrp_build fits the screen to the current users and sampled table
 * (SIGMA: the singular-value share that sets check_dim, the reference's
 * default 0.7); the screen is dropped when the table or the users change.
 * build_ms (may be NULL) receives the host time of the build. */
int rkr_qsrp_build(rkr_ctx *ctx, float sigma, double *build_ms);
/* check_dim of the built screen, and the users whose minimax line was not
 * decreasing and took the reference's fallback line.  RKR_ERR_STATE when
 * none is built. */
int rkr_qsrp_info(const rkr_ctx *ctx, int32_t *check_dim, int64_t *n_fallback_fit);
/* The screen's bucket bounds of every (query, user), before any exact score:
 * out[q * n_user + u] = b_lo | b_hi << 16 with b_hi <= #{T[u][j] > <u,q>} <=
 * b_lo over the searched table T (the query table when there is one).  For
 * checking the screen; rkr_query runs it inside the batch. */
int rkr_qsrp_screen(rkr_ctx *ctx, const float *queries, int32_t n_query, uint32_t *out);
/* The device's CDF (A&S 7.1.26, CDFPhi of NormalLinearRegressionLP.hpp:125-
 * 147, in FP32) at n points, for checking the error bound the screen assumes. */
int rkr_qsrp_phi(rkr_ctx *ctx, const float *z, int64_t n, float *out);
/* Host helper (no device): the minimax line through n points (x_i, y_i) that
 * rkr_qsrp_build fits per user -- the optimum of NormalLinearRegressionLP's LP
 * (minimise e subject to |a x_i + b - y_i| <= e).  RKR_ERR_INVALID when the
 * points have fewer than two distinct x. */
int rkr_qsrp_fit_line(const double *x, const double *y, int32_t n, double *a, double *b, double *e);

/* ---- updates (exact output only) -------------------------------------------
 * The table keeps describing the item set it was built from; each user keeps
 * sorted side arrays of its scores against inserted items and deleted base
 * items, and query bounds shift by #{inserted > q} - #{deleted > q}
 * (SampleSearchUpdate, src/include/alg/RankBoundRefinement/SampleSearchUpdate.hpp:72-166,212-279,
 * behind BaseUpdateIndex, src/include/struct/MethodBaseUpdate.hpp:15-40).
 * Answers stay exact; rkr_build_table over the current items folds the side
 * arrays into a new table.  Current item IDs: base items not deleted (in base
 * order) followed by inserted items not deleted (in insertion order). */

/* Appends n items (row-major FP32).  Replaces SampleSearchUpdate::InsertItem (:111-141). */
int rkr_insert_items(rkr_ctx *ctx, const float *items, int64_t n);
/* Removes items by current ID.  Replaces SampleSearchUpdate::DeleteItem (:142-166). */
int rkr_delete_items(rkr_ctx *ctx, const int64_t *item_ids, int64_t n);
/* Appends n users (normalised when the data were); their table rows are built
 * on the device.  Replaces SampleSearchUpdate::InsertUser (:72-87). */
int rkr_insert_users(rkr_ctx *ctx, const float *users, int64_t n);
/* Removes users by ID; later users shift down.  Replaces SampleSearchUpdate::DeleteUser (:89-109). */
int rkr_delete_users(rkr_ctx *ctx, const int64_t *user_ids, int64_t n);
/* Current sizes: users, items, pending inserted and deleted items. */
int rkr_get_counts(const rkr_ctx *ctx, int64_t *n_user, int64_t *n_item, int64_t *n_inserted, int64_t *n_deleted);
/* Sizes the update path's storage for up to n_user users and n_inserted
 * pending inserted items (and the items they add), so that updates within
 * these sizes neither allocate nor copy the large arrays (a serving loop
 * then never stalls on device allocation).  Never shrinks.  No counterpart in
 * the reference, whose updates reallocate host vectors. */
int rkr_reserve(rkr_ctx *ctx, int64_t n_user, int64_t n_inserted);

/* ---- serving loop -------------------------------------------------------------
 * Continuous batching on one context: requests (queries with their own k, and
 * user / item updates) enter one FIFO; worker threads, each with its own HIP
 * stream and scratch, batch the queries at the front of the queue, so batches
 * overlap on the device and requests that arrive meanwhile join the next one.
 * Updates are barriers (each query sees exactly the updates submitted before
 * it) or, in snapshot mode, are prepared while batches run (see
 * rkr_server_config).  Answers are identical to answering each query alone
 * against the state it reads.  The reference has
 * no serving loop (its drivers run queries one after another, US.hpp:117);
 * this replaces that loop for the serving form. */
typedef struct rkr_server rkr_server;

typedef struct rkr_server_config {
    int32_t max_batch;   /* queries per device batch (0 = 256); memory may split it */
    int32_t n_slots;     /* concurrent batches / worker threads (0 = 2, max 4)       */
    int32_t update_mode; /* RKR_SERVE_BARRIER (0) or RKR_SERVE_SNAPSHOT              */
    int32_t update_cus;  /* snapshot mode: CUs the update kernels may run on, so
                            batches keep the rest (0 = a quarter, < 0 = all)       */
    int32_t rebuild_pending; /* pending inserted + deleted items at which the update
                                side rebuilds the table over the current items (new
                                uniform sample ranks, same n_sample) so inserts keep
                                flowing past the side arrays' cap: 0 = 12288 (three
                                quarters of the 16384 cap), < 0 = never.  With
                                rebuilds on, an item update that would outgrow the
                                side arrays' reserved length (rkr_reserve) also
                                rebuilds first, so the arrays never grow while
                                serving                                            */
} rkr_server_config;

/* Barrier: an update waits for the batches in flight and blocks new ones, so
 * each query sees exactly the updates submitted before it.
 * Snapshot: updates are applied in submission order by their own thread while
 * batches keep running; a batch reads the state of the updates committed
 * before it started (an update's ticket completes at its commit), and the
 * buffers an update replaces are freed once the batches that may read them
 * have finished. */
#define RKR_SERVE_BARRIER 0
#define RKR_SERVE_SNAPSHOT 1

#define RKR_UPDATE_INSERT_USERS 1
#define RKR_UPDATE_INSERT_ITEMS 2
#define RKR_UPDATE_DELETE_USERS 3
#define RKR_UPDATE_DELETE_ITEMS 4

typedef struct rkr_request_info {
    int32_t status;       /* RKR_OK or the request's error                 */
    int32_t batch_size;   /* queries in the device batch that answered it  */
    double t_submit;      /* rkr_now() seconds: admitted to the queue      */
    double t_start;       /* taken by a worker                             */
    double t_done;        /* answer ready                                  */
    rkr_query_stats stats;
    char error[256];
} rkr_request_info;

/* Monotonic clock (seconds) used by the request timestamps. */
double rkr_now(void);
int rkr_server_start(rkr_ctx *ctx, const rkr_server_config *cfg, rkr_server **out);
/* Copies q (dim floats); the answer has k entries. */
int rkr_server_submit_query(rkr_server *srv, const float *q, int32_t k, uint64_t *ticket);
/* kind = RKR_UPDATE_*; vectors (n x dim) for inserts, ids for deletes (copied). */
int rkr_server_submit_update(rkr_server *srv, int kind, const float *vectors, const int64_t *ids, int64_t n,
                             uint64_t *ticket);
/* Blocks until the request is answered; writes k users / ranks / <u,q> for a
 * query (buffers may be NULL) and releases the ticket.  Returns the request's status. */
int rkr_server_wait(rkr_server *srv, uint64_t ticket, int32_t *out_user, int32_t *out_rank, float *out_qip,
                    rkr_request_info *info);
int rkr_server_poll(rkr_server *srv, uint64_t ticket, int *done);
int rkr_server_counters(rkr_server *srv, int64_t *n_batches, int64_t *n_queries, int64_t *n_updates);
/* Table rebuilds the serving loop has made (rkr_server_config.rebuild_pending). */
int rkr_server_rebuilds(rkr_server *srv, int64_t *n_rebuilds);
/* Answers everything queued, joins the workers and frees the server. */
int rkr_server_stop(rkr_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* RKR_H_ */
