// rkr_internal.hpp -- context layout and kernel launchers shared by the host
// code (rkr_api.cpp) and the HIP kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cmath>
#include <mutex>
#include <string>
#include <unordered_map>
#include <memory>
#include <vector>

#include "../../include/rkr.h"

namespace rkr {

// Device storage pads each vector to a multiple of 4 floats so rows can be
// read as float4; the score loops still run over exactly `dim` terms.
inline int32_t padded_dim(int32_t d) { return (d + 3) & ~3; }

// Largest n_sample the query kernels accept: bucket indices are packed as
// two uint16 values and the per-query histograms live in LDS.
constexpr int32_t kMaxSample = 4096;
// Largest k a query may ask for (top-k selection sorts <= kMaxTopk keys in LDS).
constexpr int32_t kMaxTopk = 4096;

struct DeviceBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// One in-flight query batch: its stream, timing events and scratch buffers.
// rkr_query uses the context's main slot; the serving loop gives each worker
// its own slot so batches overlap on the device.
// Page-locked host memory for a slot's staging copies (asynchronous DMA, no
// bounce buffer); grown between calls only.
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev[16] = {};
    DeviceBuf q_in, qip, bk, qinfo, cand, refine_idx, counts, out, recheck, small, hist;
    DeviceBuf q_split, coarse_info, coarse_list;  // coarse bounds
    PinnedBuf stage;                               // host side of the batch copies
};

struct Ctx {
    int device = 0;
    std::mutex mem_mu;       // guards budget accounting (worker threads allocate)
    std::mutex err_mu;
    size_t budget = 0;      // bytes the context may hold on the device
    size_t held = 0;        // bytes currently held
    hipStream_t stream = nullptr;
    hipEvent_t ev[16] = {};
    std::string err;

    int form = RKR_FORM_PLAIN;
    int output = RKR_OUTPUT_EXACT;
    int64_t query_batch = 0;
    int64_t build_batch = 0;
    int32_t screen_chunk = 0;   // items per screened work-unit chunk (0 = no limit)
    int32_t screen_rows = 0;    // refine rows per wavefront of the screen (0 = 32)
    int dedup = 1;              // answer byte-equal queries of one call once
    int coarse = 1;             // coarse bounds on matrix cores (RKR_OPT_COARSE)
    int query_slots = 2;        // rkr_query halves on main + aux (RKR_OPT_QUERY_SLOTS)
    int32_t ins_gmax[8] = {};   // see DataSnap::ins_gmax (kept by the update path)

    // data
    int64_t n_user = 0, n_item = 0;
    int32_t dim = 0, ld = 0;
    float *users = nullptr;     // n_user x ld, row-major (normalised)
    float *usersT = nullptr;    // panels of 64 users: [panel][ld][64]
    float *items = nullptr;     // n_item x ld, row-major (original order)
    std::vector<float> h_users; // host copy (normalised)
    double max_abs_user = 1.0;  // bound on |u_i| after normalisation

    // update path: the table describes the base item set; inserted items and
    // deleted base items are tracked per user in sorted side arrays
    bool normalize_users = true;
    float *items_base = nullptr;       // aliases `items` until the first item update
    int64_t n_item_base = 0;
    float *items_ins = nullptr;        // inserted items [n_ins][ld]
    float *items_del = nullptr;        // deleted base items [n_del][ld]
    int32_t n_ins = 0, n_del = 0;
    // side arrays [u_cap][ld_*], row u holds n_* scores, non-increasing; each
    // has a spare of the same shape so an update writes the next version
    // while batches read the current one
    float *side_ins = nullptr, *side_ins_spare = nullptr;
    float *side_del = nullptr, *side_del_spare = nullptr;
    int32_t ld_ins = 0, ld_del = 0;
    std::vector<int64_t> cur_src;      // current item -> base id (>= 0) or -(1 + inserted index)
    std::vector<uint8_t> base_deleted; // per base item
    // capacities (rows): users, panels, user norms, table rows and side-array
    // rows hold u_cap users; current items and their BF16 rows hold i_cap
    // items once an item update gave them their own buffers; items_ins holds
    // ins_cap rows.  New rows are appended in place: a batch only reads the
    // rows of its snapshot's counts.
    int64_t u_cap = 0, i_cap = 0, ins_cap = 0, del_cap = 0;
    DeviceBuf upd_scratch;             // score blocks of updates

    // table
    int32_t n_sample = 0;
    int32_t *ranks = nullptr;   // n_sample
    float *table = nullptr;     // n_user x n_sample
    float table_eps = 0.0f;
    std::vector<int32_t> h_ranks;
    // query table of exact output (built with the table): order statistics at
    // the sample ranks and at dense ranks of every row's top (mode 2, the
    // default: every rank up to the 4096-column cap; mode 1: all of 0..255,
    // every 8th to 1023, every 32nd to 4095).  Bounds from it settle most
    // high-ranking users without refinement; the sampled table is the subset
    // of its columns at qcol.  Absent (n_qs = 0) for installed tables.
    int32_t n_qs = 0;
    int32_t *qranks = nullptr;  // n_qs
    int32_t *qcol = nullptr;    // n_sample
    float *qtable = nullptr;    // [u_cap][n_qs]
    std::vector<int32_t> h_qranks;
    // guard rows of the coarse bounds: [u_cap][kGuardRow], the values of the
    // searched table at the columns of every guard set, -inf past each set's
    // length; kept with the table
    float *gvals = nullptr;
    int dense_top = 2;          // query table mode (RKR_OPT_DENSE_TOP)

    // screened form: hi/lo BF16 item rows (n_item x kp), norm bounds
    int32_t kp = 0;
    uint16_t *item_hi = nullptr, *item_lo = nullptr;
    float *item_norm = nullptr, *user_norm = nullptr;
    uint16_t *user_hi = nullptr, *user_lo = nullptr;  // BF16 hi/lo user rows ([u_cap] x kp)
    double max_abs_item = 0.0, max_abs_user_raw = 0.0;
    float screen_C = 0.0f, screen_abs = 0.0f;
    // the coarse first pass (one BF16 product): the FP32 accumulation and
    // rule rounding factors of its error scale, and (device, 4 bytes, for the
    // context's life) rho >= ||u - bf16(u)|| / ||u|| over every user loaded
    double coarse_g1 = 0.0, coarse_grule = 0.0;
    float *rho = nullptr;
    int n_cu = 256;

    // QSRP screen (rkr_qsrp_build): users rotated into the SVD basis, the query
    // rotation, and per-user norms and regression parameters (kQsrp* rows of
    // n_user floats); built for the current table and users, dropped when
    // either changes
    int qsrp = 0;                   // use the screen in exact queries (RKR_OPT_QSRP)
    float *qsrp_users = nullptr;    // n_user x qsrp_ldr
    float *qsrp_rot = nullptr;      // dim rows of ld: row r = s_r w_r
    float *qsrp_par = nullptr;      // [kQsrpPar][n_user]
    int32_t qsrp_cd = 0, qsrp_ldr = 0;
    int64_t qsrp_n_user = 0, qsrp_fallback = 0;
    float qsrp_sigma = 0.0f, qsrp_gcd = 0.0f, qsrp_crot = 0.0f;

    // query scratch of rkr_query (grown on demand) and build scratch
    Slot main;
    Slot aux;                   // second stream of split rkr_query calls, made on first use
    bool aux_ready = false;
    DeviceBuf misc;

    // buffers an update thread freed while batches may still read them
    // (snapshot-mode server): released once every batch that started before
    // the update committed has finished (guarded by mem_mu)
    struct Retired {
        void *p;
        size_t bytes;
    };
    std::vector<Retired> retired;
    std::unordered_map<void *, size_t> alloc_bytes;  // size of every live allocation (mem_mu)
};

// Rows of Ctx::qsrp_par: regression line a, b, the score mean and 1/sigma of
// its normal CDF argument, the bound's error in sample positions, the row's
// largest and smallest sampled scores, and the norms of the rotated user's
// left and right parts and of the user.
enum {
    kQsrpA = 0, kQsrpB, kQsrpMu, kQsrpInvSigma, kQsrpErr, kQsrpTmax, kQsrpTmin, kQsrpNrmL, kQsrpNrmR, kQsrpNrmU,
    kQsrpPar
};
// Bound on |Phi_device(z) - Phi(z)| for the FP32 evaluation of the A&S CDF,
// argument rounding included (checked against FP64 erf by the GPU tests).
constexpr double kQsrpPhiDevErr = 2e-6;
// Frees the QSRP screen's buffers (the table or the users changed).
void free_qsrp(Ctx *c);

// Frees made by this thread go to Ctx::retired instead of hipFree.
void set_defer_frees(bool on);
// Releases the retired buffers (call only when no batch can still read them).
void release_retired(Ctx *c);

// Largest |x| over a row-major matrix; -1 when a value is NaN or infinite.
inline double max_abs(const float *x, size_t n) {
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const float v = x[i];
        if (!std::isfinite(v)) return -1.0;
        const double a = std::fabs((double)v);
        if (a > m) m = a;
    }
    return m;
}

// Update-path rank bound values stay below kMaxBoundValue (two-level 11-bit
// histograms); kNoBound marks a (query, user) pair with no bound computed,
// one the coarse bounds already ruled out.
constexpr int64_t kMaxBoundValue = 1 << 22;
constexpr int32_t kNoBound = (1 << 22) - 1;

// Every partial sum of a score is bounded by dim * max|a| * max|b|; keep it
// well inside FP32 range so no score overflows.
constexpr double kScoreLimit = 1e37;

// Absolute error floor of the screened form: products a matrix core may
// flush as denormals, each < 2^-126 * max|other factor|, over 3*kp terms.
inline float screen_abs_floor(int32_t kp, double max_abs_user, double max_abs_item) {
    const double absf = 2.0 * 3.0 * kp * std::ldexp(1.0, -126) * (1.0 + max_abs_user) * (1.0 + max_abs_item);
    return std::nextafter((float)absf, INFINITY);
}

struct GuardCols {
    int32_t col[8];
    int32_t n;
};
// Guard sets of the coarse bounds: set 0 holds the first pass's columns, set
// 1 + beta the second pass's columns inside guard bin beta (empty when the bin
// has no column of its own).  Each user keeps the values of all of them in one
// row of kGuardRow floats, set s at floats 8 s .. 8 s + 7 (-inf past its n).
constexpr int kGuardSets = 8;
// Workgroups per CU of the coarse first pass (256 threads, users' BF16 hi
// halves in VGPRs) for Kp = 32 ks: its launch bound and its persistent grid.
// (Five fit at ks <= 5 only by spilling the staged query pieces: slower.)
constexpr int coarse_first_pass_wgs(int ks) { return ks >= 7 ? 3 : 4; }
constexpr int kGuardRow = 8 * kGuardSets;
struct GuardSets {
    GuardCols set[kGuardSets];
};


// Everything a query batch reads from the data and table state, captured
// when the batch starts.  Outside a snapshot-mode server it is the context's
// current state; in one, the last committed update's state, so updates can
// be prepared while batches run.
struct DataSnap {
    int64_t n_user = 0, n_item = 0, n_item_base = 0;
    const float *users = nullptr, *usersT = nullptr, *items = nullptr;
    const float *user_norm = nullptr, *item_norm = nullptr;
    const float *rho = nullptr;      // >= ||u - bf16(u)|| / ||u|| over the users (device)
    const uint16_t *user_hi = nullptr, *user_lo = nullptr;
    const uint16_t *item_hi = nullptr, *item_lo = nullptr;
    const float *table = nullptr, *side_ins = nullptr, *side_del = nullptr;
    int32_t n_ins = 0, n_del = 0, ld_ins = 0, ld_del = 0;
    double max_abs_user = 1.0;
    float screen_abs = 0.0f;
    const float *qtable = nullptr;   // query table (exact output), or null
    const int32_t *qranks = nullptr;
    int32_t n_qs = 0;
    int32_t n_guard = 0;             // guard columns of the coarse bounds
    int32_t gcol[8] = {};
    const float *gvals = nullptr;    // guard values per user (kGuardRow floats, all sets), or null
    // the sampled table's ranks (device), count and eps, the query table's
    // column of each sampled column, and the searched table's ranks on the
    // host (the guard placement): captured with the table, so a batch never
    // mixes a table with the ranks of one rebuilt after it started
    const int32_t *ranks = nullptr;
    int32_t n_sample = 0;
    float table_eps = 0.0f;
    const int32_t *qcol = nullptr;
    std::shared_ptr<const std::vector<int32_t>> hrk;
    GuardSets gsets{};               // the coarse guard sets of that table
    // pending inserts: the most inserted items any user has above its value
    // at guard column j (a bound on the rank shift at that guard)
    int32_t ins_gmax[8] = {};
    // QSRP screen (null when not built)
    const float *qsrp_users = nullptr, *qsrp_rot = nullptr, *qsrp_par = nullptr;
    int32_t qsrp_cd = 0, qsrp_ldr = 0;
    float qsrp_gcd = 0.0f, qsrp_crot = 0.0f;
};
DataSnap snap_of(const Ctx *c);
GuardSets guard_sets(const Ctx *c);
// Guard columns of the coarse bounds for the context's searched table.
void guard_columns(const Ctx *c, int32_t *col, int32_t &n);
// The table over the current items at sampled ranks S (build_table), and the
// same at uniform ranks of the current n_sample (rebuild_table): the pending
// item updates fold into the new table; keep_update_buffers leaves their
// buffers allocated for the updates that follow.
int build_table(Ctx *c, const int32_t *ranks, int32_t S, double *build_ms, bool keep_update_buffers);
int rebuild_table(Ctx *c);
// Fills Ctx::gvals rows [u0, u0 + n) from the searched table.
int refresh_guard_values(Ctx *c, int64_t u0, int64_t n);
// Recomputes Ctx::ins_gmax after inserted items or users changed.
int refresh_ins_guard_max(Ctx *c);
// Ranks of the query table: the sample ranks and the dense top ranks below
// n_item, ascending; empty when they would exceed kMaxSample.
// mode 1: all of 0..255, every 8th to 1023, every 32nd to 4095; mode 2: every
// rank below the largest D that keeps the table within kMaxSample columns.
std::vector<int32_t> query_table_ranks(const std::vector<int32_t> &sample, int64_t n_item, int mode);

// ---- error and memory helpers (rkr_api.cpp) --------------------------------
int set_err(Ctx *c, int code, const std::string &msg);
int dev_alloc(Ctx *c, void **p, size_t bytes, const char *what);
void dev_free(Ctx *c, void *p, size_t bytes);
int ensure(Ctx *c, DeviceBuf &b, size_t bytes, const char *what);
int ensure_pinned(Ctx *c, PinnedBuf &b, size_t bytes);
void free_pinned(PinnedBuf &b);
// cu_mask (optional): the CUs the slot's stream may use
int slot_init(Ctx *c, Slot &sl, bool own_stream, const std::vector<uint32_t> *cu_mask = nullptr);
void slot_free(Ctx *c, Slot &sl);
// Runs validated queries through one slot (batches sized by free memory).
int query_on_slot(Ctx *c, Slot &sl, const DataSnap &d, const float *queries, int32_t n_query, const int32_t *topk,
                  int32_t max_k, int32_t *out_user, int32_t *out_rank, float *out_qip, rkr_query_stats *stats,
                  rkr_batch_timing *timing);
int validate_queries(Ctx *c, const DataSnap &d, const float *queries, int32_t n_query, const int32_t *topk,
                     int32_t max_k);
template <class T>
int dev_alloc_t(Ctx *c, T **p, size_t n, const char *what) {
    return dev_alloc(c, reinterpret_cast<void **>(p), n * sizeof(T), what);
}
#define RKR_HIP(ctx, call)                                                            \
    do {                                                                              \
        hipError_t e__ = (call);                                                      \
        if (e__ != hipSuccess)                                                        \
            return ::rkr::set_err((ctx), RKR_ERR_HIP,                                 \
                                  std::string(#call) + ": " + hipGetErrorString(e__)); \
    } while (0)

// ---- kernel launchers (k_gemm.hip) ----------------------------------------
// C = A * B^T with the sequential FP32 score rule.  A rows are gathered
// through a_rows when non-null.  Store epilogue: C[n * ldc_n + m] (query-major
// when A = users, B = queries).
hipError_t launch_scores_store(hipStream_t s, const float *A, const int32_t *a_rows, int64_t M, int32_t lda,
                               const float *B, int64_t N, int32_t ldb, int32_t K, float *C, int64_t ldc,
                               bool c_row_major_m);
// Small-batch <u,q>: qip[b * n_user + u] for b < nq (nq <= 16), users in
// 64-user panels.
hipError_t launch_gemv_panels(hipStream_t s, const float *usersT, int64_t n_user, int32_t ld, int32_t K,
                              const float *q, int32_t nq, float *qip);
// Refinement count: cnt[r] = #{item p : score(user(r), p) > thr(r)} for every
// refine row r < *n_rows_dev (device-side count, persistent grid).
hipError_t launch_refine_count(hipStream_t s, const float *users, const float *items, int64_t n_item,
                               int32_t ld, int32_t K, const int32_t *rows_user, const float *rows_thr,
                               const int32_t *n_rows_dev, int64_t max_rows, int32_t *cnt);
// Transpose users into 64-user panels.
hipError_t launch_make_panels(hipStream_t s, const float *users, int64_t n_user, int32_t ld, float *usersT);

// ---- screened refinement (k_screen.hip) -------------------------------------
struct ScreenLaunch {
    const float *users, *user_norm;
    const uint16_t *user_hi, *user_lo;  // BF16 split of the user rows (n_user x Kp)
    const uint16_t *item_hi, *item_lo;
    const float *item_norm;
    int64_t n_item;
    int32_t ld, K, Kp;
    const int32_t *rows_user;
    const float *rows_thr;
    const int32_t *n_rows_dev;
    int64_t max_rows;
    float C, abs_floor;
    int32_t *cnt;
    int32_t *recheck_n;
    void *recheck;          // int4 records (row, item, screened score bits, 0)
    int64_t recheck_cap;
    int64_t target_units;   // work units wanted to fill the device
    int64_t grid_cap;       // workgroups of the persistent grid
    int32_t chunk_items;    // most items per work-unit chunk (L2-resident slice)
    int32_t rows_per_wave;  // 32 or 64 refine rows per wavefront
};
// BF16 hi/lo rows (Kp columns, zero padded) and, when norm is non-null, row
// norm upper bounds; used for items, queries and users.  With cs (queries of
// the coarse first pass): each row's error scale of the one-product
// estimate, given rho (device) >= ||u - bf16(u)|| / ||u|| over the users and
// the accumulation and rule rounding factors g1, g_rule.
hipError_t launch_prep_items(hipStream_t s, const float *items, int64_t n_item, int32_t ld, int32_t K, int32_t Kp,
                             uint16_t *hi, uint16_t *lo, float *norm, float *cs = nullptr, const float *rho = nullptr,
                             double g1 = 0.0, double g_rule = 0.0);
// Users' norm upper bounds; with rho (device), raises *rho to at least every
// user's ||u - bf16(u)|| / ||u||.
hipError_t launch_user_norms(hipStream_t s, const float *users, int64_t n_user, int32_t ld, int32_t K, float *norm,
                             float *rho);
hipError_t launch_refine_screen(hipStream_t s, const ScreenLaunch &L);
// Guard columns of the coarse bounds: n (1..7) ascending table columns.

// Coarse buckets of a batch (screened <u,q> against the guard columns of
// every user's table row): cb[row(b) * ldc + u] = c_lo | c_hi << 4, with
// row(b) = qmap[b] (or b when qmap is null), which also picks the query rows.
struct CoarseLaunch {
    const float *users, *user_norm, *table;
    const float *gvals = nullptr;       // n_user x kGuardRow guard values, or null: read the table
    const uint16_t *user_hi, *user_lo;  // BF16 split of the user rows (n_user x Kp)
    int64_t n_user;
    int32_t S;
    GuardCols gc;
    const uint16_t *q_hi, *q_lo;
    const float *q_norm;
    const int32_t *qmap;
    int32_t nq, ld, K, Kp;
    float C, abs_floor;
    uint8_t *cb;            // rows of ldc bytes, ldc % 128 == 0
    int64_t ldc;
    int64_t spare_row;      // a row no query owns (written by padding lanes)
    int64_t grid_cap;
    // second pass: query b uses sets[qset[b]] (<= 7 sets) instead of gc;
    // set_id[i] is the guard-row set (1 + beta) of launch set i
    const GuardCols *sets = nullptr;
    const int32_t *qset = nullptr;
    int32_t n_sets = 0;
    int32_t set_id[7] = {};
};
hipError_t launch_ip_coarse(hipStream_t s, const CoarseLaunch &L);
hipError_t launch_recheck(hipStream_t s, const float *users, const float *items, int32_t ld, int32_t K,
                          const int32_t *rows_user, const float *rows_thr, const int32_t *rows_query,
                          const int32_t *recheck_n, const void *recheck, int64_t cap, int32_t *cnt,
                          int32_t *q_recheck, uint32_t *max_err_bits);
int screen_max_kp();

// ---- selection (k_select.hip) ---------------------------------------------
// table[(u0 + b) * S + j] = r_j-th largest of scores[b * n_item + ...], b < nb.
hipError_t launch_select_rows(hipStream_t s, const float *scores, int64_t nb, int64_t n_item,
                              const int32_t *ranks, int32_t S, float *table_out);

// ---- full sorted rows (k_sort.hip) ------------------------------------------
// out[b * ld + i] = i-th largest of src[b * ld_src + 0..n_item), b < nb; tmp is
// scratch of the same shape as out.  out may be src (ld == ld_src).  At most
// grid_cap workgroups (one fits per CU), each taking rows in turn.
hipError_t launch_sort_rows(hipStream_t s, const float *src, int64_t ld_src, int64_t nb, int64_t n_item, float *out,
                            float *tmp, int64_t ld, int grid_cap);

// ---- QS sample-rank training (k_train.hip) ----------------------------------
// hist[s * (n_item + 1) + #{i : sorted[u * ld + i] > q[u * ldq + s]}] += 1 for
// u < nb, s < ns (sorted rows descending).
hipError_t launch_rank_hist(hipStream_t s, const float *sorted, int64_t ld, int64_t nb, int64_t n_item, const float *q,
                            int64_t ldq, int32_t ns, int32_t *hist);
// Transit-cost triangle T[lb * (lb - 1) / 2 + ub], 0 <= ub < lb < R, from the
// query positions pos[j], End(j) and the prefix sums kept at the breakpoints:
// C[col[j'] * R + j] = sum over queries q < pos[j'] of accu[q][rank_j].
hipError_t launch_transit(hipStream_t s, const int64_t *C, int32_t Q, const int32_t *pos, const int32_t *col,
                          const int64_t *end_cost, int32_t R, int64_t *T);
// One DP layer: dp_cur[r] = min(dp_prev[r], min_{p<r} dp_prev[p] + T[r][p]).
hipError_t launch_dp_layer(hipStream_t s, const int64_t *T, const int64_t *dp_prev, const int32_t *pos_prev, int32_t R,
                           int64_t *dp_cur, int32_t *pos_cur);

// ---- query pipeline (k_query.hip) -------------------------------------------
struct QueryInfo {          // one per query of a device batch
    int32_t topk;
    int32_t beta;           // bucket index of the k-th smallest lower bound
    int32_t gamma;          // largest upper-bound bucket that stays a candidate
    int32_t n_cand;         // candidates kept
    int32_t n_settled;      // reference output: users settled by the bounds
    int32_t n_refine;       // candidates whose rank is counted
    int32_t mode;           // classification case (reference output)
    int32_t pad;
    int32_t fill;           // candidates written so far (chunked compaction)
    int32_t n_screen;       // users a screen kept and scored exactly (QSRP)
    int64_t cand_off;       // offset of this query's candidates
};

// Table rows staged per workgroup by the batched bounds kernels: 16 rows of
// S + 1 floats (<= 32 KB at S = 511), fewer for long rows; 0 when one row
// does not fit in 64 KB.
inline int bounds_rows_per_wg(int32_t S) {
    int ut = 16;
    while (ut > 1 && (int64_t)ut * (S + 1) * 4 > 40960) ut >>= 1;
    return (int64_t)ut * (S + 1) * 4 <= 65536 ? ut : 0;
}

#ifdef __HIPCC__
// Copies nu rows of S floats (source row stride src_ld) to LDS rows of
// stride S + 1, 16-byte loads where the rows allow.
__device__ __forceinline__ void stage_rows(float *rows, const float *__restrict__ src, int64_t nu, int32_t S,
                                           int64_t src_ld) {
    if ((S & 3) == 0 && (src_ld & 3) == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
        const int32_t q = S >> 2;
        for (int64_t i = threadIdx.x; i < nu * q; i += blockDim.x) {
            const int64_t r = i / q, c = (i % q) * 4;
            const float4 v = *reinterpret_cast<const float4 *>(src + r * src_ld + c);
            float *d = rows + r * (S + 1) + c;
            d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
        }
    } else {
        for (int64_t i = threadIdx.x; i < nu * S; i += blockDim.x) rows[(i / S) * (S + 1) + i % S] = src[(i / S) * src_ld + i % S];
    }
}
#endif

// Users per chunk of the chunked prune kernels: enough workgroups per query
// to fill the device for small batches, one chunk per query for large ones.
inline int64_t chunk_users(int64_t n_user, int32_t nq, int n_cu) {
    const int64_t target = (int64_t)(n_cu > 0 ? n_cu : 1) * 4;
    int64_t n_chunk = (target + nq - 1) / (nq > 0 ? nq : 1);
    n_chunk = n_chunk < 1 ? 1 : n_chunk;
    const int64_t max_chunk = (n_user + 4095) / 4096;
    if (n_chunk > max_chunk) n_chunk = max_chunk < 1 ? 1 : max_chunk;
    int64_t cu = (n_user + n_chunk - 1) / n_chunk;
    cu = (cu + 1023) / 1024 * 1024;
    return cu < 1024 ? 1024 : cu;
}
// uint32 words of the global histograms of a batch of nq queries
inline size_t hist_words(int32_t nq, int32_t S) {
    const size_t a = (size_t)nq * 2 * (size_t)(S + 1), b = (size_t)nq * 4 * 2048;
    return a > b ? a : b;
}

hipError_t launch_bounds(hipStream_t s, const float *qip, const float *table, int64_t n_user, int32_t S,
                         int32_t nq, float eps, int ref_mode, uint32_t *bk);
hipError_t launch_hist_kth(hipStream_t s, const uint32_t *bk, int64_t n_user, int32_t S, const int32_t *ranks,
                           int64_t n_item, int32_t nq, int output_mode, int same, int64_t cu, uint32_t *ghist,
                           QueryInfo *qi);
// Coarse-bound pipeline of exact output (k_query.hip): guard values per
// user, the coarse k-th bound and list size per query (cq[b] = {beta_c,
// count, offset, fill}), the listed users, their exact scores and bucket
// pairs with the exact k-th bound (into qi), and the final candidates.
// cb rows are ldc bytes (ldc % 16 == 0, bytes past n_user unread); cu % 1024 == 0
// hcnt[b * 8 + c] = #{users with c_hi == c}, c < 7
hipError_t launch_coarse_kth(hipStream_t s, const uint8_t *cb, int64_t ldc, int64_t n_user, int32_t nq, int64_t cu,
                             int32_t n_guard, const int32_t *qmap, uint32_t *part, const QueryInfo *qi, int32_t *cq,
                             int32_t *hcnt);
// Update path: exact rank bound values of the listed pairs (lbs / ubs /
// qip at b * n_user + u; other pairs keep what the caller filled in).
hipError_t launch_coarse_exact_upd(hipStream_t s, const float *users, const float *queries, int32_t ld, int32_t dim,
                                   const float *table, int32_t S, const int32_t *ranks, int64_t n_item_base,
                                   float eps, const GuardCols *gsets, const int32_t *gset, const int32_t *cl_u,
                                   const int32_t *cl_b, int64_t n, const float *ins, int32_t n_ins, int32_t ld_ins,
                                   const float *del, int32_t n_del, int32_t ld_del, int64_t n_user, int32_t *lbs,
                                   int32_t *ubs, float *qip);
// max over users of #{inserted scores > T[u][gc.col[j]]} into out[j] (zeroed here)
hipError_t launch_ins_guard_max(hipStream_t s, const float *table, int64_t n_user, int32_t S, const GuardCols &gc,
                                const float *ins, int32_t n_ins, int32_t ld_ins, int32_t *out);
// chunk-major list: loc (n_chunk x nq), ctot and cbase (n_chunk) are scratch;
// total (optional) receives the list's length
hipError_t launch_coarse_compact(hipStream_t s, const uint8_t *cb, int64_t ldc, int64_t n_user, int32_t nq,
                                 int64_t cu, const uint32_t *part, const int32_t *cq, int32_t *loc, int32_t *ctot,
                                 int32_t *cbase, int32_t *cl_u, int32_t *cl_b, int32_t *total);
// cq[b] = {-1, 0, -, 0} for a query with no k-th bound, a count over cap or a
// count over the bound in field 2 (the full path answers it)
hipError_t launch_coarse_mark(hipStream_t s, int32_t *cq, int32_t nq, int64_t cap);
hipError_t launch_coarse_exact(hipStream_t s, const float *users, const float *queries, int32_t ld, int32_t dim,
                               const float *table, int32_t S, const int32_t *ranks, int64_t n_item, float eps,
                               const GuardCols *gsets, const int32_t *gset, int32_t nq, const int32_t *cl_u,
                               const int32_t *cl_b, int64_t n, const int32_t *n_dev, uint32_t *bk_c, float *s_c,
                               uint32_t *ghist, QueryInfo *qi);
// n_dev (optional): the list's length on the device (at most n; n sizes the grid)
hipError_t launch_coarse_final(hipStream_t s, const int32_t *cl_u, const int32_t *cl_b, const uint32_t *bk_c,
                               const float *s_c, int64_t n, const int32_t *n_dev, int32_t S, const int32_t *ranks,
                               int64_t n_item,
                               int64_t n_user, QueryInfo *qi, int32_t *cand_user, int32_t *cand_rank,
                               int32_t *refine_total, int32_t *refine_cand, int32_t *refine_user,
                               int32_t *refine_query, float *refine_thr, float *qip);
hipError_t launch_offsets(hipStream_t s, QueryInfo *qi, int32_t nq, int32_t *refine_total);
hipError_t launch_compact(hipStream_t s, const uint32_t *bk, const float *qip, int64_t n_user, int32_t S,
                          const int32_t *ranks, int64_t n_item, int32_t nq, int output_mode, int64_t cu,
                          QueryInfo *qi, int32_t *cand_user, int32_t *cand_rank, int32_t *refine_total,
                          int32_t *refine_cand, int32_t *refine_user, int32_t *refine_query, float *refine_thr);
hipError_t launch_scatter_counts(hipStream_t s, const int32_t *refine_total, const int32_t *refine_cand,
                                 const int32_t *cnt, int64_t max_rows, int32_t *cand_rank);
hipError_t launch_topk(hipStream_t s, const QueryInfo *qi, int32_t nq, const int32_t *cand_user,
                       const int32_t *cand_rank, const float *qip, int64_t n_user, int32_t max_k,
                       int32_t *out_user, int32_t *out_rank, float *out_qip, int64_t *scratch_keys);

// ---- update path (k_update.hip) ---------------------------------------------
// QSRP screen kernels (k_qsrp.hip).  qn holds 3 floats per query: upper bounds
// of the norms of the rotated query's left and right parts and of the query.
hipError_t launch_phi_eval(hipStream_t s, const float *z, int64_t n, float *out);
hipError_t launch_qsrp_qnorms(hipStream_t s, const float *qrot, int32_t ldr, const float *q, int32_t ld, int32_t d,
                              int32_t cd, int32_t nq, float *qn);
hipError_t launch_qsrp_bounds(hipStream_t s, const float *left, int64_t n_user, int32_t nq, const float *par,
                              const float *qn, float gcd, float crot, float absf, int32_t S, const int32_t *qcol,
                              int32_t S_T, uint32_t *bk);
hipError_t launch_qsrp_exact(hipStream_t s, const float *users, const float *q_in, int32_t ld, int32_t d,
                             const float *T, int32_t S, float eps, int64_t n_user, int32_t nq, QueryInfo *qi,
                             float *qip, uint32_t *bk);

// gv[u * kGuardRow + 8 s + j] = T[u][gs.set[s].col[j]] (-inf for j >= its n),
// rows [u0, u0 + n_rows).
hipError_t launch_gather_guards(hipStream_t s, const float *T, int64_t S, const GuardSets &gs, int64_t u0,
                                int64_t n_rows, float *gv);

// The sort scratch of rkr_sorted_scores_noalloc: looks up its device address
// on `device` (current) once per device.
hipError_t noalloc_init(int device);

hipError_t launch_sort_rows_desc(hipStream_t s, float *x, int64_t n_rows, int32_t n, int64_t ld);
hipError_t launch_merge_rows(hipStream_t s, const float *a, int64_t lda, int32_t na, const float *b, int64_t ldb,
                             int32_t nb, float *out, int64_t ldo, int64_t n_rows);
hipError_t launch_subtract_rows(hipStream_t s, const float *a, int64_t lda, int32_t na, const float *b, int64_t ldb,
                                int32_t nb, float *out, int64_t ldo, int64_t n_rows);
hipError_t launch_bounds_upd(hipStream_t s, const float *qip, const float *table, int64_t n_user, int32_t S,
                             const int32_t *ranks, int64_t n_item_base, float eps, const float *ins, int32_t n_ins,
                             int32_t ld_ins, const float *del, int32_t n_del, int32_t ld_del, int32_t nq,
                             int32_t *lbs, int32_t *ubs);
hipError_t launch_kth_values(hipStream_t s, const int32_t *lbs, const int32_t *ubs, int64_t n_user, int32_t nq,
                             int64_t cu, uint32_t *ghist, QueryInfo *qi);
hipError_t launch_compact_values(hipStream_t s, const int32_t *lbs, const int32_t *ubs, const float *qip,
                                 int64_t n_user, int32_t nq, int64_t cu, QueryInfo *qi, int32_t *cand_user,
                                 int32_t *cand_rank, int32_t *refine_total, int32_t *refine_cand,
                                 int32_t *refine_user, int32_t *refine_query, float *refine_thr);
// dst[r][j] = src[r][cols[j]] for r < n_rows, j < n_cols
hipError_t launch_gather_cols(hipStream_t s, const float *src, int64_t ld_src, const int32_t *cols, int32_t n_cols,
                              int64_t n_rows, float *dst, int64_t ld_dst);
hipError_t launch_gather_rows(hipStream_t s, const float *src, int64_t lds, const int64_t *idx, int64_t n_rows,
                              int64_t width, float *dst, int64_t ldd);

}  // namespace rkr
