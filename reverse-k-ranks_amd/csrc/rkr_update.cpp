// rkr_update.cpp -- the update path of the C ABI: inserting and deleting
// users and items while the sampled table stays resident.
//
// The table keeps describing the base item set it was built from.  Per user,
// sorted side arrays hold the scores of inserted items and of deleted base
// items (SampleSearchUpdate.hpp:72-166), and the query path shifts every rank
// bound by #{inserted > q} - #{deleted > q} (k_update.hip).  Refinement counts
// over the current item set, which is rebuilt as
//   [base items not deleted, in base order] ++ [inserted items, in order]
// after every item update; item IDs given to rkr_delete_items refer to that
// order.  rkr_build_table over the current items folds everything back into
// a fresh base.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rkr_internal.hpp"

using namespace rkr;

struct rkr_ctx : public rkr::Ctx {};

namespace {

constexpr int64_t kSortChunk = 1024;        // items scored per side-array merge
constexpr int64_t kMaxSideItems = 16384;    // per-user side array length before a rebuild is required
constexpr int64_t kMaxBoundValue = 1 << 22; // bound values must stay below this (k_kth_values)

int upload_padded(Ctx *c, const float *src, int64_t n, float **dst, const char *what) {
    const int32_t ld = c->ld, d = c->dim;
    int rc = dev_alloc_t(c, dst, (size_t)std::max<int64_t>(n, 1) * ld, what);
    if (rc) return rc;
    std::vector<float> tmp((size_t)n * ld, 0.0f);
    for (int64_t r = 0; r < n; ++r) std::memcpy(&tmp[(size_t)r * ld], src + (size_t)r * d, sizeof(float) * d);
    if (n) RKR_HIP(c, hipMemcpy(*dst, tmp.data(), sizeof(float) * (size_t)n * ld, hipMemcpyHostToDevice));
    return RKR_OK;
}

int check_updatable(Ctx *c) {
    if (!c->users) return set_err(c, RKR_ERR_STATE, "update: no data");
    if (!c->table) return set_err(c, RKR_ERR_STATE, "update: no table (build or load one first)");
    return RKR_OK;
}

// scores [nu][n] of users [u0, u0+nu) against the rows of X ([n][ld]),
// rows sorted non-increasing.
int score_sorted_block(Ctx *c, int64_t u0, int64_t nu, const float *X, int64_t n, float *out) {
    RKR_HIP(c, launch_scores_store(c->stream, c->users + u0 * c->ld, nullptr, nu, c->ld, X, n, c->ld, c->dim, out, n,
                                   true));
    RKR_HIP(c, launch_sort_rows_desc(c->stream, out, nu, (int32_t)n, n));
    return RKR_OK;
}

// Adds (sign > 0) or removes (sign < 0) the scores of item rows X ([n][ld])
// to every user's side array `side` ([U][n_side]) and returns the new array.
int update_side(Ctx *c, float **side, int32_t *n_side, const float *X, int64_t n, int sign) {
    const int64_t U = c->n_user;
    for (int64_t x0 = 0; x0 < n; x0 += kSortChunk) {
        const int64_t m = std::min<int64_t>(kSortChunk, n - x0);
        const int64_t n_new = sign > 0 ? (int64_t)*n_side + m : (int64_t)*n_side - m;
        float *block = nullptr, *next = nullptr;
        int rc;
        if ((rc = dev_alloc_t(c, &block, (size_t)U * m, "update score block"))) return rc;
        if ((rc = dev_alloc_t(c, &next, (size_t)U * std::max<int64_t>(n_new, 1), "side array"))) {
            dev_free(c, block, sizeof(float) * (size_t)U * m);
            return rc;
        }
        rc = score_sorted_block(c, 0, U, X + x0 * c->ld, m, block);
        if (!rc) {
            hipError_t e = sign > 0 ? launch_merge_rows(c->stream, *side, *n_side, *n_side, block, m, (int32_t)m, next,
                                                        n_new, U)
                                    : launch_subtract_rows(c->stream, *side, *n_side, *n_side, block, m, (int32_t)m,
                                                           next, n_new, U);
            if (e != hipSuccess) rc = set_err(c, RKR_ERR_HIP, std::string("side array update: ") + hipGetErrorString(e));
        }
        if (!rc) {
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = set_err(c, RKR_ERR_HIP, hipGetErrorString(e));
        }
        dev_free(c, block, sizeof(float) * (size_t)U * m);
        if (rc) {
            dev_free(c, next, sizeof(float) * (size_t)U * std::max<int64_t>(n_new, 1));
            return rc;
        }
        dev_free(c, *side, sizeof(float) * (size_t)U * std::max<int32_t>(*n_side, 1));
        *side = next;
        *n_side = (int32_t)n_new;
    }
    return RKR_OK;
}

// Rebuilds the current item set and its screened-form arrays from the base
// items, the deleted flags and the inserted items.
int rebuild_current_items(Ctx *c) {
    const int32_t ld = c->ld;
    if (c->cur_src.empty()) {  // first update: current = base, in order
        c->cur_src.resize((size_t)c->n_item_base);
        for (int64_t i = 0; i < c->n_item_base; ++i) c->cur_src[i] = i;
        c->base_deleted.assign((size_t)c->n_item_base, 0);
    }
    std::vector<int64_t> src;
    for (int64_t i = 0; i < c->n_item_base; ++i)
        if (!c->base_deleted[i]) src.push_back(i);
    const int64_t n_base_kept = (int64_t)src.size();
    for (int32_t j = 0; j < c->n_ins; ++j) src.push_back(-(1 + (int64_t)j));
    const int64_t n_cur = (int64_t)src.size();
    if (n_cur <= 1) return set_err(c, RKR_ERR_INVALID, "update: at least two items must remain");
    float *cur = nullptr;
    int rc = dev_alloc_t(c, &cur, (size_t)n_cur * ld, "current items");
    if (rc) return rc;
    int64_t *d_idx = nullptr;
    if ((rc = dev_alloc_t(c, &d_idx, (size_t)std::max<int64_t>(n_cur, 1), "gather index"))) {
        dev_free(c, cur, sizeof(float) * (size_t)n_cur * ld);
        return rc;
    }
    std::vector<int64_t> idx((size_t)n_cur);
    for (int64_t i = 0; i < n_base_kept; ++i) idx[i] = src[i];
    RKR_HIP(c, hipMemcpy(d_idx, idx.data(), sizeof(int64_t) * n_base_kept, hipMemcpyHostToDevice));
    RKR_HIP(c, launch_gather_rows(c->stream, c->items_base, ld, d_idx, n_base_kept, ld, cur, ld));
    if (c->n_ins) {
        RKR_HIP(c, hipMemcpyAsync(cur + n_base_kept * ld, c->items_ins, sizeof(float) * (size_t)c->n_ins * ld,
                                  hipMemcpyDeviceToDevice, c->stream));
    }
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(c, d_idx, sizeof(int64_t) * (size_t)std::max<int64_t>(n_cur, 1));
    // swap in the new current set
    if (c->items != c->items_base) dev_free(c, c->items, sizeof(float) * (size_t)c->n_item * ld);
    if (c->kp) {
        dev_free(c, c->item_hi, sizeof(uint16_t) * (size_t)c->n_item * c->kp);
        dev_free(c, c->item_lo, sizeof(uint16_t) * (size_t)c->n_item * c->kp);
        dev_free(c, c->item_norm, sizeof(float) * (size_t)c->n_item);
        c->item_hi = c->item_lo = nullptr;
        c->item_norm = nullptr;
    }
    c->items = cur;
    c->n_item = n_cur;
    c->cur_src = src;
    if (c->kp) {
        if ((rc = dev_alloc_t(c, &c->item_hi, (size_t)n_cur * c->kp, "item hi")) ||
            (rc = dev_alloc_t(c, &c->item_lo, (size_t)n_cur * c->kp, "item lo")) ||
            (rc = dev_alloc_t(c, &c->item_norm, (size_t)n_cur, "item norms")))
            return rc;
        RKR_HIP(c, launch_prep_items(c->stream, c->items, n_cur, ld, c->dim, c->kp, c->item_hi, c->item_lo,
                                     c->item_norm));
        RKR_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RKR_OK;
}

// Appends rows to a [n_old][w] device array (reallocates).
int append_rows(Ctx *c, float **arr, int64_t n_old, int64_t w, const float *dev_rows, int64_t n_new, const char *what) {
    float *next = nullptr;
    int rc = dev_alloc_t(c, &next, (size_t)(n_old + n_new) * w, what);
    if (rc) return rc;
    // on the context stream, after the kernels that produced dev_rows
    if (n_old)
        RKR_HIP(c, hipMemcpyAsync(next, *arr, sizeof(float) * (size_t)n_old * w, hipMemcpyDeviceToDevice, c->stream));
    if (n_new)
        RKR_HIP(c, hipMemcpyAsync(next + n_old * w, dev_rows, sizeof(float) * (size_t)n_new * w,
                                  hipMemcpyDeviceToDevice, c->stream));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(c, *arr, sizeof(float) * (size_t)std::max<int64_t>(n_old, 1) * w);
    *arr = next;
    return RKR_OK;
}

// Keeps rows idx of a [n_old][w] device array (reallocates).
int keep_rows(Ctx *c, float **arr, int64_t n_old, int64_t w, const std::vector<int64_t> &keep, const char *what) {
    const int64_t n = (int64_t)keep.size();
    float *next = nullptr;
    int64_t *d_idx = nullptr;
    int rc;
    if ((rc = dev_alloc_t(c, &next, (size_t)std::max<int64_t>(n, 1) * w, what))) return rc;
    if ((rc = dev_alloc_t(c, &d_idx, (size_t)std::max<int64_t>(n, 1), "gather index"))) {
        dev_free(c, next, sizeof(float) * (size_t)std::max<int64_t>(n, 1) * w);
        return rc;
    }
    if (n) RKR_HIP(c, hipMemcpy(d_idx, keep.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
    RKR_HIP(c, launch_gather_rows(c->stream, *arr, w, d_idx, n, w, next, w));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(c, d_idx, sizeof(int64_t) * (size_t)std::max<int64_t>(n, 1));
    dev_free(c, *arr, sizeof(float) * (size_t)std::max<int64_t>(n_old, 1) * w);
    *arr = next;
    return RKR_OK;
}

}  // namespace

extern "C" {

int rkr_insert_items(rkr_ctx *c, const float *items, int64_t n) {
    if (!c || (n > 0 && !items) || n < 0) return c ? set_err(c, RKR_ERR_INVALID, "insert_items: bad arguments") : RKR_ERR_INVALID;
    int rc = check_updatable(c);
    if (rc) return rc;
    if (c->output == RKR_OUTPUT_REFERENCE)
        return set_err(c, RKR_ERR_STATE, "insert_items: the reference's US method has no update path; use exact output");
    if (n == 0) return RKR_OK;
    const double mp = max_abs(items, (size_t)n * c->dim);
    if (mp < 0) return set_err(c, RKR_ERR_INVALID, "insert_items: items must be finite");
    const double mp_new = std::max(c->max_abs_item, mp);
    if ((double)c->dim * c->max_abs_user * mp_new > kScoreLimit)
        return set_err(c, RKR_ERR_INVALID, "insert_items: magnitudes would overflow FP32 scores");
    if ((int64_t)c->n_ins + n > kMaxSideItems)
        return set_err(c, RKR_ERR_NOMEM, "insert_items: too many pending inserts; rebuild the table");
    if (c->n_item_base + c->n_ins + n >= kMaxBoundValue) return set_err(c, RKR_ERR_INVALID, "insert_items: too many items");
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    c->max_abs_item = mp_new;
    if (c->kp) c->screen_abs = screen_abs_floor(c->kp, c->max_abs_user, c->max_abs_item);
    float *X = nullptr;
    if ((rc = upload_padded(c, items, n, &X, "inserted items"))) return rc;
    rc = update_side(c, &c->side_ins, &c->n_ins, X, n, +1);
    if (!rc) {
        const int32_t n_old = c->n_ins - (int32_t)n;
        rc = append_rows(c, &c->items_ins, n_old, c->ld, X, n, "inserted items");
    }
    dev_free(c, X, sizeof(float) * (size_t)n * c->ld);
    if (rc) return rc;
    return rebuild_current_items(c);
}

int rkr_delete_items(rkr_ctx *c, const int64_t *ids, int64_t n) {
    if (!c || (n > 0 && !ids) || n < 0) return c ? set_err(c, RKR_ERR_INVALID, "delete_items: bad arguments") : RKR_ERR_INVALID;
    int rc = check_updatable(c);
    if (rc) return rc;
    if (c->output == RKR_OUTPUT_REFERENCE)
        return set_err(c, RKR_ERR_STATE, "delete_items: the reference's US method has no update path; use exact output");
    if (n == 0) return RKR_OK;
    if (c->cur_src.empty()) {
        c->cur_src.resize((size_t)c->n_item_base);
        for (int64_t i = 0; i < c->n_item_base; ++i) c->cur_src[i] = i;
        c->base_deleted.assign((size_t)c->n_item_base, 0);
    }
    std::vector<uint8_t> seen((size_t)c->n_item, 0);
    std::vector<int64_t> del_base, del_ins;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t id = ids[i];
        if (id < 0 || id >= c->n_item || seen[id]) return set_err(c, RKR_ERR_INVALID, "delete_items: bad or repeated item id");
        seen[id] = 1;
        const int64_t s = c->cur_src[id];
        if (s >= 0) del_base.push_back(s);
        else del_ins.push_back(-s - 1);
    }
    if (c->n_item - n < 2) return set_err(c, RKR_ERR_INVALID, "delete_items: at least two items must remain");
    if ((int64_t)c->n_del + (int64_t)del_base.size() > kMaxSideItems)
        return set_err(c, RKR_ERR_NOMEM, "delete_items: too many pending deletes; rebuild the table");
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    const int32_t ld = c->ld;
    if (!del_base.empty()) {
        // gather the deleted base vectors, add their scores to the deleted side arrays
        float *X = nullptr;
        int64_t *d_idx = nullptr;
        const int64_t m = (int64_t)del_base.size();
        if ((rc = dev_alloc_t(c, &X, (size_t)m * ld, "deleted items"))) return rc;
        if ((rc = dev_alloc_t(c, &d_idx, (size_t)m, "gather index"))) { dev_free(c, X, sizeof(float) * m * ld); return rc; }
        RKR_HIP(c, hipMemcpy(d_idx, del_base.data(), sizeof(int64_t) * m, hipMemcpyHostToDevice));
        RKR_HIP(c, launch_gather_rows(c->stream, c->items_base, ld, d_idx, m, ld, X, ld));
        RKR_HIP(c, hipStreamSynchronize(c->stream));
        dev_free(c, d_idx, sizeof(int64_t) * m);
        rc = update_side(c, &c->side_del, &c->n_del, X, m, +1);
        if (!rc) rc = append_rows(c, &c->items_del, c->n_del - (int32_t)m, ld, X, m, "deleted items");
        dev_free(c, X, sizeof(float) * (size_t)m * ld);
        if (rc) return rc;
        for (int64_t s : del_base) c->base_deleted[s] = 1;
    }
    if (!del_ins.empty()) {
        // remove their scores from the inserted side arrays and the rows from the inserted set
        float *X = nullptr;
        int64_t *d_idx = nullptr;
        const int64_t m = (int64_t)del_ins.size();
        if ((rc = dev_alloc_t(c, &X, (size_t)m * ld, "removed inserts"))) return rc;
        if ((rc = dev_alloc_t(c, &d_idx, (size_t)m, "gather index"))) { dev_free(c, X, sizeof(float) * m * ld); return rc; }
        RKR_HIP(c, hipMemcpy(d_idx, del_ins.data(), sizeof(int64_t) * m, hipMemcpyHostToDevice));
        RKR_HIP(c, launch_gather_rows(c->stream, c->items_ins, ld, d_idx, m, ld, X, ld));
        RKR_HIP(c, hipStreamSynchronize(c->stream));
        dev_free(c, d_idx, sizeof(int64_t) * m);
        rc = update_side(c, &c->side_ins, &c->n_ins, X, m, -1);
        dev_free(c, X, sizeof(float) * (size_t)m * ld);
        if (rc) return rc;
        std::vector<uint8_t> gone((size_t)c->n_ins + m, 0);
        for (int64_t j : del_ins) gone[j] = 1;
        std::vector<int64_t> keep;
        for (int64_t j = 0; j < (int64_t)gone.size(); ++j)
            if (!gone[j]) keep.push_back(j);
        if ((rc = keep_rows(c, &c->items_ins, (int64_t)gone.size(), ld, keep, "inserted items"))) return rc;
    }
    return rebuild_current_items(c);
}

int rkr_insert_users(rkr_ctx *c, const float *users, int64_t n) {
    if (!c || (n > 0 && !users) || n < 0) return c ? set_err(c, RKR_ERR_INVALID, "insert_users: bad arguments") : RKR_ERR_INVALID;
    int rc = check_updatable(c);
    if (rc) return rc;
    if (n == 0) return RKR_OK;
    const double mu = max_abs(users, (size_t)n * c->dim);
    if (mu < 0) return set_err(c, RKR_ERR_INVALID, "insert_users: users must be finite");
    if (c->n_user + n > INT32_MAX) return set_err(c, RKR_ERR_INVALID, "insert_users: too many users");
    // normalised rows have |u_i| <= 1
    const double mu_raw = std::max(c->max_abs_user_raw, mu);
    const double mu_new = c->normalize_users ? c->max_abs_user : mu_raw;
    if ((double)c->dim * mu_new * c->max_abs_item > kScoreLimit)
        return set_err(c, RKR_ERR_INVALID, "insert_users: magnitudes would overflow FP32 scores");
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    c->max_abs_user_raw = mu_raw;
    c->max_abs_user = mu_new;
    if (c->kp) c->screen_abs = screen_abs_floor(c->kp, c->max_abs_user, c->max_abs_item);
    std::vector<float> h(users, users + (size_t)n * c->dim);
    if (c->normalize_users) rkr_normalize_users(h.data(), n, c->dim);  // as readData does on load
    const int64_t U0 = c->n_user, ld = c->ld, S = c->n_sample;
    float *X = nullptr;
    if ((rc = upload_padded(c, h.data(), n, &X, "inserted users"))) return rc;
    rc = append_rows(c, &c->users, U0, ld, X, n, "users");
    dev_free(c, X, sizeof(float) * (size_t)n * ld);
    if (rc) return rc;
    c->h_users.insert(c->h_users.end(), h.begin(), h.end());
    c->n_user = U0 + n;
    // panels and norms over all users
    dev_free(c, c->usersT, sizeof(float) * (size_t)((U0 + 63) / 64) * 64 * ld);
    c->usersT = nullptr;
    if ((rc = dev_alloc_t(c, &c->usersT, (size_t)((c->n_user + 63) / 64) * 64 * ld, "user panels"))) return rc;
    RKR_HIP(c, launch_make_panels(c->stream, c->users, c->n_user, (int32_t)ld, c->usersT));
    if (c->kp) {
        dev_free(c, c->user_norm, sizeof(float) * (size_t)U0);
        c->user_norm = nullptr;
        if ((rc = dev_alloc_t(c, &c->user_norm, (size_t)c->n_user, "user norms"))) return rc;
        RKR_HIP(c, launch_user_norms(c->stream, c->users, c->n_user, (int32_t)ld, c->dim, c->user_norm));
    }
    // table rows of the new users against the base items
    float *rows = nullptr, *block = nullptr;
    if ((rc = dev_alloc_t(c, &rows, (size_t)n * S, "new table rows"))) return rc;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)((1ull << 31) / (sizeof(float) * c->n_item_base))));
    if ((rc = dev_alloc_t(c, &block, (size_t)chunk * c->n_item_base, "score block"))) { dev_free(c, rows, sizeof(float) * n * S); return rc; }
    for (int64_t u0 = 0; u0 < n; u0 += chunk) {
        const int64_t m = std::min<int64_t>(chunk, n - u0);
        RKR_HIP(c, launch_scores_store(c->stream, c->users + (U0 + u0) * ld, nullptr, m, (int32_t)ld, c->items_base,
                                       c->n_item_base, (int32_t)ld, c->dim, block, c->n_item_base, true));
        RKR_HIP(c, launch_select_rows(c->stream, block, m, c->n_item_base, c->ranks, (int32_t)S, rows + u0 * S));
    }
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(c, block, sizeof(float) * (size_t)chunk * c->n_item_base);
    rc = append_rows(c, &c->table, U0, S, rows, n, "table");
    dev_free(c, rows, sizeof(float) * (size_t)n * S);
    if (rc) return rc;
    // side rows of the new users
    for (int side = 0; side < 2; ++side) {
        float **arr = side == 0 ? &c->side_ins : &c->side_del;
        const int32_t ns = side == 0 ? c->n_ins : c->n_del;
        const float *X2 = side == 0 ? c->items_ins : c->items_del;
        if (!ns) continue;
        float *blk = nullptr;
        if ((rc = dev_alloc_t(c, &blk, (size_t)n * ns, "side rows"))) return rc;
        rc = score_sorted_block(c, U0, n, X2, ns, blk);
        if (!rc) rc = append_rows(c, arr, U0, ns, blk, n, "side array");
        dev_free(c, blk, sizeof(float) * (size_t)n * ns);
        if (rc) return rc;
    }
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    return RKR_OK;
}

int rkr_delete_users(rkr_ctx *c, const int64_t *ids, int64_t n) {
    if (!c || (n > 0 && !ids) || n < 0) return c ? set_err(c, RKR_ERR_INVALID, "delete_users: bad arguments") : RKR_ERR_INVALID;
    int rc = check_updatable(c);
    if (rc) return rc;
    if (n == 0) return RKR_OK;
    const int64_t U0 = c->n_user, ld = c->ld;
    std::vector<uint8_t> gone((size_t)U0, 0);
    for (int64_t i = 0; i < n; ++i) {
        if (ids[i] < 0 || ids[i] >= U0 || gone[ids[i]]) return set_err(c, RKR_ERR_INVALID, "delete_users: bad or repeated user id");
        gone[ids[i]] = 1;
    }
    if (U0 - n < 1) return set_err(c, RKR_ERR_INVALID, "delete_users: at least one user must remain");
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<int64_t> keep;
    keep.reserve((size_t)(U0 - n));
    for (int64_t u = 0; u < U0; ++u)
        if (!gone[u]) keep.push_back(u);
    // SampleSearchUpdate::DeleteUser (SampleSearchUpdate.hpp:89-109): later users shift down
    if ((rc = keep_rows(c, &c->users, U0, ld, keep, "users"))) return rc;
    if ((rc = keep_rows(c, &c->table, U0, c->n_sample, keep, "table"))) return rc;
    if (c->n_ins && (rc = keep_rows(c, &c->side_ins, U0, c->n_ins, keep, "side array"))) return rc;
    if (c->n_del && (rc = keep_rows(c, &c->side_del, U0, c->n_del, keep, "side array"))) return rc;
    std::vector<float> hu((size_t)keep.size() * c->dim);
    for (size_t i = 0; i < keep.size(); ++i)
        std::memcpy(&hu[i * c->dim], &c->h_users[(size_t)keep[i] * c->dim], sizeof(float) * c->dim);
    c->h_users.swap(hu);
    c->n_user = (int64_t)keep.size();
    dev_free(c, c->usersT, sizeof(float) * (size_t)((U0 + 63) / 64) * 64 * ld);
    c->usersT = nullptr;
    if ((rc = dev_alloc_t(c, &c->usersT, (size_t)((c->n_user + 63) / 64) * 64 * ld, "user panels"))) return rc;
    RKR_HIP(c, launch_make_panels(c->stream, c->users, c->n_user, (int32_t)ld, c->usersT));
    if (c->kp) {
        dev_free(c, c->user_norm, sizeof(float) * (size_t)U0);
        c->user_norm = nullptr;
        if ((rc = dev_alloc_t(c, &c->user_norm, (size_t)c->n_user, "user norms"))) return rc;
        RKR_HIP(c, launch_user_norms(c->stream, c->users, c->n_user, (int32_t)ld, c->dim, c->user_norm));
    }
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    return RKR_OK;
}

int rkr_get_counts(const rkr_ctx *c, int64_t *n_user, int64_t *n_item, int64_t *n_inserted, int64_t *n_deleted) {
    if (!c) return RKR_ERR_INVALID;
    if (n_user) *n_user = c->n_user;
    if (n_item) *n_item = c->n_item;
    if (n_inserted) *n_inserted = c->n_ins;
    if (n_deleted) *n_deleted = c->n_del;
    return RKR_OK;
}

}  // extern "C"
