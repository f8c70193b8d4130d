// rkr_update.cpp -- the update path of the C ABI: inserting and deleting
// users and items while the sampled table stays resident.
//
// The table keeps describing the base item set it was built from.  Per user,
// sorted side arrays hold the scores of inserted items and of deleted base
// items (SampleSearchUpdate.hpp:72-166), and the query path shifts every rank
// bound by #{inserted > q} - #{deleted > q} (k_update.hip).  Refinement counts
// over the current item set,
//   [base items not deleted, in base order] ++ [inserted items, in order];
// item IDs given to rkr_delete_items refer to that order.  rkr_build_table
// over the current items folds everything back into a fresh base.
//
// Storage is sized by capacity so that steady-state updates neither copy nor
// free the large arrays, and so that a serving loop can run query batches
// while an update is applied (snapshot mode):
//  - per-user arrays (users, panels, norms, table rows, side rows) hold
//    Ctx::u_cap users; inserted users are written past the current count;
//  - the current items and their BF16 rows hold Ctx::i_cap items; inserts
//    append (while no item was deleted, the current set is base ++ inserted
//    and shares the base buffer);
//  - each side array has a spare of the same shape: an item update merges
//    the current rows with the new scores into the spare and swaps, so the
//    rows a running batch reads are never written.
// Anything an update replaces goes through dev_free, which a snapshot-mode
// server defers until the batches that may read it have finished.
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

int check_updatable(Ctx *c) {
    if (!c->users) return set_err(c, RKR_ERR_STATE, "update: no data");
    if (!c->table) return set_err(c, RKR_ERR_STATE, "update: no table (build or load one first)");
    free_qsrp(c);  // the QSRP screen describes the users and items it was built for
    return RKR_OK;
}

// Host rows (n x dim) -> device rows of stride ld at dst (pad columns zero).
int upload_rows(Ctx *c, const float *src, int64_t n, float *dst) {
    if (n <= 0) return RKR_OK;
    const int32_t ld = c->ld, d = c->dim;
    std::vector<float> tmp((size_t)n * ld, 0.0f);
    for (int64_t r = 0; r < n; ++r) std::memcpy(&tmp[(size_t)r * ld], src + (size_t)r * d, sizeof(float) * d);
    RKR_HIP(c, hipMemcpyAsync(dst, tmp.data(), sizeof(float) * (size_t)n * ld, hipMemcpyHostToDevice, c->stream));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    return RKR_OK;
}

// n rows of w floats between row-major arrays, on the context stream.
int copy_rows(Ctx *c, float *dst, int64_t ld_dst, const float *src, int64_t ld_src, int64_t n, int64_t w) {
    if (n <= 0 || w <= 0) return RKR_OK;
    RKR_HIP(c, hipMemcpy2DAsync(dst, sizeof(float) * ld_dst, src, sizeof(float) * ld_src, sizeof(float) * w, n,
                                hipMemcpyDeviceToDevice, c->stream));
    return RKR_OK;
}

int64_t grown(int64_t cap, int64_t need) { return std::max<int64_t>(need, cap + cap / 4 + 1024); }

// Per-user arrays for at least `need` users.  Existing rows are copied once;
// the replaced buffers go through dev_free.
int reserve_users(Ctx *c, int64_t need) {
    if (need <= c->u_cap) return RKR_OK;
    const int64_t cap = grown(c->u_cap, need), U = c->n_user;
    const int32_t ld = c->ld, S = c->n_sample;
    const int32_t QS = c->n_qs;
    float *users = nullptr, *panels = nullptr, *unorm = nullptr, *table = nullptr, *qtable = nullptr, *si = nullptr,
          *sd = nullptr, *gv = nullptr;
    int rc = RKR_OK;
    if ((c->gvals && (rc = dev_alloc_t(c, &gv, (size_t)cap * kGuardRow, "guard values"))) ||
        (rc = dev_alloc_t(c, &users, (size_t)cap * ld, "users")) ||
        (rc = dev_alloc_t(c, &panels, (size_t)((cap + 63) / 64) * 64 * ld, "user panels")) ||
        (c->kp && (rc = dev_alloc_t(c, &unorm, (size_t)cap, "user norms"))) ||
        (rc = dev_alloc_t(c, &table, (size_t)cap * S, "table")) ||
        (QS && (rc = dev_alloc_t(c, &qtable, (size_t)cap * QS, "query table"))) ||
        (c->ld_ins && (rc = dev_alloc_t(c, &si, (size_t)cap * c->ld_ins, "side array"))) ||
        (c->ld_del && (rc = dev_alloc_t(c, &sd, (size_t)cap * c->ld_del, "side array")))) {
        for (float *p : {users, panels, unorm, table, qtable, si, sd, gv}) dev_free(c, p, 0);
        return rc;
    }
    uint16_t *uhi = nullptr, *ulo = nullptr;
    if (c->kp && ((rc = dev_alloc_t(c, &uhi, (size_t)cap * c->kp, "user hi")) ||
                  (rc = dev_alloc_t(c, &ulo, (size_t)cap * c->kp, "user lo")))) {
        for (float *p : {users, panels, unorm, table, qtable, si, sd, gv}) dev_free(c, p, 0);
        dev_free(c, uhi, 0);
        dev_free(c, ulo, 0);
        return rc;
    }
    if (c->kp && U > 0) {
        RKR_HIP(c, hipMemcpyAsync(uhi, c->user_hi, sizeof(uint16_t) * (size_t)U * c->kp, hipMemcpyDeviceToDevice,
                                  c->stream));
        RKR_HIP(c, hipMemcpyAsync(ulo, c->user_lo, sizeof(uint16_t) * (size_t)U * c->kp, hipMemcpyDeviceToDevice,
                                  c->stream));
    }
    if ((rc = copy_rows(c, users, ld, c->users, ld, U, ld)) ||
        (rc = copy_rows(c, panels, 64 * ld, c->usersT, 64 * ld, (U + 63) / 64, 64 * ld)) ||
        (c->kp && (rc = copy_rows(c, unorm, 1, c->user_norm, 1, U, 1))) ||
        (rc = copy_rows(c, table, S, c->table, S, U, S)) ||
        (QS && (rc = copy_rows(c, qtable, QS, c->qtable, QS, U, QS))) ||
        (gv && (rc = copy_rows(c, gv, kGuardRow, c->gvals, kGuardRow, U, kGuardRow))) ||
        (si && (rc = copy_rows(c, si, c->ld_ins, c->side_ins, c->ld_ins, U, c->n_ins))) ||
        (sd && (rc = copy_rows(c, sd, c->ld_del, c->side_del, c->ld_del, U, c->n_del))))
        return rc;
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    for (float *p : {c->users, c->usersT, c->user_norm, c->table, c->qtable, c->side_ins, c->side_del,
                     c->side_ins_spare, c->side_del_spare, c->gvals})
        dev_free(c, p, 0);
    c->gvals = gv;
    dev_free(c, c->user_hi, 0);
    dev_free(c, c->user_lo, 0);
    c->user_hi = uhi;
    c->user_lo = ulo;
    c->users = users;
    c->usersT = panels;
    c->user_norm = unorm;
    c->table = table;
    c->qtable = qtable;
    c->side_ins = si;
    c->side_del = sd;
    c->side_ins_spare = c->side_del_spare = nullptr;
    c->u_cap = cap;
    return RKR_OK;
}

// Current-item arrays (items, BF16 rows, norms) for at least `need` items.
// While the current set is base ++ inserted, the base buffer grows with it.
int reserve_items(Ctx *c, int64_t need) {
    if (need <= c->i_cap) return RKR_OK;
    const int64_t cap = grown(c->i_cap, need), I = c->n_item;
    const int32_t ld = c->ld, kp = c->kp;
    float *items = nullptr, *inorm = nullptr;
    uint16_t *hi = nullptr, *lo = nullptr;
    int rc = RKR_OK;
    if ((rc = dev_alloc_t(c, &items, (size_t)cap * ld, "current items")) ||
        (kp && ((rc = dev_alloc_t(c, &hi, (size_t)cap * kp, "item hi")) ||
                (rc = dev_alloc_t(c, &lo, (size_t)cap * kp, "item lo")) ||
                (rc = dev_alloc_t(c, &inorm, (size_t)cap, "item norms"))))) {
        dev_free(c, items, 0);
        dev_free(c, hi, 0);
        dev_free(c, lo, 0);
        dev_free(c, inorm, 0);
        return rc;
    }
    if ((rc = copy_rows(c, items, ld, c->items, ld, I, ld))) return rc;
    if (kp) {
        RKR_HIP(c, hipMemcpyAsync(hi, c->item_hi, sizeof(uint16_t) * (size_t)I * kp, hipMemcpyDeviceToDevice, c->stream));
        RKR_HIP(c, hipMemcpyAsync(lo, c->item_lo, sizeof(uint16_t) * (size_t)I * kp, hipMemcpyDeviceToDevice, c->stream));
        if ((rc = copy_rows(c, inorm, 1, c->item_norm, 1, I, 1))) return rc;
    }
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    const bool shared_base = c->items == c->items_base;
    dev_free(c, c->items, 0);
    dev_free(c, c->item_hi, 0);
    dev_free(c, c->item_lo, 0);
    dev_free(c, c->item_norm, 0);
    c->items = items;
    if (shared_base) c->items_base = items;  // base rows keep their offsets
    c->item_hi = hi;
    c->item_lo = lo;
    c->item_norm = inorm;
    c->i_cap = cap;
    return RKR_OK;
}

// A [cap][ld] list of item vectors (inserted or deleted items) with room for
// `need` rows, `n` of them in use.
int reserve_list(Ctx *c, float **arr, int64_t *cap, int64_t n, int64_t need, const char *what) {
    if (need <= *cap) return RKR_OK;
    const int64_t nc = grown(*cap, need);
    float *next = nullptr;
    int rc = dev_alloc_t(c, &next, (size_t)nc * c->ld, what);
    if (rc) return rc;
    if ((rc = copy_rows(c, next, c->ld, *arr, c->ld, n, c->ld))) return rc;
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(c, *arr, 0);
    *arr = next;
    *cap = nc;
    return RKR_OK;
}

// Merges (sign > 0) or removes (sign < 0) the scores of item rows X ([n][ld])
// into one side array of every user.  The result is written to the spare (or
// a fresh buffer when the rows must grow), never to the rows batches read.
int update_side(Ctx *c, bool ins_side, const float *X, int64_t n, int sign) {
    float **side = ins_side ? &c->side_ins : &c->side_del;
    float **spare = ins_side ? &c->side_ins_spare : &c->side_del_spare;
    int32_t *ld = ins_side ? &c->ld_ins : &c->ld_del;
    int32_t *ns = ins_side ? &c->n_ins : &c->n_del;
    const int64_t U = c->n_user;
    float *const published = *side;
    const int32_t ld_published = *ld;
    for (int64_t x0 = 0; x0 < n; x0 += kSortChunk) {
        const int64_t m = std::min<int64_t>(kSortChunk, n - x0);
        const int64_t n_new = sign > 0 ? (int64_t)*ns + m : (int64_t)*ns - m;
        int32_t ld_t = *ld;
        float *target = nullptr;
        int rc;
        if (n_new > *ld) {  // longer rows: a fresh pair of buffers
            ld_t = (int32_t)((std::max<int64_t>({n_new, (int64_t)*ld + *ld / 2, 256}) + 63) / 64 * 64);
            dev_free(c, *spare, 0);
            *spare = nullptr;
        } else if (*side == published && *spare) {
            target = *spare;
            *spare = nullptr;
        }
        if (!target && (rc = dev_alloc_t(c, &target, (size_t)c->u_cap * ld_t, "side array"))) return rc;
        if ((rc = ensure(c, c->upd_scratch, sizeof(float) * (size_t)U * m, "update score block"))) {
            dev_free(c, target, 0);
            return rc;
        }
        float *block = reinterpret_cast<float *>(c->upd_scratch.p);
        RKR_HIP(c, launch_scores_store(c->stream, c->users, nullptr, U, c->ld, X + x0 * c->ld, m, c->ld, c->dim, block,
                                       m, true));
        RKR_HIP(c, launch_sort_rows_desc(c->stream, block, U, (int32_t)m, m));
        if (sign > 0)
            RKR_HIP(c, launch_merge_rows(c->stream, *side, *ld, *ns, block, m, (int32_t)m, target, ld_t, U));
        else
            RKR_HIP(c, launch_subtract_rows(c->stream, *side, *ld, *ns, block, m, (int32_t)m, target, ld_t, U));
        RKR_HIP(c, hipStreamSynchronize(c->stream));
        if (*side != published) dev_free(c, *side, 0);  // an intermediate of this call
        *side = target;
        *ld = ld_t;
        *ns = (int32_t)n_new;
    }
    // the rows batches read until this update commits become the next spare
    if (published && published != *side) {
        if (ld_published == *ld && !*spare) *spare = published;
        else dev_free(c, published, 0);
    }
    return RKR_OK;
}

// Item rows X ([ids.size()][ld]) gathered by index from a device list.
int gather_items(Ctx *c, const float *src, const std::vector<int64_t> &ids, float **X) {
    const int64_t m = (int64_t)ids.size();
    int64_t *d_idx = nullptr;
    int rc;
    if ((rc = dev_alloc_t(c, X, (size_t)std::max<int64_t>(m, 1) * c->ld, "item rows"))) return rc;
    if (m == 0) return RKR_OK;
    if ((rc = dev_alloc_t(c, &d_idx, (size_t)m, "gather index"))) {
        dev_free(c, *X, 0);
        return rc;
    }
    RKR_HIP(c, hipMemcpyAsync(d_idx, ids.data(), sizeof(int64_t) * m, hipMemcpyHostToDevice, c->stream));
    RKR_HIP(c, launch_gather_rows(c->stream, src, c->ld, d_idx, m, c->ld, *X, c->ld));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(c, d_idx, 0);
    return RKR_OK;
}

// Rebuilds the current item set into new buffers after deletes:
// [base not deleted] ++ [inserted].
int rebuild_current_items(Ctx *c) {
    const int32_t ld = c->ld, kp = c->kp;
    std::vector<int64_t> src;
    for (int64_t i = 0; i < c->n_item_base; ++i)
        if (!c->base_deleted[i]) src.push_back(i);
    const int64_t n_base_kept = (int64_t)src.size();
    for (int32_t j = 0; j < c->n_ins; ++j) src.push_back(-(1 + (int64_t)j));
    const int64_t n_cur = (int64_t)src.size();
    if (n_cur <= 1) return set_err(c, RKR_ERR_INVALID, "update: at least two items must remain");
    const int64_t cap = std::max<int64_t>(n_cur, c->i_cap);
    float *cur = nullptr, *inorm = nullptr;
    uint16_t *hi = nullptr, *lo = nullptr;
    int64_t *d_idx = nullptr;
    int rc;
    if ((rc = dev_alloc_t(c, &cur, (size_t)cap * ld, "current items")) ||
        (rc = dev_alloc_t(c, &d_idx, (size_t)std::max<int64_t>(n_base_kept, 1), "gather index")) ||
        (kp && ((rc = dev_alloc_t(c, &hi, (size_t)cap * kp, "item hi")) ||
                (rc = dev_alloc_t(c, &lo, (size_t)cap * kp, "item lo")) ||
                (rc = dev_alloc_t(c, &inorm, (size_t)cap, "item norms"))))) {
        for (void *p : {(void *)cur, (void *)d_idx, (void *)hi, (void *)lo, (void *)inorm}) dev_free(c, p, 0);
        return rc;
    }
    if (n_base_kept) {
        RKR_HIP(c, hipMemcpyAsync(d_idx, src.data(), sizeof(int64_t) * n_base_kept, hipMemcpyHostToDevice, c->stream));
        RKR_HIP(c, launch_gather_rows(c->stream, c->items_base, ld, d_idx, n_base_kept, ld, cur, ld));
    }
    if ((rc = copy_rows(c, cur + n_base_kept * ld, ld, c->items_ins, ld, c->n_ins, ld))) return rc;
    if (kp) RKR_HIP(c, launch_prep_items(c->stream, cur, n_cur, ld, c->dim, kp, hi, lo, inorm));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(c, d_idx, 0);
    if (c->items != c->items_base) dev_free(c, c->items, 0);
    dev_free(c, c->item_hi, 0);
    dev_free(c, c->item_lo, 0);
    dev_free(c, c->item_norm, 0);
    c->items = cur;
    c->item_hi = hi;
    c->item_lo = lo;
    c->item_norm = inorm;
    c->i_cap = cap;
    c->n_item = n_cur;
    c->cur_src = src;
    return RKR_OK;
}

// current item -> base id map, on the first item update
void init_item_map(Ctx *c) {
    if (!c->cur_src.empty() || c->n_ins > 0) return;
    c->cur_src.resize((size_t)c->n_item_base);
    for (int64_t i = 0; i < c->n_item_base; ++i) c->cur_src[i] = i;
    c->base_deleted.assign((size_t)c->n_item_base, 0);
}

// Row compaction of a per-user array ([u_cap][ld], w floats used per row)
// into a new buffer of the same shape.
int keep_user_rows(Ctx *c, float **arr, int64_t w, int64_t ld, const int64_t *d_keep, int64_t n_keep,
                   const char *what) {
    if (!*arr) return RKR_OK;
    float *next = nullptr;
    int rc = dev_alloc_t(c, &next, (size_t)c->u_cap * ld, what);
    if (rc) return rc;
    if (n_keep && w) RKR_HIP(c, launch_gather_rows(c->stream, *arr, ld, d_keep, n_keep, w, next, ld));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(c, *arr, 0);
    *arr = next;
    return RKR_OK;
}

}  // namespace

int rkr::refresh_ins_guard_max(Ctx *c) {
    for (int j = 0; j < 8; ++j) c->ins_gmax[j] = 0;
    const float *T = c->qtable ? c->qtable : c->table;
    const int32_t S = c->qtable ? c->n_qs : c->n_sample;
    if (!c->n_ins || !T || c->n_user <= 0) return RKR_OK;
    GuardCols gc{};
    guard_columns(c, gc.col, gc.n);
    if (gc.n <= 0) return RKR_OK;
    int rc;
    if ((rc = ensure(c, c->upd_scratch, std::max<size_t>(c->upd_scratch.bytes, 64), "update scratch"))) return rc;
    int32_t *d = reinterpret_cast<int32_t *>(c->upd_scratch.p);
    RKR_HIP(c, launch_ins_guard_max(c->stream, T, c->n_user, S, gc, c->side_ins, c->n_ins, c->ld_ins, d));
    RKR_HIP(c, hipMemcpyAsync(c->ins_gmax, d, sizeof(int32_t) * 8, hipMemcpyDeviceToHost, c->stream));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    return RKR_OK;
}

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
    if (c->n_item_base + c->n_ins + n >= kMaxBoundValue - 1)
        return set_err(c, RKR_ERR_INVALID, "insert_items: too many items");
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    init_item_map(c);
    const int64_t I0 = c->n_item;
    if ((rc = reserve_items(c, I0 + n)) ||
        (rc = reserve_list(c, &c->items_ins, &c->ins_cap, c->n_ins, (int64_t)c->n_ins + n, "inserted items")))
        return rc;
    // new rows past the current count: batches in flight never read them
    float *X = c->items + I0 * c->ld;
    if ((rc = upload_rows(c, items, n, X))) return rc;
    if ((rc = copy_rows(c, c->items_ins + (int64_t)c->n_ins * c->ld, c->ld, X, c->ld, n, c->ld))) return rc;
    if (c->kp)
        RKR_HIP(c, launch_prep_items(c->stream, X, n, c->ld, c->dim, c->kp, c->item_hi + I0 * c->kp,
                                     c->item_lo + I0 * c->kp, c->item_norm + I0));
    const int32_t ins0 = c->n_ins;
    if ((rc = update_side(c, true, X, n, +1))) return rc;
    c->max_abs_item = mp_new;
    if (c->kp) c->screen_abs = screen_abs_floor(c->kp, c->max_abs_user, c->max_abs_item);
    for (int64_t j = 0; j < n; ++j) c->cur_src.push_back(-(1 + (int64_t)ins0 + j));
    c->n_item = I0 + n;
    return rkr::refresh_ins_guard_max(c);
}

int rkr_delete_items(rkr_ctx *c, const int64_t *ids, int64_t n) {
    if (!c || (n > 0 && !ids) || n < 0) return c ? set_err(c, RKR_ERR_INVALID, "delete_items: bad arguments") : RKR_ERR_INVALID;
    int rc = check_updatable(c);
    if (rc) return rc;
    if (c->output == RKR_OUTPUT_REFERENCE)
        return set_err(c, RKR_ERR_STATE, "delete_items: the reference's US method has no update path; use exact output");
    if (n == 0) return RKR_OK;
    init_item_map(c);
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
        // their scores join the deleted side arrays, their vectors the deleted list
        float *X = nullptr;
        const int64_t m = (int64_t)del_base.size();
        if ((rc = gather_items(c, c->items_base, del_base, &X))) return rc;
        if (!(rc = reserve_list(c, &c->items_del, &c->del_cap, c->n_del, (int64_t)c->n_del + m, "deleted items")) &&
            !(rc = copy_rows(c, c->items_del + (int64_t)c->n_del * ld, ld, X, ld, m, ld)))
            rc = update_side(c, false, X, m, +1);
        dev_free(c, X, 0);
        if (rc) return rc;
        for (int64_t s : del_base) c->base_deleted[s] = 1;
    }
    if (!del_ins.empty()) {
        // their scores leave the inserted side arrays, their rows the inserted list
        float *X = nullptr;
        const int64_t m = (int64_t)del_ins.size();
        const int64_t n_ins_old = c->n_ins;
        if ((rc = gather_items(c, c->items_ins, del_ins, &X))) return rc;
        rc = update_side(c, true, X, m, -1);
        dev_free(c, X, 0);
        if (rc) return rc;
        std::vector<uint8_t> gone((size_t)n_ins_old, 0);
        for (int64_t j : del_ins) gone[j] = 1;
        std::vector<int64_t> keep;
        for (int64_t j = 0; j < n_ins_old; ++j)
            if (!gone[j]) keep.push_back(j);
        // the kept rows go to a new list: the old one may still be read
        float *next = nullptr;
        if ((rc = dev_alloc_t(c, &next, (size_t)std::max<int64_t>(c->ins_cap, 1) * ld, "inserted items"))) return rc;
        if (!keep.empty()) {
            float *kept = nullptr;
            if ((rc = gather_items(c, c->items_ins, keep, &kept))) {
                dev_free(c, next, 0);
                return rc;
            }
            rc = copy_rows(c, next, ld, kept, ld, (int64_t)keep.size(), ld);
            if (!rc) {
                hipError_t e = hipStreamSynchronize(c->stream);
                if (e != hipSuccess) rc = set_err(c, RKR_ERR_HIP, hipGetErrorString(e));
            }
            dev_free(c, kept, 0);
            if (rc) {
                dev_free(c, next, 0);
                return rc;
            }
        }
        dev_free(c, c->items_ins, 0);
        c->items_ins = next;
        c->ins_cap = std::max<int64_t>(c->ins_cap, 1);
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
    const int64_t U0 = c->n_user, U1 = U0 + n;
    const int32_t ld = c->ld, S = c->n_sample;
    if ((rc = reserve_users(c, U1))) return rc;
    std::vector<float> h(users, users + (size_t)n * c->dim);
    if (c->normalize_users) rkr_normalize_users(h.data(), n, c->dim);  // as readData does on load
    // rows past the current count: batches in flight never read them
    if ((rc = upload_rows(c, h.data(), n, c->users + U0 * ld))) return rc;
    const int64_t p0 = U0 / 64;  // first panel that changes (lanes of earlier users are rewritten unchanged)
    RKR_HIP(c, launch_make_panels(c->stream, c->users + p0 * 64 * ld, U1 - p0 * 64, ld, c->usersT + p0 * 64 * ld));
    if (c->kp) {
        RKR_HIP(c, launch_user_norms(c->stream, c->users + U0 * ld, n, ld, c->dim, c->user_norm + U0, c->rho));
        RKR_HIP(c, launch_prep_items(c->stream, c->users + U0 * ld, n, ld, c->dim, c->kp, c->user_hi + U0 * c->kp,
                                     c->user_lo + U0 * c->kp, nullptr));
    }
    // table rows of the new users against the base items
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)((1ull << 31) / (sizeof(float) * c->n_item_base))));
    if ((rc = ensure(c, c->upd_scratch, sizeof(float) * (size_t)chunk * c->n_item_base, "score block"))) return rc;
    float *block = reinterpret_cast<float *>(c->upd_scratch.p);
    for (int64_t u0 = 0; u0 < n; u0 += chunk) {
        const int64_t m = std::min<int64_t>(chunk, n - u0);
        RKR_HIP(c, launch_scores_store(c->stream, c->users + (U0 + u0) * ld, nullptr, m, ld, c->items_base,
                                       c->n_item_base, ld, c->dim, block, c->n_item_base, true));
        if (c->qtable) {  // query table rows, then their sampled columns
            const int32_t QS = c->n_qs;
            RKR_HIP(c, launch_select_rows(c->stream, block, m, c->n_item_base, c->qranks, QS,
                                          c->qtable + (U0 + u0) * QS));
            RKR_HIP(c, launch_gather_cols(c->stream, c->qtable + (U0 + u0) * QS, QS, c->qcol, S, m,
                                          c->table + (U0 + u0) * S, S));
        } else {
            RKR_HIP(c, launch_select_rows(c->stream, block, m, c->n_item_base, c->ranks, S, c->table + (U0 + u0) * S));
        }
    }
    if ((rc = refresh_guard_values(c, U0, n))) return rc;
    // side rows of the new users, sorted in place at the side arrays' stride
    for (int side = 0; side < 2; ++side) {
        float *arr = side == 0 ? c->side_ins : c->side_del;
        const int32_t ns = side == 0 ? c->n_ins : c->n_del, lds = side == 0 ? c->ld_ins : c->ld_del;
        const float *X = side == 0 ? c->items_ins : c->items_del;
        if (!ns) continue;
        float *rows = arr + U0 * lds;
        RKR_HIP(c, launch_scores_store(c->stream, c->users + U0 * ld, nullptr, n, ld, X, ns, ld, c->dim, rows, lds, true));
        RKR_HIP(c, launch_sort_rows_desc(c->stream, rows, n, ns, lds));
    }
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    c->h_users.insert(c->h_users.end(), h.begin(), h.end());
    c->max_abs_user_raw = mu_raw;
    c->max_abs_user = mu_new;
    if (c->kp) c->screen_abs = screen_abs_floor(c->kp, c->max_abs_user, c->max_abs_item);
    c->n_user = U1;
    return rkr::refresh_ins_guard_max(c);
}

int rkr_delete_users(rkr_ctx *c, const int64_t *ids, int64_t n) {
    if (!c || (n > 0 && !ids) || n < 0) return c ? set_err(c, RKR_ERR_INVALID, "delete_users: bad arguments") : RKR_ERR_INVALID;
    int rc = check_updatable(c);
    if (rc) return rc;
    if (n == 0) return RKR_OK;
    const int64_t U0 = c->n_user;
    const int32_t ld = c->ld;
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
    const int64_t nk = (int64_t)keep.size();
    int64_t *d_keep = nullptr;
    if ((rc = dev_alloc_t(c, &d_keep, (size_t)nk, "gather index"))) return rc;
    RKR_HIP(c, hipMemcpyAsync(d_keep, keep.data(), sizeof(int64_t) * nk, hipMemcpyHostToDevice, c->stream));
    // SampleSearchUpdate::DeleteUser (SampleSearchUpdate.hpp:89-109): later users shift down
    rc = keep_user_rows(c, &c->users, ld, ld, d_keep, nk, "users");
    if (!rc) rc = keep_user_rows(c, &c->table, c->n_sample, c->n_sample, d_keep, nk, "table");
    if (!rc) rc = keep_user_rows(c, &c->qtable, c->n_qs, c->n_qs, d_keep, nk, "query table");
    if (!rc) rc = keep_user_rows(c, &c->gvals, kGuardRow, kGuardRow, d_keep, nk, "guard values");
    if (!rc) rc = keep_user_rows(c, &c->side_ins, c->n_ins, c->ld_ins, d_keep, nk, "side array");
    if (!rc) rc = keep_user_rows(c, &c->side_del, c->n_del, c->ld_del, d_keep, nk, "side array");
    dev_free(c, d_keep, 0);
    if (rc) return rc;
    std::vector<float> hu((size_t)nk * c->dim);
    for (int64_t i = 0; i < nk; ++i)
        std::memcpy(&hu[(size_t)i * c->dim], &c->h_users[(size_t)keep[i] * c->dim], sizeof(float) * c->dim);
    c->h_users.swap(hu);
    c->n_user = nk;
    float *panels = nullptr, *unorm = nullptr;
    uint16_t *uhi = nullptr, *ulo = nullptr;
    if ((rc = dev_alloc_t(c, &panels, (size_t)((c->u_cap + 63) / 64) * 64 * ld, "user panels"))) return rc;
    RKR_HIP(c, launch_make_panels(c->stream, c->users, c->n_user, ld, panels));
    if (c->kp) {
        if ((rc = dev_alloc_t(c, &unorm, (size_t)c->u_cap, "user norms")) ||
            (rc = dev_alloc_t(c, &uhi, (size_t)c->u_cap * c->kp, "user hi")) ||
            (rc = dev_alloc_t(c, &ulo, (size_t)c->u_cap * c->kp, "user lo"))) {
            dev_free(c, panels, 0);
            dev_free(c, unorm, 0);
            dev_free(c, uhi, 0);
            return rc;
        }
        RKR_HIP(c, launch_user_norms(c->stream, c->users, c->n_user, ld, c->dim, unorm, c->rho));
        RKR_HIP(c, launch_prep_items(c->stream, c->users, c->n_user, ld, c->dim, c->kp, uhi, ulo, nullptr));
    }
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(c, c->usersT, 0);
    dev_free(c, c->user_norm, 0);
    dev_free(c, c->user_hi, 0);
    dev_free(c, c->user_lo, 0);
    c->usersT = panels;
    c->user_norm = unorm;
    c->user_hi = uhi;
    c->user_lo = ulo;
    return RKR_OK;
}

int rkr_reserve(rkr_ctx *c, int64_t n_user, int64_t n_inserted) {
    if (!c || n_user < 0 || n_inserted < 0 || n_inserted > kMaxSideItems)
        return c ? set_err(c, RKR_ERR_INVALID, "reserve: sizes out of range") : RKR_ERR_INVALID;
    int rc = check_updatable(c);
    if (rc) return rc;
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    if ((rc = reserve_users(c, std::max(n_user, c->n_user)))) return rc;
    const int64_t extra = std::max<int64_t>(0, n_inserted - c->n_ins);
    if (extra > 0) {
        init_item_map(c);
        if ((rc = reserve_items(c, c->n_item + extra)) ||
            (rc = reserve_list(c, &c->items_ins, &c->ins_cap, c->n_ins, n_inserted, "inserted items")))
            return rc;
    }
    // both side buffers at the reserved row length
    const int32_t ld = (int32_t)((n_inserted + 63) / 64 * 64);
    if (ld > c->ld_ins) {
        float *a = nullptr, *b = nullptr;
        if ((rc = dev_alloc_t(c, &a, (size_t)c->u_cap * ld, "side array")) ||
            (rc = dev_alloc_t(c, &b, (size_t)c->u_cap * ld, "side array"))) {
            dev_free(c, a, 0);
            return rc;
        }
        if ((rc = copy_rows(c, a, ld, c->side_ins, c->ld_ins, c->n_user, c->n_ins))) return rc;
        RKR_HIP(c, hipStreamSynchronize(c->stream));
        dev_free(c, c->side_ins, 0);
        dev_free(c, c->side_ins_spare, 0);
        c->side_ins = a;
        c->side_ins_spare = b;
        c->ld_ins = ld;
    } else if (!c->side_ins_spare && c->ld_ins > 0) {
        if ((rc = dev_alloc_t(c, &c->side_ins_spare, (size_t)c->u_cap * c->ld_ins, "side array"))) return rc;
    }
    // the merge score block of a full sort chunk over the reserved users
    const size_t blk = sizeof(float) * (size_t)c->u_cap * (size_t)std::min<int64_t>(kSortChunk, std::max<int64_t>(n_inserted, 1));
    return ensure(c, c->upd_scratch, blk, "update score block");
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
