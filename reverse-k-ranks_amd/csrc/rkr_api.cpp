// rkr_api.cpp -- the C ABI of librkr.so (include/rkr.h): context, resident
// device data, table build and the batched query pipeline.  Host-side logic
// only; every score, bound, count and selection runs in the HIP kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <string_view>
#include <unordered_map>
#include <thread>
#include <vector>

#include "rkr_internal.hpp"

namespace rkr {

int set_err(Ctx *c, int code, const std::string &msg) {
    if (c) {
        std::lock_guard<std::mutex> g(c->err_mu);
        c->err = msg;
    }
    return code;
}

const char *kVersion = "rkr 0.1 (gfx950)";

int dev_alloc(Ctx *c, void **p, size_t bytes, const char *what) {
    *p = nullptr;
    if (bytes == 0) return RKR_OK;
    std::lock_guard<std::mutex> g(c->mem_mu);
    if (c->held + bytes > c->budget)
        return set_err(c, RKR_ERR_NOMEM, std::string("device memory budget exceeded allocating ") + what + " (" +
                                             std::to_string(bytes) + " bytes, budget " + std::to_string(c->budget) +
                                             ", held " + std::to_string(c->held) + ")");
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(c, RKR_ERR_NOMEM, std::string("hipMalloc ") + what + ": " + hipGetErrorString(e));
    }
    c->held += bytes;
    c->alloc_bytes[*p] = bytes;
    return RKR_OK;
}

namespace {
thread_local bool tl_defer_frees = false;
}

void set_defer_frees(bool on) { tl_defer_frees = on; }

// The size recorded at allocation is what returns to the budget; `bytes` is
// only used for a pointer this context did not allocate.
void dev_free(Ctx *c, void *p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> g(c->mem_mu);
    auto it = c->alloc_bytes.find(p);
    if (it != c->alloc_bytes.end()) {
        bytes = it->second;
        c->alloc_bytes.erase(it);
    }
    if (tl_defer_frees) {
        c->retired.push_back({p, bytes});
        return;
    }
    (void)hipFree(p);
    c->held -= std::min(c->held, bytes);
}

void release_retired(Ctx *c) {
    std::lock_guard<std::mutex> g(c->mem_mu);
    for (const auto &r : c->retired) {
        (void)hipFree(r.p);
        c->held -= std::min(c->held, r.bytes);
    }
    c->retired.clear();
}

DataSnap snap_of(const Ctx *c) {
    DataSnap d;
    d.n_user = c->n_user;
    d.n_item = c->n_item;
    d.n_item_base = c->n_item_base;
    d.users = c->users;
    d.usersT = c->usersT;
    d.items = c->items;
    d.user_norm = c->user_norm;
    d.rho = c->rho;
    d.user_hi = c->user_hi;
    d.user_lo = c->user_lo;
    d.item_norm = c->item_norm;
    d.item_hi = c->item_hi;
    d.item_lo = c->item_lo;
    d.table = c->table;
    d.side_ins = c->side_ins;
    d.side_del = c->side_del;
    d.n_ins = c->n_ins;
    d.n_del = c->n_del;
    d.ld_ins = c->ld_ins;
    d.ld_del = c->ld_del;
    d.max_abs_user = c->max_abs_user;
    d.screen_abs = c->screen_abs;
    d.qtable = c->qtable;
    d.qranks = c->qranks;
    d.n_qs = c->n_qs;
    guard_columns(c, d.gcol, d.n_guard);
    d.gvals = c->gvals;
    d.ranks = c->ranks;
    d.n_sample = c->n_sample;
    d.table_eps = c->table_eps;
    d.qcol = c->qcol;
    d.hrk = std::make_shared<const std::vector<int32_t>>(c->qtable ? c->h_qranks : c->h_ranks);
    if (d.n_guard > 0) d.gsets = guard_sets(c);
    for (int j = 0; j < 8; ++j) d.ins_gmax[j] = c->ins_gmax[j];
    if (c->qsrp_users && c->qsrp_n_user == c->n_user) {
        d.qsrp_users = c->qsrp_users;
        d.qsrp_rot = c->qsrp_rot;
        d.qsrp_par = c->qsrp_par;
        d.qsrp_cd = c->qsrp_cd;
        d.qsrp_ldr = c->qsrp_ldr;
        d.qsrp_gcd = c->qsrp_gcd;
        d.qsrp_crot = c->qsrp_crot;
    }
    return d;
}

// 64-bit hash of n bytes salted with `salt`: four independent multiply-add
// lanes over 8-byte words (zero-padded tail), then an xorshift-multiply fold.
// The dedup key of a query row; equal hashes are confirmed by a byte compare.
uint64_t row_hash(const void *p, size_t n, uint32_t salt) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    const uint64_t m = 0xBF58476D1CE4E5B9ull;
    uint64_t h[4] = {0x9E3779B97F4A7C15ull ^ n, 0xC2B2AE3D27D4EB4Full ^ salt, 0x165667B19E3779F9ull,
                     0xD6E8FEB86659FD93ull};
    size_t i = 0;
    for (; i + 32 <= n; i += 32) {
        uint64_t w[4];
        std::memcpy(w, b + i, 32);
        for (int l = 0; l < 4; ++l) h[l] = (h[l] + w[l]) * m;
    }
    if (i < n) {
        uint64_t w[4] = {0, 0, 0, 0};
        std::memcpy(w, b + i, n - i);
        for (int l = 0; l < 4; ++l) h[l] = (h[l] + w[l]) * m;
    }
    uint64_t x = (h[0] ^ (h[0] >> 29)) + 3 * (h[1] ^ (h[1] >> 31)) + 5 * (h[2] ^ (h[2] >> 27)) +
                 7 * (h[3] ^ (h[3] >> 33));
    x ^= x >> 29;
    x *= 0x94D049BB133111EBull;
    return x ^ (x >> 32);
}

// Guard columns of the coarse bounds: the table columns nearest below ranks
// 4^j - 1 (3, 15, ..., 16383), at most 7, ascending.
void guard_columns(const Ctx *c, int32_t *col, int32_t &n) {
    const std::vector<int32_t> &hr = c->qtable ? c->h_qranks : c->h_ranks;
    n = 0;
    int32_t last = -1;
    for (int64_t target = 3; target <= 16383 && n < 7; target = target * 4 + 3) {
        const auto it = std::upper_bound(hr.begin(), hr.end(), (int32_t)target);
        const int32_t j = (int32_t)(it - hr.begin()) - 1;
        if (j > last) {
            col[n++] = j;
            last = j;
        }
    }
}

// Set 0: the guard columns.  Set 1 + beta: up to 7 columns inside guard bin
// beta (the columns after guard beta - 1 up to guard beta, the bin's last
// column included), spaced geometrically in rank, for the second coarse pass
// of queries whose k-th bound fell in that bin.
GuardSets guard_sets(const Ctx *c) {
    GuardSets gs{};
    GuardCols &gc = gs.set[0];
    guard_columns(c, gc.col, gc.n);
    const std::vector<int32_t> &hrk = c->qtable ? c->h_qranks : c->h_ranks;
    for (int32_t beta = 0; beta < gc.n; ++beta) {
        GuardCols &g2 = gs.set[1 + beta];
        const int32_t lo_col = beta == 0 ? 0 : gc.col[beta - 1] + 1, hi_col = gc.col[beta];
        if (hi_col <= lo_col) continue;
        const double ra = (double)hrk[lo_col] + 1.0, rb = (double)hrk[hi_col] + 1.0;
        int32_t last = lo_col - 1;
        for (int m = 1; m <= 6; ++m) {
            const double target = ra * std::pow(rb / ra, m / 7.0) - 1.0;
            int32_t j = last + 1;
            while (j + 1 < hi_col && (double)hrk[j + 1] <= target) ++j;
            if (j < hi_col && j > last && (double)hrk[j] <= target) {
                g2.col[g2.n++] = j;
                last = j;
            }
        }
        g2.col[g2.n++] = hi_col;
    }
    return gs;
}

int refresh_guard_values(Ctx *c, int64_t u0, int64_t n) {
    if (!c->gvals || n <= 0) return RKR_OK;
    const GuardSets gs = guard_sets(c);
    const float *T = c->qtable ? c->qtable : c->table;
    const int64_t S = c->qtable ? c->n_qs : c->n_sample;
    RKR_HIP(c, launch_gather_guards(c->stream, T, S, gs, u0, n, c->gvals));
    return RKR_OK;
}

std::vector<int32_t> query_table_ranks(const std::vector<int32_t> &sample, int64_t n_item, int mode) {
    std::vector<int32_t> r(sample);
    if (mode == 2) {
        // every rank below D, with D as large as the column cap allows
        int64_t lo = 0, hi = std::min<int64_t>(n_item, kMaxSample);
        while (lo < hi) {
            const int64_t D = (lo + hi + 1) / 2;
            int64_t above = 0;
            for (int32_t x : sample) above += x >= D ? 1 : 0;
            if (D + above <= kMaxSample) lo = D; else hi = D - 1;
        }
        for (int64_t x = 0; x < lo; ++x) r.push_back((int32_t)x);
    } else {
        for (int64_t x = 0; x < 256; ++x) r.push_back((int32_t)x);
        for (int64_t x = 256; x < 1024; x += 8) r.push_back((int32_t)x);
        for (int64_t x = 1024; x < 4096; x += 32) r.push_back((int32_t)x);
    }
    std::vector<int32_t> out;
    std::sort(r.begin(), r.end());
    for (int32_t x : r)
        if (x < n_item && (out.empty() || x != out.back())) out.push_back(x);
    if ((int64_t)out.size() > kMaxSample) out.clear();
    return out;
}

// Grow-only scratch buffer.
int ensure(Ctx *c, DeviceBuf &b, size_t bytes, const char *what) {
    if (b.bytes >= bytes) return RKR_OK;
    dev_free(c, b.p, b.bytes);
    b.p = nullptr;
    b.bytes = 0;
    size_t want = std::max(bytes, b.bytes * 3 / 2);
    int rc = dev_alloc(c, &b.p, want, what);
    if (rc != RKR_OK) {
        rc = dev_alloc(c, &b.p, bytes, what);
        if (rc != RKR_OK) return rc;
        want = bytes;
    }
    b.bytes = want;
    return RKR_OK;
}

int ensure_pinned(Ctx *c, PinnedBuf &b, size_t bytes) {
    if (b.bytes >= bytes) return RKR_OK;
    free_pinned(b);
    const size_t want = std::max(bytes, b.bytes + b.bytes / 2);
    if (hipHostMalloc(&b.p, want, hipHostMallocDefault) != hipSuccess) {
        b.p = nullptr;
        return set_err(c, RKR_ERR_NOMEM, "pinned host staging (" + std::to_string(want) + " bytes)");
    }
    b.bytes = want;
    return RKR_OK;
}

void free_pinned(PinnedBuf &b) {
    if (b.p) (void)hipHostFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

void free_qsrp(Ctx *c) {
    for (float *p : {c->qsrp_users, c->qsrp_rot, c->qsrp_par}) dev_free(c, p, 0);
    c->qsrp_users = c->qsrp_rot = c->qsrp_par = nullptr;
    c->qsrp_cd = c->qsrp_ldr = 0;
    c->qsrp_n_user = c->qsrp_fallback = 0;
}

namespace {

// Forgets the pending item updates.  keep_buffers leaves the side arrays and
// item lists allocated at their row lengths and capacities for the inserts
// that follow a rebuild (a serving loop's memory stays flat); with two buffers
// per side the next update writes the one no batch of the new state reads.
void free_update_state(Ctx *c, bool keep_buffers = false) {
    if (!keep_buffers) {
        for (float *p : {c->items_ins, c->items_del, c->side_ins, c->side_del, c->side_ins_spare, c->side_del_spare})
            dev_free(c, p, 0);
        c->items_ins = c->items_del = c->side_ins = c->side_del = nullptr;
        c->side_ins_spare = c->side_del_spare = nullptr;
        c->ld_ins = c->ld_del = 0;
        c->ins_cap = c->del_cap = 0;
    }
    c->n_ins = c->n_del = 0;
    for (int j = 0; j < 8; ++j) c->ins_gmax[j] = 0;
    if (c->items_base && c->items_base != c->items)
        dev_free(c, c->items_base, sizeof(float) * (size_t)c->n_item_base * c->ld);
    c->items_base = c->items;
    c->n_item_base = c->n_item;
    c->cur_src.clear();
    c->base_deleted.clear();
}

void free_table(Ctx *c) {
    free_qsrp(c);  // fitted to this table
    dev_free(c, c->ranks, 0);
    dev_free(c, c->table, 0);
    dev_free(c, c->qranks, 0);
    dev_free(c, c->qcol, 0);
    dev_free(c, c->qtable, 0);
    dev_free(c, c->gvals, 0);
    c->ranks = c->qranks = c->qcol = nullptr;
    c->table = c->qtable = c->gvals = nullptr;
    c->n_sample = c->n_qs = 0;
    c->h_ranks.clear();
    c->h_qranks.clear();
    c->table_eps = 0.0f;
}

void free_data(Ctx *c) {
    free_update_state(c);
    free_table(c);
    dev_free(c, c->upd_scratch.p, 0);
    c->upd_scratch = DeviceBuf{};
    c->u_cap = c->i_cap = 0;
    const size_t ld = (size_t)c->ld;
    dev_free(c, c->users, sizeof(float) * (size_t)c->n_user * ld);
    dev_free(c, c->usersT, sizeof(float) * (size_t)((c->n_user + 63) / 64) * 64 * ld);
    dev_free(c, c->items, sizeof(float) * (size_t)c->n_item * ld);
    dev_free(c, c->item_hi, sizeof(uint16_t) * (size_t)c->n_item * c->kp);
    dev_free(c, c->item_lo, sizeof(uint16_t) * (size_t)c->n_item * c->kp);
    dev_free(c, c->item_norm, sizeof(float) * (size_t)c->n_item);
    dev_free(c, c->user_norm, sizeof(float) * (size_t)c->n_user);
    dev_free(c, c->user_hi, 0);
    dev_free(c, c->user_lo, 0);
    c->user_hi = c->user_lo = nullptr;
    c->item_hi = c->item_lo = nullptr;
    c->item_norm = c->user_norm = nullptr;
    c->kp = 0;
    c->users = c->usersT = c->items = nullptr;
    c->n_user = c->n_item = 0;
    c->dim = c->ld = 0;
    c->h_users.clear();
}

void free_bufs(Ctx *c, std::initializer_list<DeviceBuf *> bufs) {
    for (DeviceBuf *b : bufs) {
        dev_free(c, b->p, b->bytes);
        b->p = nullptr;
        b->bytes = 0;
    }
}

void free_scratch(Ctx *c) {
    Slot &s = c->main;
    free_bufs(c, {&s.q_in, &s.qip, &s.bk, &s.qinfo, &s.cand, &s.refine_idx, &s.counts, &s.out, &s.recheck, &s.small,
                  &c->misc});
}

}  // namespace

int slot_init(Ctx *c, Slot &sl, bool own_stream, const std::vector<uint32_t> *cu_mask) {
    sl.own_stream = own_stream;
    if (own_stream && cu_mask) {
        RKR_HIP(c, hipExtStreamCreateWithCUMask(&sl.stream, (uint32_t)cu_mask->size(), cu_mask->data()));
    } else if (own_stream) {
        RKR_HIP(c, hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
    } else {
        sl.stream = c->stream;
    }
    for (auto &e : sl.ev) RKR_HIP(c, hipEventCreate(&e));
    return RKR_OK;
}

void slot_free(Ctx *c, Slot &sl) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    free_bufs(c, {&sl.q_in, &sl.qip, &sl.bk, &sl.qinfo, &sl.cand, &sl.refine_idx, &sl.counts, &sl.out, &sl.recheck,
                  &sl.small, &sl.hist, &sl.q_split, &sl.coarse_info, &sl.coarse_list});
    free_pinned(sl.stage);
    for (auto &e : sl.ev)
        if (e) { (void)hipEventDestroy(e); e = nullptr; }
    if (sl.own_stream && sl.stream) (void)hipStreamDestroy(sl.stream);
    sl.stream = nullptr;
}

namespace {

// FP32 normalisation exactly as VectorMatrix::vectorNormalize
// (src/include/struct/VectorMatrix.hpp:51-66).  This file is compiled with
// -ffp-contract=off, so x*x is rounded before the add.
void normalize_rows(float *u, int64_t n, int32_t d) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float *row = u + i * d;
        float l2 = 0.0f;
        for (int32_t j = 0; j < d; ++j) {
            const float sq = row[j] * row[j];
            l2 = l2 + sq;
        }
        l2 = std::sqrt(l2);
        if ((double)l2 <= 0.001) continue;
        for (int32_t j = 0; j < d; ++j) row[j] = row[j] / l2;
    }
}

bool ranks_valid(const int32_t *r, int32_t S, int64_t n_item) {
    if (S <= 0 || (int64_t)S >= n_item) return false;
    for (int32_t j = 0; j < S; ++j) {
        if (r[j] < 0 || (int64_t)r[j] >= n_item) return false;
        if (j > 0 && r[j] <= r[j - 1]) return false;
    }
    return true;
}

double ev_ms(hipEvent_t *ev, int a, int b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev[a], ev[b]) != hipSuccess) return 0.0;
    return ms;
}

}  // namespace
}  // namespace rkr

using namespace rkr;

struct rkr_ctx : public rkr::Ctx {};

extern "C" {

const char *rkr_version(void) { return kVersion; }

int rkr_create(int device, size_t mem_budget_bytes, rkr_ctx **out) {
    if (!out) return RKR_ERR_INVALID;
    *out = nullptr;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= 0) return RKR_ERR_NODEVICE;
    if (device < 0 || device >= n_dev) return RKR_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return RKR_ERR_HIP;
    rkr_ctx *c = new (std::nothrow) rkr_ctx();
    if (!c) return RKR_ERR_NOMEM;
    c->device = device;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { delete c; return RKR_ERR_HIP; }
    c->budget = mem_budget_bytes ? std::min(mem_budget_bytes, free_b) : (size_t)(0.9 * (double)free_b);
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && n_cu > 0)
        c->n_cu = n_cu;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return RKR_ERR_HIP; }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) { rkr_destroy(c); return RKR_ERR_HIP; }
    if (slot_init(c, c->main, false) != RKR_OK) { rkr_destroy(c); return RKR_ERR_HIP; }
    if (hipMalloc(&c->rho, 256) != hipSuccess) { c->rho = nullptr; rkr_destroy(c); return RKR_ERR_NOMEM; }
    // the no-allocation entry point's scratch: its address on this device,
    // looked up once per device (a failure is reported by that entry point)
    (void)noalloc_init(device);
    *out = c;
    return RKR_OK;
}

void rkr_destroy(rkr_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_scratch(c);
    slot_free(c, c->main);
    if (c->aux_ready) slot_free(c, c->aux);
    free_data(c);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->rho) (void)hipFree(c->rho);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *rkr_last_error(const rkr_ctx *c) { return c ? c->err.c_str() : "null context"; }

int rkr_set_option(rkr_ctx *c, int option, int64_t value) {
    if (!c) return RKR_ERR_INVALID;
    switch (option) {
        case RKR_OPT_FORM:
            if (value != RKR_FORM_PLAIN && value != RKR_FORM_SCREENED)
                return set_err(c, RKR_ERR_INVALID, "unknown form");
            c->form = (int)value;
            return RKR_OK;
        case RKR_OPT_OUTPUT:
            if (value != RKR_OUTPUT_EXACT && value != RKR_OUTPUT_REFERENCE)
                return set_err(c, RKR_ERR_INVALID, "unknown output order");
            c->output = (int)value;
            return RKR_OK;
        case RKR_OPT_QUERY_BATCH:
            if (value < 0) return set_err(c, RKR_ERR_INVALID, "query batch must be >= 0");
            c->query_batch = value;
            return RKR_OK;
        case RKR_OPT_BUILD_BATCH:
            if (value < 0) return set_err(c, RKR_ERR_INVALID, "build batch must be >= 0");
            c->build_batch = value;
            return RKR_OK;
        case RKR_OPT_SCREEN_CHUNK:
            if (value < 0 || value > (1 << 30)) return set_err(c, RKR_ERR_INVALID, "screen chunk must be in [0, 2^30]");
            c->screen_chunk = (int32_t)value;
            return RKR_OK;
        case RKR_OPT_DENSE_TOP:
            if (value < 0 || value > 2) return set_err(c, RKR_ERR_INVALID, "dense top must be 0, 1 or 2");
            c->dense_top = (int)value;
            return RKR_OK;
        case RKR_OPT_SCREEN_ROWS:
            if (value != 0 && value != 32 && value != 64) return set_err(c, RKR_ERR_INVALID, "screen rows must be 0, 32 or 64");
            c->screen_rows = (int32_t)value;
            return RKR_OK;
        case RKR_OPT_DEDUP:
            if (value != 0 && value != 1) return set_err(c, RKR_ERR_INVALID, "dedup must be 0 or 1");
            c->dedup = (int)value;
            return RKR_OK;
        case RKR_OPT_COARSE:
            if (value != 0 && value != 1) return set_err(c, RKR_ERR_INVALID, "coarse must be 0 or 1");
            c->coarse = (int)value;
            return RKR_OK;
        case RKR_OPT_QUERY_SLOTS:
            if (value != 1 && value != 2) return set_err(c, RKR_ERR_INVALID, "query slots must be 1 or 2");
            c->query_slots = (int)value;
            return RKR_OK;
        case RKR_OPT_QSRP:
            if (value != 0 && value != 1) return set_err(c, RKR_ERR_INVALID, "qsrp must be 0 or 1");
            c->qsrp = (int)value;
            return RKR_OK;
        default:
            return set_err(c, RKR_ERR_INVALID, "unknown option");
    }
}

int rkr_get_option(const rkr_ctx *c, int option, int64_t *value) {
    if (!c || !value) return RKR_ERR_INVALID;
    switch (option) {
        case RKR_OPT_FORM: *value = c->form; return RKR_OK;
        case RKR_OPT_OUTPUT: *value = c->output; return RKR_OK;
        case RKR_OPT_QUERY_BATCH: *value = c->query_batch; return RKR_OK;
        case RKR_OPT_BUILD_BATCH: *value = c->build_batch; return RKR_OK;
        case RKR_OPT_SCREEN_CHUNK: *value = c->screen_chunk; return RKR_OK;
        case RKR_OPT_DEDUP: *value = c->dedup; return RKR_OK;
        case RKR_OPT_SCREEN_ROWS: *value = c->screen_rows; return RKR_OK;
        case RKR_OPT_DENSE_TOP: *value = c->dense_top; return RKR_OK;
        case RKR_OPT_COARSE: *value = c->coarse; return RKR_OK;
        case RKR_OPT_QUERY_SLOTS: *value = c->query_slots; return RKR_OK;
        case RKR_OPT_QSRP: *value = c->qsrp; return RKR_OK;
        default: return RKR_ERR_INVALID;
    }
}

int rkr_sync(rkr_ctx *c) {
    if (!c) return RKR_ERR_INVALID;
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    return RKR_OK;
}

int rkr_normalize_users(float *users, int64_t n_user, int32_t dim) {
    if (!users || n_user < 0 || dim <= 0) return RKR_ERR_INVALID;
    normalize_rows(users, n_user, dim);
    return RKR_OK;
}

int rkr_uniform_sample_ranks(int64_t n_item, int32_t n_sample, int32_t *out) {
    // FindUniformSample, UniformSample.hpp:17-25
    if (!out || n_sample <= 0 || (int64_t)n_sample >= n_item) return RKR_ERR_INVALID;
    const int end_rank = (int)n_item - 1;
    const double delta = (end_rank - 0) * 1.0 / n_sample;
    for (int32_t j = 0; j < n_sample; ++j) out[j] = (int32_t)std::floor(j * delta);
    return RKR_OK;
}

int rkr_set_data(rkr_ctx *c, const float *users, int64_t n_user, const float *items, int64_t n_item, int32_t dim,
                 int normalize_users) {
    if (!c) return RKR_ERR_INVALID;
    if (!users || !items || n_user <= 0 || n_item <= 0 || dim <= 0)
        return set_err(c, RKR_ERR_INVALID, "set_data: empty or null input");
    if (n_user > INT32_MAX || n_item > INT32_MAX) return set_err(c, RKR_ERR_INVALID, "set_data: too many rows");
    if (dim > 4096) return set_err(c, RKR_ERR_INVALID, "set_data: dim > 4096 unsupported");
    {
        const double mu = max_abs(users, (size_t)n_user * dim), mp = max_abs(items, (size_t)n_item * dim);
        if (mu < 0 || mp < 0) return set_err(c, RKR_ERR_INVALID, "set_data: users and items must be finite");
        const double mu_n = normalize_users ? 1.0 : mu;  // normalised rows have |u_i| <= 1
        c->max_abs_user = mu_n;
        c->max_abs_item = mp;
        c->max_abs_user_raw = mu;
        if ((double)dim * mu_n * mp > kScoreLimit)
            return set_err(c, RKR_ERR_INVALID, "set_data: magnitudes would overflow FP32 scores");
    }
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    free_scratch(c);
    free_data(c);
    const int32_t ld = padded_dim(dim);
    c->n_user = n_user;
    c->n_item = n_item;
    c->u_cap = n_user;
    c->i_cap = n_item;
    c->normalize_users = normalize_users != 0;
    c->dim = dim;
    c->ld = ld;
    // host staging with padding
    c->h_users.assign(users, users + (size_t)n_user * dim);
    if (normalize_users) normalize_rows(c->h_users.data(), n_user, dim);
    // upload row-major with rows padded to ld floats (pad columns zero)
    auto upload = [&](const float *src, int64_t n, float **dst, const char *what) -> int {
        int rc = dev_alloc_t(c, dst, (size_t)n * ld, what);
        if (rc) return rc;
        const int64_t chunk = 65536;
        std::vector<float> tmp((size_t)std::min<int64_t>(chunk, n) * ld, 0.0f);
        for (int64_t r0 = 0; r0 < n; r0 += chunk) {
            const int64_t nr = std::min<int64_t>(chunk, n - r0);
            for (int64_t r = 0; r < nr; ++r)
                std::memcpy(&tmp[(size_t)r * ld], src + (size_t)(r0 + r) * dim, sizeof(float) * dim);
            RKR_HIP(c, hipMemcpy(*dst + r0 * ld, tmp.data(), sizeof(float) * (size_t)nr * ld, hipMemcpyHostToDevice));
        }
        return RKR_OK;
    };
    int rc = upload(c->h_users.data(), n_user, &c->users, "users");
    if (rc) { free_data(c); return rc; }
    rc = upload(items, n_item, &c->items, "items");
    if (rc) { free_data(c); return rc; }
    rc = dev_alloc_t(c, &c->usersT, (size_t)((n_user + 63) / 64) * 64 * ld, "user panels");
    if (rc) { free_data(c); return rc; }
    RKR_HIP(c, launch_make_panels(c->stream, c->users, n_user, ld, c->usersT));
    // screened form: hi/lo BF16 item rows and norm bounds (d <= 256)
    const int32_t kp = (dim + 31) / 32 * 32;
    if (kp <= screen_max_kp()) {
        c->kp = kp;
        // |s^ - s| <= C ||u|| ||p|| + abs: dropped split terms (3.2 * 2^-16),
        // FP32 accumulation of 3*kp exact products at <= 2^-23 each (x1.03 for
        // the split magnitudes), the sequential rule's own rounding
        // (gamma_dim at 2^-24), times a 1% margin that also absorbs the FP32
        // evaluation of the bound itself.
        const double n3 = 3.0 * kp, e23 = std::ldexp(1.0, -23), e24 = std::ldexp(1.0, -24);
        const double g_mfma = n3 * e23 / (1.0 - n3 * e23), g_rule = dim * e24 / (1.0 - dim * e24);
        const double C = (3.2 * std::ldexp(1.0, -16) + 1.03 * g_mfma + g_rule) * 1.01;
        c->screen_C = std::nextafter((float)C, INFINITY);
        // the coarse first pass's single product u_hi q_hi: its error scale
        // per query (k_prep_items) takes the users' largest BF16 residual
        // ratio rho (k_user_norms), the FP32 accumulation of kp exact
        // products and the rule's own rounding
        const double n1 = (double)kp;
        c->coarse_g1 = n1 * e23 / (1.0 - n1 * e23);
        c->coarse_grule = g_rule;
        c->screen_abs = screen_abs_floor(kp, c->max_abs_user, c->max_abs_item);
        if ((rc = dev_alloc_t(c, &c->item_hi, (size_t)n_item * kp, "item hi")) ||
            (rc = dev_alloc_t(c, &c->item_lo, (size_t)n_item * kp, "item lo")) ||
            (rc = dev_alloc_t(c, &c->item_norm, (size_t)n_item, "item norms")) ||
            (rc = dev_alloc_t(c, &c->user_norm, (size_t)n_user, "user norms")) ||
            (rc = dev_alloc_t(c, &c->user_hi, (size_t)n_user * kp, "user hi")) ||
            (rc = dev_alloc_t(c, &c->user_lo, (size_t)n_user * kp, "user lo"))) {
            free_data(c);
            return rc;
        }
        RKR_HIP(c, launch_prep_items(c->stream, c->items, n_item, ld, dim, kp, c->item_hi, c->item_lo, c->item_norm));
        RKR_HIP(c, hipMemsetAsync(c->rho, 0, sizeof(float), c->stream));
        RKR_HIP(c, launch_user_norms(c->stream, c->users, n_user, ld, dim, c->user_norm, c->rho));
        RKR_HIP(c, launch_prep_items(c->stream, c->users, n_user, ld, dim, kp, c->user_hi, c->user_lo, nullptr));
    }
    c->items_base = c->items;
    c->n_item_base = n_item;
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    return RKR_OK;
}

int rkr_build_table(rkr_ctx *c, const int32_t *ranks, int32_t S, double *build_ms) {
    if (!c) return RKR_ERR_INVALID;
    return rkr::build_table(c, ranks, S, build_ms, false);
}

}  // extern "C"

namespace rkr {

int rebuild_table(Ctx *c) {
    if (!c->table) return set_err(c, RKR_ERR_STATE, "rebuild: no table");
    std::vector<int32_t> ranks((size_t)c->n_sample);
    const int rc = rkr_uniform_sample_ranks(c->n_item, c->n_sample, ranks.data());
    if (rc) return set_err(c, rc, "rebuild: n_sample does not fit the current items");
    return build_table(c, ranks.data(), c->n_sample, nullptr, true);
}

int build_table(Ctx *c, const int32_t *ranks, int32_t S, double *build_ms, bool keep_update_buffers) {
    if (!c->users) return set_err(c, RKR_ERR_STATE, "build_table: no data");
    if (!ranks || S > kMaxSample || !ranks_valid(ranks, S, c->n_item))
        return set_err(c, RKR_ERR_INVALID, "build_table: ranks must be strictly increasing in [0, n_item), "
                                           "0 < n_sample < n_item, n_sample <= " + std::to_string(kMaxSample));
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    free_table(c);
    free_update_state(c, keep_update_buffers);  // the new table describes the current item set
    int rc = dev_alloc_t(c, &c->ranks, (size_t)S, "ranks");
    if (rc) return rc;
    rc = dev_alloc_t(c, &c->table, (size_t)std::max(c->u_cap, c->n_user) * S, "table");
    if (!rc) rc = dev_alloc_t(c, &c->gvals, (size_t)std::max(c->u_cap, c->n_user) * kGuardRow, "guard values");
    if (rc) { free_table(c); return rc; }
    c->n_sample = S;
    c->h_ranks.assign(ranks, ranks + S);
    RKR_HIP(c, hipMemcpyAsync(c->ranks, ranks, sizeof(int32_t) * S, hipMemcpyHostToDevice, c->stream));
    if (c->dense_top) {
        // the query table: select at its ranks, then copy the sampled columns
        c->h_qranks = query_table_ranks(c->h_ranks, c->n_item, c->dense_top);
        const int32_t nqs = (int32_t)c->h_qranks.size();
        if (nqs > S) {
            std::vector<int32_t> col((size_t)S);
            for (int32_t j = 0, q = 0; j < S; ++j) {
                while (c->h_qranks[q] != ranks[j]) ++q;
                col[j] = q;
            }
            if ((rc = dev_alloc_t(c, &c->qranks, (size_t)nqs, "query table ranks")) ||
                (rc = dev_alloc_t(c, &c->qcol, (size_t)S, "query table columns")) ||
                (rc = dev_alloc_t(c, &c->qtable, (size_t)std::max(c->u_cap, c->n_user) * nqs, "query table"))) {
                free_table(c);
                return rc;
            }
            c->n_qs = nqs;
            RKR_HIP(c, hipMemcpyAsync(c->qranks, c->h_qranks.data(), sizeof(int32_t) * nqs, hipMemcpyHostToDevice,
                                      c->stream));
            RKR_HIP(c, hipMemcpyAsync(c->qcol, col.data(), sizeof(int32_t) * S, hipMemcpyHostToDevice, c->stream));
            RKR_HIP(c, hipStreamSynchronize(c->stream));
        } else {
            c->h_qranks.clear();
        }
    }

    // user batch: the score block [nb x n_item] lives in HBM between the
    // GEMM and the selection; size it from the remaining budget (the
    // reference sizes its batch from free memory the same way,
    // GPUComputeSort.hpp:112-135).
    const size_t row_bytes = sizeof(float) * (size_t)c->n_item;
    size_t avail = c->budget > c->held ? c->budget - c->held : 0;
    int64_t nb = (int64_t)std::min<size_t>(avail / 2 / row_bytes, (size_t)(8ull << 30) / row_bytes);
    if (c->build_batch > 0) nb = std::min<int64_t>(nb, c->build_batch);
    nb = std::min<int64_t>(nb, c->n_user);
    if (nb > 128) nb = nb / 128 * 128;
    if (nb <= 0) { free_table(c); return set_err(c, RKR_ERR_NOMEM, "build_table: no memory for a score block"); }
    DeviceBuf &blk = c->misc;
    rc = ensure(c, blk, (size_t)nb * row_bytes, "score block");
    if (rc) { free_table(c); return rc; }
    float *scores = reinterpret_cast<float *>(blk.p);
    RKR_HIP(c, hipEventRecord(c->ev[0], c->stream));
    for (int64_t u0 = 0; u0 < c->n_user; u0 += nb) {
        const int64_t n = std::min<int64_t>(nb, c->n_user - u0);
        RKR_HIP(c, launch_scores_store(c->stream, c->users + u0 * c->ld, nullptr, n, c->ld, c->items, c->n_item, c->ld,
                                       c->dim, scores, c->n_item, true));
        if (c->qtable) {
            RKR_HIP(c, launch_select_rows(c->stream, scores, n, c->n_item, c->qranks, c->n_qs,
                                          c->qtable + u0 * c->n_qs));
            RKR_HIP(c, launch_gather_cols(c->stream, c->qtable + u0 * c->n_qs, c->n_qs, c->qcol, S, n,
                                          c->table + u0 * S, S));
        } else {
            RKR_HIP(c, launch_select_rows(c->stream, scores, n, c->n_item, c->ranks, S, c->table + u0 * S));
        }
    }
    if ((rc = refresh_guard_values(c, 0, c->n_user))) { free_table(c); return rc; }
    RKR_HIP(c, hipEventRecord(c->ev[1], c->stream));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    if (build_ms) *build_ms = ev_ms(c->ev, 0, 1);
    return RKR_OK;
}

}  // namespace rkr

extern "C" {

int rkr_set_table(rkr_ctx *c, const int32_t *ranks, int32_t S, const float *table, float eps) {
    if (!c) return RKR_ERR_INVALID;
    if (!c->users) return set_err(c, RKR_ERR_STATE, "set_table: no data");
    if (!ranks || !table || S > kMaxSample || !ranks_valid(ranks, S, c->n_item))
        return set_err(c, RKR_ERR_INVALID, "set_table: bad ranks");
    if (!(eps >= 0.0f) || std::isinf(eps)) return set_err(c, RKR_ERR_INVALID, "set_table: eps must be finite, >= 0");
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    free_table(c);
    free_update_state(c);  // the installed table describes the current item set
    int rc = dev_alloc_t(c, &c->ranks, (size_t)S, "ranks");
    if (rc) return rc;
    rc = dev_alloc_t(c, &c->table, (size_t)std::max(c->u_cap, c->n_user) * S, "table");
    if (!rc) rc = dev_alloc_t(c, &c->gvals, (size_t)std::max(c->u_cap, c->n_user) * kGuardRow, "guard values");
    if (rc) { free_table(c); return rc; }
    c->n_sample = S;
    c->h_ranks.assign(ranks, ranks + S);
    c->table_eps = eps;
    RKR_HIP(c, hipMemcpy(c->ranks, ranks, sizeof(int32_t) * S, hipMemcpyHostToDevice));
    RKR_HIP(c, hipMemcpy(c->table, table, sizeof(float) * (size_t)c->n_user * S, hipMemcpyHostToDevice));
    if ((rc = refresh_guard_values(c, 0, c->n_user))) { free_table(c); return rc; }
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    return RKR_OK;
}

extern "C++" {
namespace {
// A device allocation freed when the scope ends.
struct ScopedDev {
    rkr::Ctx *c;
    void *p = nullptr;
    size_t bytes = 0;
    int alloc(size_t n, const char *what) {
        rkr::dev_free(c, p, bytes);
        p = nullptr;
        bytes = n;
        return rkr::dev_alloc(c, &p, n, what);
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
    ~ScopedDev() { rkr::dev_free(c, p, bytes); }
};

// Runs fn(i0, i1) over [0, n) in up to 16 ranges, one host thread per range and
// no range under min_per units; a range whose thread cannot start runs here.
template <class F>
void host_ranges(int64_t n, int64_t min_per, F &&fn) {
    if (n <= 0) return;
    const int64_t nt = std::min<int64_t>({(int64_t)std::max(1u, std::thread::hardware_concurrency()), 16,
                                          std::max<int64_t>(1, n / std::max<int64_t>(1, min_per))});
    std::vector<std::thread> th;
    int64_t launched = 1;
    try {
        th.reserve((size_t)nt);
        for (; launched < nt; ++launched) th.emplace_back(fn, n * launched / nt, n * (launched + 1) / nt);
    } catch (...) {
    }
    fn(0, n / nt);
    for (int64_t t = launched; t < nt; ++t) fn(n * t / nt, n * (t + 1) / nt);
    for (auto &x : th) x.join();
}

// Device-to-host copies through two pinned staging buffers: while one chunk
// crosses the link, the host consumes the previous one.  A plain copy into
// pageable memory runs at a fraction of the link's rate.
constexpr size_t kStageBytes = (size_t)32 << 20;
struct Staging {
    hipStream_t s;
    size_t chunk = 0;
    void *p[2] = {nullptr, nullptr};
    hipEvent_t e[2] = {nullptr, nullptr};
    explicit Staging(hipStream_t st) : s(st) {}
    Staging(const Staging &) = delete;
    Staging &operator=(const Staging &) = delete;
    ~Staging() {
        (void)hipStreamSynchronize(s);  // no copy into a buffer outlives it
        for (int k = 0; k < 2; ++k) {
            if (p[k]) {
                if (pinned) (void)hipHostFree(p[k]);
                else std::free(p[k]);
            }
            if (e[k]) (void)hipEventDestroy(e[k]);
        }
    }
    bool pinned = true;
    // Pinned buffers when the runtime grants them; otherwise ordinary pageable
    // buffers, which the same asynchronous copies fill at the pageable rate.
    hipError_t init(size_t bytes) {
        chunk = std::max<size_t>(bytes, 1);
        for (int k = 0; k < 2 && pinned; ++k)
            if (hipHostMalloc(&p[k], chunk, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                for (int j = 0; j < k; ++j) (void)hipHostFree(p[j]);
                p[0] = p[1] = nullptr;
                pinned = false;
            }
        for (int k = 0; k < 2; ++k) {
            if (!pinned && !(p[k] = std::malloc(chunk))) return hipErrorOutOfMemory;
            const hipError_t r = hipEventCreateWithFlags(&e[k], hipEventDisableTiming);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }
    // Chunk k of n_chunk: issue(k, buf) enqueues its copy into buf on the stream
    // (at most `chunk` bytes); take(k, buf) consumes it once it has arrived and
    // returns false to stop.
    template <class Issue, class Take>
    hipError_t run(int64_t n_chunk, Issue &&issue, Take &&take) {
        auto put = [&](int64_t k) {
            const hipError_t r = issue(k, p[k & 1]);
            return r == hipSuccess ? hipEventRecord(e[k & 1], s) : r;
        };
        hipError_t r = n_chunk > 0 ? put(0) : hipSuccess;
        if (r == hipSuccess && n_chunk > 1) r = put(1);
        for (int64_t k = 0; r == hipSuccess && k < n_chunk; ++k) {
            r = hipEventSynchronize(e[k & 1]);
            if (r != hipSuccess || !take(k, static_cast<const char *>(p[k & 1]))) break;
            if (k + 2 < n_chunk) r = put(k + 2);  // the buffer just consumed takes the chunk after next
        }
        return r;
    }
    // dst[0, bytes) = src[0, bytes), src in device memory; host threads copy out.
    hipError_t to_host(void *dst, const void *src, size_t bytes) {
        const int64_t n_chunk = (int64_t)((bytes + chunk - 1) / chunk);
        return run(
            n_chunk,
            [&](int64_t k, void *buf) {
                const size_t off = (size_t)k * chunk, len = std::min(chunk, bytes - off);
                return hipMemcpyAsync(buf, static_cast<const char *>(src) + off, len, hipMemcpyDeviceToHost, s);
            },
            [&](int64_t k, const char *buf) {
                const size_t off = (size_t)k * chunk, len = std::min(chunk, bytes - off);
                char *d = static_cast<char *>(dst) + off;
                host_ranges((int64_t)len, (int64_t)1 << 20,
                            [d, buf](int64_t a, int64_t b) { std::memcpy(d + a, buf + a, (size_t)(b - a)); });
                return true;
            });
    }
};
}  // namespace

// Calls sink(rows, user, m) for users [u0, u0 + n) in batches of m rows: rows
// (device memory, stride n_item) holds each user's scores against the current
// items, sorted in descending order.  The block size follows the memory budget
// as the table build's does.
template <class Sink>
static int sorted_score_batches(rkr_ctx *c, int64_t u0, int64_t n, const char *what, Sink &&sink,
                                double *gemm_ms = nullptr, double *sort_ms = nullptr) {
    const int64_t I = c->n_item;
    const size_t row_bytes = sizeof(float) * (size_t)I;
    const size_t avail = c->budget > c->held ? c->budget - c->held : 0;
    int64_t nb = (int64_t)std::min<size_t>(avail / 3 / row_bytes, (size_t)(2ull << 30) / row_bytes);
    if (c->build_batch > 0) nb = std::min<int64_t>(nb, c->build_batch);
    nb = std::min<int64_t>(nb, n);
    if (nb <= 0) return set_err(c, RKR_ERR_NOMEM, std::string(what) + ": no memory for a score block");
    ScopedDev blk_buf{c}, tmp_buf{c};
    int rc = blk_buf.alloc((size_t)nb * row_bytes, "score block");
    if (!rc) rc = tmp_buf.alloc((size_t)nb * row_bytes, "sort scratch");
    if (rc) return rc;
    float *blk = blk_buf.as<float>(), *tmp = tmp_buf.as<float>();
    for (int64_t b0 = 0; b0 < n; b0 += nb) {
        const int64_t m = std::min<int64_t>(nb, n - b0);
        RKR_HIP(c, hipEventRecord(c->ev[0], c->stream));
        RKR_HIP(c, launch_scores_store(c->stream, c->users + (u0 + b0) * c->ld, nullptr, m, c->ld, c->items, I, c->ld,
                                       c->dim, blk, I, true));
        RKR_HIP(c, hipEventRecord(c->ev[1], c->stream));
        RKR_HIP(c, launch_sort_rows(c->stream, blk, I, m, I, blk, tmp, I, c->n_cu));
        RKR_HIP(c, hipEventRecord(c->ev[2], c->stream));
        if ((rc = sink(blk, u0 + b0, m))) return rc;  // the sink synchronises the stream
        if (gemm_ms) *gemm_ms += ev_ms(c->ev, 0, 1);
        if (sort_ms) *sort_ms += ev_ms(c->ev, 1, 2);
    }
    return RKR_OK;
}
}  // extern "C++"

int rkr_get_table(rkr_ctx *c, int64_t user_begin, int64_t n_rows, float *out) {
    if (!c || !out) return RKR_ERR_INVALID;
    if (!c->table) return set_err(c, RKR_ERR_STATE, "get_table: no table");
    if (user_begin < 0 || n_rows < 0 || user_begin + n_rows > c->n_user)
        return set_err(c, RKR_ERR_INVALID, "get_table: row range out of bounds");
    if (n_rows == 0) return RKR_OK;
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    const size_t bytes = sizeof(float) * (size_t)n_rows * c->n_sample;
    Staging stg{c->stream};
    RKR_HIP(c, stg.init(std::min(kStageBytes, bytes)));
    RKR_HIP(c, stg.to_host(out, c->table + user_begin * c->n_sample, bytes));
    return RKR_OK;
}

int rkr_sorted_scores(rkr_ctx *c, int64_t user_begin, int64_t n_rows, float *out) {
    // ComputeScoreTableBatch::ComputeSortItemsBatch (ComputeScoreTableBatch.hpp:79-92)
    if (!c || (n_rows > 0 && !out)) return c ? set_err(c, RKR_ERR_INVALID, "sorted_scores: bad arguments") : RKR_ERR_INVALID;
    if (!c->users) return set_err(c, RKR_ERR_STATE, "sorted_scores: no data");
    if (user_begin < 0 || n_rows < 0 || user_begin + n_rows > c->n_user)
        return set_err(c, RKR_ERR_INVALID, "sorted_scores: row range out of bounds");
    if (n_rows == 0) return RKR_OK;
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t I = c->n_item;
    Staging stg{c->stream};
    RKR_HIP(c, stg.init(std::min(kStageBytes, sizeof(float) * (size_t)(n_rows * I))));
    return sorted_score_batches(c, user_begin, n_rows, "sorted_scores", [&](const float *rows, int64_t u, int64_t m) {
        RKR_HIP(c, stg.to_host(out + (u - user_begin) * I, rows, sizeof(float) * (size_t)(m * I)));
        return (int)RKR_OK;
    });
}

int rkr_save_score_table(rkr_ctx *c, const char *path, int64_t n_rows, double *timing_ms) {
    // ReadAll::BuildIndexLoop layout (ReadAll.hpp:80-87) as BuildScoreTableBatch.cpp:69-144 writes it
    if (!c || !path) return c ? set_err(c, RKR_ERR_INVALID, "save_score_table: bad arguments") : RKR_ERR_INVALID;
    if (!c->users) return set_err(c, RKR_ERR_STATE, "save_score_table: no data");
    if (n_rows < 0) n_rows = c->n_user;
    if (n_rows > c->n_user) return set_err(c, RKR_ERR_INVALID, "save_score_table: more rows than users");
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t I = c->n_item;
    // the file is written straight from the pinned staging buffers, one chunk
    // while the next crosses the link
    Staging stg{c->stream};
    RKR_HIP(c, stg.init(std::min(kStageBytes, sizeof(float) * (size_t)(n_rows * I))));
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_err(c, RKR_ERR_IO, std::string("save_score_table: cannot open ") + path);
    double gemm_ms = 0, sort_ms = 0, write_s = 0;
    const double t0 = rkr_now();
    int rc = n_rows == 0 ? (int)RKR_OK
                         : sorted_score_batches(c, 0, n_rows, "save_score_table",
                                                [&](const float *rows, int64_t, int64_t m) {
        const size_t bytes = sizeof(float) * (size_t)(m * I);
        bool ok = true;
        RKR_HIP(c, stg.run(
                       (int64_t)((bytes + stg.chunk - 1) / stg.chunk),
                       [&](int64_t k, void *buf) {
                           const size_t off = (size_t)k * stg.chunk, len = std::min(stg.chunk, bytes - off);
                           return hipMemcpyAsync(buf, reinterpret_cast<const char *>(rows) + off, len,
                                                 hipMemcpyDeviceToHost, c->stream);
                       },
                       [&](int64_t k, const char *buf) {
                           const size_t off = (size_t)k * stg.chunk, len = std::min(stg.chunk, bytes - off);
                           const double w0 = rkr_now();
                           ok = std::fwrite(buf, 1, len, f) == len;
                           write_s += rkr_now() - w0;
                           return ok;
                       }));
        if (!ok) return set_err(c, RKR_ERR_IO, std::string("save_score_table: write failed on ") + path);
        return (int)RKR_OK;
    }, &gemm_ms, &sort_ms);
    const double w0 = rkr_now();
    if (std::fclose(f) != 0 && !rc) rc = set_err(c, RKR_ERR_IO, std::string("save_score_table: close failed on ") + path);
    write_s += rkr_now() - w0;
    if (timing_ms) {
        timing_ms[0] = gemm_ms;
        timing_ms[1] = sort_ms;
        timing_ms[2] = (rkr_now() - t0 - write_s) * 1e3;
        timing_ms[3] = write_s * 1e3;
    }
    return rc;
}

int rkr_query_rank_distribution(rkr_ctx *c, const int32_t *item_ids, int32_t n_sample_item, int32_t *accu,
                                double *device_ms) {
    // ComputeQueryRankByComputation (src/app/build-index/QueryRankByComputation.hpp:122-242): for every sample
    // item s and user u, rank(u, s) = #{items scoring above <u, s>}, the lower_bound with greater<float> in
    // u's sorted row (QueryRankGPU.hpp:63-92); accu[s][r] = #{u : rank(u, s) <= r}, r = 0..n_item.
    if (!c || !item_ids || !accu || n_sample_item <= 0)
        return c ? set_err(c, RKR_ERR_INVALID, "query_rank_distribution: bad arguments") : RKR_ERR_INVALID;
    if (!c->users) return set_err(c, RKR_ERR_STATE, "query_rank_distribution: no data");
    const int64_t I = c->n_item, W = I + 1, ns = n_sample_item;
    std::vector<int64_t> ids((size_t)ns);
    for (int64_t s = 0; s < ns; ++s) {
        if (item_ids[s] < 0 || item_ids[s] >= I) return set_err(c, RKR_ERR_INVALID, "query_rank_distribution: bad item id");
        ids[s] = item_ids[s];
    }
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    ScopedDev smat{c}, qbuf{c}, dids{c}, hist{c};
    int rc = smat.alloc(sizeof(float) * (size_t)(ns * c->ld), "sample items");
    if (!rc) rc = dids.alloc(sizeof(int64_t) * (size_t)ns, "sample ids");
    if (!rc) rc = hist.alloc(sizeof(int32_t) * (size_t)(ns * W), "rank histogram");
    if (rc) return rc;
    int32_t *h = hist.as<int32_t>();
    RKR_HIP(c, hipMemcpyAsync(dids.p, ids.data(), dids.bytes, hipMemcpyHostToDevice, c->stream));
    RKR_HIP(c, hipMemsetAsync(hist.p, 0, hist.bytes, c->stream));
    RKR_HIP(c, launch_gather_rows(c->stream, c->items, c->ld, dids.as<const int64_t>(), ns, c->ld, smat.as<float>(),
                                  c->ld));
    double ms = 0;
    rc = sorted_score_batches(c, 0, c->n_user, "query_rank_distribution", [&](const float *rows, int64_t u, int64_t m) {
        if (qbuf.bytes < sizeof(float) * (size_t)(m * ns)) {  // the first batch is the largest
            const int r2 = qbuf.alloc(sizeof(float) * (size_t)(m * ns), "sample scores");
            if (r2) return r2;
        }
        RKR_HIP(c, hipEventRecord(c->ev[3], c->stream));
        RKR_HIP(c, launch_scores_store(c->stream, c->users + u * c->ld, nullptr, m, c->ld, smat.as<float>(), ns, c->ld,
                                       c->dim, qbuf.as<float>(), ns, true));
        RKR_HIP(c, launch_rank_hist(c->stream, rows, I, m, I, qbuf.as<float>(), ns, (int32_t)ns, h));
        RKR_HIP(c, hipEventRecord(c->ev[4], c->stream));
        RKR_HIP(c, hipEventSynchronize(c->ev[4]));
        ms += ev_ms(c->ev, 0, 2) + ev_ms(c->ev, 3, 4);
        return (int)RKR_OK;
    });
    if (rc) return rc;
    // The counts leave through the pinned staging buffers in chunks of whole
    // rows; host threads turn each chunk's counts per rank into users with
    // rank <= r, written into the caller's rows, while the next chunk copies.
    const int64_t chunk_rows = std::min<int64_t>(ns, std::max<int64_t>(1, (int64_t)(kStageBytes / sizeof(int32_t)) / W));
    Staging stg{c->stream};
    RKR_HIP(c, stg.init(sizeof(int32_t) * (size_t)(chunk_rows * W)));
    RKR_HIP(c, stg.run(
                   (ns + chunk_rows - 1) / chunk_rows,
                   [&](int64_t k, void *buf) {
                       const int64_t r0 = k * chunk_rows, nr = std::min(chunk_rows, ns - r0);
                       return hipMemcpyAsync(buf, h + r0 * W, sizeof(int32_t) * (size_t)(nr * W),
                                             hipMemcpyDeviceToHost, c->stream);
                   },
                   [&](int64_t k, const char *buf) {
                       const int64_t r0 = k * chunk_rows, nr = std::min(chunk_rows, ns - r0);
                       const int32_t *src = reinterpret_cast<const int32_t *>(buf);
                       host_ranges(nr, std::max<int64_t>(1, ((int64_t)1 << 16) / W), [src, accu, r0, W](int64_t s0, int64_t s1) {
                           for (int64_t s = s0; s < s1; ++s) {
                               const int32_t *x = src + s * W;
                               int32_t *row = accu + (r0 + s) * W;
                               int32_t run = 0;
                               for (int64_t r = 0; r < W; ++r) row[r] = run += x[r];
                           }
                       });
                       return true;
                   }));
    if (device_ms) *device_ms = ms;
    return RKR_OK;
}

int rkr_get_ranks(const rkr_ctx *c, int32_t *out, int32_t *n_sample) {
    if (!c) return RKR_ERR_INVALID;
    if (n_sample) *n_sample = c->n_sample;
    if (out && c->n_sample) std::memcpy(out, c->h_ranks.data(), sizeof(int32_t) * c->n_sample);
    return RKR_OK;
}

int rkr_save_table(rkr_ctx *c, const char *path) {
    // SampleSearch::SaveIndex layout (SampleSearch.hpp:186-194)
    if (!c || !path) return RKR_ERR_INVALID;
    if (!c->table) return set_err(c, RKR_ERR_STATE, "save_table: no table");
    // The table describes the item set it was built from; with pending inserts
    // or deletes a saved file would pair those rows with the current n_item and
    // lose the updates' rank offsets when it is loaded back.
    if (c->n_ins > 0 || c->n_del > 0)
        return set_err(c, RKR_ERR_STATE,
                       "save_table: item updates are pending; rebuild the table (rkr_build_table) before saving");
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    // the table is written straight from the pinned staging buffers, one chunk
    // while the next crosses the link
    const size_t bytes = sizeof(float) * (size_t)c->n_user * c->n_sample;
    Staging stg{c->stream};
    RKR_HIP(c, stg.init(std::min(kStageBytes, std::max<size_t>(bytes, 1))));
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_err(c, RKR_ERR_IO, std::string("save_table: cannot open ") + path);
    const size_t hdr[3] = {(size_t)c->n_sample, (size_t)c->n_item, (size_t)c->n_user};
    bool ok = std::fwrite(hdr, sizeof(size_t), 3, f) == 3;
    ok = ok && std::fwrite(c->h_ranks.data(), sizeof(int32_t), c->n_sample, f) == (size_t)c->n_sample;
    const hipError_t e = !ok ? hipSuccess : stg.run(
        (int64_t)((bytes + stg.chunk - 1) / stg.chunk),
        [&](int64_t k, void *buf) {
            const size_t off = (size_t)k * stg.chunk, len = std::min(stg.chunk, bytes - off);
            return hipMemcpyAsync(buf, reinterpret_cast<const char *>(c->table) + off, len, hipMemcpyDeviceToHost,
                                  c->stream);
        },
        [&](int64_t k, const char *buf) {
            const size_t off = (size_t)k * stg.chunk, len = std::min(stg.chunk, bytes - off);
            return ok = std::fwrite(buf, 1, len, f) == len;
        });
    ok = (std::fclose(f) == 0) && ok;
    if (e != hipSuccess) return set_err(c, RKR_ERR_HIP, std::string("save_table: ") + hipGetErrorString(e));
    if (!ok) return set_err(c, RKR_ERR_IO, std::string("save_table: short write to ") + path);
    return RKR_OK;
}

int rkr_load_table(rkr_ctx *c, const char *path, float eps) {
    // SampleSearch::LoadIndex layout (SampleSearch.hpp:229-241)
    if (!c || !path) return RKR_ERR_INVALID;
    if (!c->users) return set_err(c, RKR_ERR_STATE, "load_table: no data");
    FILE *f = std::fopen(path, "rb");
    if (!f) return set_err(c, RKR_ERR_IO, std::string("load_table: cannot open ") + path);
    size_t hdr[3];
    if (std::fread(hdr, sizeof(size_t), 3, f) != 3) { std::fclose(f); return set_err(c, RKR_ERR_IO, "load_table: short header"); }
    const size_t S = hdr[0], I = hdr[1], U = hdr[2];
    if ((int64_t)I != c->n_item || (int64_t)U != c->n_user || S == 0 || S > (size_t)kMaxSample) {
        std::fclose(f);
        return set_err(c, RKR_ERR_IO, "load_table: header (n_sample " + std::to_string(S) + ", n_item " +
                                          std::to_string(I) + ", n_user " + std::to_string(U) +
                                          ") does not match the data");
    }
    std::vector<int32_t> r(S);
    std::vector<float> t(S * U);
    bool ok = std::fread(r.data(), sizeof(int32_t), S, f) == S;
    ok = ok && std::fread(t.data(), sizeof(float), t.size(), f) == t.size();
    std::fclose(f);
    if (!ok) return set_err(c, RKR_ERR_IO, "load_table: file too short");
    return rkr_set_table(c, r.data(), (int32_t)S, t.data(), eps);
}

int64_t rkr_index_size_bytes(const rkr_ctx *c) {
    if (!c) return 0;
    return (int64_t)sizeof(int32_t) * c->n_sample + (int64_t)sizeof(float) * c->n_user * c->n_sample;
}

int32_t rkr_query_table_cols(const rkr_ctx *c) { return c && c->qtable ? c->n_qs : 0; }

}  // extern "C"

namespace {

constexpr int32_t kSplitMinQueries = 512;

// Answers queries [0, n) on the main slot, or -- with two query slots, at
// least kSplitMinQueries queries and two or more device batches -- the first
// half of the batches on the main slot and the rest on the aux slot from a
// second host thread.  Each slot synchronises its own stream between kernels
// while the other's kernels run.  A single batch is not split: two half
// batches repeat the per-batch passes over every user (measured slower at
// 1,000 queries over 1M users, no faster over 100k).
int query_split(rkr_ctx *c, const DataSnap &snap, const float *q, int32_t n, const int32_t *k, int32_t max_k,
                int32_t *ou, int32_t *orank, float *oqip, rkr_query_stats *st, rkr_batch_timing *timing) {
    const int64_t nqb = c->query_batch > 0 ? c->query_batch : 1024;
    const int64_t nb = (n + nqb - 1) / nqb;
    if (c->query_slots < 2 || n < kSplitMinQueries || nb < 2)
        return query_on_slot(c, c->main, snap, q, n, k, max_k, ou, orank, oqip, st, timing);
    if (!c->aux_ready) {
        if (int rc = slot_init(c, c->aux, true, nullptr)) {
            slot_free(c, c->aux);
            return rc;
        }
        c->aux_ready = true;
    }
    const int32_t n1 = (int32_t)std::min<int64_t>(n - 1, (nb + 1) / 2 * nqb);
    const int32_t n2 = n - n1;
    const size_t o = (size_t)n1 * max_k;
    rkr_batch_timing t1{}, t2{};
    int rc2 = RKR_OK;
    const auto t0 = std::chrono::steady_clock::now();
    std::thread th([&] {
        if (hipSetDevice(c->device) != hipSuccess) {
            rc2 = set_err(c, RKR_ERR_HIP, "query: hipSetDevice failed on the second query thread");
            return;
        }
        rc2 = query_on_slot(c, c->aux, snap, q + (size_t)n1 * c->dim, n2, k + n1, max_k, ou + o, orank + o,
                            oqip ? oqip + o : nullptr, st ? st + n1 : nullptr, &t2);
    });
    const int rc1 = query_on_slot(c, c->main, snap, q, n1, k, max_k, ou, orank, oqip, st, &t1);
    th.join();
    if (rc1) return rc1;
    if (rc2) return rc2;
    if (st)
        for (int32_t i = 0; i < n2; ++i) st[n1 + i].query_id += n1;
    if (timing) {
        rkr_batch_timing &t = *timing;
        t = t1;
        t.ip_ms += t2.ip_ms; t.bound_ms += t2.bound_ms; t.prune_ms += t2.prune_ms; t.refine_ms += t2.refine_ms;
        t.recheck_ms += t2.recheck_ms; t.topk_ms += t2.topk_ms; t.h2d_ms += t2.h2d_ms; t.d2h_ms += t2.d2h_ms;
        t.n_batches += t2.n_batches; t.n_refined += t2.n_refined; t.n_rechecked += t2.n_rechecked;
        t.n_screened += t2.n_screened; t.max_screen_err = std::max(t.max_screen_err, t2.max_screen_err);
        t.n_listed += t2.n_listed; t.n_fallback += t2.n_fallback; t.n_coarse_refined += t2.n_coarse_refined;
        t.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return RKR_OK;
}

}  // namespace

extern "C" {

int rkr_query(rkr_ctx *c, const float *queries, int32_t n_query, const int32_t *topk, int32_t max_k,
              int32_t *out_user, int32_t *out_rank, float *out_qip, rkr_query_stats *stats,
              rkr_batch_timing *timing) {
    if (!c) return RKR_ERR_INVALID;
    const DataSnap snap = snap_of(c);
    int rc = validate_queries(c, snap, queries, n_query, topk, max_k);
    if (rc) return rc;
    if (!out_user || !out_rank) return set_err(c, RKR_ERR_INVALID, "query: null output buffer");
    if (timing) std::memset(timing, 0, sizeof(*timing));
    if (n_query == 0) return RKR_OK;
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    if (!c->dedup || n_query < 2)
        return query_split(c, snap, queries, n_query, topk, max_k, out_user, out_rank, out_qip, stats, timing);

    // Identical query vectors (byte-equal) are answered once, with the largest
    // k asked for them: under exact output the answer for k is the first k
    // entries of the answer for any larger k, since (rank, userID) is a total
    // order.  Reference output lists the users its bounds settle first, and that
    // set depends on k (PruneCandidateByBound.hpp:35-77 keeps k-entry heaps), so
    // there only queries equal in vector and k share a run.
    const size_t row = sizeof(float) * (size_t)c->dim;
    const bool by_k = c->output == RKR_OUTPUT_REFERENCE;
    // open addressing over a hash of the row's bytes (and k), byte compare on
    // a hash match: no per-query key copies
    size_t cap = 16;
    while (cap < 2 * (size_t)n_query) cap <<= 1;
    std::vector<int32_t> slot(cap, -1);  // index into reps
    std::vector<int32_t> rep_of((size_t)n_query), reps;
    std::vector<uint64_t> rep_hash;
    for (int32_t i = 0; i < n_query; ++i) {
        const float *qi = queries + (size_t)i * c->dim;
        const uint64_t h = row_hash(qi, row, by_k ? (uint32_t)topk[i] : 0u);
        size_t pos = (size_t)h & (cap - 1);
        int32_t found = -1;
        for (; slot[pos] >= 0; pos = (pos + 1) & (cap - 1)) {
            const int32_t r = slot[pos];
            if (rep_hash[r] == h && std::memcmp(queries + (size_t)reps[r] * c->dim, qi, row) == 0 &&
                (!by_k || topk[reps[r]] == topk[i])) {
                found = r;
                break;
            }
        }
        if (found < 0) {
            found = (int32_t)reps.size();
            slot[pos] = found;
            reps.push_back(i);
            rep_hash.push_back(h);
        }
        rep_of[i] = found;
    }
    const int32_t n_rep = (int32_t)reps.size();
    if (n_rep == n_query)
        return query_split(c, snap, queries, n_query, topk, max_k, out_user, out_rank, out_qip, stats, timing);
    std::vector<float> rq((size_t)n_rep * c->dim);
    std::vector<int32_t> rk((size_t)n_rep, 0);
    for (int32_t j = 0; j < n_rep; ++j) std::memcpy(&rq[(size_t)j * c->dim], queries + (size_t)reps[j] * c->dim, row);
    for (int32_t i = 0; i < n_query; ++i) rk[rep_of[i]] = std::max(rk[rep_of[i]], topk[i]);
    std::vector<int32_t> ru((size_t)n_rep * max_k), rr((size_t)n_rep * max_k);
    std::vector<float> rqip((size_t)n_rep * max_k);
    std::vector<rkr_query_stats> rst(stats ? (size_t)n_rep : 0);
    rc = query_split(c, snap, rq.data(), n_rep, rk.data(), max_k, ru.data(), rr.data(), rqip.data(),
                     stats ? rst.data() : nullptr, timing);
    if (rc) return rc;
    for (int32_t i = 0; i < n_query; ++i) {
        const size_t src = (size_t)rep_of[i] * max_k, dst = (size_t)i * max_k;
        for (int32_t j = 0; j < max_k; ++j) {
            const bool keep = j < topk[i];
            out_user[dst + j] = keep ? ru[src + j] : -1;
            out_rank[dst + j] = keep ? rr[src + j] : -1;
            if (out_qip) out_qip[dst + j] = keep ? rqip[src + j] : -1.0f;
        }
        if (stats) {
            stats[i] = rst[rep_of[i]];
            stats[i].query_id = i;
        }
    }
    return RKR_OK;
}

}  // extern "C"

namespace rkr {

int validate_queries(Ctx *c, const DataSnap &d, const float *queries, int32_t n_query, const int32_t *topk,
                     int32_t max_k) {
    if (!d.table) return set_err(c, RKR_ERR_STATE, "query: no table (build or load one first)");
    if (n_query < 0 || (n_query > 0 && (!queries || !topk)))
        return set_err(c, RKR_ERR_INVALID, "query: null buffer");
    if (max_k <= 0 || max_k > kMaxTopk) return set_err(c, RKR_ERR_INVALID, "query: max_k must be in [1, 4096]");
    for (int32_t i = 0; i < n_query; ++i)
        if (topk[i] <= 0 || topk[i] > max_k || (int64_t)topk[i] > d.n_user)
            return set_err(c, RKR_ERR_INVALID, "query: topk[" + std::to_string(i) + "] = " + std::to_string(topk[i]) +
                                                   " outside [1, min(n_user, max_k)]");
    if (n_query > 0) {
        const double mq = max_abs(queries, (size_t)n_query * c->dim);
        if (mq < 0) return set_err(c, RKR_ERR_INVALID, "query: query vectors must be finite");
        if ((double)c->dim * mq * d.max_abs_user > kScoreLimit)
            return set_err(c, RKR_ERR_INVALID, "query: magnitudes would overflow FP32 scores");
    }
    if (c->output == RKR_OUTPUT_REFERENCE && (d.n_ins > 0 || d.n_del > 0))
        return set_err(c, RKR_ERR_STATE, "query: reference output has no update path; rebuild the table or use exact output");
    return RKR_OK;
}

namespace {

// a coarse list longer than max(kCoarseRefineMin, U / 64) gets a second pass:
// the pass costs about a user row per query, an exact-scored entry about a
// user row per entry, and the pass narrows a list several times over
constexpr int64_t kCoarseRefineMin = 4096;
// users per chunk of the coarse histograms and of the list's chunk-major
// order: at 1M x 50k (1000 queries) 4096, 8192, 16384 and 65536 ran at 258k,
// 280k, 285k and 279k queries/s (small chunks cost histogram workgroups,
// large ones the exact pass's reuse of user rows within a chunk)
constexpr int64_t kCoarseChunk = 16384;
constexpr int32_t kCoarseMinBatch = 8;

// Device batches of query_on_slot.  allow_coarse: screened-form exact batches
// take the coarse bounds; queries whose coarse list would be empty of a k-th
// bound or too long are appended to *fallback (indices into queries) with
// their outputs left unwritten.
int run_batches(Ctx *c, Slot &sl, const DataSnap &dv, const float *queries, int32_t n_query, const int32_t *topk,
                int32_t max_k, int32_t *out_user, int32_t *out_rank, float *out_qip, rkr_query_stats *stats,
                rkr_batch_timing &tm, bool allow_coarse, std::vector<int32_t> *fallback) {
    const int64_t U = dv.n_user, I = dv.n_item;
    const int32_t ld = c->ld, d = c->dim;
    const int mode = c->output;
    // exact output searches the query table when there is one (exact under
    // the score rule, eps 0); reference output keeps the sampled table
    const bool use_q = mode == RKR_OUTPUT_EXACT && dv.qtable;
    const float *T = use_q ? dv.qtable : dv.table;
    const int32_t *R = use_q ? dv.qranks : dv.ranks;
    const int32_t S = use_q ? dv.n_qs : dv.n_sample;
    const float eps = mode == RKR_OUTPUT_EXACT && !use_q ? dv.table_eps : 0.0f;
    // pending item inserts/deletes: bounds become per-user values (k_update.hip)
    const bool upd = dv.n_ins > 0 || dv.n_del > 0;
    // the QSRP screen (RKR_OPT_QSRP, built by rkr_qsrp_build) replaces the bound
    // stage of exact batches without pending updates
    const bool qsrp = c->qsrp && dv.qsrp_users && mode == RKR_OUTPUT_EXACT && !upd && U > 0;
    const bool coarse = !qsrp && allow_coarse && c->coarse && c->form == RKR_FORM_SCREENED &&
                        mode == RKR_OUTPUT_EXACT && c->kp > 0 && dv.user_norm && dv.n_guard > 0 && U > 0;
    const size_t pair_bytes = upd ? 12 : 8;  // <u,q> + buckets, or + two bound values

    // queries per device batch
    int64_t nqb = c->query_batch > 0 ? c->query_batch : 1024;
    {
        size_t avail = c->budget > c->held ? c->budget - c->held : 0;
        const int64_t by_mem = (int64_t)(avail / 4 / ((size_t)U * pair_bytes + 1));
        nqb = std::max<int64_t>(1, std::min<int64_t>(nqb, by_mem));
        nqb = std::min<int64_t>(nqb, n_query);
        nqb = std::min<int64_t>(nqb, 65535);
    }
    int rc;
    if ((rc = ensure(c, sl.q_in, sizeof(float) * (size_t)nqb * ld, "query vectors"))) return rc;
    if ((rc = ensure(c, sl.qip, sizeof(float) * (size_t)nqb * U, "query scores"))) return rc;
    // coarse bucket rows: whole row blocks of 128 users, plus a spare row
    const int64_t ldc = (U + 127) / 128 * 128;
    if ((rc = ensure(c, sl.bk, std::max((pair_bytes - 4) * (size_t)U * nqb, (size_t)ldc * (nqb + 1)),
                     "bucket pairs")))
        return rc;
    if ((rc = ensure(c, sl.qinfo, sizeof(QueryInfo) * (size_t)nqb + 64, "query info"))) return rc;
    if ((rc = ensure(c, sl.out, (sizeof(int32_t) * 2 + sizeof(float)) * (size_t)nqb * max_k, "outputs"))) return rc;
    if ((rc = ensure(c, sl.hist, sizeof(uint32_t) * hist_words((int32_t)nqb, S), "bound histograms"))) return rc;
    if (qsrp && (rc = ensure(c, sl.q_split, sizeof(float) * (size_t)nqb * (dv.qsrp_ldr + 3), "QSRP query rows")))
        return rc;
    GuardCols gc{};
    if (coarse) {
        if ((rc = ensure(c, sl.q_split, (sizeof(uint16_t) * 2 * c->kp + 2 * sizeof(float)) * (size_t)nqb, "query BF16 rows")))
            return rc;
        for (int j = 0; j < dv.n_guard; ++j) gc.col[j] = dv.gcol[j];
        gc.n = dv.n_guard;
    }
    // ranks of the searched table's columns (guard placement; the coarse
    // bounds and the update path run in exact output only, which searches the
    // query table when there is one)
    const std::vector<int32_t> &hrk = *dv.hrk;
    // coarse lists longer than refine_min get a second, finer pass; a list
    // longer than coarse_cap takes the full path instead
    const int64_t coarse_cap = std::max<int64_t>(4096, U / 4);
    const int64_t refine_min = std::max<int64_t>(kCoarseRefineMin, U / 64);

    // pinned staging: queries | query info | outputs (the device layout of
    // sl.out) | coarse rows, bin counts, guard sets, maps | recheck counts,
    // scalars; sections 256-byte aligned
    const size_t out_words = 3 * (size_t)nqb * max_k;
    size_t sz[10] = {sizeof(float) * (size_t)nqb * ld, sizeof(QueryInfo) * (size_t)nqb, 4 * out_words,
                     sizeof(int32_t) * 4 * (size_t)nqb, sizeof(int32_t) * 8 * (size_t)nqb, sizeof(GuardCols) * 8,
                     sizeof(int32_t) * 3 * (size_t)nqb, sizeof(int32_t) * (size_t)nqb, 64, 0};
    size_t off[10], tot = 0;
    for (int i = 0; i < 10; ++i) {
        off[i] = tot;
        tot += (sz[i] + 255) & ~(size_t)255;
    }
    if ((rc = ensure_pinned(c, sl.stage, tot))) return rc;
    char *stg = reinterpret_cast<char *>(sl.stage.p);
    float *hq = reinterpret_cast<float *>(stg + off[0]);
    QueryInfo *hqi = reinterpret_cast<QueryInfo *>(stg + off[1]);
    int32_t *ho_user = reinterpret_cast<int32_t *>(stg + off[2]);
    int32_t *ho_rank = ho_user + (size_t)nqb * max_k;
    float *ho_qip = reinterpret_cast<float *>(ho_rank + (size_t)nqb * max_k);
    int32_t *hcq = reinterpret_cast<int32_t *>(stg + off[3]);
    int32_t *hhc = reinterpret_cast<int32_t *>(stg + off[4]);
    GuardCols *hsets = reinterpret_cast<GuardCols *>(stg + off[5]);
    int32_t *hgset = reinterpret_cast<int32_t *>(stg + off[6]);  // then qmap, qset
    int32_t *h_qrecheck = reinterpret_cast<int32_t *>(stg + off[7]);
    int32_t *h_scalars = reinterpret_cast<int32_t *>(stg + off[8]);  // refine total, recheck count, error bits

    for (int64_t q0 = 0; q0 < n_query; q0 += nqb) {
        const int32_t nq = (int32_t)std::min<int64_t>(nqb, n_query - q0);
        float *q_in = reinterpret_cast<float *>(sl.q_in.p);
        float *qip = reinterpret_cast<float *>(sl.qip.p);
        uint32_t *bk = reinterpret_cast<uint32_t *>(sl.bk.p);
        QueryInfo *qi = reinterpret_cast<QueryInfo *>(sl.qinfo.p);
        int32_t *refine_total = reinterpret_cast<int32_t *>(reinterpret_cast<char *>(sl.qinfo.p) + sizeof(QueryInfo) * nqb);
        int32_t *o_user = reinterpret_cast<int32_t *>(sl.out.p);
        int32_t *o_rank = o_user + (size_t)nqb * max_k;
        float *o_qip = reinterpret_cast<float *>(o_rank + (size_t)nqb * max_k);

        float q_amax = 0.0f;  // largest |query value| of the batch (validated finite)
        for (int32_t i = 0; i < nq; ++i) {
            const float *src = queries + (size_t)(q0 + i) * d;
            float *dst = &hq[(size_t)i * ld];
            for (int32_t k = 0; k < d; ++k) {
                dst[k] = src[k];
                q_amax = std::max(q_amax, std::fabs(src[k]));
            }
            for (int32_t k = d; k < ld; ++k) dst[k] = 0.0f;
            hqi[i] = QueryInfo{};
            hqi[i].topk = topk[q0 + i];
        }
        RKR_HIP(c, hipEventRecord(sl.ev[0], sl.stream));
        RKR_HIP(c, hipMemcpyAsync(q_in, hq, sizeof(float) * (size_t)nq * ld, hipMemcpyHostToDevice, sl.stream));
        RKR_HIP(c, hipMemcpyAsync(qi, hqi, sizeof(QueryInfo) * nq, hipMemcpyHostToDevice, sl.stream));
        RKR_HIP(c, hipEventRecord(sl.ev[1], sl.stream));
        int32_t *lbv = reinterpret_cast<int32_t *>(bk);
        int32_t *ubv = lbv + (size_t)nq * U;
        const int64_t cu = chunk_users(U, nq, c->n_cu);
        uint32_t *ghist = reinterpret_cast<uint32_t *>(sl.hist.p);
        // a few queries are answered faster by the full path: the coarse
        // kernel converts every user row once per launch
        const bool coarse_b = coarse && nq >= kCoarseMinBatch;
        int32_t *cl_u = nullptr, *cl_b = nullptr;
        uint32_t *bk_c = nullptr;
        float *s_c = nullptr;
        int64_t n_listed = 0;             // the list length, or a bound on it when dev_lists
        bool dev_lists = false;           // the lists' true length and fallbacks are on the device
        int32_t *d_list_n = nullptr;      // the list length (device), when dev_lists
        int32_t *cq_dev = nullptr;        // the coarse per-query records (device)
        std::vector<int32_t> deep_check;  // update path: queries listed without the deepest bin
        if (coarse_b) {
            // coarse buckets on the matrix cores, the coarse k-th bound and
            // list sizes; exact scores and bucket pairs of the listed users
            uint16_t *q_hi = reinterpret_cast<uint16_t *>(sl.q_split.p);
            uint16_t *q_lo = q_hi + (size_t)nqb * c->kp;
            float *qn = reinterpret_cast<float *>(q_lo + (size_t)nqb * c->kp);
            float *qcs = qn + nqb;  // the first pass's error scales
            RKR_HIP(c, launch_prep_items(sl.stream, q_in, nq, ld, d, c->kp, q_hi, q_lo, qn, qcs, dv.rho, c->coarse_g1,
                                         c->coarse_grule));
            const double mq = q_amax;
            const double absf = ((double)screen_abs_floor(c->kp, dv.max_abs_user, mq) + (double)eps) *
                                (1.0 + std::ldexp(1.0, -18));
            uint8_t *cb = reinterpret_cast<uint8_t *>(bk);
            CoarseLaunch L{};
            L.users = dv.users; L.user_norm = dv.user_norm; L.table = T; L.gvals = dv.gvals; L.n_user = U; L.S = S;
            L.gc = gc;
            L.user_hi = dv.user_hi; L.user_lo = dv.user_lo;
            // first pass: e = C ||u|| cs + floor, C = 1 + 2^-20 covering the
            // FP32 evaluation of that product and sum
            L.q_hi = q_hi; L.q_lo = q_lo; L.q_norm = qcs; L.qmap = nullptr; L.nq = nq; L.ld = ld; L.K = d;
            L.Kp = c->kp; L.C = 1.0f + 0x1p-20f; L.abs_floor = std::nextafter((float)absf, INFINITY); L.cb = cb;
            L.ldc = ldc; L.spare_row = nqb;
            L.grid_cap = (int64_t)c->n_cu * coarse_first_pass_wgs(c->kp / 32);
            RKR_HIP(c, launch_ip_coarse(sl.stream, L));
            RKR_HIP(c, hipEventRecord(sl.ev[2], sl.stream));
            // coarse chunks: a fixed user count, so launches over few queries
            // still fill the device and the list stays chunk-major
            const int64_t ccu = kCoarseChunk;
            const int64_t n_chunk = (U + ccu - 1) / ccu;
            // part | cq | loc | ctot | cbase | guard sets | gset | qmap | qset
            const size_t w_sets = sizeof(GuardCols) * 8 / 4;
            if ((rc = ensure(c, sl.coarse_info, sizeof(uint32_t) * (17 * (size_t)nq * n_chunk + 15 * (size_t)nqb +
                                                                   2 * (size_t)n_chunk + w_sets + 64),
                             "coarse counts")))
                return rc;
            uint32_t *part = reinterpret_cast<uint32_t *>(sl.coarse_info.p);
            int32_t *cq = reinterpret_cast<int32_t *>(part + 16 * (size_t)nq * n_chunk);
            cq_dev = cq;
            int32_t *loc = cq + 4 * (size_t)nqb;
            int32_t *ctot = loc + (size_t)nq * n_chunk;
            int32_t *cbase = ctot + n_chunk;
            GuardCols *gsets = reinterpret_cast<GuardCols *>(cbase + n_chunk);
            int32_t *gset = reinterpret_cast<int32_t *>(gsets + 8);
            int32_t *qmap = gset + nqb;  // then qset (nqb)
            int32_t *hcnt = qmap + 2 * (size_t)nqb;
            RKR_HIP(c, launch_coarse_kth(sl.stream, cb, ldc, U, nq, ccu, dv.n_guard, nullptr, part, qi, cq, hcnt));
            RKR_HIP(c, hipMemcpyAsync(hcq, cq, sizeof(int32_t) * 4 * nq, hipMemcpyDeviceToHost, sl.stream));
            if (upd)
                RKR_HIP(c, hipMemcpyAsync(hhc, hcnt, sizeof(int32_t) * 8 * nq, hipMemcpyDeviceToHost, sl.stream));
            RKR_HIP(c, hipStreamSynchronize(sl.stream));
            if (upd) {
                // pending inserts move ranks: a user at or above guard m has
                // rank <= R[j_m] + ins_gmax[m], so the k-th bound is at most
                // V = R[j_beta] + ins_gmax[beta]; a user below guard c_hi - 1
                // has rank >= R[j_{c_hi - 1}] + 1 - n_del, and is listed when
                // that can be <= V
                for (int32_t i = 0; i < nq; ++i) {
                    int32_t *e = &hcq[(size_t)i * 4];
                    if (e[0] < 0) continue;
                    const int64_t V = (int64_t)hrk[gc.col[e[0]]] + dv.ins_gmax[e[0]];
                    int32_t bl = 0;
                    while (bl < gc.n && (int64_t)hrk[gc.col[bl]] + 1 - dv.n_del <= V) ++bl;
                    if (bl >= gc.n) {
                        // users below the deepest guard are left out; the
                        // exact k-th value then has to stay below their
                        // lowest possible rank (checked after the selection)
                        bl = gc.n - 1;
                        deep_check.push_back(i);
                    }
                    int64_t n = 0;
                    for (int32_t cc = 0; cc <= bl; ++cc) n += hhc[(size_t)i * 8 + cc];
                    e[0] = bl;
                    e[1] = (int32_t)std::min<int64_t>(n, INT32_MAX);
                }
            }
            // second pass for long lists: the queries whose k-th bound fell in
            // guard bin beta are bucketed again against up to 7 columns inside
            // that bin (the bin's last column included), spaced geometrically
            // in rank
            std::vector<GuardCols> sets{gc};
            std::vector<int32_t> hqmap;
            for (int32_t i = 0; i < nq; ++i) hgset[i] = 0;
            std::vector<std::pair<int32_t, int32_t>> groups;  // (set, first index in hqmap)
            // (not with pending updates: the rank shift bounds are kept for
            // the first pass's guards only)
            std::vector<int32_t> set_ids;  // guard-row set of each second-pass set
            for (int32_t beta = 0; beta < gc.n && !upd; ++beta) {
                const GuardCols &g2 = dv.gsets.set[1 + beta];
                if (g2.n == 0) continue;
                const size_t first = hqmap.size();
                for (int32_t i = 0; i < nq; ++i)
                    if (hcq[(size_t)i * 4] == beta && hcq[(size_t)i * 4 + 1] > refine_min) hqmap.push_back(i);
                if (hqmap.size() == first) continue;
                for (size_t t = first; t < hqmap.size(); ++t) hgset[hqmap[t]] = (int32_t)sets.size();
                groups.push_back({(int32_t)sets.size(), (int32_t)first});
                sets.push_back(g2);
                set_ids.push_back(1 + beta);
            }
            // With a second pass the lists stay on the device: the batch goes
            // on without reading the new counts.  A second-pass list is a
            // subset of the query's first-pass list (its guards lie inside the
            // bin that held the k-th bound), so the first-pass lengths bound
            // the list; k_coarse_mark sends a query whose list exceeds its
            // bound or the cap to the full path, and the final counts and
            // those queries are read with the outputs.
            dev_lists = !groups.empty();
            if (dev_lists) {
                const int32_t nr = (int32_t)hqmap.size();
                for (int32_t i = 0; i < nq; ++i) {
                    int32_t *e = &hcq[(size_t)i * 4];
                    const int32_t b1 = e[0] >= 0 ? (int32_t)std::min<int64_t>(e[1], coarse_cap) : 0;
                    e[2] = b1;  // field 2 carries the bound to k_coarse_mark
                    n_listed += b1;
                }
                // one launch for every refined query: set of each (local)
                // query in qset; every set is uploaded once, set 0 (the first
                // pass's guards) first, and the second pass reads from set 1
                int32_t *h_map = hgset + nqb, *h_set = h_map + nr;
                for (int32_t t = 0; t < nr; ++t) {
                    h_map[t] = hqmap[t];
                    h_set[t] = hgset[hqmap[t]] - 1;
                }
                for (size_t t = 0; t < sets.size(); ++t) hsets[t] = sets[t];
                int32_t *qset = qmap + nr;
                RKR_HIP(c, hipMemcpyAsync(cq, hcq, sizeof(int32_t) * 4 * nq, hipMemcpyHostToDevice, sl.stream));
                RKR_HIP(c, hipMemcpyAsync(qmap, h_map, sizeof(int32_t) * 2 * nr, hipMemcpyHostToDevice, sl.stream));
                RKR_HIP(c, hipMemcpyAsync(gsets, hsets, sizeof(GuardCols) * sets.size(), hipMemcpyHostToDevice,
                                          sl.stream));
                CoarseLaunch L2 = L;
                L2.C = c->screen_C;  // the second pass's three-product estimate
                L2.q_norm = qn;
                L2.grid_cap = (int64_t)c->n_cu * 3;
                L2.qmap = qmap;
                L2.nq = nr;
                L2.sets = gsets + 1;
                L2.qset = qset;
                L2.n_sets = (int32_t)sets.size() - 1;
                for (int32_t t = 0; t < L2.n_sets; ++t) L2.set_id[t] = set_ids[t];
                RKR_HIP(c, launch_ip_coarse(sl.stream, L2));
                for (size_t gi = 0; gi < groups.size(); ++gi) {
                    const int32_t first = groups[gi].second;
                    const int32_t end = gi + 1 < groups.size() ? groups[gi + 1].second : nr;
                    RKR_HIP(c, launch_coarse_kth(sl.stream, cb, ldc, U, end - first, ccu, sets[groups[gi].first].n,
                                                 qmap + first, part, qi, cq, hcnt));
                }
                RKR_HIP(c, launch_coarse_mark(sl.stream, cq, nq, coarse_cap));
                tm.n_coarse_refined += (int64_t)hqmap.size();
            } else {
                for (int32_t i = 0; i < nq; ++i) {
                    int32_t *e = &hcq[(size_t)i * 4];
                    if (e[0] < 0 || e[1] > coarse_cap) {
                        if (fallback) fallback->push_back((int32_t)(q0 + i));
                        tm.n_fallback += 1;
                        e[0] = -1;
                        e[1] = 0;
                    }
                    e[2] = (int32_t)n_listed;
                    e[3] = 0;
                    n_listed += e[1];
                }
                tm.n_listed += n_listed;
            }
            if (n_listed >= INT32_MAX) return set_err(c, RKR_ERR_NOMEM, "query: coarse list exceeds 2^31; lower the query batch");
            if ((rc = ensure(c, sl.coarse_list, 16 * (size_t)std::max<int64_t>(n_listed, 1), "coarse list"))) return rc;
            cl_u = reinterpret_cast<int32_t *>(sl.coarse_list.p);
            cl_b = cl_u + std::max<int64_t>(n_listed, 1);
            bk_c = reinterpret_cast<uint32_t *>(cl_b + std::max<int64_t>(n_listed, 1));
            s_c = reinterpret_cast<float *>(bk_c + std::max<int64_t>(n_listed, 1));
            if (!dev_lists) {
                for (size_t t = 0; t < sets.size(); ++t) hsets[t] = sets[t];
                RKR_HIP(c, hipMemcpyAsync(cq, hcq, sizeof(int32_t) * 4 * nq, hipMemcpyHostToDevice, sl.stream));
                RKR_HIP(c, hipMemcpyAsync(gsets, hsets, sizeof(GuardCols) * sets.size(), hipMemcpyHostToDevice,
                                          sl.stream));
            }
            RKR_HIP(c, hipMemcpyAsync(gset, hgset, sizeof(int32_t) * nq, hipMemcpyHostToDevice, sl.stream));
            d_list_n = dev_lists ? hcnt + 8 * (size_t)nqb : nullptr;  // in the spare words of coarse_info
            RKR_HIP(c, launch_coarse_compact(sl.stream, cb, ldc, U, nq, ccu, part, cq, loc, ctot, cbase, cl_u, cl_b,
                                             d_list_n));
            if (upd) {
                // bound values of the listed pairs only (the rest are ruled
                // out), then the value selection of the update path; the
                // bound arrays reuse the coarse rows, read by now
                RKR_HIP(c, hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(lbv), kNoBound, 2 * (size_t)nq * U,
                                             sl.stream));
                RKR_HIP(c, launch_coarse_exact_upd(sl.stream, dv.users, q_in, ld, d, T, S, R, dv.n_item_base, eps, gsets,
                                                   gset, cl_u, cl_b, n_listed, dv.side_ins, dv.n_ins, dv.ld_ins,
                                                   dv.side_del, dv.n_del, dv.ld_del, U, lbv, ubv, qip));
                RKR_HIP(c, launch_kth_values(sl.stream, lbv, ubv, U, nq, cu, ghist, qi));
            } else {
                RKR_HIP(c, launch_coarse_exact(sl.stream, dv.users, q_in, ld, d, T, S, R, I, eps, gsets, gset, nq, cl_u,
                                               cl_b, n_listed, d_list_n, bk_c, s_c, ghist, qi));
            }
            RKR_HIP(c, hipEventRecord(sl.ev[3], sl.stream));
        } else if (qsrp) {
            // QSRP screen: rotated queries and their part norms, the left-part
            // scores over the first check_dim dimensions, bounds from the
            // partial-dimension interval and the regression, the k-th bound
            // over those, then the exact score and table search for the users
            // they keep, and the k-th bound again (the exact path's from here)
            float *qrot = reinterpret_cast<float *>(sl.q_split.p);
            float *qn = qrot + (size_t)nqb * dv.qsrp_ldr;
            RKR_HIP(c, launch_scores_store(sl.stream, q_in, nullptr, nq, ld, dv.qsrp_rot, d, ld, d, qrot,
                                           dv.qsrp_ldr, true));
            RKR_HIP(c, launch_qsrp_qnorms(sl.stream, qrot, dv.qsrp_ldr, q_in, ld, d, dv.qsrp_cd, nq, qn));
            RKR_HIP(c, launch_scores_store(sl.stream, dv.qsrp_users, nullptr, U, dv.qsrp_ldr, qrot, nq, dv.qsrp_ldr,
                                           dv.qsrp_cd, qip, U, false));
            RKR_HIP(c, hipEventRecord(sl.ev[2], sl.stream));
            // eps widens the interval as the exact path widens its search; the
            // floor covers gradual underflow in the d products of a score
            const float absf = std::nextafter((float)((double)eps + (double)d * std::ldexp(1.0, -140)), INFINITY);
            RKR_HIP(c, launch_qsrp_bounds(sl.stream, qip, U, nq, dv.qsrp_par, qn, dv.qsrp_gcd, dv.qsrp_crot, absf,
                                          dv.n_sample, use_q ? dv.qcol : nullptr, S, bk));
            RKR_HIP(c, launch_hist_kth(sl.stream, bk, U, S, R, I, nq, mode, 0, cu, ghist, qi));
            RKR_HIP(c, launch_qsrp_exact(sl.stream, dv.users, q_in, ld, d, T, S, eps, U, nq, qi, qip, bk));
            RKR_HIP(c, hipEventRecord(sl.ev[3], sl.stream));
            RKR_HIP(c, launch_hist_kth(sl.stream, bk, U, S, R, I, nq, mode, 0, cu, ghist, qi));
        } else {
            // <u, q> for every (query, user): qip[b * U + u]
            if (nq <= 16 && (size_t)nq * d * sizeof(float) <= 65536)
                RKR_HIP(c, launch_gemv_panels(sl.stream, dv.usersT, U, ld, d, q_in, nq, qip));
            else RKR_HIP(c, launch_scores_store(sl.stream, dv.users, nullptr, U, ld, q_in, nq, ld, d, qip, U, false));
            RKR_HIP(c, hipEventRecord(sl.ev[2], sl.stream));
            if (upd)
                RKR_HIP(c, launch_bounds_upd(sl.stream, qip, T, U, S, R, dv.n_item_base, eps, dv.side_ins,
                                             dv.n_ins, dv.ld_ins, dv.side_del, dv.n_del, dv.ld_del, nq, lbv, ubv));
            else
                RKR_HIP(c, launch_bounds(sl.stream, qip, T, U, S, nq, eps, mode == RKR_OUTPUT_REFERENCE ? 1 : 0, bk));
            RKR_HIP(c, hipEventRecord(sl.ev[3], sl.stream));
            if (upd) RKR_HIP(c, launch_kth_values(sl.stream, lbv, ubv, U, nq, cu, ghist, qi));
            else
                RKR_HIP(c, launch_hist_kth(sl.stream, bk, U, S, R, I, nq, mode,
                                           (mode == RKR_OUTPUT_EXACT && eps == 0.0f) ? 1 : 0, cu, ghist, qi));
        }
        RKR_HIP(c, launch_offsets(sl.stream, qi, nq, refine_total));
        // Coarse bounds without pending updates: every candidate is a listed
        // pair, so the list length bounds the candidate and refine arrays and
        // the batch goes on without waiting for the counts (the refinement and
        // top-k read them on the device).  Otherwise the counts size them.
        const bool sized_by_list = coarse_b && !upd;
        int64_t total = n_listed, total_cand = n_listed;
        if (!sized_by_list) {
            RKR_HIP(c, hipMemcpyAsync(hqi, qi, sizeof(QueryInfo) * nq, hipMemcpyDeviceToHost, sl.stream));
            RKR_HIP(c, hipStreamSynchronize(sl.stream));
        }
        if (!deep_check.empty()) {
            // a user below the deepest guard has rank >= r[j_last] + 1 - n_del:
            // the selection over the listed users is exact when the k-th value
            // L (QueryInfo.gamma of the value selection) is below that
            const int64_t floor_rank = (int64_t)hrk[gc.col[gc.n - 1]] + 1 - dv.n_del;
            for (int32_t i : deep_check)
                if ((int64_t)hqi[i].gamma >= floor_rank) {
                    if (fallback) fallback->push_back((int32_t)(q0 + i));
                    tm.n_fallback += 1;
                }
        }
        if (!sized_by_list) {
            total = hqi[nq - 1].cand_off + hqi[nq - 1].n_settled + hqi[nq - 1].n_cand;
            total_cand = 0;
            for (int32_t i = 0; i < nq; ++i) total_cand += hqi[i].n_cand;
        }
        if (total >= INT32_MAX) return set_err(c, RKR_ERR_NOMEM, "query: candidate count exceeds 2^31; lower the query batch");
        const size_t cap = (size_t)std::max<int64_t>(total, 1);
        const size_t rcap = (size_t)std::max<int64_t>(total_cand, 1);
        if ((rc = ensure(c, sl.cand, sizeof(int32_t) * 2 * cap, "candidates"))) return rc;
        if ((rc = ensure(c, sl.refine_idx, (sizeof(int32_t) * 3 + sizeof(float)) * rcap, "refine rows"))) return rc;
        if ((rc = ensure(c, sl.counts, sizeof(int32_t) * rcap, "refine counts"))) return rc;
        if ((rc = ensure(c, sl.small, sizeof(int32_t) * ((size_t)nqb + 8), "query counters"))) return rc;
        int32_t *cand_user = reinterpret_cast<int32_t *>(sl.cand.p);
        int32_t *cand_rank = cand_user + cap;
        int32_t *refine_cand = reinterpret_cast<int32_t *>(sl.refine_idx.p);
        int32_t *refine_user = refine_cand + rcap;
        int32_t *refine_query = refine_user + rcap;
        float *refine_thr = reinterpret_cast<float *>(refine_query + rcap);
        int32_t *cnt = reinterpret_cast<int32_t *>(sl.counts.p);
        int32_t *q_recheck = reinterpret_cast<int32_t *>(sl.small.p);  // per query
        int32_t *recheck_n = q_recheck + nqb;
        uint32_t *max_err_bits = reinterpret_cast<uint32_t *>(recheck_n + 1);
        const bool screened = c->form == RKR_FORM_SCREENED;

        if (coarse_b && !upd)
            RKR_HIP(c, launch_coarse_final(sl.stream, cl_u, cl_b, bk_c, s_c, n_listed, d_list_n, S, R, I, U, qi, cand_user,
                                           cand_rank, refine_total, refine_cand, refine_user, refine_query, refine_thr,
                                           qip));
        else if (upd)
            RKR_HIP(c, launch_compact_values(sl.stream, lbv, ubv, qip, U, nq, cu, qi, cand_user, cand_rank,
                                             refine_total, refine_cand, refine_user, refine_query, refine_thr));
        else
            RKR_HIP(c, launch_compact(sl.stream, bk, qip, U, S, R, I, nq, mode, cu, qi, cand_user, cand_rank,
                                      refine_total, refine_cand, refine_user, refine_query, refine_thr));
        int32_t &h_refine_total = h_scalars[0], &h_recheck_n = h_scalars[1];
        uint32_t &h_err_bits = reinterpret_cast<uint32_t *>(h_scalars)[2];
        h_refine_total = h_recheck_n = 0;
        h_err_bits = 0;
        for (int32_t i = 0; i < nq; ++i) h_qrecheck[i] = 0;
        // refinement + top-k; re-run when the recheck list overflowed
        for (int attempt = 0; attempt < 3; ++attempt) {
            int64_t rec_cap = 0;
            if (screened) {
                const int64_t want = std::max<int64_t>(1 << 20, (int64_t)sl.recheck.bytes / 16);
                if ((rc = ensure(c, sl.recheck, 16 * (size_t)want, "recheck list"))) return rc;
                rec_cap = (int64_t)(sl.recheck.bytes / 16);
                RKR_HIP(c, hipMemsetAsync(sl.small.p, 0, sizeof(int32_t) * ((size_t)nqb + 8), sl.stream));
            }
            RKR_HIP(c, hipMemsetAsync(cnt, 0, sizeof(int32_t) * rcap, sl.stream));
            RKR_HIP(c, hipEventRecord(sl.ev[4], sl.stream));
            // events 4..8 bracket exactly the refinement (or screening) kernel
            if (screened) {
                ScreenLaunch L{};
                L.users = dv.users; L.user_norm = dv.user_norm;
                L.user_hi = dv.user_hi; L.user_lo = dv.user_lo;
                L.item_hi = dv.item_hi; L.item_lo = dv.item_lo; L.item_norm = dv.item_norm;
                L.n_item = I; L.ld = ld; L.K = d; L.Kp = c->kp;
                L.rows_user = refine_user; L.rows_thr = refine_thr; L.n_rows_dev = refine_total;
                L.max_rows = (int64_t)rcap; L.C = c->screen_C; L.abs_floor = dv.screen_abs;
                L.cnt = cnt; L.recheck_n = recheck_n; L.recheck = sl.recheck.p; L.recheck_cap = rec_cap;
                L.target_units = (int64_t)c->n_cu * 8;
                L.grid_cap = (int64_t)c->n_cu * 4;
                L.chunk_items = c->screen_chunk > 0 ? c->screen_chunk : (1 << 30);
                L.rows_per_wave = c->screen_rows > 0 ? c->screen_rows : (c->kp / 32 >= 6 ? 64 : 32);  // d > 160: 32 rows spill
                RKR_HIP(c, launch_refine_screen(sl.stream, L));
                RKR_HIP(c, hipEventRecord(sl.ev[8], sl.stream));
                RKR_HIP(c, launch_recheck(sl.stream, dv.users, dv.items, ld, d, refine_user, refine_thr, refine_query,
                                          recheck_n, sl.recheck.p, rec_cap, cnt, q_recheck, max_err_bits));
            } else {
                RKR_HIP(c, launch_refine_count(sl.stream, dv.users, dv.items, I, ld, d, refine_user, refine_thr,
                                               refine_total, (int64_t)rcap, cnt));
                RKR_HIP(c, hipEventRecord(sl.ev[8], sl.stream));
            }
            RKR_HIP(c, hipEventRecord(sl.ev[5], sl.stream));
            RKR_HIP(c, launch_scatter_counts(sl.stream, refine_total, refine_cand, cnt, (int64_t)rcap, cand_rank));
            RKR_HIP(c, launch_topk(sl.stream, qi, nq, cand_user, cand_rank, qip, U, max_k, o_user, o_rank, o_qip,
                                   nullptr));
            RKR_HIP(c, hipEventRecord(sl.ev[6], sl.stream));
            // outputs in one copy (the device layout), then the counters
            RKR_HIP(c, hipMemcpyAsync(ho_user, o_user, 4 * out_words, hipMemcpyDeviceToHost, sl.stream));
            RKR_HIP(c, hipMemcpyAsync(hqi, qi, sizeof(QueryInfo) * nq, hipMemcpyDeviceToHost, sl.stream));
            RKR_HIP(c, hipMemcpyAsync(&h_refine_total, refine_total, sizeof(int32_t), hipMemcpyDeviceToHost, sl.stream));
            if (dev_lists) {
                // the lists' length and the queries k_coarse_mark sent to the full path
                RKR_HIP(c, hipMemcpyAsync(hcq, cq_dev, sizeof(int32_t) * 4 * nq, hipMemcpyDeviceToHost, sl.stream));
                RKR_HIP(c, hipMemcpyAsync(&h_scalars[3], d_list_n, sizeof(int32_t), hipMemcpyDeviceToHost, sl.stream));
            }
            if (screened) {
                // q_recheck, recheck_n, max_err_bits are adjacent in sl.small
                RKR_HIP(c, hipMemcpyAsync(h_qrecheck, q_recheck, sizeof(int32_t) * nq, hipMemcpyDeviceToHost, sl.stream));
                RKR_HIP(c, hipMemcpyAsync(&h_recheck_n, recheck_n, sizeof(int32_t) * 2, hipMemcpyDeviceToHost,
                                          sl.stream));
            }
            RKR_HIP(c, hipEventRecord(sl.ev[7], sl.stream));
            RKR_HIP(c, hipStreamSynchronize(sl.stream));
            if (!screened || (int64_t)h_recheck_n <= rec_cap) break;
            // the list overflowed: grow it and run the refinement again
            const size_t need = 16 * ((size_t)h_recheck_n + (size_t)h_recheck_n / 4 + 1024);
            dev_free(c, sl.recheck.p, sl.recheck.bytes);
            sl.recheck.p = nullptr;
            sl.recheck.bytes = 0;
            if ((rc = ensure(c, sl.recheck, need, "recheck list"))) return rc;
            if (attempt == 2) return set_err(c, RKR_ERR_NOMEM, "query: recheck list kept overflowing");
        }
        if (dev_lists) {
            for (int32_t i = 0; i < nq; ++i)
                if (hcq[(size_t)i * 4] < 0) {
                    if (fallback) fallback->push_back((int32_t)(q0 + i));
                    tm.n_fallback += 1;
                }
            tm.n_listed += h_scalars[3];
        }
        if (qsrp)
            for (int32_t i = 0; i < nq; ++i) tm.n_listed += hqi[i].n_screen;

        std::memcpy(out_user + (size_t)q0 * max_k, ho_user, sizeof(int32_t) * (size_t)nq * max_k);
        std::memcpy(out_rank + (size_t)q0 * max_k, ho_rank, sizeof(int32_t) * (size_t)nq * max_k);
        if (out_qip) std::memcpy(out_qip + (size_t)q0 * max_k, ho_qip, sizeof(float) * (size_t)nq * max_k);
        if (stats) {
            for (int32_t i = 0; i < nq; ++i) {
                rkr_query_stats &st = stats[q0 + i];
                std::memset(&st, 0, sizeof(st));
                st.query_id = (int32_t)(q0 + i);
                st.n_candidate = hqi[i].n_cand;
                st.n_refine_user = hqi[i].n_refine;
                if (mode == RKR_OUTPUT_EXACT) {
                    st.n_result_user = hqi[i].n_cand - hqi[i].n_refine;
                    st.n_prune_user = (int32_t)(U - hqi[i].n_cand);
                } else {
                    st.n_result_user = hqi[i].n_settled;
                    st.n_prune_user = (int32_t)(U - hqi[i].n_settled - hqi[i].n_cand);
                }
                // inner products: every user's (QSRP: the screen's survivors), then the refinement's
                st.ip_cost = (qsrp ? (int64_t)hqi[i].n_screen : U) + (int64_t)hqi[i].n_refine * I;
                if (screened) {
                    st.n_screened = (int64_t)hqi[i].n_refine * I;
                    st.n_rechecked = h_qrecheck[i];
                }
            }
        }
        tm.h2d_ms += ev_ms(sl.ev, 0, 1);
        tm.ip_ms += ev_ms(sl.ev, 1, 2);
        tm.bound_ms += ev_ms(sl.ev, 2, 3);
        tm.prune_ms += ev_ms(sl.ev, 3, 4);
        tm.refine_ms += ev_ms(sl.ev, 4, 8);
        tm.recheck_ms += ev_ms(sl.ev, 8, 5);
        tm.topk_ms += ev_ms(sl.ev, 5, 6);
        tm.d2h_ms += ev_ms(sl.ev, 6, 7);
        tm.n_batches += 1;
        tm.n_refined += h_refine_total;
        if (screened) {
            tm.n_rechecked += h_recheck_n;
            tm.n_screened += (int64_t)h_refine_total * I;
            float e;
            std::memcpy(&e, &h_err_bits, sizeof(e));
            tm.max_screen_err = std::max(tm.max_screen_err, (double)e);
        }
    }
    return RKR_OK;
}

}  // namespace

int query_on_slot(Ctx *c, Slot &sl, const DataSnap &dv, const float *queries, int32_t n_query, const int32_t *topk,
                  int32_t max_k, int32_t *out_user, int32_t *out_rank, float *out_qip, rkr_query_stats *stats,
                  rkr_batch_timing *timing) {
    using clock = std::chrono::steady_clock;
    const auto t_start = clock::now();
    rkr_batch_timing tm{};
    std::vector<int32_t> fb;
    int rc = run_batches(c, sl, dv, queries, n_query, topk, max_k, out_user, out_rank, out_qip, stats, tm, true, &fb);
    if (rc) return rc;
    if (!fb.empty()) {
        // queries the coarse bounds could not narrow enough: full path
        const int32_t n = (int32_t)fb.size(), d = c->dim;
        std::vector<float> fq((size_t)n * d);
        std::vector<int32_t> fk(n), fu((size_t)n * max_k), fr((size_t)n * max_k);
        std::vector<float> fqip((size_t)n * max_k);
        std::vector<rkr_query_stats> fs(stats ? n : 0);
        for (int32_t i = 0; i < n; ++i) {
            std::memcpy(&fq[(size_t)i * d], queries + (size_t)fb[i] * d, sizeof(float) * d);
            fk[i] = topk[fb[i]];
        }
        rc = run_batches(c, sl, dv, fq.data(), n, fk.data(), max_k, fu.data(), fr.data(), fqip.data(),
                         stats ? fs.data() : nullptr, tm, false, nullptr);
        if (rc) return rc;
        for (int32_t i = 0; i < n; ++i) {
            const size_t o = (size_t)fb[i] * max_k, f = (size_t)i * max_k;
            std::memcpy(out_user + o, &fu[f], sizeof(int32_t) * max_k);
            std::memcpy(out_rank + o, &fr[f], sizeof(int32_t) * max_k);
            if (out_qip) std::memcpy(out_qip + o, &fqip[f], sizeof(float) * max_k);
            if (stats) {
                stats[fb[i]] = fs[i];
                stats[fb[i]].query_id = fb[i];
            }
        }
    }
    tm.total_ms = std::chrono::duration<double, std::milli>(clock::now() - t_start).count();
    if (timing) *timing = tm;
    return RKR_OK;
}

}  // namespace rkr
