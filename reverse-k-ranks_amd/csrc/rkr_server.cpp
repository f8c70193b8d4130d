// rkr_server.cpp -- the serving loop: continuous batching of reverse k-ranks
// queries (mixed k) interleaved with user / item updates, on one context.
//
// Queries enter a FIFO.  Worker threads, each owning a Slot (its own HIP
// stream and scratch buffers), take the longest run of queries at the front
// of the queue (up to max_batch), so requests that arrive while a batch is on
// the device are admitted by the next free worker and the batches overlap on
// the GPU.  Batches are further split by free device memory inside
// query_on_slot.  Answers do not depend on how queries are batched (every
// kernel decides each (query, user) pair on its own), which the tests check
// by comparing answers alone and in the mix.
//
// Updates, in one of two modes:
//  - barrier (default): an update waits in the same FIFO, starts when no
//    batch is in flight, and no batch starts while it runs, so every query
//    sees exactly the updates submitted before it;
//  - snapshot: an updater thread applies updates in submission order while
//    batches keep running.  A batch reads the state the last committed update
//    left (captured when it starts); the update's frees are deferred until
//    every batch that started before its commit has finished.  A query then
//    sees every update that committed before its batch started.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rkr_internal.hpp"

using namespace rkr;

struct rkr_ctx : public rkr::Ctx {};

namespace {

enum Kind { Q_QUERY = 0, Q_INSERT_USERS = 1, Q_INSERT_ITEMS = 2, Q_DELETE_USERS = 3, Q_DELETE_ITEMS = 4 };

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id = 0;
    int kind = Q_QUERY;
    std::vector<float> data;
    std::vector<int64_t> ids;
    int64_t n = 0;
    int32_t k = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = RKR_OK;
    int32_t batch_size = 0;
    bool done = false;
    std::string err;
    std::vector<int32_t> users, ranks;
    std::vector<float> qips;
    rkr_query_stats stats{};
};

}  // namespace

struct rkr_server {
    Ctx *c = nullptr;
    rkr_server_config cfg{};
    int64_t rebuild_at = 12288;  // pending item updates that trigger a table rebuild (< 0: never)
    std::vector<std::unique_ptr<Slot>> slots;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_upd;
    std::deque<Request *> queue;   // queries (and updates in barrier mode)
    std::deque<Request *> uqueue;  // updates (snapshot mode)
    std::thread updater;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1;
    int busy = 0;            // query batches in flight
    bool exclusive = false;  // an update is running (barrier mode)
    bool stopping = false;
    int64_t n_batches = 0, n_queries = 0, n_updates = 0, n_rebuilds = 0;
    // snapshot mode: the committed state, its version, and the versions the
    // batches in flight started from (version -> batches)
    DataSnap published;
    uint64_t version = 0;
    std::map<uint64_t, int> inflight;
    // snapshot mode: the update thread's stream, limited to a subset of the
    // CUs; the context's own stream is restored at stop
    hipStream_t upd_stream = nullptr, ctx_stream = nullptr;

    bool trace = false;      // RKR_SERVER_TRACE=1: one stderr line per snapshot update

    bool snapshot_mode() const { return cfg.update_mode == RKR_SERVE_SNAPSHOT; }

    bool can_start() const {
        if (exclusive || queue.empty()) return false;
        return queue.front()->kind == Q_QUERY || busy == 0;
    }

    // no batch in flight started before version v
    bool drained_before(uint64_t v) const { return inflight.empty() || inflight.begin()->first >= v; }

    // Applies one update; when the pending inserted and deleted items then reach
    // the rebuild threshold, rebuilds the table over the current items (uniform
    // ranks, same n_sample), which folds the side arrays into a fresh base, so an
    // endless insert stream never meets the side arrays' cap.  An item update
    // that would outgrow the side arrays' reserved length rebuilds first, so
    // the arrays never grow while serving (inserts that queued up behind a long
    // rebuild then fit).  In snapshot mode this runs on the update thread and
    // its CUs: batches keep reading the previous state until the commit, and
    // the replaced buffers are retired.
    int rebuild() {
        const int rc = rebuild_table(c);
        if (!rc) {
            std::lock_guard<std::mutex> g(mu);
            ++n_rebuilds;
        }
        return rc;
    }

    int apply_update(Request *r) {
        int rc;
        if (rebuild_at >= 0 && c->table &&
            ((r->kind == Q_INSERT_ITEMS && c->n_ins > 0 && (int64_t)c->n_ins + r->n > c->ld_ins) ||
             (r->kind == Q_DELETE_ITEMS && c->n_del > 0 && (int64_t)c->n_del + r->n > c->ld_del)) &&
            (rc = rebuild()))
            return rc;
        switch (r->kind) {
            case Q_INSERT_USERS: rc = rkr_insert_users(static_cast<rkr_ctx *>(c), r->data.data(), r->n); break;
            case Q_INSERT_ITEMS: rc = rkr_insert_items(static_cast<rkr_ctx *>(c), r->data.data(), r->n); break;
            case Q_DELETE_USERS: rc = rkr_delete_users(static_cast<rkr_ctx *>(c), r->ids.data(), r->n); break;
            case Q_DELETE_ITEMS: rc = rkr_delete_items(static_cast<rkr_ctx *>(c), r->ids.data(), r->n); break;
            default: return RKR_ERR_INVALID;
        }
        if (rc || rebuild_at < 0 || (int64_t)c->n_ins + c->n_del < rebuild_at) return rc;
        return rebuild();
    }

    void run_batch(Slot &sl, const DataSnap &snap, std::vector<Request *> &all) {
        // a request that fails validation is answered alone with its error,
        // so it cannot fail the other queries of its batch
        std::vector<Request *> batch;
        for (Request *r : all) {
            const int32_t k = r->k;
            const int rc = validate_queries(c, snap, r->data.data(), 1, &k, std::max(k, 1));
            if (!rc) {
                batch.push_back(r);
                continue;
            }
            std::string err;
            {
                std::lock_guard<std::mutex> g(c->err_mu);
                err = c->err;
            }
            std::lock_guard<std::mutex> g(mu);
            r->status = rc;
            r->err = err;
            r->batch_size = 1;
            r->t_done = now_s();
            r->done = true;
        }
        if (batch.empty()) return;
        const int32_t nq = (int32_t)batch.size();
        const int32_t d = c->dim;
        int32_t max_k = 1;
        std::vector<float> q((size_t)nq * d);
        std::vector<int32_t> ks((size_t)nq);
        for (int32_t i = 0; i < nq; ++i) {
            std::memcpy(&q[(size_t)i * d], batch[i]->data.data(), sizeof(float) * d);
            ks[i] = batch[i]->k;
            max_k = std::max(max_k, ks[i]);
        }
        std::vector<int32_t> ou((size_t)nq * max_k), orank((size_t)nq * max_k);
        std::vector<float> oq((size_t)nq * max_k);
        std::vector<rkr_query_stats> st((size_t)nq);
        rkr_batch_timing tm{};
        const int rc = query_on_slot(c, sl, snap, q.data(), nq, ks.data(), max_k, ou.data(), orank.data(), oq.data(),
                                     st.data(), &tm);
        std::string err;
        if (rc) {
            std::lock_guard<std::mutex> g(c->err_mu);
            err = c->err;
        }
        const double t = now_s();
        std::lock_guard<std::mutex> g(mu);
        for (int32_t i = 0; i < nq; ++i) {
            Request *r = batch[i];
            r->status = rc;
            r->err = err;
            r->batch_size = nq;
            if (!rc) {
                r->users.assign(ou.begin() + (size_t)i * max_k, ou.begin() + (size_t)i * max_k + r->k);
                r->ranks.assign(orank.begin() + (size_t)i * max_k, orank.begin() + (size_t)i * max_k + r->k);
                r->qips.assign(oq.begin() + (size_t)i * max_k, oq.begin() + (size_t)i * max_k + r->k);
                r->stats = st[i];
            }
            r->t_done = t;
            r->done = true;
        }
        ++n_batches;
        n_queries += nq;
    }

    void worker(int w) {
        (void)hipSetDevice(c->device);
        // snapshot mode: scratch a batch outgrows is released by the update
        // thread, so a worker never waits in hipFree behind update kernels
        if (snapshot_mode()) set_defer_frees(true);
        Slot &sl = *slots[w];
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            cv_work.wait(lk, [&] { return (stopping && queue.empty()) || can_start(); });
            if (stopping && queue.empty()) break;
            Request *front = queue.front();
            if (front->kind != Q_QUERY) {  // barrier update
                queue.pop_front();
                exclusive = true;
                front->t_start = now_s();
                lk.unlock();
                const int rc = apply_update(front);
                std::string err;
                if (rc) {
                    std::lock_guard<std::mutex> g(c->err_mu);
                    err = c->err;
                }
                lk.lock();
                front->status = rc;
                front->err = err;
                front->t_done = now_s();
                front->done = true;
                exclusive = false;
                ++n_updates;
                cv_done.notify_all();
                cv_work.notify_all();
                continue;
            }
            std::vector<Request *> batch;
            const double t = now_s();
            while (!queue.empty() && queue.front()->kind == Q_QUERY && (int32_t)batch.size() < cfg.max_batch) {
                queue.front()->t_start = t;
                batch.push_back(queue.front());
                queue.pop_front();
            }
            ++busy;
            const uint64_t v = version;
            const DataSnap snap = snapshot_mode() ? published : snap_of(c);
            if (snapshot_mode()) ++inflight[v];
            lk.unlock();
            run_batch(sl, snap, batch);
            lk.lock();
            --busy;
            if (snapshot_mode() && --inflight[v] == 0) inflight.erase(v);
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }

    // Snapshot mode: applies updates in order while batches run, commits each
    // by publishing the new state, then frees what it replaced once the
    // batches that may read the old state have finished.
    void update_loop() {
        (void)hipSetDevice(c->device);
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            cv_upd.wait(lk, [&] { return stopping || !uqueue.empty(); });
            if (uqueue.empty()) break;  // stopping with nothing left to apply
            Request *r = uqueue.front();
            uqueue.pop_front();
            r->t_start = now_s();
            lk.unlock();
            set_defer_frees(true);
            const int rc = apply_update(r);
            set_defer_frees(false);
            const double t_applied = now_s();
            std::string err;
            if (rc) {
                std::lock_guard<std::mutex> g(c->err_mu);
                err = c->err;
            }
            lk.lock();
            published = snap_of(c);
            const uint64_t v = ++version;
            // r belongs to its waiter once done is set
            const int kind = r->kind;
            const long long n_rows = (long long)r->n;
            const double t_started = r->t_start;
            r->status = rc;
            r->err = err;
            r->t_done = now_s();
            r->done = true;
            ++n_updates;
            cv_done.notify_all();
            cv_done.wait(lk, [&] { return drained_before(v); });
            const double t_drained = now_s();
            lk.unlock();
            size_t freed = 0;
            {
                std::lock_guard<std::mutex> g(c->mem_mu);
                for (const auto &x : c->retired) freed += x.bytes;
            }
            release_retired(c);
            if (trace)
                std::fprintf(stderr, "[rkr-server] update kind=%d n=%lld apply_ms=%.2f drain_ms=%.2f free_ms=%.2f freed_mb=%.1f\n",
                             kind, n_rows, 1e3 * (t_applied - t_started), 1e3 * (t_drained - t_applied),
                             1e3 * (now_s() - t_drained), freed / 1048576.0);
            lk.lock();
        }
    }

    int submit(std::unique_ptr<Request> r, uint64_t *ticket) {
        std::lock_guard<std::mutex> g(mu);
        if (stopping) return RKR_ERR_STATE;
        r->id = next_id++;
        r->t_submit = now_s();
        *ticket = r->id;
        if (snapshot_mode() && r->kind != Q_QUERY) {
            uqueue.push_back(r.get());
            cv_upd.notify_one();
        } else {
            queue.push_back(r.get());
            cv_work.notify_one();
        }
        reqs.emplace(r->id, std::move(r));
        return RKR_OK;
    }
};

extern "C" {

double rkr_now(void) { return now_s(); }

int rkr_server_start(rkr_ctx *c, const rkr_server_config *cfg, rkr_server **out) {
    if (!c || !out) return RKR_ERR_INVALID;
    *out = nullptr;
    if (!c->table) return set_err(c, RKR_ERR_STATE, "server: no table (build or load one first)");
    auto *s = new (std::nothrow) rkr_server();
    if (!s) return RKR_ERR_NOMEM;
    s->c = c;
    s->cfg.max_batch = cfg && cfg->max_batch > 0 ? cfg->max_batch : 256;
    s->cfg.n_slots = cfg && cfg->n_slots > 0 ? std::min(cfg->n_slots, 4) : 2;
    s->cfg.update_mode = cfg ? cfg->update_mode : RKR_SERVE_BARRIER;
    s->cfg.rebuild_pending = cfg ? cfg->rebuild_pending : 0;
    s->rebuild_at = s->cfg.rebuild_pending == 0 ? 12288 : s->cfg.rebuild_pending < 0 ? -1 : s->cfg.rebuild_pending;
    if (s->cfg.update_mode != RKR_SERVE_BARRIER && s->cfg.update_mode != RKR_SERVE_SNAPSHOT) {
        delete s;
        return set_err(c, RKR_ERR_INVALID, "server: unknown update mode");
    }
    s->published = snap_of(c);
    const char *tr = std::getenv("RKR_SERVER_TRACE");
    s->trace = tr && tr[0] == '1';
    (void)hipSetDevice(c->device);
    // snapshot mode: updates run on the last k CUs and batches on the others
    std::vector<uint32_t> upd_mask, query_mask;
    if (s->snapshot_mode()) {
        const int n_cu = std::max(c->n_cu, 1);
        int k = cfg && cfg->update_cus != 0 ? cfg->update_cus : n_cu / 4;
        if (k < 0 || k > n_cu) k = n_cu;
        k = std::max(k, 1);
        if (k < n_cu) {
            upd_mask.assign((size_t)(n_cu + 31) / 32, 0u);
            query_mask.assign((size_t)(n_cu + 31) / 32, 0u);
            for (int cu = 0; cu < n_cu; ++cu) (cu >= n_cu - k ? upd_mask : query_mask)[cu / 32] |= 1u << (cu % 32);
        }
    }
    for (int i = 0; i < s->cfg.n_slots; ++i) {
        s->slots.emplace_back(new Slot());
        const int rc = slot_init(c, *s->slots.back(), true, query_mask.empty() ? nullptr : &query_mask);
        if (rc) {
            for (auto &sl : s->slots) slot_free(c, *sl);
            delete s;
            return rc;
        }
    }
    for (int i = 0; i < s->cfg.n_slots; ++i) s->workers.emplace_back(&rkr_server::worker, s, i);
    if (s->snapshot_mode()) {
        s->ctx_stream = c->stream;
        if (!upd_mask.empty()) {
            if (hipExtStreamCreateWithCUMask(&s->upd_stream, (uint32_t)upd_mask.size(), upd_mask.data()) == hipSuccess)
                c->stream = s->upd_stream;
            else
                s->upd_stream = nullptr;  // updates then share all CUs
        }
        s->updater = std::thread(&rkr_server::update_loop, s);
    }
    *out = s;
    return RKR_OK;
}

int rkr_server_submit_query(rkr_server *s, const float *q, int32_t k, uint64_t *ticket) {
    if (!s || !q || !ticket || k <= 0) return RKR_ERR_INVALID;
    auto r = std::make_unique<Request>();
    r->kind = Q_QUERY;
    r->data.assign(q, q + s->c->dim);
    r->k = k;
    return s->submit(std::move(r), ticket);
}

int rkr_server_submit_update(rkr_server *s, int kind, const float *vectors, const int64_t *ids, int64_t n,
                             uint64_t *ticket) {
    if (!s || !ticket || n < 0) return RKR_ERR_INVALID;
    auto r = std::make_unique<Request>();
    r->n = n;
    switch (kind) {
        case RKR_UPDATE_INSERT_USERS: r->kind = Q_INSERT_USERS; break;
        case RKR_UPDATE_INSERT_ITEMS: r->kind = Q_INSERT_ITEMS; break;
        case RKR_UPDATE_DELETE_USERS: r->kind = Q_DELETE_USERS; break;
        case RKR_UPDATE_DELETE_ITEMS: r->kind = Q_DELETE_ITEMS; break;
        default: return RKR_ERR_INVALID;
    }
    if (r->kind == Q_INSERT_USERS || r->kind == Q_INSERT_ITEMS) {
        if (n > 0 && !vectors) return RKR_ERR_INVALID;
        r->data.assign(vectors, vectors + (size_t)n * s->c->dim);
    } else {
        if (n > 0 && !ids) return RKR_ERR_INVALID;
        r->ids.assign(ids, ids + n);
    }
    return s->submit(std::move(r), ticket);
}

int rkr_server_wait(rkr_server *s, uint64_t ticket, int32_t *out_user, int32_t *out_rank, float *out_qip,
                    rkr_request_info *info) {
    if (!s) return RKR_ERR_INVALID;
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return RKR_ERR_INVALID;
        Request *p = it->second.get();
        s->cv_done.wait(lk, [&] { return p->done; });
        r = std::move(it->second);
        s->reqs.erase(it);
    }
    if (!r->status && r->kind == Q_QUERY) {
        if (out_user) std::memcpy(out_user, r->users.data(), sizeof(int32_t) * r->users.size());
        if (out_rank) std::memcpy(out_rank, r->ranks.data(), sizeof(int32_t) * r->ranks.size());
        if (out_qip) std::memcpy(out_qip, r->qips.data(), sizeof(float) * r->qips.size());
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->status = r->status;
        info->batch_size = r->batch_size;
        info->t_submit = r->t_submit;
        info->t_start = r->t_start;
        info->t_done = r->t_done;
        info->stats = r->stats;
        std::strncpy(info->error, r->err.c_str(), sizeof(info->error) - 1);
    }
    return r->status;
}

int rkr_server_poll(rkr_server *s, uint64_t ticket, int *done) {
    if (!s || !done) return RKR_ERR_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return RKR_ERR_INVALID;
    *done = it->second->done ? 1 : 0;
    return RKR_OK;
}

int rkr_server_counters(rkr_server *s, int64_t *n_batches, int64_t *n_queries, int64_t *n_updates) {
    if (!s) return RKR_ERR_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    if (n_batches) *n_batches = s->n_batches;
    if (n_queries) *n_queries = s->n_queries;
    if (n_updates) *n_updates = s->n_updates;
    return RKR_OK;
}

int rkr_server_rebuilds(rkr_server *s, int64_t *n_rebuilds) {
    if (!s || !n_rebuilds) return RKR_ERR_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    *n_rebuilds = s->n_rebuilds;
    return RKR_OK;
}

int rkr_server_stop(rkr_server *s) {
    if (!s) return RKR_ERR_INVALID;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stopping = true;
        s->cv_work.notify_all();
        s->cv_upd.notify_all();
    }
    for (auto &t : s->workers) t.join();
    if (s->updater.joinable()) s->updater.join();
    if (s->upd_stream) {
        (void)hipStreamSynchronize(s->upd_stream);
        s->c->stream = s->ctx_stream;
        (void)hipStreamDestroy(s->upd_stream);
    }
    release_retired(s->c);
    for (auto &sl : s->slots) slot_free(s->c, *sl);
    delete s;
    return RKR_OK;
}

}  // extern "C"
