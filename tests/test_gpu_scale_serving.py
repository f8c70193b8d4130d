"""GPU parity of the serving form at its full size: 2M users x 100k items x
d=128, S=512, queries with k uniform in [10, 200] through the continuous-
batching server, and an insert round of 1% new users (20k) and 0.5% new items
(500) applied beside running batches (snapshot mode).

Checks, all against the oracle on the CPU:
  - answers in the mix equal the same queries answered alone, and a query
    submitted together with the insert round equals its answer either before
    or after the round commits;
  - every answer is ordered by (rank, userID), and sampled (query, user) pairs
    carry the strict count over the items current when the query ran
    (RefineCPU.hpp:33-55; the inserted items follow SampleSearchUpdate's
    rank offsets, SampleSearchUpdate.hpp:212-279, through exact counts);
  - answers given with the inserts pending equal the answers after the table
    is rebuilt over the current items, and for two queries those equal the
    oracle's exact query through the rebuilt table, whose sampled rows equal
    the oracle's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _order_ok(u, r):
    key = r.astype(np.int64) * (1 << 32) + u.astype(np.int64)
    return len(np.unique(u)) == len(u) and np.all(u >= 0) and np.all(np.diff(key) > 0)


@pytest.mark.timeout(1200)
def test_serving_2m_x_100k_with_inserts(orc):
    import rkr
    from rkr import synth
    _case("scale: serving 2M x 100k d=128 S=512 with an insert round")
    wl = synth.make("serving", seed=0)
    U0, I0, d = wl.users.shape[0], wl.items.shape[0], wl.users.shape[1]
    rng = np.random.default_rng(55)
    idx = rkr.Index(0)
    try:
        idx.set_data(wl.users, wl.items)
        ranks = rkr.Index.uniform_sample_ranks(I0, wl.n_sample)
        idx.build_table(ranks)
        n_new_u, n_new_i = U0 // 100, I0 // 200
        idx.reserve(U0 + n_new_u, n_new_i)
        new_u = rng.standard_normal((n_new_u, d), dtype=np.float32) / np.float32(np.sqrt(d))
        new_i = rng.standard_normal((n_new_i, d), dtype=np.float32) / np.float32(np.sqrt(d))
        # before the inserts (a), submitted with them (b), after them (c; some are inserted items)
        q_a, k_a = wl.queries[:160], wl.topk[:160]
        q_b, k_b = wl.queries[160:320], wl.topk[160:320]
        q_post = np.concatenate([wl.queries[320:460], new_i[:20]])
        k_post = rng.integers(10, 201, len(q_post)).astype(np.int32)
        ans_a = idx.query(q_a, k_a, max_k=200)
        ans_b0 = idx.query(q_b, k_b, max_k=200)

        with rkr.Server(idx, max_batch=64, n_slots=2, update_mode="snapshot") as srv:
            mix_pre = [srv.wait(t, int(k))[:2] for t, k in
                       zip([srv.submit_query(q, int(k)) for q, k in zip(q_a, k_a)], k_a)]
            t_u = srv.submit_update("insert_users", vectors=new_u)
            t_i = srv.submit_update("insert_items", vectors=new_i)
            t_b = [srv.submit_query(q, int(k)) for q, k in zip(q_b, k_b)]
            mix_b = [srv.wait(t, int(k))[:2] for t, k in zip(t_b, k_b)]
            srv.wait(t_u)
            srv.wait(t_i)
            mix_post = [srv.wait(t, int(k))[:2] for t, k in
                        zip([srv.submit_query(q, int(k)) for q, k in zip(q_post, k_post)], k_post)]
        for i, (u, r) in enumerate(mix_pre):
            k = int(k_a[i])
            assert np.array_equal(u, ans_a.users[i, :k]) and np.array_equal(r, ans_a.ranks[i, :k]), \
                f"pre-insert query {i}: mix differs from the batch answer"
        assert idx.counts()["n_inserted"] == n_new_i and idx.counts()["n_user"] == U0 + n_new_u
        _case("  pre-insert answers in the mix equal the batch answers")
        ans_b1 = idx.query(q_b, k_b, max_k=200)
        n_old = 0
        for i, (u, r) in enumerate(mix_b):
            k = int(k_b[i])
            old = np.array_equal(u, ans_b0.users[i, :k]) and np.array_equal(r, ans_b0.ranks[i, :k])
            new = np.array_equal(u, ans_b1.users[i, :k]) and np.array_equal(r, ans_b1.ranks[i, :k])
            assert old or new, f"query {i} submitted with the inserts matches neither state"
            n_old += old and not new
        _case(f"  queries submitted with the inserts: {n_old} answered before the commit, "
              f"{len(mix_b) - n_old} after (or alike in both)")

        post_batch = idx.query(q_post, k_post, max_k=200)
        for i, (u, r) in enumerate(mix_post):
            k = int(k_post[i])
            assert np.array_equal(u, post_batch.users[i, :k]) and np.array_equal(r, post_batch.ranks[i, :k]), \
                f"post-insert query {i}: mix differs from the batch answer"
        for i in rng.choice(len(q_post), 24, replace=False):
            one = idx.query(q_post[i:i + 1], k_post[i:i + 1], max_k=200)
            assert np.array_equal(one.users[0], post_batch.users[i]) and \
                np.array_equal(one.ranks[0], post_batch.ranks[i]), f"post-insert query {i}: alone differs"
        _case("  post-insert answers: mix, batch and alone equal")

        users_cur = np.concatenate([orc.normalize_users(wl.users), orc.normalize_users(new_u)])
        items_cur = np.concatenate([wl.items, new_i])
        un0 = users_cur[:U0]
        # (query, user) pairs: k-th user of every query and every user of 4 queries per phase
        for phase, (qs, ks, ans, users, items) in enumerate((
                (q_a, k_a, mix_pre, un0, wl.items), (q_post, k_post, mix_post, users_cur, items_cur))):
            qid, uid, rk = [], [], []
            pick = set(rng.choice(len(qs), 4, replace=False).tolist())
            for i, (u, r) in enumerate(ans):
                assert _order_ok(u, r), f"phase {phase} query {i}: not ordered by (rank, userID)"
                js = range(len(u)) if i in pick else [len(u) - 1]
                for j in js:
                    qid.append(i), uid.append(int(u[j])), rk.append(int(r[j]))
            got, _ = orc.refine_ranks(users, qs, items, np.array(uid), np.array(qid))
            assert np.array_equal(got, np.array(rk, np.int32)), f"phase {phase}: ranks differ from the oracle"
            _case(f"  phase {phase}: {len(qid)} (query, user) ranks equal the oracle's strict count")

        # rebuild over the current items: the exact answers do not depend on the table
        ranks2 = rkr.Index.uniform_sample_ranks(len(items_cur), wl.n_sample)
        idx.build_table(ranks2)
        after = idx.query(q_post, k_post, max_k=200)
        for i, (u, r) in enumerate(mix_post):
            k = int(k_post[i])
            assert np.array_equal(after.users[i, :k], u) and np.array_equal(after.ranks[i, :k], r), \
                f"post-insert query {i}: differs from the answer after the rebuild"
        rows = np.sort(rng.choice(len(users_cur), 256, replace=False))
        T_rows = np.concatenate([idx.get_table(int(r), 1) for r in rows])
        assert np.array_equal(T_rows, orc.build_table(users_cur[rows], items_cur, ranks2))
        T = idx.get_table()
        for i in (int(np.argmax(k_post)), len(q_post) - 1):
            k = int(k_post[i])
            ou, orank, _ = orc.exact_query(users_cur, items_cur, T, ranks2, q_post[i:i + 1], k)
            assert np.array_equal(after.users[i, :k], ou[0]) and np.array_equal(after.ranks[i, :k], orank[0]), \
                f"post-insert query {i}: differs from the oracle's exact answer"
        _case("  answers with pending inserts equal the rebuilt table's and the oracle's")
    finally:
        idx.close()
