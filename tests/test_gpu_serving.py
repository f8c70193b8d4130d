"""GPU tests of the serving loop: answers are the same alone and in the mix,
updates are barriers, and every answer equals brute force over the users and
items current when the query was admitted."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _setup(seed=41, U=4000, I=2500, d=32, S=48):
    import rkr
    rng = np.random.default_rng(seed)
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.lognormal(0, 0.4, (I, 1)).astype(np.float32)
    idx = rkr.Index(0)
    idx.set_data(users, items)
    idx.build_table(rkr.Index.uniform_sample_ranks(I, S))
    return rng, users, items, idx


@pytest.mark.parametrize("form", ["plain", "screened"])
def test_alone_equals_mix(orc, form):
    """Each query answered alone (one request in flight) equals its answer when
    submitted together with 200 others of mixed k from two threads."""
    _case(f"serving alone vs mix ({form})")
    import rkr
    rng, users, items, idx = _setup()
    idx.set_form(form)
    qs = items[rng.choice(len(items), 120, replace=False)]
    ks = rng.integers(1, 120, len(qs)).astype(np.int32)
    with rkr.Server(idx, max_batch=37, n_slots=2) as srv:
        alone = []
        for q, k in zip(qs, ks):
            u, r, _, info = srv.wait(srv.submit_query(q, int(k)), int(k))
            assert info["batch_size"] == 1
            alone.append((u.copy(), r.copy()))
        tickets = [None] * len(qs)

        def submit(part):
            for i in part:
                tickets[i] = srv.submit_query(qs[i], int(ks[i]))
        th = [threading.Thread(target=submit, args=(range(j, len(qs), 2),)) for j in range(2)]
        [t.start() for t in th]
        [t.join() for t in th]
        sizes = []
        for i, t in enumerate(tickets):
            u, r, _, info = srv.wait(t, int(ks[i]))
            sizes.append(info["batch_size"])
            assert np.array_equal(u, alone[i][0]) and np.array_equal(r, alone[i][1]), f"query {i} differs in the mix"
        assert max(sizes) > 1  # they were batched
        cnt = srv.counters()
        assert cnt["n_queries"] == 2 * len(qs)
    un = orc.normalize_users(users)
    bu, br, _ = orc.bruteforce(un, items, qs[:10], int(ks[:10].max()))
    for i in range(10):
        assert np.array_equal(alone[i][0], bu[i, :ks[i]]) and np.array_equal(alone[i][1], br[i, :ks[i]])


def test_updates_are_barriers(orc):
    """Queries before an update see the old data, queries after it the new data."""
    _case("serving updates as barriers")
    import rkr
    rng, users, items, idx = _setup(seed=42, U=3000, I=2000)
    qs = items[:8]
    new_users = rng.standard_normal((300, users.shape[1]), dtype=np.float32)
    new_items = rng.standard_normal((120, items.shape[1]), dtype=np.float32) * np.float32(1.3)
    k = 40
    with rkr.Server(idx, max_batch=16, n_slots=2) as srv:
        t_before = [srv.submit_query(q, k) for q in qs]
        t_u = srv.submit_update("insert_users", vectors=new_users)
        t_mid = [srv.submit_query(q, k) for q in qs]
        t_i = srv.submit_update("insert_items", vectors=new_items)
        t_after = [srv.submit_query(q, k) for q in qs]
        before = [srv.wait(t, k)[:2] for t in t_before]
        srv.wait(t_u)
        mid = [srv.wait(t, k)[:2] for t in t_mid]
        srv.wait(t_i)
        after = [srv.wait(t, k)[:2] for t in t_after]
    un0 = orc.normalize_users(users)
    un1 = np.concatenate([un0, orc.normalize_users(new_users)])
    items2 = np.concatenate([items, new_items])
    for states, (U_, I_) in ((before, (un0, items)), (mid, (un1, items)), (after, (un1, items2))):
        bu, br, _ = orc.bruteforce(U_, I_, qs, k)
        for i in range(len(qs)):
            assert np.array_equal(states[i][0], bu[i]) and np.array_equal(states[i][1], br[i])


def test_bad_request_does_not_stop_the_loop():
    _case("serving error isolation")
    import rkr
    rng, users, items, idx = _setup(seed=43, U=500, I=800)
    with rkr.Server(idx, max_batch=8) as srv:
        bad = srv.submit_query(items[0], 5000)  # k > n_user
        ok = srv.submit_query(items[1], 5)
        with pytest.raises(rkr.RkrError):
            srv.wait(bad, 5000)
        u, r, _, info = srv.wait(ok, 5)
        assert info["status"] == 0 and len(u) == 5


def test_snapshot_updates(orc):
    """Snapshot mode: queries keep running while an update is applied; each
    answer equals brute force over the state before or after the update, and
    queries submitted after the update committed see it."""
    _case("serving snapshot updates")
    import rkr
    rng, users, items, idx = _setup(seed=44, U=5000, I=2200)
    qs = items[rng.choice(len(items), 24, replace=False)]
    k = 35
    new_users = rng.standard_normal((800, users.shape[1]), dtype=np.float32)
    new_items = rng.standard_normal((300, items.shape[1]), dtype=np.float32) * np.float32(1.2)
    un0 = orc.normalize_users(users)
    un1 = np.concatenate([un0, orc.normalize_users(new_users)])
    items2 = np.concatenate([items, new_items])
    states = [(un0, items), (un1, items), (un1, items2)]
    truth = []
    for U_, I_ in states:
        bu, br, _ = orc.bruteforce(U_, I_, qs, k)
        truth.append((bu, br))

    def which(u, r, i):
        return [s for s, (bu, br) in enumerate(truth) if np.array_equal(u, bu[i]) and np.array_equal(r, br[i])]

    with rkr.Server(idx, max_batch=8, n_slots=2, update_mode="snapshot") as srv:
        t_u = srv.submit_update("insert_users", vectors=new_users)
        during = [srv.submit_query(q, k) for q in qs]
        t_i = srv.submit_update("insert_items", vectors=new_items)
        during += [srv.submit_query(q, k) for q in qs]
        for j, t in enumerate(during):
            u, r, _, info = srv.wait(t, k)
            assert info["status"] == 0
            assert which(u, r, j % len(qs)), f"query {j}: answer matches no committed state"
        assert srv.wait(t_u)[3]["status"] == 0 and srv.wait(t_i)[3]["status"] == 0
        after = [srv.submit_query(q, k) for q in qs]
        for j, t in enumerate(after):
            u, r, _, _ = srv.wait(t, k)
            assert which(u, r, j) == [2], f"query {j} after both commits does not see them"
        cnt = srv.counters()
        assert cnt["n_updates"] == 2
    assert idx.counts()["n_user"] == 5800 and idx.counts()["n_item"] == 2500


@pytest.mark.parametrize("mode", ["barrier", "snapshot"])
def test_inserts_past_the_side_cap_rebuild_the_table(orc, mode):
    """A steady stream of item inserts with a small rebuild threshold: the loop
    rebuilds the table over the current items mid-stream (folding the side
    arrays into a fresh base) and keeps accepting inserts; every answer equals
    brute force over a state the stream passed through (barrier mode: exactly
    the updates submitted before the query)."""
    _case(f"serving rebuilds under inserts ({mode})")
    import rkr
    rng, users, items, idx = _setup(seed=45, U=3000, I=1500, d=24, S=32)
    k = 25
    qs = items[rng.choice(len(items), 6, replace=False)]
    rounds = [rng.standard_normal((40, items.shape[1]), dtype=np.float32) * np.float32(1.1) for _ in range(7)]
    un = orc.normalize_users(users)
    states = [items]
    for r in rounds:
        states.append(np.concatenate([states[-1], r]))
    truth = [orc.bruteforce(un, st, qs, k)[:2] for st in states]
    with rkr.Server(idx, max_batch=8, n_slots=2, update_mode=mode, rebuild_pending=100) as srv:
        tickets = []
        for r in rounds:
            tickets.append(("q", [srv.submit_query(q, k) for q in qs]))
            tickets.append(("u", srv.submit_update("insert_items", vectors=r)))
        tickets.append(("q", [srv.submit_query(q, k) for q in qs]))
        n_upd = 0
        for kind, t in tickets:
            if kind == "u":
                assert srv.wait(t)[3]["status"] == 0
                n_upd += 1
                continue
            for j, tq in enumerate(t):
                u, r, _, info = srv.wait(tq, k)
                assert info["status"] == 0
                ok = [s for s, (bu, br) in enumerate(truth) if np.array_equal(u, bu[j]) and np.array_equal(r, br[j])]
                assert ok, f"query {j} after {n_upd} insert rounds matches no state"
                if mode == "barrier":
                    assert n_upd in ok
        assert srv.rebuilds() >= 2  # 280 inserted items, a rebuild at every 100 pending
    c = idx.counts()
    assert c["n_item"] == 1500 + 280 and c["n_inserted"] < 100
    bu, br = truth[-1]
    res = idx.query(qs, k)
    assert np.array_equal(res.users, bu) and np.array_equal(res.ranks, br)
