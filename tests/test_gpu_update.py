"""GPU parity of the update path: inserts and deletes of items and users keep
every answer equal to brute force over the current users and items."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


class Mirror:
    """Host copy of the current user and item sets, in the library's order:
    base items not deleted, then inserted items not deleted."""

    def __init__(self, orc, users, items):
        self.orc = orc
        self.users = orc.normalize_users(users)
        self.base = items.copy()
        self.base_alive = np.ones(len(items), bool)
        self.ins = np.zeros((0, items.shape[1]), np.float32)

    def items(self):
        return np.concatenate([self.base[self.base_alive], self.ins])

    def insert_items(self, x):
        self.ins = np.concatenate([self.ins, x])

    def delete_items(self, ids):
        cur_base = np.flatnonzero(self.base_alive)
        nb = len(cur_base)
        ins_keep = np.ones(len(self.ins), bool)
        for i in ids:
            if i < nb:
                self.base_alive[cur_base[i]] = False
            else:
                ins_keep[i - nb] = False
        self.ins = self.ins[ins_keep]

    def insert_users(self, u):
        self.users = np.concatenate([self.users, self.orc.normalize_users(u)])

    def delete_users(self, ids):
        keep = np.ones(len(self.users), bool)
        keep[ids] = False
        self.users = self.users[keep]

    def check(self, idx, queries, k, form="plain"):
        idx.set_form(form)
        res = idx.query(queries, k)
        items = self.items()
        assert idx.counts()["n_item"] == len(items) and idx.counts()["n_user"] == len(self.users)
        bu, br, bq = self.orc.bruteforce(self.users, items, queries, k)
        for i in range(len(queries)):
            assert np.array_equal(res.users[i], bu[i]), f"query {i}: users differ ({form})"
            assert np.array_equal(res.ranks[i], br[i]), f"query {i}: ranks differ ({form})"
        return res


def _data(seed, U=3000, I=2000, d=24):
    rng = np.random.default_rng(seed)
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.lognormal(0, 0.4, (I, 1)).astype(np.float32)
    return rng, users, items


@pytest.mark.parametrize("form", ["plain", "screened"])
def test_insert_and_delete_items(orc, gpu_index_factory, form):
    _case(f"item inserts/deletes ({form})")
    rng, users, items = _data(31)
    idx, _ = gpu_index_factory(users, items, 40)
    m = Mirror(orc, users, items)
    queries = items[rng.choice(len(items), 12, replace=False)]
    m.check(idx, queries, 30, form)
    new = rng.standard_normal((150, items.shape[1]), dtype=np.float32) * np.float32(1.5)
    idx.insert_items(new)
    m.insert_items(new)
    q2 = np.concatenate([queries, new[:6]])  # queries equal to inserted items
    m.check(idx, q2, 30, form)
    # delete base and inserted items in one call
    dels = np.array([0, 5, 17, len(items) + 3, len(items) + 40], np.int64)
    idx.delete_items(dels)
    m.delete_items(dels)
    m.check(idx, q2, 30, form)
    # a second insert batch merges into the sorted side arrays
    new2 = rng.standard_normal((1100, items.shape[1]), dtype=np.float32)
    idx.insert_items(new2)
    m.insert_items(new2)
    m.check(idx, q2, 60, form)


def test_insert_and_delete_users(orc, gpu_index_factory):
    _case("user inserts/deletes")
    rng, users, items = _data(32)
    idx, _ = gpu_index_factory(users, items, 40)
    m = Mirror(orc, users, items)
    new_items = rng.standard_normal((70, items.shape[1]), dtype=np.float32)
    idx.insert_items(new_items)
    m.insert_items(new_items)
    idx.delete_items(np.array([2, 3], np.int64))
    m.delete_items([2, 3])
    queries = items[rng.choice(len(items), 10, replace=False)]
    new_users = rng.standard_normal((500, items.shape[1]), dtype=np.float32)
    new_users[0] *= 0  # a zero user (left unnormalised) has rank 0 everywhere
    idx.insert_users(new_users)
    m.insert_users(new_users)
    m.check(idx, queries, 25)
    m.check(idx, queries, 25, "screened")
    gone = np.array([0, 1, 999, 3100, 3499], np.int64)
    idx.delete_users(gone)
    m.delete_users(gone)
    m.check(idx, queries, 25)


@pytest.mark.parametrize("nq,inserts", [(1, True), (12, True), (6, False)])
def test_updates_many_users_chunked(orc, gpu_index_factory, nq, inserts):
    """Enough users that the prune kernels split each query over several
    chunks, with deletes that make some shifted upper bounds negative (and a
    delete-only case: side arrays of deletes without inserts)."""
    _case(f"updates, chunked prune, nq={nq}, inserts={inserts}")
    rng, users, items = _data(35, U=13000, I=900, d=16)
    idx, _ = gpu_index_factory(users, items, 24)
    m = Mirror(orc, users, items)
    new = rng.standard_normal((40, items.shape[1]), dtype=np.float32) * np.float32(1.4)
    if inserts:
        idx.insert_items(new)
        m.insert_items(new)
    top = np.argsort(-np.linalg.norm(items, axis=1))[:25].astype(np.int64)
    idx.delete_items(np.sort(top))
    m.delete_items(np.sort(top))
    queries = m.items()[rng.choice(len(m.items()), nq, replace=False)]
    m.check(idx, queries, 60)
    m.check(idx, queries, 60, "screened")


def test_update_rounds_reuse_buffers(orc, gpu_index_factory):
    """Several rounds of user and item inserts (the side arrays alternate with
    their spares, capacities grow), then a user delete and an item delete;
    every round's answers equal brute force."""
    _case("update rounds, spare rotation and capacity growth")
    rng, users, items = _data(36, U=2500, I=1200, d=20)
    idx, _ = gpu_index_factory(users, items, 32)
    m = Mirror(orc, users, items)
    queries = items[rng.choice(len(items), 9, replace=False)]
    for r in range(4):
        nu = rng.standard_normal((700 + 300 * r, users.shape[1]), dtype=np.float32)
        ni = rng.standard_normal((25 + 10 * r, items.shape[1]), dtype=np.float32) * np.float32(1.1)
        idx.insert_users(nu)
        m.insert_users(nu)
        idx.insert_items(ni)
        m.insert_items(ni)
        m.check(idx, queries, 45, "screened" if r % 2 else "plain")
    gone = np.sort(rng.choice(idx.counts()["n_user"], 400, replace=False)).astype(np.int64)
    idx.delete_users(gone)
    m.delete_users(gone)
    dels = np.array([1, 7, len(items) + 2, len(items) + 30], np.int64)
    idx.delete_items(dels)
    m.delete_items(dels)
    m.check(idx, queries, 45)
    ni = rng.standard_normal((15, items.shape[1]), dtype=np.float32)
    idx.insert_items(ni)
    m.insert_items(ni)
    m.check(idx, queries, 45, "screened")


def test_reserved_updates(orc, gpu_index_factory):
    """Updates within reserved capacity (no reallocation) stay exact."""
    _case("updates within reserved capacity")
    rng, users, items = _data(37, U=3000, I=1000, d=16)
    idx, _ = gpu_index_factory(users, items, 24)
    idx.reserve(5000, 400)
    m = Mirror(orc, users, items)
    queries = items[rng.choice(len(items), 8, replace=False)]
    for r in range(3):
        nu = rng.standard_normal((500, users.shape[1]), dtype=np.float32)
        ni = rng.standard_normal((100, items.shape[1]), dtype=np.float32)
        idx.insert_items(ni)
        m.insert_items(ni)
        idx.insert_users(nu)
        m.insert_users(nu)
        m.check(idx, queries, 40, "screened")
    m.check(idx, queries, 40)


def test_rebuild_after_updates_matches_oracle_table(orc, gpu_index_factory):
    _case("rebuild table after updates")
    rng, users, items = _data(33, U=800, I=1500)
    idx, _ = gpu_index_factory(users, items, 30)
    m = Mirror(orc, users, items)
    new = rng.standard_normal((300, items.shape[1]), dtype=np.float32)
    idx.insert_items(new)
    m.insert_items(new)
    idx.delete_items(np.array([10, 11, 1600], np.int64))
    m.delete_items([10, 11, 1600])
    cur = m.items()
    ranks = idx.uniform_sample_ranks(len(cur), 30)
    idx.build_table(ranks)
    assert idx.counts()["n_inserted"] == 0 and idx.counts()["n_deleted"] == 0
    assert np.array_equal(idx.get_table(), orc.build_table(m.users, cur, ranks))
    m.check(idx, cur[:8], 20)


def test_save_table_refuses_pending_updates(orc, gpu_index_factory, tmp_path):
    """A saved table must describe the item count in its header: with pending
    inserts or deletes the save is refused, and after a rebuild the file
    round-trips and answers stay exact."""
    _case("save_table with pending updates")
    import rkr
    from rkr import formats
    rng, users, items = _data(35, U=500, I=700)
    idx, _ = gpu_index_factory(users, items, 20)
    m = Mirror(orc, users, items)
    new = rng.standard_normal((40, items.shape[1]), dtype=np.float32)
    idx.insert_items(new)
    m.insert_items(new)
    p = str(tmp_path / "t.index")
    with pytest.raises(rkr.RkrError):
        idx.save_table(p)
    idx.delete_items(np.array([5], np.int64))
    m.delete_items([5])
    with pytest.raises(rkr.RkrError):
        idx.save_table(p)
    m.check(idx, m.items()[:6], 15)  # the refusals left the context usable
    cur = m.items()
    ranks = idx.uniform_sample_ranks(len(cur), 20)
    idx.build_table(ranks)
    idx.save_table(p)
    f_ranks, f_table, f_items = formats.read_table(p)
    assert f_items == len(cur) and np.array_equal(f_ranks, ranks)
    assert np.array_equal(f_table, orc.build_table(m.users, cur, ranks))
    idx.load_table(p)
    m.check(idx, cur[:6], 15)


def test_reference_output_rejects_updates(gpu_index_factory):
    _case("reference output + updates")
    import rkr
    rng, users, items = _data(34, U=200, I=300)
    idx, _ = gpu_index_factory(users, items, 10)
    idx.set_output("reference")
    with pytest.raises(rkr.RkrError):
        idx.insert_items(items[:3])
    idx.set_output("exact")
    idx.insert_items(items[:3])
    idx.set_output("reference")
    with pytest.raises(rkr.RkrError):
        idx.query(items[:2], 5)


@pytest.mark.parametrize("deletes", [False, True])
def test_coarse_bounds_with_updates(orc, gpu_index_factory, deletes):
    """Screened form with pending inserts (and deletes): the coarse bounds
    widen the list by the largest rank shift any user has at the k-th guard,
    and the listed users get the update path's exact bound values; answers
    and prune counts equal the full update path's and brute force."""
    _case(f"coarse bounds with updates, deletes={deletes}")
    rng, users, items = _data(77, U=9000, I=2500, d=32)
    idx, _ = gpu_index_factory(users, items, 64)
    m = Mirror(orc, users, items)
    for r in range(3):
        new = rng.standard_normal((60, items.shape[1]), dtype=np.float32) * np.float32(1.2)
        idx.insert_items(new)
        m.insert_items(new)
        nu = rng.standard_normal((150, users.shape[1]), dtype=np.float32)
        idx.insert_users(nu)
        m.insert_users(nu)
    if deletes:
        gone = np.sort(rng.choice(len(m.items()), 40, replace=False)).astype(np.int64)
        idx.delete_items(gone)
        m.delete_items(gone)
    cur = m.items()
    nq = 48
    queries = np.concatenate([cur[rng.choice(len(cur), nq // 2, replace=False)],
                              rng.standard_normal((nq // 2, cur.shape[1]), dtype=np.float32)])
    ks = [int(k) for k in rng.integers(1, 150, nq)]
    idx.set_form("screened")
    idx.set_option(9, 0)  # RKR_OPT_COARSE off: the full update path
    full = idx.query(queries, ks)
    idx.set_option(9, 1)
    coarse = idx.query(queries, ks)
    assert coarse.timing["n_listed"] > 0
    bu, br, bq = orc.bruteforce(m.users, cur, queries, max(ks))
    for i, k in enumerate(ks):
        for res in (full, coarse):
            assert np.array_equal(res.users[i, :k], bu[i, :k]), f"query {i}: users differ"
            assert np.array_equal(res.ranks[i, :k], br[i, :k]), f"query {i}: ranks differ"
        assert np.array_equal(full.qips[i, :k], coarse.qips[i, :k])
        for f in ("n_candidate", "n_refine_user", "n_prune_user", "n_result_user"):
            assert full.stats[i][f] == coarse.stats[i][f], f"query {i}: {f}"
    # weak queries (negated items) with large k put the k-th bound at the
    # deepest guard: users below it are left out of the list and the k-th
    # value is checked against their lowest possible rank afterwards
    weak = -cur[rng.choice(len(cur), 24, replace=False)]
    kw = [int(k) for k in rng.integers(200, 1500, len(weak))]
    idx.set_option(9, 0)
    full = idx.query(weak, kw)
    idx.set_option(9, 1)
    coarse = idx.query(weak, kw)
    bu, br, _ = orc.bruteforce(m.users, cur, weak, max(kw))
    for i, k in enumerate(kw):
        for res in (full, coarse):
            assert np.array_equal(res.users[i, :k], bu[i, :k]), f"weak query {i}: users differ"
            assert np.array_equal(res.ranks[i, :k], br[i, :k]), f"weak query {i}: ranks differ"


@pytest.mark.parametrize("form", ["plain", "screened"])
def test_two_query_slots_with_pending_updates(orc, gpu_index_factory, form):
    """A call split over two streams (device batches of 200, 650 queries) with pending item
    inserts and deletes and inserted users: both parts read the same snapshot of the side
    arrays, so the answers equal the one-stream run and brute force."""
    _case(f"two query slots with pending updates ({form})")
    rng, users, items = _data(35)
    idx, _ = gpu_index_factory(users, items, 40)
    m = Mirror(orc, users, items)
    new = rng.standard_normal((120, items.shape[1]), dtype=np.float32)
    idx.insert_items(new)
    m.insert_items(new)
    idx.delete_items(np.array([1, 4, 9, 2100], np.int64))
    m.delete_items([1, 4, 9, 2100])
    nu = rng.standard_normal((300, items.shape[1]), dtype=np.float32)
    idx.insert_users(nu)
    m.insert_users(nu)
    cur = m.items()
    queries = cur[rng.integers(0, len(cur), 650)]
    idx.set_option(3, 200)  # RKR_OPT_QUERY_BATCH: four device batches
    res = m.check(idx, queries, 20, form)
    idx.set_option(10, 1)   # RKR_OPT_QUERY_SLOTS: one stream
    one = idx.query(queries, 20)
    assert np.array_equal(res.users, one.users) and np.array_equal(res.ranks, one.ranks)
    for bad in (0, 3):
        with pytest.raises(Exception):
            idx.set_option(10, bad)


@pytest.mark.parametrize("dataset,utype", [("fake-normal-30d-update-item", "data_item"),
                                           ("fake-normal-30d-update-user", "user")])
def test_cli_rriu_reference_update_fixtures(orc, tmp_path, dataset, utype):
    """The reference's update fixtures (dataset/fake-normal-30d-update-{item,user}: 400 items or 100
    users to insert) through bsibc and rriu (ReverseRankIndexUpdate.cpp:81-227): for updateID 0, 2 and
    4, inserts of the first updateID quarters and deletes of as many vectors from the front, the
    answers written to the result CSVs equal brute force over the updated sets."""
    import os
    import subprocess
    from conftest import GOLDEN, ROOT
    from rkr import formats
    _case(f"cli rriu {dataset}")
    cli = os.path.join(ROOT, "reverse-k-ranks_amd", "rkr_cli")
    idir, rdir = str(tmp_path / "index"), str(tmp_path / "result")
    common = ["--dataset_dir", GOLDEN, "--dataset_name", dataset, "--index_dir", idir, "--n_sample", "128"]
    r = subprocess.run([cli, "bsibc"] + common, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    users, items, queries = formats.read_dataset(GOLDEN, dataset)
    upd = formats.read_fvecs(os.path.join(GOLDEN, dataset, f"{dataset}_{utype}_update.fvecs"))
    un = orc.normalize_users(users)
    for op in ("insert", "delete"):
        for uid in (0, 2, 4):
            if op == "delete" and uid == 0:
                continue
            r = subprocess.run([cli, "rriu"] + common + ["--topk", "10", "--result_dir", rdir, "--update_type", utype,
                                                          "--update_operator", op, "--updateID", str(uid)],
                               capture_output=True, text=True, timeout=100)
            assert r.returncode == 0, (op, uid, r.stderr)
            n = len(upd) // 4 * uid
            cur_u, cur_i = un, items
            if utype == "user":
                cur_u = np.concatenate([un, orc.normalize_users(upd[:n])]) if op == "insert" else un[n:]
            else:
                cur_i = np.concatenate([items, upd[:n]]) if op == "insert" else items[n:]
            stem = os.path.join(rdir, "rank", f"{dataset}-US-top10-n_sample_128-n_thread_-1-updateID_{uid}-{utype}-{op}")
            got_u = np.loadtxt(f"{stem}-userID.csv", delimiter=",", dtype=np.int64)
            got_r = np.loadtxt(f"{stem}-rank.csv", delimiter=",", dtype=np.int64)
            bu, br, _ = orc.bruteforce(cur_u, cur_i, queries, 10)
            assert np.array_equal(got_u, bu) and np.array_equal(got_r, br), (op, uid)
