"""GPU parity: librkr.so (HIP, gfx950) against the CPU oracle.

Integer outputs (user IDs, ranks) must be bit-exact; scores (table entries,
<u,q>) follow the reference's sequential FP32 InnerProduct rule and must be
bit-identical too (compared as floats, so -0.0 == +0.0).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _normalized(orc, users):
    return orc.normalize_users(users)


def test_table_fake30d_matches_oracle(orc, fake30d, gpu_index_factory):
    users, items, _ = fake30d
    _case("table fake-normal-30d S=128")
    idx, ranks = gpu_index_factory(users, items, 128)
    T = idx.get_table()
    T_ref = orc.build_table(_normalized(orc, users), items, ranks)
    assert T.shape == T_ref.shape
    assert np.array_equal(T, T_ref)


@pytest.mark.parametrize("U,I,d,S,seed", [
    (300, 1000, 30, 16, 1),      # small
    (257, 4099, 7, 100, 2),      # ragged: I, d not multiples of 4/128
    (130, 20000, 150, 512, 3),   # plain-GPU row width
    (64, 50001, 33, 512, 4),     # row larger than one LDS gather pass
    (64, 50000, 33, 512, 8),     # 50k-value rows held in registers by the selection
    (48, 60000, 20, 512, 9),     # aligned rows too long for registers: read from memory
    (16, 65536, 12, 256, 10),    # the largest rows whose target ranks the selection keeps in LDS
    (16, 70000, 12, 256, 11),    # ranks read from memory
    (5, 3, 4, 2, 5),             # tiny
    (20000, 2000, 8, 512, 6),    # 41 MB table: two 32 MB host staging chunks, the second partial
])
def test_table_random_matches_oracle(orc, gpu_index_factory, U, I, d, S, seed, tmp_path):
    _case(f"table U={U} I={I} d={d} S={S}")
    from rkr import formats
    rng = np.random.default_rng(seed)
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32)
    idx, ranks = gpu_index_factory(users, items, S)
    T = idx.get_table()
    T_ref = orc.build_table(_normalized(orc, users), items, ranks)
    assert np.array_equal(T, T_ref)
    p = str(tmp_path / "t.index")
    idx.save_table(p)
    f_ranks, f_table, f_items = formats.read_table(p)
    assert f_items == I and np.array_equal(f_ranks, ranks) and np.array_equal(f_table, T)


def test_table_ties_and_degenerate_rows(orc, gpu_index_factory):
    """Duplicate items (score ties), a zero user (all scores 0), a one-hot user
    (mostly equal scores) and heavy value clustering."""
    _case("table ties")
    rng = np.random.default_rng(7)
    d, I = 12, 3000
    base = rng.standard_normal((40, d), dtype=np.float32)
    items = base[rng.integers(0, 40, I)]                     # only 40 distinct items
    items[:500] = 0.0                                        # many zero scores
    users = rng.standard_normal((50, d), dtype=np.float32)
    users[0] = 0.0                                           # zero user: norm <= 0.001, left unnormalised
    users[1] = 0.0
    users[1, 3] = 1.0                                        # one-hot
    idx, ranks = gpu_index_factory(users, items, 300)
    T = idx.get_table()
    T_ref = orc.build_table(_normalized(orc, users), items, ranks)
    assert np.array_equal(T, T_ref)


def _check_exact(orc, users, items, queries, res, ks):
    ub, rb, qb = orc.bruteforce(users, items, queries, int(max(ks)))
    for i, k in enumerate(ks):
        assert np.array_equal(res.users[i, :k], ub[i, :k]), f"query {i}: users differ"
        assert np.array_equal(res.ranks[i, :k], rb[i, :k]), f"query {i}: ranks differ"
        assert np.array_equal(res.qips[i, :k], qb[i, :k]), f"query {i}: <u,q> differ"


def test_query_exact_fake30d(orc, fake30d, gpu_index_factory):
    users, items, queries = fake30d
    _case("query exact fake-normal-30d k=10")
    idx, ranks = gpu_index_factory(users, items, 128)
    res = idx.query(queries, 10)
    _check_exact(orc, _normalized(orc, users), items, queries, res, [10] * len(queries))


def test_query_reference_output_fake30d(orc, fake30d, gpu_index_factory):
    """RKR_OUTPUT_REFERENCE reproduces US::Index::Retrieval's answers."""
    users, items, queries = fake30d
    _case("query reference-output fake-normal-30d k=10")
    idx, ranks = gpu_index_factory(users, items, 128)
    idx.set_output("reference")
    res = idx.query(queries, 10)
    un = _normalized(orc, users)
    T = orc.build_table(un, items, ranks)
    ou, orank, oq = orc.us_retrieval(un, items, T, ranks, queries, 10)
    assert np.array_equal(res.users, ou)
    assert np.array_equal(res.ranks, orank)
    assert np.array_equal(res.qips, oq)


def test_query_reference_output_dedup_mixed_k(orc, fake30d, gpu_index_factory):
    """Reference output with byte-equal queries of different k: the users the
    bounds settle depend on k, so each (vector, k) must be answered as the
    reference answers it alone (US.hpp:173-186), with deduplication on."""
    users, items, queries = fake30d
    _case("query reference-output duplicates with mixed k")
    idx, ranks = gpu_index_factory(users, items, 128)
    idx.set_output("reference")
    pick = np.array([0, 1, 0, 2, 1, 0, 2, 0, 1, 1])
    ks = np.array([3, 10, 25, 7, 40, 3, 60, 11, 10, 2], np.int32)
    res = idx.query(queries[pick], ks, max_k=60)
    un = _normalized(orc, users)
    T = orc.build_table(un, items, ranks)
    for i, (qi, k) in enumerate(zip(pick, ks)):
        ou, orank, oq = orc.us_retrieval(un, items, T, ranks, queries[qi:qi + 1], int(k))
        assert np.array_equal(res.users[i, :k], ou[0]), (i, k)
        assert np.array_equal(res.ranks[i, :k], orank[0]), (i, k)
        assert np.array_equal(res.qips[i, :k], oq[0]), (i, k)
        assert np.all(res.users[i, k:] == -1)


@pytest.mark.parametrize("nq", [1, 3, 16, 17, 100])
def test_query_batch_shapes_agree(orc, fake30d, gpu_index_factory, nq):
    """The GEMV (<= 16 queries) and GEMM paths give identical answers."""
    users, items, queries = fake30d
    _case(f"query batch nq={nq}")
    idx, _ = gpu_index_factory(users, items, 128)
    res = idx.query(queries[:nq], 25)
    _check_exact(orc, _normalized(orc, users), items, queries[:nq], res, [25] * nq)


def test_query_mixed_k_and_query_batches(orc, gpu_index_factory):
    _case("query mixed k, device batches of 7")
    rng = np.random.default_rng(11)
    U, I, d = 3000, 2500, 24
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32)
    queries = items[rng.choice(I, 30, replace=False)]
    ks = rng.integers(1, 300, 30).astype(np.int32)
    idx, _ = gpu_index_factory(users, items, 64)
    idx.set_option(3, 7)  # RKR_OPT_QUERY_BATCH
    res = idx.query(queries, ks, max_k=300)
    _check_exact(orc, _normalized(orc, users), items, queries, res, ks)
    for i, k in enumerate(ks):
        assert np.all(res.users[i, k:] == -1) and np.all(res.ranks[i, k:] == -1)


def test_query_duplicates_dedup(orc, gpu_index_factory):
    """Byte-equal queries with different k are answered once (largest k) and
    sliced; the answers equal the ones computed without deduplication and the
    brute-force oracle.  U spans several prune chunks."""
    _case("query duplicates, dedup on/off")
    rng = np.random.default_rng(15)
    U, I, d = 10000, 1500, 20
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32)
    pick = rng.integers(0, 5, 40)  # 40 queries over 5 distinct vectors
    queries = items[pick]
    ks = rng.integers(1, 150, 40).astype(np.int32)
    idx, _ = gpu_index_factory(users, items, 48)
    on = idx.query(queries, ks, max_k=150)
    idx.set_option(6, 0)  # RKR_OPT_DEDUP off
    off = idx.query(queries, ks, max_k=150)
    assert np.array_equal(on.users, off.users) and np.array_equal(on.ranks, off.ranks)
    assert np.array_equal(on.qips, off.qips)
    assert [s["query_id"] for s in on.stats] == list(range(40))
    _check_exact(orc, _normalized(orc, users), items, queries, on, ks)


@pytest.mark.parametrize("dense", [0, 1])
def test_query_many_candidates(orc, gpu_index_factory, dense):
    """Coarse table (S=4) so thousands of users survive pruning: exercises the
    radix-select top-k path (> 4096 candidates) and large refine batches.
    With the dense query table the same queries refine far fewer users."""
    _case(f"query many candidates dense={dense}")
    rng = np.random.default_rng(12)
    U, I, d = 20000, 600, 16
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32)
    queries = items[:4]
    idx, _ = gpu_index_factory(users, items, 4, dense_top=dense)
    res = idx.query(queries, 200)
    if not dense:
        assert max(s["n_candidate"] for s in res.stats) > 4096
    _check_exact(orc, _normalized(orc, users), items, queries, res, [200] * 4)


def test_query_ties(orc, gpu_index_factory):
    """Queries equal to items, duplicate items and zero users: the strict
    count and (rank, userID) order decide every tie."""
    _case("query ties")
    rng = np.random.default_rng(13)
    d, I, U = 10, 800, 600
    base = rng.standard_normal((60, d), dtype=np.float32)
    items = base[rng.integers(0, 60, I)]
    users = rng.standard_normal((U, d), dtype=np.float32)
    users[:5] = 0.0
    queries = np.concatenate([items[:6], np.zeros((1, d), np.float32)])
    idx, _ = gpu_index_factory(users, items, 50)
    res = idx.query(queries, 40)
    _check_exact(orc, _normalized(orc, users), items, queries, res, [40] * len(queries))


def test_query_k_equals_n_user(orc, gpu_index_factory):
    _case("query k = n_user")
    rng = np.random.default_rng(14)
    users = rng.standard_normal((150, 8), dtype=np.float32)
    items = rng.standard_normal((400, 8), dtype=np.float32)
    idx, _ = gpu_index_factory(users, items, 20)
    res = idx.query(items[:3], 150)
    _check_exact(orc, _normalized(orc, users), items, items[:3], res, [150] * 3)


def test_loaded_table_with_eps(orc, tmp_path, gpu_index_factory):
    """A table perturbed within eps and loaded with that eps still gives exact answers."""
    _case("query perturbed table with eps")
    rng = np.random.default_rng(15)
    U, I, d = 2000, 3000, 20
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32)
    idx, ranks = gpu_index_factory(users, items, 64)
    T = idx.get_table()
    eps = 1e-3
    Tp = np.sort(T + rng.uniform(-eps / 2, eps / 2, T.shape).astype(np.float32), axis=1)[:, ::-1].copy()
    from rkr import formats
    p = str(tmp_path / "t.index")
    formats.write_table(p, ranks, Tp, I)
    idx.load_table(p, eps=eps)
    queries = items[:20]
    res = idx.query(queries, 30)
    _check_exact(orc, _normalized(orc, users), items, queries, res, [30] * 20)


def test_save_load_roundtrip(tmp_path, gpu_index_factory):
    _case("table save/load")
    rng = np.random.default_rng(16)
    users = rng.standard_normal((500, 12), dtype=np.float32)
    items = rng.standard_normal((900, 12), dtype=np.float32)
    idx, ranks = gpu_index_factory(users, items, 32)
    T = idx.get_table()
    p = str(tmp_path / "US-x-n_sample_32.index")
    idx.save_table(p)
    from rkr import formats
    r2, T2, n_item = formats.read_table(p)
    assert n_item == 900 and np.array_equal(r2, ranks) and np.array_equal(T2, T)
    idx.load_table(p)
    assert np.array_equal(idx.get_table(), T)


@pytest.mark.parametrize("form,batch", [("screened", 300), ("plain", 300), ("screened", 150)])
def test_two_query_slots_agree(orc, gpu_index_factory, form, batch):
    """Calls of >= 512 queries run in two halves on two streams and host threads
    (RKR_OPT_QUERY_SLOTS = 2, the default): answers, per-query counters and query
    IDs equal the one-stream run and brute force, with duplicates, mixed k and
    device batches of 300 or 150 (the split falls on a batch boundary)."""
    _case(f"two query slots form={form} batch={batch}")
    rng = np.random.default_rng(31)
    U, I, d = 4000, 2500, 24
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32)
    queries = items[rng.integers(0, I, 700)]          # some byte-equal repeats
    ks = rng.integers(1, 120, 700).astype(np.int32)
    idx, _ = gpu_index_factory(users, items, 64)
    idx.set_form(form)
    idx.set_option(3, batch)  # RKR_OPT_QUERY_BATCH
    two = idx.query(queries, ks, max_k=120)
    idx.set_option(10, 1)  # RKR_OPT_QUERY_SLOTS: one stream
    one = idx.query(queries, ks, max_k=120)
    assert np.array_equal(two.users, one.users) and np.array_equal(two.ranks, one.ranks)
    assert np.array_equal(two.qips, one.qips)
    assert [s["query_id"] for s in two.stats] == list(range(700))
    for key in ("n_candidate", "n_refine_user", "n_result_user", "n_prune_user"):
        assert [s[key] for s in two.stats] == [s[key] for s in one.stats], key
    _check_exact(orc, _normalized(orc, users), items, queries, two, ks)


# ---- full score table (ComputeSortItemsBatch rows, ReadAll's index file) --------

@pytest.mark.parametrize("U,I,d,seed", [
    (300, 1000, 30, 21),      # one LDS tile
    (64, 8192, 16, 22),       # exactly one tile
    (70, 8193, 9, 23),        # one tile and one value: a single merge pass
    (40, 20000, 150, 24),     # plain-GPU row width: two merge passes, a partial run
    (9, 50001, 33, 25),       # precision row width: three merge passes
    (3, 1, 4, 26),            # one item
    (50, 9999, 20, 31),       # an odd row width: the bucket path's 4-byte loads, one row per workgroup
    (1000, 20000, 8, 28),     # 80 MB of rows: three 32 MB host staging chunks, the last one partial
])
def test_sorted_scores_match_oracle(orc, U, I, d, seed, tmp_path):
    _case(f"sorted scores U={U} I={I} d={d}")
    import rkr
    from rkr import formats
    rng = np.random.default_rng(seed)
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32)
    idx = rkr.Index(0)
    idx.set_data(users, items)
    un = _normalized(orc, users)
    got = idx.sorted_scores()
    assert got.shape == (U, I)
    assert np.array_equal(got, orc.sorted_score_rows(un, items))
    u0, n = U // 3, max(1, U // 4)
    assert np.array_equal(idx.sorted_scores(u0, n), got[u0:u0 + n])
    p = formats.score_table_path(str(tmp_path), "t")
    idx.save_score_table(p)
    assert np.array_equal(formats.read_score_table(p, I, mmap=False), got)


def test_sorted_scores_ties_batches_and_table(orc, tmp_path):
    """Tied and zero scores, a zero user, rows split over several device batches;
    the sampled table is the sorted rows at the sample ranks; the file holds the rows."""
    _case("sorted scores ties + file")
    import rkr
    from rkr import formats
    rng = np.random.default_rng(27)
    d, I, U = 12, 9000, 37
    base = rng.standard_normal((50, d), dtype=np.float32)
    items = base[rng.integers(0, 50, I)]
    items[:700] = 0.0
    users = rng.standard_normal((U, d), dtype=np.float32)
    users[0] = 0.0
    idx = rkr.Index(0)
    idx.set_option(4, 8)  # RKR_OPT_BUILD_BATCH: 8 users per device batch
    idx.set_data(users, items)
    got = idx.sorted_scores()
    assert np.array_equal(got, orc.sorted_score_rows(_normalized(orc, users), items))
    ranks = rkr.Index.uniform_sample_ranks(I, 300)
    idx.build_table(ranks)
    assert np.array_equal(idx.get_table(), got[:, ranks])
    p = formats.score_table_path(str(tmp_path), "ties")
    idx.save_score_table(p)
    assert os.path.getsize(p) == U * I * 4
    assert np.array_equal(formats.read_score_table(p, I), got)
    assert np.array_equal(formats.read_score_rows(p, I, 5, 100, 30), got[5, 100:130])
    idx.save_score_table(p, 10)
    assert np.array_equal(formats.read_score_table(p, I, mmap=False), got[:10])


@pytest.mark.parametrize("case", ["moderate_ties", "outlier", "lds_max", "lds_max_plus_one"])
def test_sorted_scores_row_shapes(orc, case):
    """Rows the LDS bucket sort takes (ties of a few dozen copies, up to 28672
    values) and rows it hands to the merge path (an outlier that crowds every
    other value into one bin, 28673 values) sort to the oracle's rows."""
    _case(f"sorted scores {case}")
    import rkr
    rng = np.random.default_rng(29)
    d = 10
    if case == "moderate_ties":
        base = rng.standard_normal((400, d), dtype=np.float32)
        items = base[rng.integers(0, 400, 8000)]
    elif case == "outlier":
        items = rng.standard_normal((6000, d), dtype=np.float32)
        items[17] *= np.float32(1e5)
    else:
        items = rng.standard_normal((28672 if case == "lds_max" else 28673, d), dtype=np.float32)
    users = rng.standard_normal((12, d), dtype=np.float32)
    idx = rkr.Index(0)
    try:
        idx.set_data(users, items)
        got = idx.sorted_scores()
        assert np.array_equal(got, orc.sorted_score_rows(_normalized(orc, users), items))
    finally:
        idx.close()


def test_sorted_scores_follow_updates(orc):
    """After item inserts/deletes and user deletes the rows cover the current sets."""
    _case("sorted scores after updates")
    import rkr
    rng = np.random.default_rng(28)
    users = rng.standard_normal((200, 16), dtype=np.float32)
    items = rng.standard_normal((3000, 16), dtype=np.float32)
    new = rng.standard_normal((120, 16), dtype=np.float32)
    idx = rkr.Index(0)
    idx.set_data(users, items)
    idx.build_table(rkr.Index.uniform_sample_ranks(3000, 64))
    idx.insert_items(new)
    idx.delete_items(np.array([0, 5, 2999, 3001], np.int64))
    idx.delete_users(np.array([3, 150], np.int64))
    cur_items = np.concatenate([items, new])
    cur_items = np.delete(cur_items, [0, 5, 2999, 3001], axis=0)
    cur_users = np.delete(_normalized(orc, users), [3, 150], axis=0)
    got = idx.sorted_scores()
    assert got.shape == (198, 3116)
    assert np.array_equal(got, orc.sorted_score_rows(cur_users, cur_items))


# ---- screened form: BF16 MFMA screen + exact recheck -------------------------

def _screened(idx):
    idx.set_form("screened")
    return idx


def test_screened_fake30d_matches_bruteforce(orc, fake30d, gpu_index_factory):
    users, items, queries = fake30d
    _case("screened fake-normal-30d k=10")
    idx, _ = gpu_index_factory(users, items, 128, dense_top=0)  # sampled table only: users get refined
    res = _screened(idx).query(queries, 10)
    _check_exact(orc, _normalized(orc, users), items, queries, res, [10] * len(queries))
    # queries are not items here, so only near-threshold items are rechecked
    assert res.timing["n_screened"] > 0
    assert res.timing["n_rechecked"] < 0.01 * res.timing["n_screened"]


@pytest.mark.parametrize("nq,k,dense", [(1, 50, 0), (16, 7, 0), (40, 120, 0), (40, 120, 1)])
def test_screened_equals_plain(orc, gpu_index_factory, nq, k, dense):
    _case(f"screened vs plain nq={nq} k={k} dense={dense}")
    rng = np.random.default_rng(21 + nq)
    U, I, d = 6000, 3000, 150
    users = rng.standard_normal((U, d), dtype=np.float32) / np.float32(np.sqrt(d))
    items = (rng.standard_normal((I, d), dtype=np.float32) / np.float32(np.sqrt(d))
             * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32))
    queries = items[rng.choice(I, nq, replace=False)]
    idx, _ = gpu_index_factory(users, items, 64, dense_top=dense)
    plain = idx.query(queries, k)
    scr = _screened(idx).query(queries, k)
    assert np.array_equal(plain.users, scr.users) and np.array_equal(plain.ranks, scr.ranks)
    _check_exact(orc, _normalized(orc, users), items, queries, scr, [k] * nq)
    if dense:
        return  # the statistics below need refined users
    # q is an item: <u,q> ties with that item's own score, so it must be rechecked
    assert scr.timing["n_rechecked"] >= plain.timing["n_refined"]
    # the observed screening error stays inside the analytic bound's scale
    assert 0 < scr.timing["max_screen_err"] < 1e-3


_PRUNE_FIELDS = ("n_candidate", "n_refine_user", "n_prune_user", "n_result_user")


def _same_answers(a, b, ks, stats=True):
    for i, k in enumerate(ks):
        assert np.array_equal(a.users[i, :k], b.users[i, :k]), f"query {i}: users differ"
        assert np.array_equal(a.ranks[i, :k], b.ranks[i, :k]), f"query {i}: ranks differ"
        assert np.array_equal(a.qips[i, :k], b.qips[i, :k]), f"query {i}: <u,q> differ"
        if stats:
            for f in _PRUNE_FIELDS:
                assert a.stats[i][f] == b.stats[i][f], f"query {i}: {f} {a.stats[i][f]} != {b.stats[i][f]}"


@pytest.mark.parametrize("d,dense,nq", [(150, 1, 70), (150, 0, 70), (37, 1, 70), (20, 0, 33), (150, 1, 1),
                                        (64, 1, 8), (64, 0, 9)])
def test_coarse_bounds(orc, gpu_index_factory, d, dense, nq):
    """Coarse bounds (screened <u,q> on BF16 matrix cores against guard
    columns, exact scores and searches for the listed users only) keep the
    full path's candidates, refinements and answers, bit for bit."""
    _case(f"coarse bounds d={d} dense={dense} nq={nq}")
    rng = np.random.default_rng(90 + d + dense + nq)
    U, I = 5003, 2100
    users = rng.standard_normal((U, d), dtype=np.float32) * rng.lognormal(0, 0.3, (U, 1)).astype(np.float32)
    items = (rng.standard_normal((I, d), dtype=np.float32) / np.float32(np.sqrt(d))
             * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32))
    # half the queries are items (ties with their own score), half are not
    queries = np.concatenate([items[rng.choice(I, (nq + 1) // 2, replace=False)],
                              rng.standard_normal((nq // 2, d), dtype=np.float32) / np.float32(np.sqrt(d))])
    idx, _ = gpu_index_factory(users, items, 64, dense_top=dense)
    ks = [int(k) for k in rng.integers(1, 150, nq)]
    plain = idx.query(queries, ks)
    idx.set_form("screened")
    idx.set_option(9, 0)  # RKR_OPT_COARSE
    full = idx.query(queries, ks)
    idx.set_option(9, 1)
    coarse = idx.query(queries, ks)
    _same_answers(full, coarse, ks)
    _same_answers(plain, coarse, ks)
    _check_exact(orc, _normalized(orc, users), items, queries, coarse, ks)


@pytest.mark.parametrize("dense", [1, 0])
def test_coarse_bounds_second_pass(orc, gpu_index_factory, dense):
    """Long coarse lists get a second pass against guard columns inside the
    bin of their k-th bound (one launch, a guard set per query); answers and
    prune counts stay those of the full path."""
    _case(f"coarse bounds second pass dense={dense}")
    rng = np.random.default_rng(31 + dense)
    U, I, d, nq = 40000, 3000, 32, 40
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = (rng.standard_normal((I, d), dtype=np.float32) / np.float32(np.sqrt(d))
             * rng.lognormal(0, 0.4, (I, 1)).astype(np.float32))
    queries = np.concatenate([items[rng.choice(I, nq // 2, replace=False)],
                              rng.standard_normal((nq // 2, d), dtype=np.float32) / np.float32(np.sqrt(d))])
    ks = [int(k) for k in rng.integers(50, 400, nq)]
    idx, _ = gpu_index_factory(users, items, 128, dense_top=dense)
    idx.set_form("screened")
    idx.set_option(9, 0)
    full = idx.query(queries, ks)
    idx.set_option(9, 1)
    coarse = idx.query(queries, ks)
    assert coarse.timing["n_coarse_refined"] > 0
    _same_answers(full, coarse, ks)
    _check_exact(orc, _normalized(orc, users), items, queries, coarse, ks)


def test_coarse_bounds_fallback(orc, gpu_index_factory):
    """Queries the guards cannot narrow (k beyond every guard's reach, or a
    list longer than the cap) take the full path inside the same call."""
    _case("coarse bounds fallback")
    rng = np.random.default_rng(7)
    U, I, d = 5003, 2100, 48
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) / np.float32(np.sqrt(d))
    # negated items rank near the bottom for most users; k up to 4000
    queries = np.concatenate([-items[:6], items[6:12], rng.standard_normal((6, d), dtype=np.float32)])
    ks = [4000, 3000, 2500, 50, 10, 1, 4000, 3500, 200, 100, 20, 5, 4000, 2000, 900, 30, 3, 1]
    idx, _ = gpu_index_factory(users, items, 64)
    plain = idx.query(queries, ks)
    idx.set_form("screened")
    idx.set_option(9, 0)
    full = idx.query(queries, ks)
    idx.set_option(9, 1)
    coarse = idx.query(queries, ks)
    _same_answers(full, coarse, ks)
    _same_answers(plain, coarse, ks)
    _check_exact(orc, _normalized(orc, users), items, queries, coarse, ks)
    # positive users and items, negative queries: every query ranks last for
    # every user, below all guards, so no coarse k-th bound exists
    users, items = np.abs(users[:800]), np.abs(items[:300])
    queries = -np.abs(rng.standard_normal((20, d), dtype=np.float32))
    ks = [int(k) for k in rng.integers(1, 60, 20)]
    idx, _ = gpu_index_factory(users, items, 32)
    idx.set_form("screened")
    coarse = idx.query(queries, ks)
    _check_exact(orc, _normalized(orc, users), items, queries, coarse, ks)
    assert all(coarse.ranks[i, 0] == 300 for i in range(20))


@pytest.mark.parametrize("d", [150, 64, 20])
def test_screened_64_rows_per_wave(orc, gpu_index_factory, d):
    """The 64-rows-per-wavefront screen tile gives the same answers as the
    32-row tile and the plain form (several row blocks, ragged last block)."""
    _case(f"screened 64 rows per wave d={d}")
    rng = np.random.default_rng(40 + d)
    U, I = 7000, 2600
    users = rng.standard_normal((U, d), dtype=np.float32) / np.float32(np.sqrt(d))
    items = (rng.standard_normal((I, d), dtype=np.float32) / np.float32(np.sqrt(d))
             * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32))
    queries = items[rng.choice(I, 24, replace=False)]
    idx, _ = gpu_index_factory(users, items, 48)
    plain = idx.query(queries, 60)
    idx.set_form("screened")
    idx.set_option(7, 64)  # RKR_OPT_SCREEN_ROWS
    wide = idx.query(queries, 60)
    idx.set_option(7, 32)
    narrow = idx.query(queries, 60)
    for res in (wide, narrow):
        assert np.array_equal(plain.users, res.users) and np.array_equal(plain.ranks, res.ranks)
    _check_exact(orc, _normalized(orc, users), items, queries, wide, [60] * 24)


@pytest.mark.parametrize("form", ["plain", "screened"])
def test_dense_query_table(orc, form):
    """The query table (dense top ranks) leaves the sampled table and every
    answer unchanged and refines fewer users."""
    _case(f"dense query table ({form})")
    import rkr
    rng = np.random.default_rng(50)
    U, I, d = 20000, 3000, 24
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32)
    queries = items[rng.choice(I, 30, replace=False)]
    ks = rng.integers(1, 120, 30).astype(np.int32)
    ranks = rkr.Index.uniform_sample_ranks(I, 40)
    out = {}
    for dense in (0, 1, 2):
        idx = rkr.Index(0)
        idx.set_option(8, dense)  # RKR_OPT_DENSE_TOP
        idx.set_form(form)
        idx.set_data(users, items)
        idx.build_table(ranks)
        res = idx.query(queries, ks, max_k=120)
        out[dense] = (idx.get_table(), res)
        idx.close()
    refined = [sum(s["n_refine_user"] for s in out[x][1].stats) for x in (0, 1, 2)]
    for dense in (1, 2):
        assert np.array_equal(out[0][0], out[dense][0])
        assert np.array_equal(out[0][1].users, out[dense][1].users)
        assert np.array_equal(out[0][1].ranks, out[dense][1].ranks)
        assert refined[dense] < refined[0]
        _check_exact(orc, _normalized(orc, users), items, queries, out[dense][1], ks)
    # every rank up to the cap (here all 3000 items) is a column: nothing is refined
    assert refined[2] <= refined[1] and refined[2] == 0


@pytest.mark.parametrize("dense", [0, 1])
def test_screened_ties_and_many_candidates(orc, gpu_index_factory, dense):
    _case(f"screened ties + many candidates dense={dense}")
    rng = np.random.default_rng(23)
    d, I, U = 10, 800, 20000
    base = rng.standard_normal((60, d), dtype=np.float32)
    items = base[rng.integers(0, 60, I)]
    users = rng.standard_normal((U, d), dtype=np.float32)
    users[:5] = 0.0
    queries = np.concatenate([items[:6], np.zeros((1, d), np.float32)])
    idx, _ = gpu_index_factory(users, items, 8, dense_top=dense)
    res = _screened(idx).query(queries, 300)
    if not dense:
        assert max(s["n_candidate"] for s in res.stats) > 4096
    _check_exact(orc, _normalized(orc, users), items, queries, res, [300] * len(queries))


def test_screened_reference_output(orc, fake30d, gpu_index_factory):
    users, items, queries = fake30d
    _case("screened reference-output")
    idx, ranks = gpu_index_factory(users, items, 128)
    idx.set_output("reference")
    res = _screened(idx).query(queries, 10)
    un = _normalized(orc, users)
    T = orc.build_table(un, items, ranks)
    ou, orank, oq = orc.us_retrieval(un, items, T, ranks, queries, 10)
    assert np.array_equal(res.users, ou) and np.array_equal(res.ranks, orank)


def test_screened_large_magnitudes(orc, gpu_index_factory):
    """Unnormalised users and items with large and tiny magnitudes (denormal
    BF16 splits): the bound's absolute floor and scaling keep answers exact."""
    _case("screened magnitudes")
    rng = np.random.default_rng(24)
    U, I, d = 2000, 1500, 40
    users = rng.standard_normal((U, d), dtype=np.float32) * np.float32(1e3)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.choice(
        np.array([1e-20, 1e-3, 1.0, 1e4], np.float32), (I, 1))
    queries = items[:12]
    idx = __import__("rkr").Index(0)
    idx.set_data(users, items, normalize_users=False)
    ranks = idx.uniform_sample_ranks(I, 40)
    idx.build_table(ranks)
    res = _screened(idx).query(queries, 25)
    _check_exact(orc, users, items, queries, res, [25] * 12)


def test_cli_pipeline_fake30d(orc, fake30d, tmp_path):
    """fsr -> bsibc -> rri through the C++ driver, as the reference's run_rkr.py
    chains its binaries; the rank CSVs hold the exact answers."""
    import subprocess
    from conftest import GOLDEN, ROOT
    _case("cli fsr/bsibc/rri")
    cli = os.path.join(ROOT, "reverse-k-ranks_amd", "rkr_cli")
    common = ["--dataset_dir", GOLDEN, "--dataset_name", "fake-normal-30d", "--index_dir", str(tmp_path / "index"),
              "--method_name", "US", "--n_sample", "128"]
    for cmd in (["fsr"], ["bsibc"], ["rri", "--topk", "10", "--result_dir", str(tmp_path / "result")]):
        r = subprocess.run([cli] + cmd + common, capture_output=True, text=True, timeout=100)
        assert r.returncode == 0, r.stderr
    stem = tmp_path / "result" / "rank" / "fake-normal-30d-US-top10-n_sample_128-n_thread_-1"
    users = np.loadtxt(f"{stem}-userID.csv", delimiter=",", dtype=np.int64)
    ranks = np.loadtxt(f"{stem}-rank.csv", delimiter=",", dtype=np.int64)
    u, _, q = fake30d
    bu, br, _ = orc.bruteforce(orc.normalize_users(u), fake30d[1], q, 10)
    assert np.array_equal(users, bu) and np.array_equal(ranks, br)


def test_cli_bst_fake30d(orc, fake30d, tmp_path):
    """bst writes ReadAll's index file: every user's sorted scores, or the rows
    that fit eval_size_gb."""
    import subprocess
    from conftest import GOLDEN, ROOT
    from rkr import formats
    _case("cli bst")
    cli = os.path.join(ROOT, "reverse-k-ranks_amd", "rkr_cli")
    common = ["--dataset_dir", GOLDEN, "--dataset_name", "fake-normal-30d", "--index_dir", str(tmp_path / "index")]
    r = subprocess.run([cli, "bst"] + common, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    u, items, _ = fake30d
    rows = formats.read_score_table(formats.score_table_path(str(tmp_path / "index"), "fake-normal-30d"),
                                    items.shape[0])
    assert rows.shape == (u.shape[0], items.shape[0])
    assert np.array_equal(rows, orc.sorted_score_rows(orc.normalize_users(u), items))


def test_query_into_reused_outputs(orc, gpu_index_factory):
    """Index.query(..., out=previous) writes the same answers and stats into the
    previous call's arrays; a different batch shape allocates new ones."""
    _case("query outputs reused")
    rng = np.random.default_rng(77)
    U, I, d = 3000, 1500, 24
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32)
    idx, _ = gpu_index_factory(users, items, 64)
    q1, q2 = items[:40], items[40:80]
    ks = rng.integers(1, 60, 40).astype(np.int32)
    a = idx.query(q1, ks, max_k=60)
    first = a.users.copy(), a.ranks.copy()
    b = idx.query(q2, ks, max_k=60, out=a)
    assert b.users is a.users  # the same arrays, written over
    fresh = idx.query(q2, ks, max_k=60)
    assert np.array_equal(b.users, fresh.users) and np.array_equal(b.ranks, fresh.ranks)
    assert np.array_equal(b.stats["n_candidate"], fresh.stats["n_candidate"])
    c = idx.query(q1, ks, max_k=60, out=b)
    assert np.array_equal(c.users, first[0]) and np.array_equal(c.ranks, first[1])
    d = idx.query(q1[:5], ks[:5], max_k=60, out=c)
    assert d.users is not c.users and d.users.shape == (5, 60)
    assert np.array_equal(d.users, first[0][:5])
