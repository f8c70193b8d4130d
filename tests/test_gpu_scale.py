"""GPU parity at the full sizes of the benchmark configurations.

Each configuration builds its table on the device and answers its own query
workload (synthetic, seeded: `rkr.synth`).  The checks, all against the
oracle on the CPU:
  - every answer has k distinct users ordered by (rank, userID);
  - each checked (query, user) pair has the rank RefineCPU::RefineRank gives
    (RefineCPU.hpp:33-55, the strict count over every item) and the <u,q> of
    the reference's GEMV loop (US.hpp:124-136): the k-th user of every query
    and every user of a sample of queries;
  - for a few queries the whole answer equals the oracle's exact query through
    the device-built sampled table (US.hpp:96-213 semantics with exact ranks);
  - 256 random rows of the device-built sampled table equal the oracle's rows
    (BuildSampleIndexByComputation.cpp:111-144 gathered per SampleSearch.hpp:80-85);
  - the other forms and options (plain FP32, the sampled table without the
    query table, deduplication off, one query at a time) return the same answers.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _check_order(res, ks):
    """k distinct users per query, ordered by (rank, userID), -1 past k."""
    for i, k in enumerate(ks):
        u, r = res.users[i, :k].astype(np.int64), res.ranks[i, :k].astype(np.int64)
        assert np.all(u >= 0) and np.all(r >= 0), f"query {i}: missing entries"
        assert len(np.unique(u)) == k, f"query {i}: repeated users"
        key = r * (1 << 32) + u
        assert np.all(np.diff(key) > 0), f"query {i}: not ordered by (rank, userID)"
        assert np.all(res.users[i, k:] == -1)


def _check_pairs(orc, un, items, queries, res, ks, sample_q, rng):
    """Ranks and <u,q> of the k-th user of every query and of every user of
    sample_q random queries, against the oracle's strict count."""
    qid, uid, pos = [], [], []
    for i, k in enumerate(ks):
        qid.append(i), uid.append(int(res.users[i, k - 1])), pos.append(k - 1)
    for i in rng.choice(len(ks), min(sample_q, len(ks)), replace=False):
        for j in range(int(ks[i]) - 1):
            qid.append(int(i)), uid.append(int(res.users[i, j])), pos.append(j)
    qid, uid, pos = np.array(qid), np.array(uid), np.array(pos)
    r, q = orc.refine_ranks(un, queries, items, uid, qid)
    bad = np.flatnonzero(r != res.ranks[qid, pos])
    assert len(bad) == 0, f"{len(bad)} ranks differ, first (query, user) {qid[bad[0]]}, {uid[bad[0]]}"
    assert np.array_equal(q, res.qips[qid, pos]), "<u,q> differs from the sequential rule"
    return len(qid)


def _check_table_rows(orc, idx, un, items, ranks, rng, n=256):
    rows = np.sort(rng.choice(un.shape[0], n, replace=False))
    T = np.concatenate([idx.get_table(int(r), 1) for r in rows])
    assert np.array_equal(T, orc.build_table(un[rows], items, ranks)), "table rows differ from the oracle"


def _check_exact(orc, idx, un, items, ranks, queries, res, ks, which):
    T = idx.get_table()
    for i in which:
        k = int(ks[i])
        ou, orank, oq = orc.exact_query(un, items, T, ranks, queries[i:i + 1], k)
        assert np.array_equal(res.users[i, :k], ou[0]), f"query {i}: users differ from the oracle"
        assert np.array_equal(res.ranks[i, :k], orank[0]), f"query {i}: ranks differ from the oracle"
        assert np.array_equal(res.qips[i, :k], oq[0]), f"query {i}: <u,q> differ from the oracle"


def _same(a, b):
    return np.array_equal(a.users, b.users) and np.array_equal(a.ranks, b.ranks) and np.array_equal(a.qips, b.qips)


def _index(wl, form="screened", dense_top=2):
    import rkr
    idx = rkr.Index(0)
    idx.set_form(form)
    idx.set_option(8, dense_top)  # RKR_OPT_DENSE_TOP
    idx.set_data(wl.users, wl.items)
    ranks = rkr.Index.uniform_sample_ranks(wl.items.shape[0], wl.n_sample)
    idx.build_table(ranks)
    return idx, ranks


@pytest.mark.timeout(900)
@pytest.mark.parametrize("k", [10, 50, 100])
def test_plain_gpu_100k_x_20k(orc, k):
    """100k users x 20k items x d=150 (lognormal item norms), S=512, 1000
    queries uniform over the items."""
    from rkr import synth
    _case(f"scale: 100k x 20k d=150 S=512, 1000 queries, k={k}")
    wl = synth.make("plain_gpu", seed=0)
    idx, ranks = _index(wl)
    try:
        un = orc.normalize_users(wl.users)
        ks = np.full(len(wl.queries), k, np.int32)
        res = idx.query(wl.queries, ks)
        _check_order(res, ks)
        rng = np.random.default_rng(100 + k)
        n = _check_pairs(orc, un, wl.items, wl.queries, res, ks, 20, rng)
        _case(f"  {n} (query, user) ranks equal the oracle's strict count")
        _check_exact(orc, idx, un, wl.items, ranks, wl.queries, res, ks, rng.choice(len(ks), 5, replace=False))
        _check_table_rows(orc, idx, un, wl.items, ranks, rng)
        idx.set_form("plain")
        assert _same(idx.query(wl.queries, ks), res), "plain FP32 form differs"
        _case("  plain FP32 form equal; oracle exact answers equal; table rows equal")
    finally:
        idx.close()


@pytest.fixture(scope="module")
def one_m():
    """1M users x 50k items x d=150, S=512 (the precision and batching
    configurations share users and items)."""
    from rkr import synth
    import oracle
    _case("scale: building 1M x 50k d=150 S=512 (query table: every rank to 4096 columns)")
    wl = synth.make("precision", seed=0)
    idx, ranks = _index(wl)
    un = oracle.normalize_users(wl.users)
    yield wl, idx, ranks, un
    idx.close()


@pytest.mark.timeout(900)
def test_precision_1m_x_50k_k50(orc, one_m):
    """1000 queries, k=50, BF16-screened form with exact recheck; then the
    plain FP32 form, the sampled table alone (S=512 columns, guard columns to
    the deepest) and the plain form on it."""
    wl, idx, ranks, un = one_m
    ks = np.full(len(wl.queries), 50, np.int32)
    res = idx.query(wl.queries, ks)
    t = res.timing
    _case(f"  screened: rechecked {t['n_rechecked']} of {t['n_screened']}, listed {t['n_listed']}, "
          f"second pass {t['n_coarse_refined']}, fallback {t['n_fallback']}")
    _check_order(res, ks)
    rng = np.random.default_rng(7)
    n = _check_pairs(orc, un, wl.items, wl.queries, res, ks, 30, rng)
    _case(f"  {n} (query, user) ranks equal the oracle's strict count")
    _check_exact(orc, idx, un, wl.items, ranks, wl.queries, res, ks, rng.choice(len(ks), 4, replace=False))
    _check_table_rows(orc, idx, un, wl.items, ranks, rng)
    _case("  oracle exact answers equal; table rows equal")
    idx.set_form("plain")
    assert _same(idx.query(wl.queries, ks), res), "plain FP32 form differs"
    idx.set_option(8, 0)
    idx.build_table(ranks)
    try:
        assert idx.query_table_cols() == 0
        _check_table_rows(orc, idx, un, wl.items, ranks, rng)
        idx.set_form("screened")
        s0 = idx.query(wl.queries, ks)
        assert _same(s0, res), "screened form on the sampled table differs"
        idx.set_form("plain")
        p0 = idx.query(wl.queries[:300], ks[:300])
        assert np.array_equal(p0.users, s0.users[:300]) and np.array_equal(p0.ranks, s0.ranks[:300]), \
            "plain form on the sampled table differs"
        _case("  plain FP32 and the sampled-table-only forms equal")
    finally:
        idx.set_form("screened")
        idx.set_option(8, 2)
        idx.build_table(ranks)


@pytest.mark.timeout(900)
def test_batching_1m_x_50k_zipf(orc, one_m):
    """10,000 queries drawn Zipf(1.1) over the items sorted by norm, k uniform
    in [10, 200], one call (device batches, two query streams, byte-equal
    queries answered once): equal to deduplication off and to one query at a
    time."""
    from rkr import synth
    wl0, idx, ranks, un = one_m
    wl = synth.make("batching", seed=0)
    assert np.array_equal(wl.items, wl0.items)
    ks = wl.topk
    n_distinct = len(np.unique(wl.queries, axis=0))
    _case(f"scale: 10000 Zipf queries ({n_distinct} distinct), k in [10, 200]")
    res = idx.query(wl.queries, ks, max_k=200)
    _check_order(res, ks)
    rng = np.random.default_rng(8)
    n = _check_pairs(orc, un, wl.items, wl.queries, res, ks, 12, rng)
    _case(f"  {n} (query, user) ranks equal the oracle's strict count")
    # the most frequent vector with its largest and smallest k
    vals, inv, cnt = np.unique(wl.queries, axis=0, return_inverse=True, return_counts=True)
    top = np.flatnonzero(inv.ravel() == np.argmax(cnt))
    which = [top[np.argmax(ks[top])], top[np.argmin(ks[top])]] + list(rng.choice(len(ks), 2, replace=False))
    _check_exact(orc, idx, un, wl.items, ranks, wl.queries, res, ks, which)
    _case("  oracle exact answers equal")
    idx.set_option(6, 0)  # RKR_OPT_DEDUP off
    try:
        assert _same(idx.query(wl.queries, ks, max_k=200), res), "deduplication changes answers"
    finally:
        idx.set_option(6, 1)
    for i in rng.choice(len(ks), 48, replace=False):
        one = idx.query(wl.queries[i:i + 1], ks[i:i + 1], max_k=200)
        assert np.array_equal(one.users[0], res.users[i]) and np.array_equal(one.ranks[0], res.ranks[i]), \
            f"query {i}: alone differs from the batch"
    _case("  dedup off and one-at-a-time answers equal")
