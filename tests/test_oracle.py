"""Pins the CPU oracle (oracle/rkr_oracle.c) before it is trusted as the checker.

Each oracle function is checked against (a) the reference's own routines
compiled from its headers into oracle/_ref when they are built, (b) the
golden vectors in tests/golden/ (made from those routines by
tests/golden/make_golden.py, so they also pin the oracle where
/root/reference is absent), and (c) closed forms, brute force and invariants
chosen so that a dropped term, a flipped comparison or an off-by-one fails.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _ref(orc):
    r = orc.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref not built (reference sources absent); golden-vector tests cover it")
    return r


# ---- a1 normalisation (VectorMatrix.hpp:51-66) --------------------------------

def test_normalize_matches_reference_code(orc, fake30d):
    R = _ref(orc)
    users = fake30d[0]
    a = orc.normalize_users(users)
    b = users.copy()
    R.ref_normalize(b, b.shape[0], b.shape[1])
    assert np.array_equal(a.view(np.int32), b.view(np.int32))


def test_normalize_closed_form_and_small_norm_rows(orc):
    rng = np.random.default_rng(0)
    u = rng.standard_normal((50, 9), dtype=np.float32)
    u[3] = 1e-5            # norm 3e-5 <= 0.001: row left unchanged
    u[4] = 0.0
    n = orc.normalize_users(u)
    assert np.array_equal(n[3], u[3]) and np.array_equal(n[4], u[4])
    ok = np.r_[0:3, 5:50]
    np.testing.assert_allclose(np.linalg.norm(n[ok].astype(np.float64), axis=1), 1.0, rtol=1e-6)
    # sequential FP32 sum of squares, then division by the FP32 sqrt
    row = u[7]
    s = np.float32(0)
    for x in row:
        s = np.float32(s + np.float32(x * x))
    assert np.array_equal(n[7], row / np.sqrt(s))


# ---- score rule (SpaceInnerProduct.hpp:14-20) ----------------------------------

def test_ip_matches_reference_code_and_sequential_fp32(orc):
    R = _ref(orc)
    rng = np.random.default_rng(1)
    for d in (1, 3, 30, 150, 257):
        a = rng.standard_normal(d, dtype=np.float32)
        b = rng.standard_normal(d, dtype=np.float32) * np.float32(100)
        x = orc.ip_f32(a, b)
        assert x == np.float32(R.ref_inner_product(a, b, d))
        s = np.float32(0)
        for i in range(d):
            s = np.float32(s + np.float32(a[i] * b[i]))
        assert x == s  # no FMA, no reassociation


def test_ip_not_fused(orc):
    """A case where an FMA or a pairwise sum gives a different float."""
    a = np.array([1 + 2 ** -12, 1.0, -1.0], np.float32)
    b = np.array([1 + 2 ** -12, 2 ** -23, -1.0], np.float32)
    seq = np.float32(np.float32(np.float32(a[0] * b[0]) + np.float32(a[1] * b[1])) + np.float32(a[2] * b[2]))
    fused = np.float32(np.float64(a[0]) * b[0] + np.float64(a[1]) * b[1] - 1.0)
    assert seq != fused
    assert orc.ip_f32(a, b) == seq


# ---- a2 uniform sample ranks (UniformSample.hpp:17-25) -------------------------

@pytest.mark.parametrize("I,S", [(5000, 128), (20000, 512), (50000, 512), (100000, 512), (10, 9), (3, 2)])
def test_uniform_sample_closed_form(orc, I, S):
    r = orc.uniform_sample(I, S)
    delta = (I - 1) * 1.0 / S
    assert list(r) == [int(np.floor(j * delta)) for j in range(S)]
    assert r[0] == 0 and r[-1] < I and np.all(np.diff(r) > 0)
    assert np.diff(r).max() <= np.ceil(I / S)  # UniformSample.hpp:35


def test_uniform_sample_known_values(orc):
    r = orc.uniform_sample(5000, 128)  # delta = 39.0546875
    assert list(r[:5]) == [0, 39, 78, 117, 156] and list(r[-3:]) == [4881, 4920, 4959]


@pytest.mark.parametrize("I,S", [(100, 0), (100, 100), (100, 150), (1, 1)])
def test_uniform_sample_rejects(orc, I, S):
    with pytest.raises(ValueError):
        orc.uniform_sample(I, S)


# ---- a3 score rows (CPUScoreTable.hpp:55-68) -----------------------------------

def test_score_rows_match_reference_code(orc, fake30d):
    R = _ref(orc)
    users, items, _ = fake30d
    un = orc.normalize_users(users)
    a = orc.score_rows(un[:64], items)
    b = np.empty_like(a)
    R.ref_score_list(un, items, un.shape[0], items.shape[0], un.shape[1], 0, 64, b)
    assert np.array_equal(a.view(np.int32), b.view(np.int32))


# ---- a3+a4 table (sort desc + LoopPreprocess gather) -----------------------------

def test_build_table_is_sorted_gather(orc):
    rng = np.random.default_rng(2)
    users = rng.standard_normal((40, 6), dtype=np.float32)
    items = rng.standard_normal((700, 6), dtype=np.float32)
    ranks = orc.uniform_sample(700, 33)
    T = orc.build_table(users, items, ranks)
    s = orc.score_rows(users, items)
    expect = -np.sort(-s, axis=1)[:, ranks]
    assert np.array_equal(T, expect)
    assert np.all(np.diff(T, axis=1) <= 0)


def test_sorted_score_rows(orc):
    """Full sorted rows: the same multiset as the scores, non-increasing, and equal to the
    C oracle's own qsort-based table at every rank 0..I-1."""
    rng = np.random.default_rng(12)
    users = rng.standard_normal((25, 9), dtype=np.float32)
    items = rng.standard_normal((513, 9), dtype=np.float32)
    items[:40] = items[40:80]                      # ties
    s = orc.score_rows(users, items)
    r = orc.sorted_score_rows(users, items)
    assert np.array_equal(np.sort(r, axis=1), np.sort(s, axis=1))
    assert np.all(np.diff(r, axis=1) <= 0)
    assert np.array_equal(r, orc.build_table(users, items, np.arange(513, dtype=np.int32)))
    assert np.array_equal(orc.sorted_score_rows(users, items, 5, 7), r[5:12])


# ---- a6 coarse binary search (SampleSearch.hpp:91-130) ------------------------

def _bsearch_brute(row, ranks, I, q):
    S = len(row)
    b = int(np.sum(row >= q))  # prefix length: rows are non-increasing
    tlb = I if b == S else ranks[b]
    tub = 0 if b == 0 else ranks[b - 1]
    if b == S:
        return I, tub
    if b == 0:
        return tlb, 0
    if tlb - tub <= 1:
        return tlb, tlb
    return tlb, tub


def test_coarse_bsearch_cases(orc):
    rng = np.random.default_rng(3)
    for trial in range(300):
        S = int(rng.integers(1, 40))
        I = int(rng.integers(S + 1, 200))
        ranks = np.sort(rng.choice(I, S, replace=False)).astype(np.int32)
        if trial % 3 == 0:  # adjacent sample ranks exercise the lb == ub rule
            ranks = np.arange(S, dtype=np.int32) + int(rng.integers(0, I - S + 1))
        row = -np.sort(-rng.integers(-5, 5, S).astype(np.float32))
        for q in list(row) + [row.max() + 1, row.min() - 1, 0.5]:
            assert orc.coarse_bsearch(row, ranks, I, q) == _bsearch_brute(row, ranks, I, np.float32(q))


def test_safe_bounds_contain_true_rank(orc):
    """lbs >= rank >= ubs for every user, with ties, q on table values and eps."""
    rng = np.random.default_rng(4)
    for _ in range(200):
        I = int(rng.integers(5, 120))
        S = int(rng.integers(1, I))
        ranks = orc.uniform_sample(I, S)
        scores = rng.integers(-6, 6, I).astype(np.float32)
        srt = -np.sort(-scores)
        row = srt[ranks]
        for q in list(np.unique(scores)) + [99.0, -99.0]:
            rank = int(np.sum(scores > q))
            lbs, ubs = orc.safe_bounds(row, ranks, I, q)
            assert ubs <= rank <= lbs
            # a perturbed table with eps still brackets the rank
            eps = 0.25
            rowp = -np.sort(-(row + rng.uniform(-eps, eps, S).astype(np.float32)))
            lbs, ubs = orc.safe_bounds(rowp, ranks, I, q, eps=eps + 1e-6)
            assert ubs <= rank <= lbs


# ---- a7 prune (PruneCandidateByBound.hpp:19-179) ------------------------------

def test_prune_matches_reference_code(orc):
    R = _ref(orc)
    import ctypes as C
    rng = np.random.default_rng(5)
    for trial in range(400):
        n = int(rng.integers(1, 60))
        k = int(rng.integers(1, n + 1))
        ub = rng.integers(0, 30, n).astype(np.int32)
        lb = (ub + rng.integers(0, 3, n) * rng.integers(0, 12, n)).astype(np.int32)
        a = orc.prune(lb, ub, k)
        p = np.zeros(n, np.int8)
        r = np.zeros(n, np.int8)
        rs, nr, npr = C.c_int(n), C.c_int(0), C.c_int(0)
        R.ref_prune(lb, ub, n, k, C.byref(rs), C.byref(nr), C.byref(npr), p, r)
        assert (a["refine"], a["n_result"], a["n_prune"]) == (rs.value, nr.value, npr.value), trial
        assert np.array_equal(a["prune"], p) and np.array_equal(a["result"], r), trial


def test_prune_golden_vectors(orc):
    g = np.load(os.path.join(GOLDEN, "prune_vectors.npz"))
    for i in range(int(g["n"])):
        lb, ub, k = g[f"lb{i}"], g[f"ub{i}"], int(g[f"k{i}"])
        a = orc.prune(lb, ub, k)
        assert np.array_equal(a["prune"], g[f"prune{i}"]) and np.array_equal(a["result"], g[f"result{i}"])
        assert [a["refine"], a["n_result"], a["n_prune"]] == list(g[f"counts{i}"])


# ---- a8 exact rank (RefineCPU.hpp:33-55) ---------------------------------------

def test_refine_rank_matches_reference_code(orc, fake30d):
    R = _ref(orc)
    users, items, queries = fake30d
    un = orc.normalize_users(users)
    rng = np.random.default_rng(6)
    for _ in range(50):
        u = int(rng.integers(0, un.shape[0]))
        q = queries[int(rng.integers(0, queries.shape[0]))]
        qip = orc.ip_f32(q, un[u])
        assert orc.refine_rank(un[u], items, qip) == R.ref_refine_rank(un, items, un.shape[0], items.shape[0],
                                                                         30, qip, u)


def test_refine_ranks_batch_matches_golden(orc, fake30d):
    """The batched, threaded form used by the full-size GPU tests gives the
    golden (rank, <u,q>) of every pair."""
    g = np.load(os.path.join(GOLDEN, "refine_vectors.npz"))
    users, items, queries = fake30d
    un = orc.normalize_users(users)
    r, q = orc.refine_ranks(un, queries, items, g["user"], g["query"])
    assert np.array_equal(r, g["rank"]) and np.array_equal(q, g["qip"])


def test_refine_rank_strict_count(orc):
    items = np.array([[1.0], [2.0], [2.0], [3.0]], np.float32)
    u = np.array([1.0], np.float32)
    assert orc.refine_rank(u, items, 2.0) == 1  # only 3 > 2; ties do not count
    assert orc.refine_rank(u, items, 0.5) == 4
    assert orc.refine_rank(u, items, 3.0) == 0


def test_refine_golden_vectors(orc, fake30d):
    g = np.load(os.path.join(GOLDEN, "refine_vectors.npz"))
    users, items, queries = fake30d
    un = orc.normalize_users(users)
    for u, qi, qip, rank in zip(g["user"], g["query"], g["qip"], g["rank"]):
        assert orc.ip_f32(queries[qi], un[u]) == qip
        assert orc.refine_rank(un[u], items, qip) == rank


# ---- a5-a9 end to end -----------------------------------------------------------

def test_exact_query_equals_bruteforce(orc, fake30d):
    users, items, queries = fake30d
    un = orc.normalize_users(users)
    ranks = orc.uniform_sample(items.shape[0], 128)
    T = orc.build_table(un, items, ranks)
    q = queries[:25]
    eu, er, eq = orc.exact_query(un, items, T, ranks, q, 10)
    bu, br, bq = orc.bruteforce(un, items, q, 10)
    assert np.array_equal(eu, bu) and np.array_equal(er, br) and np.array_equal(eq, bq)


def test_us_retrieval_golden(orc, fake30d):
    """The US path on the reference fixture: user sets equal the exact answer,
    ranks of refined users are exact, users settled by bounds carry rank_lb
    >= their exact rank (the reference reports the bound, US.hpp:176)."""
    users, items, queries = fake30d
    un = orc.normalize_users(users)
    ranks = orc.uniform_sample(items.shape[0], 128)
    T = orc.build_table(un, items, ranks)
    ou, orank, oq, st = orc.us_retrieval(un, items, T, ranks, queries, 10, stats=True)
    g = np.load(os.path.join(GOLDEN, "fake30d_us_top10.npz"))
    assert np.array_equal(ou, g["us_user"]) and np.array_equal(orank, g["us_rank"])
    bu, br = g["exact_user"], g["exact_rank"]
    for i in range(queries.shape[0]):
        assert set(ou[i]) == set(bu[i])
        ex = dict(zip(bu[i], br[i]))
        n_res = st[i]["n_result_user"]
        for j, (u, r) in enumerate(zip(ou[i], orank[i])):
            if j < n_res:
                assert r >= ex[u]
            else:
                assert r == ex[u]


def test_bruteforce_golden(orc, fake30d):
    users, items, queries = fake30d
    un = orc.normalize_users(users)
    g = np.load(os.path.join(GOLDEN, "fake30d_us_top10.npz"))
    q = queries[:10]
    bu, br, _ = orc.bruteforce(un, items, q, 10)
    assert np.array_equal(bu, g["exact_user"][:10]) and np.array_equal(br, g["exact_rank"][:10])


def test_table_golden_checksum(orc, fake30d):
    import hashlib
    users, items, _ = fake30d
    un = orc.normalize_users(users)
    ranks = orc.uniform_sample(items.shape[0], 128)
    T = orc.build_table(un, items, ranks)
    g = np.load(os.path.join(GOLDEN, "fake30d_us_top10.npz"))
    assert hashlib.sha256(np.ascontiguousarray(T + np.float32(0)).tobytes()).hexdigest() == str(g["table_sha256"])
    assert np.array_equal(T[:8], g["table_head"])


def test_safe_bounds_fix_rank_zero_and_sampled_positions(orc):
    """q equal to the user's top score has rank 0 and the bounds say so (the
    reference's #{T >= q} bucket would give [0, r_1]); q strictly between
    samples is bracketed by (r[b-1], r[b]]."""
    rng = np.random.default_rng(8)
    I, S = 500, 50
    ranks = orc.uniform_sample(I, S)
    scores = rng.standard_normal(I).astype(np.float32)
    srt = -np.sort(-scores)
    row = srt[ranks]
    assert orc.safe_bounds(row, ranks, I, srt[0]) == (0, 0)
    assert orc.coarse_bsearch(row, ranks, I, srt[0]) == (ranks[1], 0)
    for j in range(1, S):
        q = srt[ranks[j]]  # q ties the sampled entry j: r[j] items beat it
        lbs, ubs = orc.safe_bounds(row, ranks, I, q)
        assert lbs == ranks[j] and ubs == ranks[j - 1] + 1


# ---- QS sample-rank training (QueryRankByComputation.hpp, OptimalSampleRankFast.hpp) ----

def _train_inputs(orc, seed, U, I, ns, d=4):
    rng = np.random.default_rng(seed)
    users = orc.normalize_users(rng.standard_normal((U, d), dtype=np.float32))
    items = rng.standard_normal((I, d), dtype=np.float32)
    ids = rng.choice(I, ns, replace=False)
    return rng, users, items, ids


def test_query_rank_distribution_invariants(orc):
    """Counts of users by rank: non-decreasing rows that reach n_user by rank I-1 (an item never
    beats itself), and the rank of a sample item's own score row is the strict count."""
    _, users, items, ids = _train_inputs(orc, 3, 40, 60, 12)
    items[10] = items[11]                                   # a tie
    accu = orc.query_rank_distribution(users, items, ids)
    assert accu.shape == (12, 61)
    assert np.all(np.diff(accu, axis=1) >= 0) and np.all(accu[:, 59] == 40) and np.all(accu[:, 60] == 40)
    S = orc.score_rows(users, items)
    ranks = (S[:, None, :] > S[:, ids][:, :, None]).sum(axis=2)  # brute-force strict counts
    for s in range(12):
        assert np.array_equal(accu[s], np.cumsum(np.bincount(ranks[:, s], minlength=61)))


def test_sort_kth_rank_is_lower_bound(orc):
    accu = np.array([[0, 1, 3, 5, 5], [2, 2, 4, 5, 5], [0, 0, 0, 1, 5]], np.int32)
    kth, order = orc.sort_kth_rank(accu, 3)
    assert list(kth) == [2, 2, 4] and list(order) == [0, 1, 2]
    kth, order = orc.sort_kth_rank(accu, 2)
    assert list(kth) == [0, 2, 4] and list(order) == [1, 0, 2]


def test_optimal_sample_dp_equals_brute_force(orc):
    """The DP's minimum equals the best cost over every subset of candidate ranks (at most n_sample
    of them), with the cost taken from its meaning: users whose rank shares the query's k-th-rank
    bucket."""
    import itertools
    rng = np.random.default_rng(0)
    checked = 0
    for trial in range(25):
        U, I, ns = int(rng.integers(3, 20)), int(rng.integers(8, 15)), int(rng.integers(3, 8))
        _, users, items, ids = _train_inputs(orc, 100 + trial, U, I, ns)
        accu = orc.query_rank_distribution(users, items, ids)
        kth, order = orc.sort_kth_rank(accu, int(rng.integers(1, U + 1)))
        acc_s = accu[order]
        for method in (0, 1):
            cand = sorted(set(orc.candidate_ranks(I, kth, method)))
            for n_sample in (1, 2, 3):
                ranks, cost = orc.find_optimal_sample_fast(U, I, kth, acc_s, method, n_sample)
                assert len(ranks) == n_sample and ranks == sorted(set(ranks)) and all(0 <= r < I for r in ranks)
                if len(orc.candidate_ranks(I, kth, method)) <= n_sample:
                    assert set(cand) <= set(ranks)
                    continue
                bf = min(orc.sample_set_cost(U, kth, acc_s, c)
                         for m in range(1, n_sample + 1) for c in itertools.combinations(cand, m))
                assert cost == bf
                checked += 1
    assert checked > 50
