"""GPU parity of QS sample-rank training (qdibc + fsr --method_name QS) against the oracle.

The rank distribution is integer output and must match bit for bit; the DP's
ranks and minimum cost must equal the oracle's restatement of
FindOptimalSampleFast, whose minimum the oracle tests pin to a brute-force
search over sample sets.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _data(seed, U, I, d, dup=0):
    rng = np.random.default_rng(seed)
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32)
    if dup:
        items[:dup] = items[dup:2 * dup]  # tied scores
    return rng, users, items


@pytest.mark.parametrize("U,I,d,ns,batch,seed", [
    (200, 1000, 30, 40, 0, 1),       # one sort tile
    (64, 9000, 12, 25, 16, 2),       # rows over one tile, several device batches
    (33, 20001, 20, 7, 0, 3),        # two merge passes, odd width
    (5, 40, 3, 40, 0, 4),            # every item sampled
    (33, 20001, 20, 900, 0, 10),     # counts leave in three staging chunks of 419, 419 and 62 rows
    (9, 50001, 12, 30, 0, 11),       # a 200 KB row: more than a workgroup's LDS, searched in global memory
])
def test_rank_distribution_matches_oracle(orc, U, I, d, ns, batch, seed):
    _case(f"rank distribution U={U} I={I} ns={ns} batch={batch}")
    import rkr
    from rkr import train
    rng, users, items = _data(seed, U, I, d, dup=5)
    idx = rkr.Index(0)
    if batch:
        idx.set_option(4, batch)  # RKR_OPT_BUILD_BATCH
    idx.set_data(users, items)
    ids = rng.choice(I, ns, replace=False).astype(np.int32)
    accu, ms = train.query_rank_distribution(idx, ids)
    ref = orc.query_rank_distribution(orc.normalize_users(users), items, ids)
    assert np.array_equal(accu, ref)
    assert ms > 0


def _synthetic_distribution(rng, U, I, Q, spread):
    """Monotone per-query user counts whose k-th ranks spread over [0, spread)."""
    accu = np.zeros((Q, I + 1), np.int64)
    for q in range(Q):
        ranks = np.minimum(rng.integers(0, spread, U) + rng.integers(0, I - spread, U) * (rng.random(U) < 0.3), I - 1)
        accu[q] = np.cumsum(np.bincount(ranks, minlength=I + 1))
    return accu.astype(np.int32)


@pytest.mark.parametrize("method", ["OptimalPart", "OptimalUniform"])
@pytest.mark.parametrize("U,I,Q,topk,n_sample,seed", [
    (300, 400, 40, 20, 12, 5),     # DP path: more candidates than samples
    (300, 400, 40, 20, 3, 6),      # DP with few samples (ranks repeat, then the fill)
    (50, 300, 25, 5, 200, 7),      # every candidate fits: the fill path, n_sample_query < 500
    (20, 900, 600, 3, 700, 8),     # the fill path with n_sample_query >= 500
])
def test_find_optimal_sample_matches_oracle(orc, method, U, I, Q, topk, n_sample, seed):
    _case(f"optimal sample {method} U={U} I={I} Q={Q} S={n_sample}")
    import rkr
    from rkr import train
    rng = np.random.default_rng(seed)
    accu = _synthetic_distribution(rng, U, I, Q, max(2, I // 8))
    kth, order = train.sort_kth_rank(accu, topk)
    acc_s = accu[order]
    idx = rkr.Index(0)
    ranks, cost = train.find_optimal_sample(idx, U, I, kth, acc_s, n_sample, method)
    m = 0 if method == "OptimalPart" else 1
    oranks, ocost = orc.find_optimal_sample_fast(U, I, kth, acc_s, m, n_sample)
    assert ranks.tolist() == list(oranks) and cost == ocost
    assert np.all(np.diff(ranks) > 0) and ranks[0] >= 0 and ranks[-1] < I


def test_trained_ranks_answer_exactly(orc):
    """qdibc + fsr on the device, then a table on the trained ranks: answers equal brute force,
    and the training cost is the cost of the trained ranks."""
    _case("train QS ranks end to end")
    import rkr
    from rkr import train
    rng, users, items = _data(9, 2000, 3000, 16)
    idx = rkr.Index(0)
    idx.set_data(users, items)
    out = train.train_sample_ranks(idx, n_sample=64, n_sample_query=150, sample_topk=20)
    ranks = out["ranks"]
    assert ranks.shape == (64,) and np.all(np.diff(ranks) > 0)
    un = orc.normalize_users(users)
    ref = orc.query_rank_distribution(un, items, out["sample_item_ids"])
    assert np.array_equal(ref, out["accu_sorted"])
    # extra ranks from the fill can only split buckets, so the final set costs at most the DP minimum
    assert orc.sample_set_cost(2000, out["sort_kth_rank"], out["accu_sorted"], ranks) <= out["train_io_cost"]
    idx.build_table(ranks)
    queries = items[rng.choice(3000, 20, replace=False)]
    res = idx.query(queries, 20)
    bu, br, _ = orc.bruteforce(un, items, queries, 20)
    assert np.array_equal(res.users, bu) and np.array_equal(res.ranks, br)


def test_cli_qs_pipeline_fake30d(orc, fake30d, tmp_path):
    """qdibc -> fsr -> bsibc -> rri with --method_name QS on the reference fixture: the query
    distribution files hold the oracle's counts, fsr's ranks are the oracle DP's, and the answers
    are exact."""
    import os
    import subprocess
    from conftest import GOLDEN, ROOT
    from rkr import formats
    _case("cli qdibc/fsr/bsibc/rri QS")
    cli = os.path.join(ROOT, "reverse-k-ranks_amd", "rkr_cli")
    idir = str(tmp_path / "index")
    common = ["--dataset_dir", GOLDEN, "--dataset_name", "fake-normal-30d", "--index_dir", idir]
    qs = ["--method_name", "QS", "--n_sample", "128", "--n_sample_query", "150", "--sample_topk", "10"]
    steps = (["qdibc", "--n_sample_item", "150", "--sample_topk", "10"], ["fsr"] + qs, ["bsibc"] + qs,
             ["rri"] + qs + ["--topk", "10", "--result_dir", str(tmp_path / "result")])
    for cmd in steps:
        r = subprocess.run([cli] + cmd + common, capture_output=True, text=True, timeout=100)
        assert r.returncode == 0, (cmd[0], r.stderr)
    users, items, queries = fake30d
    un = orc.normalize_users(users)
    ids, kth, accu = formats.read_query_distribution(
        formats.query_distribution_dir(idir, "fake-normal-30d", 150, 10), items.shape[0])
    assert np.array_equal(accu, orc.query_rank_distribution(un, items, ids))
    okth, _ = orc.sort_kth_rank(accu, 10)
    assert np.array_equal(kth, okth)
    ranks, n_item, n_user = formats.read_ranks(formats.qs_ranks_path(idir, "fake-normal-30d", 128, 150, 10))
    oranks, _ = orc.find_optimal_sample_fast(1000, 5000, kth, accu, 0, 128)
    assert (n_item, n_user) == (5000, 1000) and ranks.tolist() == list(oranks)
    stem = tmp_path / "result" / "rank" / "fake-normal-30d-QS-top10-n_sample_128-n_sample_query_150-sample_topk_10-n_thread_-1"
    got_u = np.loadtxt(f"{stem}-userID.csv", delimiter=",", dtype=np.int64)
    got_r = np.loadtxt(f"{stem}-rank.csv", delimiter=",", dtype=np.int64)
    bu, br, _ = orc.bruteforce(un, items, queries, 10)
    assert np.array_equal(got_u, bu) and np.array_equal(got_r, br)


def test_training_and_score_table_errors(orc, tmp_path):
    """The new entry points fail with status codes, never exit or fault: bad row ranges, bad item
    IDs, an unwritable path, a sample_topk above the user count, a bad method."""
    _case("training and score table errors")
    import rkr
    from rkr import _lib as L
    from rkr import train
    rng, users, items = _data(41, 50, 300, 8)
    idx = rkr.Index(0)
    with pytest.raises(L.RkrError) as e:
        idx.sorted_scores()  # no data yet
    assert e.value.status == L.RKR_ERR_STATE
    idx.set_data(users, items)
    for u0, n in ((-1, 2), (49, 2), (0, 51)):
        with pytest.raises(L.RkrError) as e:
            idx.sorted_scores(u0, n)
        assert e.value.status == L.RKR_ERR_INVALID
    with pytest.raises(L.RkrError) as e:
        idx.save_score_table(str(tmp_path / "missing_dir" / "x.index"))
    assert e.value.status == L.RKR_ERR_IO
    with pytest.raises(L.RkrError) as e:
        train.query_rank_distribution(idx, np.array([0, 300], np.int32))
    assert e.value.status == L.RKR_ERR_INVALID
    accu, _ = train.query_rank_distribution(idx, np.arange(20, dtype=np.int32))
    kth, order = train.sort_kth_rank(accu, 51)  # one more than the users: no k-th rank below n_item
    assert np.all(kth == 301)
    with pytest.raises(L.RkrError) as e:
        train.find_optimal_sample(idx, 50, 300, kth, accu[order], 16)
    assert e.value.status == L.RKR_ERR_INVALID
    kth, order = train.sort_kth_rank(accu, 5)
    with pytest.raises(KeyError):
        train.find_optimal_sample(idx, 50, 300, kth, accu[order], 16, "OptimalAll")
    out = np.empty(4, np.int32)
    acc_s = np.ascontiguousarray(accu[order])
    rc = idx.lib.rkr_find_optimal_sample(idx.ctx, 50, 300, 20, kth.ctypes.data, acc_s.ctypes.data, 7, 4,
                                         out.ctypes.data, None)
    assert rc == L.RKR_ERR_INVALID
    # the context still answers after the errors
    assert np.array_equal(idx.sorted_scores(0, 3), orc.sorted_score_rows(orc.normalize_users(users[:3]), items))
