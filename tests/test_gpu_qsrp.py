"""GPU tests of the QSRP screen (rkr_qsrp_build / RKR_OPT_QSRP): the device CDF
within the error the bounds assume, the screen's bucket bounds containing the
exact bucket of every (query, user) pair, and answers with the screen equal to
brute force and to the answers without it."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _low_rank(rng, U, I, d, decay=0.6):
    """Users and items with singular values decaying as decay^k in a random basis
    (like learned embeddings): the SVD concentrates the energy in a few leading
    dimensions, so with SIGMA 0.9 the partial-dimension bound is tight."""
    Q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    B = ((decay ** np.arange(d))[:, None] * Q).astype(np.float32)
    users = rng.standard_normal((U, d)).astype(np.float32) @ B
    items = (rng.standard_normal((I, d)).astype(np.float32) @ B) * rng.lognormal(0, 0.3, (I, 1)).astype(np.float32)
    return users.astype(np.float32), items.astype(np.float32)


def _index(users, items, S, dense_top=0):
    import rkr
    idx = rkr.Index(0)
    idx.set_option(8, dense_top)  # RKR_OPT_DENSE_TOP
    idx.set_data(users, items)
    ranks = rkr.Index.uniform_sample_ranks(items.shape[0], S)
    idx.build_table(ranks)
    return idx, ranks


def test_device_cdf_within_assumed_error():
    """The bounds assume |Phi_device - Phi| <= 2e-6 (A&S 7.1.26 in FP32,
    argument rounding aside); checked against FP64 erf on a dense grid."""
    import rkr
    _case("QSRP device CDF error")
    z = np.linspace(-12, 12, 1 << 20, dtype=np.float32)
    with rkr.Index(0) as idx:
        got = idx.qsrp_phi(z).astype(np.float64)
    exact = 0.5 * (1.0 + np.array([math.erf(float(v) / math.sqrt(2.0)) for v in z]))
    err = float(np.max(np.abs(got - exact)))
    print(f"max |Phi_dev - Phi| = {err:.3e}")
    assert err <= 2e-6


@pytest.mark.parametrize("kind,eps", [("low_rank", 0.0), ("isotropic", 0.0), ("low_rank", 1e-3)])
def test_screen_contains_exact_bucket(orc, kind, eps):
    """b_hi <= #{T > s + eps} and #{T > s - eps} <= b_lo for every pair, with s the
    sequential-rule score: the partial-dimension interval and the regression
    never exclude the true bucket (eps: a table loaded with that error bound)."""
    _case(f"QSRP screen containment {kind} eps={eps}")
    rng = np.random.default_rng(31)
    U, I, d, S = 3000, 4000, 48, 64
    if kind == "low_rank":
        users, items = _low_rank(rng, U, I, d)
    else:
        users = rng.standard_normal((U, d), dtype=np.float32)
        items = rng.standard_normal((I, d), dtype=np.float32)
    idx, ranks = _index(users, items, S)
    try:
        un = orc.normalize_users(users)
        if eps > 0:
            T = orc.build_table(un, items, ranks)
            T = (T + rng.uniform(-eps, eps, T.shape).astype(np.float32) * 0.5).astype(np.float32)
            T = -np.sort(-T, axis=1)
            idx.set_table(ranks, T, eps)
        T = idx.get_table()
        ms = idx.qsrp_build(0.9 if kind == "low_rank" else 0.7)
        info = idx.qsrp_info()
        queries = np.concatenate([items[rng.choice(I, 24, replace=False)],
                                  rng.standard_normal((8, d)).astype(np.float32)])
        bk = idx.qsrp_screen(queries)
        lo, hi = (bk & 0xFFFF).astype(np.int64), (bk >> 16).astype(np.int64)
        s = orc.score_rows(un, queries)                       # [U][nq] sequential rule
        up = (T[:, None, :] > (s[:, :, None] - np.float32(eps))).sum(axis=2).T   # [nq][U]
        dn = (T[:, None, :] > (s[:, :, None] + np.float32(eps))).sum(axis=2).T
        assert np.all(hi <= dn), f"screen lower bucket above the exact one at {np.argwhere(hi > dn)[:3]}"
        assert np.all(lo >= up), f"screen upper bucket below the exact one at {np.argwhere(lo < up)[:3]}"
        width = float(np.mean(lo - hi))
        print(f"check_dim {info['check_dim']} of {d}, fallback fits {info['n_fallback_fit']}, "
              f"build {ms:.1f} ms, mean bucket width {width:.1f} of {S}")
        if kind == "low_rank":
            assert width < S / 2  # the screen narrows the buckets on low-rank data
    finally:
        idx.close()


@pytest.mark.parametrize("dense_top,form", [(0, "plain"), (2, "screened")])
def test_qsrp_answers_exact(orc, dense_top, form):
    """Answers with the screen equal brute force and the answers without it;
    on low-rank data the screen keeps only part of the users."""
    _case(f"QSRP answers dense_top={dense_top} form={form}")
    rng = np.random.default_rng(32)
    U, I, d, S = 6000, 3000, 40, 64
    users, items = _low_rank(rng, U, I, d)
    idx, ranks = _index(users, items, S, dense_top)
    try:
        idx.set_form(form)
        idx.qsrp_build(0.9)
        queries = items[rng.choice(I, 40, replace=False)]
        ks = rng.integers(1, 120, len(queries)).astype(np.int32)
        off = idx.query(queries, ks, max_k=120)
        idx.set_option(11, 1)  # RKR_OPT_QSRP
        on = idx.query(queries, ks, max_k=120)
        assert np.array_equal(on.users, off.users) and np.array_equal(on.ranks, off.ranks)
        assert np.array_equal(on.qips, off.qips)
        kept = on.timing["n_listed"]
        print(f"screen kept {kept} of {U * len(queries)} (query, user) pairs")
        assert 0 < kept < U * len(queries)
        un = orc.normalize_users(users)
        bu, br, _ = orc.bruteforce(un, items, queries[:10], int(ks[:10].max()))
        for i in range(10):
            k = int(ks[i])
            assert np.array_equal(on.users[i, :k], bu[i, :k]) and np.array_equal(on.ranks[i, :k], br[i, :k])
    finally:
        idx.close()


def test_qsrp_dropped_by_updates(orc):
    """The screen describes the users and table it was built for: an update drops
    it (queries stay exact without it) and a rebuild of the table too."""
    import rkr
    _case("QSRP dropped by updates")
    rng = np.random.default_rng(33)
    users, items = _low_rank(rng, 2000, 1500, 24)
    idx, ranks = _index(users, items, 32)
    try:
        idx.qsrp_build(0.7)
        idx.set_option(11, 1)
        idx.insert_users(users[:5] * 1.5)
        with pytest.raises(rkr.RkrError):
            idx.qsrp_info()
        un = orc.normalize_users(np.concatenate([users, users[:5] * 1.5]))
        q = items[:6]
        res = idx.query(q, 30)
        bu, br, _ = orc.bruteforce(un, items, q, 30)
        assert np.array_equal(res.users, bu) and np.array_equal(res.ranks, br)
        idx.qsrp_build(0.7)
        idx.build_table(ranks)
        with pytest.raises(rkr.RkrError):
            idx.qsrp_info()
    finally:
        idx.close()


def test_cli_qsrp_pipeline_fake30d(orc, fake30d, tmp_path):
    """qdibc -> fsr -> bsibc -> rri with --method_name QSRPNormalLP on the reference
    fixture (QSRPNormalLP runs on the QS table, NameTranslation.hpp:36-48): the
    answers written to the result CSVs are exact."""
    import os
    import subprocess
    from conftest import GOLDEN, ROOT
    _case("cli qdibc/fsr/bsibc/rri QSRPNormalLP")
    cli = os.path.join(ROOT, "reverse-k-ranks_amd", "rkr_cli")
    idir = str(tmp_path / "index")
    common = ["--dataset_dir", GOLDEN, "--dataset_name", "fake-normal-30d", "--index_dir", idir]
    m = ["--method_name", "QSRPNormalLP", "--n_sample", "128", "--n_sample_query", "150", "--sample_topk", "10"]
    steps = (["qdibc", "--n_sample_item", "150", "--sample_topk", "10"], ["fsr"] + m, ["bsibc"] + m,
             ["rri"] + m + ["--topk", "10", "--result_dir", str(tmp_path / "result")])
    for cmd in steps:
        r = subprocess.run([cli] + cmd + common, capture_output=True, text=True, timeout=100)
        assert r.returncode == 0, (cmd[0], r.stderr)
    users, items, queries = fake30d
    un = orc.normalize_users(users)
    stem = (tmp_path / "result" / "rank" /
            "fake-normal-30d-QSRPNormalLP-top10-n_sample_128-n_sample_query_150-sample_topk_10-n_thread_-1")
    got_u = np.loadtxt(f"{stem}-userID.csv", delimiter=",", dtype=np.int64)
    got_r = np.loadtxt(f"{stem}-rank.csv", delimiter=",", dtype=np.int64)
    bu, br, _ = orc.bruteforce(un, items, queries, 10)
    assert np.array_equal(got_u, bu) and np.array_equal(got_r, br)
