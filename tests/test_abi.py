"""The C ABI without a GPU: librkr.so loads, exports every entry point that
include/rkr.h declares, and its host-only helpers and argument checks behave
(no kernel runs here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "rkr.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rkr_[a-z_0-9]+)\s*\(", src)))


def test_header_lists_match_binding():
    from rkr import _lib
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from rkr import _lib
    L = _lib.load()
    for name in _declared():
        assert hasattr(L, name), name
    assert L.rkr_version().startswith(b"rkr")


def test_nm_shows_c_linkage():
    import subprocess
    from rkr import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in _declared():
        assert name in syms, name


def test_uniform_sample_ranks_host_helper(orc):
    from rkr import Index
    for I, S in [(5000, 128), (20000, 512), (50000, 512), (7, 3)]:
        assert np.array_equal(Index.uniform_sample_ranks(I, S), orc.uniform_sample(I, S))
    from rkr import RkrError
    for I, S in [(100, 0), (100, 100), (5, 9)]:
        with pytest.raises(RkrError):
            Index.uniform_sample_ranks(I, S)


def test_normalize_users_host_helper_matches_oracle(orc, fake30d):
    from rkr import _lib
    users = fake30d[0].copy()
    users[5] = 1e-6
    a = users.copy()
    assert _lib.load().rkr_normalize_users(a.ctypes.data, a.shape[0], a.shape[1]) == 0
    assert np.array_equal(a.view(np.int32), orc.normalize_users(users).view(np.int32))


def test_null_and_invalid_arguments():
    from rkr import _lib
    L = _lib.load()
    assert L.rkr_create(0, 0, None) == _lib.RKR_ERR_INVALID
    assert L.rkr_set_option(None, 1, 0) == _lib.RKR_ERR_INVALID
    assert L.rkr_sync(None) == _lib.RKR_ERR_INVALID
    assert L.rkr_last_error(None) == b"null context"
    assert L.rkr_index_size_bytes(None) == 0
    assert L.rkr_normalize_users(None, 3, 3) == _lib.RKR_ERR_INVALID
    assert L.rkr_uniform_sample_ranks(10, 3, None) == _lib.RKR_ERR_INVALID
    L.rkr_destroy(None)  # no-op


def test_create_without_gpu_reports_nodevice():
    """In a container without a GPU the library reports it instead of exiting."""
    import torch
    from rkr import _lib
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    ctx = C.c_void_p()
    assert _lib.load().rkr_create(0, 0, C.byref(ctx)) == _lib.RKR_ERR_NODEVICE
    assert not ctx.value


def test_formats_roundtrip(tmp_path):
    from rkr import formats
    rng = np.random.default_rng(0)
    x = rng.standard_normal((17, 5), dtype=np.float32)
    p = str(tmp_path / "a.fvecs")
    formats.write_fvecs(p, x)
    assert np.array_equal(formats.read_fvecs(p), x)
    ranks = np.array([0, 3, 9], np.int32)
    T = -np.sort(-rng.standard_normal((4, 3), dtype=np.float32), axis=1)
    tp = str(tmp_path / "t.index")
    formats.write_table(tp, ranks, T, 10)
    raw = open(tp, "rb").read()
    assert len(raw) == 3 * 8 + 3 * 4 + 12 * 4  # size_t x3, int32 ranks, float table
    r2, T2, n_item = formats.read_table(tp)
    assert n_item == 10 and np.array_equal(r2, ranks) and np.array_equal(T2, T)
    paths = formats.write_rank_result(str(tmp_path), "ds", "US", "top2", [[1, 2], [3, 4]], [[0, 5], [1, 1]],
                                      [[0.5, 0.25], [1.0, 2.0]])
    assert open(paths[0]).read() == "1,2\n3,4\n"
    assert open(paths[1]).read() == "0,5\n1,1\n"
    # full score table: float[n_user][n_item], no header
    S = -np.sort(-rng.standard_normal((6, 11), dtype=np.float32), axis=1)
    sp = formats.score_table_path(str(tmp_path), "ds")
    assert sp.endswith("ds.index")
    formats.write_score_table(sp, S)
    assert os.path.getsize(sp) == 6 * 11 * 4
    assert np.array_equal(formats.read_score_table(sp, 11), S)
    assert np.array_equal(formats.read_score_table(sp, 11, mmap=False), S)
    assert np.array_equal(formats.read_score_rows(sp, 11, 4, 3, 5), S[4, 3:8])
    with pytest.raises(ValueError):
        formats.read_score_table(sp, 7)


def test_reference_dataset_fixture_layout(fake30d):
    users, items, queries = fake30d
    assert users.shape == (1000, 30) and items.shape == (5000, 30) and queries.shape == (100, 30)


def test_synth_shapes_and_seeds():
    from rkr import synth
    a = synth.make("plain_cpu", seed=3)
    b = synth.make("plain_cpu", seed=3)
    assert a.users.shape == (1000, 30) and a.items.shape == (5000, 30) and a.queries.shape == (100, 30)
    assert np.array_equal(a.users, b.users) and np.array_equal(a.queries, b.queries)
    assert np.array_equal(a.queries, a.items[a.query_items])
    z = synth.make("batching", seed=1, n_user=100, n_item=2000, n_query=5000)
    assert z.topk.min() >= 10 and z.topk.max() <= 200
    norms = np.linalg.norm(z.items.astype(np.float64), axis=1)
    # Zipf over norm-sorted items: the largest-norm item is the most frequent query
    top = np.bincount(z.query_items, minlength=2000).argmax()
    assert top == int(np.argmax(norms))


def test_sample_items_host():
    """SampleItem: a prefix of one fixed shuffle of 0..I-1 (std::mt19937 seeded 0)."""
    from rkr import train
    full = train.sample_items(1000, 1000)
    assert sorted(full.tolist()) == list(range(1000))
    assert np.array_equal(train.sample_items(1000, 37), full[:37])
    assert np.array_equal(train.sample_items(1000, 37), train.sample_items(1000, 37))
    assert not np.array_equal(full[:50], np.arange(50))
    for bad in ((0, 1), (10, 0), (10, 11)):
        with pytest.raises(Exception):
            train.sample_items(*bad)


def test_sort_kth_rank_host_matches_oracle(orc):
    from rkr import train
    rng = np.random.default_rng(5)
    counts = rng.integers(0, 4, size=(300, 51))
    accu = np.cumsum(counts, axis=1).astype(np.int32)
    for topk in (1, 10, 40):
        kth, ids = train.sort_kth_rank(accu, topk)
        okth, _ = orc.sort_kth_rank(accu, topk)
        assert np.array_equal(kth, okth)
        assert sorted(ids.tolist()) == list(range(300))
        assert np.array_equal(np.array([np.searchsorted(accu[i], topk) for i in ids]), kth)


def test_query_distribution_and_rank_files_roundtrip(tmp_path):
    """qdibc's three files (FileIO.hpp:56-137) and fsr's rank file (SampleSearch.hpp:158-194)."""
    from rkr import formats
    rng = np.random.default_rng(9)
    Q, I = 7, 30
    ids = rng.choice(I, Q, replace=False).astype(np.int32)
    accu = np.cumsum(rng.integers(0, 3, size=(Q, I + 1)), axis=1).astype(np.int32)
    kth = np.sort(rng.integers(0, I, Q)).astype(np.int32)
    d = formats.query_distribution_dir(str(tmp_path), "ds", Q, 5)
    assert d.endswith(os.path.join("query_distribution", "ds-n_sample_item_7-sample_topk_5"))
    formats.write_query_distribution(d, ids, kth, accu)
    assert os.path.getsize(os.path.join(d, "sort_kth_rank_l.index")) == 4 * Q
    assert os.path.getsize(os.path.join(d, "accu_n_user_rank_l.index")) == 4 * Q * (I + 1)
    assert open(os.path.join(d, "sample_itemID_l.txt")).read().split() == [str(i) for i in ids]
    i2, k2, a2 = formats.read_query_distribution(d, I)
    assert np.array_equal(i2, ids) and np.array_equal(k2, kth) and np.array_equal(a2, accu)
    with pytest.raises(ValueError):
        formats.read_query_distribution(d, I + 1)
    rp = formats.qs_ranks_path(str(tmp_path), "ds", 4, Q, 5)
    assert os.path.basename(rp) == "QS-ds-n_sample_4-n_sample_query_7-sample_topk_5.index"
    os.makedirs(os.path.dirname(rp), exist_ok=True)
    formats.write_ranks(rp, [0, 3, 9, 20], I, 100)
    assert os.path.getsize(rp) == 3 * 8 + 4 * 4
    r, n_item, n_user = formats.read_ranks(rp)
    assert r.tolist() == [0, 3, 9, 20] and (n_item, n_user) == (I, 100)
