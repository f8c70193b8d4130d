"""GPU test of the autotuner: it leaves a configuration set whose answers
equal the plain form's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_keeps_answers(orc, gpu_index_factory):
    print("[case] autotune", flush=True)
    from rkr import tune
    rng = np.random.default_rng(61)
    U, I, d = 5000, 2000, 40
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32)
    queries = items[rng.choice(I, 40, replace=False)]
    ks = rng.integers(5, 80, 40).astype(np.int32)
    idx, _ = gpu_index_factory(users, items, 32)
    plain = idx.query(queries, ks, max_k=80)
    space = {"screen_rows": (32, 64), "screen_chunk": (0, 512), "query_batch": (16, 1024)}
    best, results = tune.autotune(idx, queries, ks, space=space, reps=1)
    assert len(results) == 8 and set(best) == set(space)
    tuned = idx.query(queries, ks, max_k=80)
    assert np.array_equal(plain.users, tuned.users) and np.array_equal(plain.ranks, tuned.ranks)
