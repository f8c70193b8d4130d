import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "reverse-k-ranks_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librkr.so on cuda:0)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def fake30d():
    """The reference's own fixture dataset fake-normal-30d (1000 x 5000, d=30)."""
    from rkr import formats
    users, items, queries = formats.read_dataset(GOLDEN, "fake-normal-30d")
    return users, items, queries


@pytest.fixture(scope="session")
def gpu_index_factory():
    import rkr

    def make(users, items, n_sample, normalize=True, dense_top=None):
        idx = rkr.Index(0)
        if dense_top is not None:
            idx.set_option(8, int(dense_top))  # RKR_OPT_DENSE_TOP: query table with dense top ranks
        idx.set_data(users, items, normalize_users=normalize)
        ranks = rkr.Index.uniform_sample_ranks(items.shape[0], n_sample)
        idx.build_table(ranks)
        return idx, ranks
    return make


def rng(seed=0):
    return np.random.default_rng(seed)
