"""Seeded synthetic workloads of the reverse k-ranks benchmark configurations.

plain_cpu   1,000 users x 5,000 items, d=30, i.i.d. N(0,1); 128 sample ranks;
            100 queries drawn from the items; k=10.
plain_gpu   100k users x 20k items, d=150; users N(0,1)/sqrt(d), items
            N(0,1)/sqrt(d) scaled by a lognormal(0, 0.5) norm; 512 ranks;
            1,000 queries uniform over items; k in {10, 50, 100}.
precision   1M x 50k, d=150, as plain_gpu; k=50.
batching    1M x 50k, d=150; 10,000 queries drawn Zipf(1.1) over items
            sorted by norm (descending); k uniform in [10, 200].
serving     2M x 100k, d=128, N(0,1)/sqrt(d).

Users are returned raw; the library normalises them on upload as the
reference's readData does.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Workload:
    name: str
    users: np.ndarray
    items: np.ndarray
    queries: np.ndarray
    query_items: np.ndarray  # item index of each query, -1 when not an item
    topk: np.ndarray
    n_sample: int


CONFIGS = {
    "plain_cpu": dict(n_user=1_000, n_item=5_000, dim=30, n_sample=128, n_query=100, dist="normal", k=(10,)),
    "plain_gpu": dict(n_user=100_000, n_item=20_000, dim=150, n_sample=512, n_query=1_000, dist="lognorm",
                      k=(10, 50, 100)),
    "precision": dict(n_user=1_000_000, n_item=50_000, dim=150, n_sample=512, n_query=1_000, dist="lognorm",
                      k=(50,)),
    "batching": dict(n_user=1_000_000, n_item=50_000, dim=150, n_sample=512, n_query=10_000, dist="lognorm",
                     k="uniform10_200", query_dist="zipf"),
    "serving": dict(n_user=2_000_000, n_item=100_000, dim=128, n_sample=512, n_query=10_000, dist="normal_scaled",
                    k="uniform10_200"),
}


def _vectors(rng, n, d, dist):
    z = rng.standard_normal((n, d), dtype=np.float32)
    if dist == "normal":
        return z
    z /= np.float32(np.sqrt(d))
    if dist == "lognorm":
        z *= rng.lognormal(0.0, 0.5, size=(n, 1)).astype(np.float32)
    return z


def zipf_indices(rng, n, n_items, s=1.1):
    """Positions in [0, n_items) with P(i) proportional to (i+1)^-s."""
    w = 1.0 / np.arange(1, n_items + 1, dtype=np.float64) ** s
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return np.minimum(np.searchsorted(cdf, rng.random(n)), n_items - 1)


def make(name: str, seed: int = 0, n_user=None, n_item=None, n_query=None, k=None) -> Workload:
    cfg = dict(CONFIGS[name])
    U = n_user or cfg["n_user"]
    I = n_item or cfg["n_item"]
    Q = n_query or cfg["n_query"]
    d = cfg["dim"]
    rng = np.random.default_rng(seed)
    user_dist = "normal" if cfg["dist"] == "normal" else "scaled"
    item_dist = cfg["dist"] if cfg["dist"] != "normal_scaled" else "scaled"
    users = _vectors(rng, U, d, user_dist)
    items = _vectors(rng, I, d, item_dist)
    if cfg.get("query_dist") == "zipf":
        order = np.argsort(-np.linalg.norm(items.astype(np.float64), axis=1), kind="stable")
        qidx = order[zipf_indices(rng, Q, I)]
    else:
        qidx = rng.choice(I, size=Q, replace=Q > I)
    queries = items[qidx].copy()
    if k is not None:
        topk = np.broadcast_to(np.asarray(k, np.int32), (Q,)).copy()
    elif cfg["k"] == "uniform10_200":
        topk = rng.integers(10, 201, size=Q).astype(np.int32)
    else:
        topk = np.full(Q, cfg["k"][0], np.int32)
    n_sample = min(cfg["n_sample"], I - 1)
    return Workload(name, users, items, queries, qidx.astype(np.int64), topk, n_sample)
