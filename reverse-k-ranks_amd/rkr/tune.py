"""Autotuner for the batched screened form.

Times the tile shape (refine rows per wavefront), the work split (items per
work-unit chunk) and the device batch size on a sample of the caller's
queries, and leaves the fastest configuration set on the index.  Every
configuration returns the same answers (the tests check both tiles against
the plain form), so only time is compared.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

from . import _lib as L

DEFAULT_SPACE = {
    "screen_rows": (32, 64),
    "screen_chunk": (0, 8192, 2048),
    "query_batch": (256, 1024, 4096),
}

_OPTION = {"screen_rows": L.RKR_OPT_SCREEN_ROWS, "screen_chunk": L.RKR_OPT_SCREEN_CHUNK,
           "query_batch": L.RKR_OPT_QUERY_BATCH}


def autotune(index, queries, topk, space=None, reps: int = 2, verbose: bool = False):
    """Sets the fastest configuration on `index` and returns
    (best_config, [(config, seconds per query), ...]).

    `queries` should be a representative sample (a few hundred to a few
    thousand rows); `topk` a k or an array of k per query."""
    space = dict(DEFAULT_SPACE if space is None else space)
    queries = np.ascontiguousarray(queries, np.float32)
    ks = np.broadcast_to(np.asarray(topk, np.int32), (len(queries),))
    max_k = int(ks.max())
    index.set_form("screened")
    names = list(space)
    results = []
    for values in itertools.product(*(space[n] for n in names)):
        cfg = dict(zip(names, values))
        for n, v in cfg.items():
            index.set_option(_OPTION[n], v)
        index.query(queries, ks, max_k=max_k)  # warm-up (allocations, code load)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            index.query(queries, ks, max_k=max_k)
            times.append(time.perf_counter() - t0)
        per_q = min(times) / len(queries)
        results.append((cfg, per_q))
        if verbose:
            print(f"{cfg} {per_q * 1e6:.1f} us/query", flush=True)
    best = min(results, key=lambda r: r[1])[0]
    for n, v in best.items():
        index.set_option(_OPTION[n], v)
    return best, results
