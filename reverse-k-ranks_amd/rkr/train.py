"""QS sample-rank training over the C ABI.

The reference trains QS's sampled ranks with two binaries (run_rkr.py chains
them before bsibc):
  qdibc: sample items (SampleItem), compute each sample item's rank over every
         user (ComputeQueryRankByComputation) and sort the samples by the rank
         of their sample_topk-th user (ComputeSortKthRank); writes the query
         distribution files (FileIO.hpp:56-137);
  fsr --method_name QS: FindOptimalSampleFast over those files.
``train_sample_ranks`` runs both steps; the pieces are exposed for the CLI and
the tests.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .index import Index, _check, _ptr

METHODS = {"OptimalPart": L.RKR_SAMPLE_OPTIMAL_PART, "OptimalUniform": L.RKR_SAMPLE_OPTIMAL_UNIFORM}


def sample_items(n_item: int, n_sample_item: int) -> np.ndarray:
    """SampleItem (src/app/build-index/QueryRankBySample.hpp:20-37)."""
    out = np.empty(max(n_sample_item, 1), np.int32)
    rc = L.load().rkr_sample_items(n_item, n_sample_item, _ptr(out))
    if rc != L.RKR_OK:
        raise L.RkrError(rc, "need 0 < n_sample_item <= n_item")
    return out[:n_sample_item]


def query_rank_distribution(idx: Index, item_ids) -> tuple[np.ndarray, float]:
    """accu[s][r] = #{users whose rank of item_ids[s] is <= r}; returns (accu, device ms)."""
    ids = np.ascontiguousarray(item_ids, np.int32)
    accu = np.empty((ids.shape[0], idx.n_item + 1), np.int32)
    ms = C.c_double()
    _check(idx.lib, idx.ctx, idx.lib.rkr_query_rank_distribution(idx.ctx, _ptr(ids), ids.shape[0], _ptr(accu),
                                                                  C.byref(ms)))
    return accu, ms.value


def sort_kth_rank(accu, sample_topk: int) -> tuple[np.ndarray, np.ndarray]:
    """ComputeSortKthRank: (sort_kth_rank, sort_sample_ids)."""
    accu = np.ascontiguousarray(accu, np.int32)
    ns, w = accu.shape
    kth = np.empty(ns, np.int32)
    ids = np.empty(ns, np.int32)
    rc = L.load().rkr_sort_kth_rank(_ptr(accu), w - 1, ns, sample_topk, _ptr(kth), _ptr(ids))
    if rc != L.RKR_OK:
        raise L.RkrError(rc, "bad accu or sample_topk")
    return kth, ids


def find_optimal_sample(idx: Index, n_user: int, n_item: int, sort_kth, accu_sorted, n_sample: int,
                        method: str = "OptimalPart") -> tuple[np.ndarray, int]:
    """FindOptimalSampleFast: (ranks, train IO cost)."""
    kth = np.ascontiguousarray(sort_kth, np.int32)
    accu = np.ascontiguousarray(accu_sorted, np.int32)
    if accu.shape != (kth.shape[0], n_item + 1):
        raise ValueError("accu must be n_sample_query x (n_item + 1)")
    out = np.empty(n_sample, np.int32)
    cost = C.c_int64()
    _check(idx.lib, idx.ctx, idx.lib.rkr_find_optimal_sample(idx.ctx, n_user, n_item, kth.shape[0], _ptr(kth),
                                                              _ptr(accu), METHODS[method], n_sample, _ptr(out),
                                                              C.byref(cost)))
    return out, cost.value


def train_sample_ranks(idx: Index, n_sample: int, n_sample_query: int, sample_topk: int,
                       method: str = "OptimalPart") -> dict:
    """qdibc + fsr for the index's users and items; returns the ranks and the intermediate arrays."""
    ids = sample_items(idx.n_item, n_sample_query)
    accu, ms = query_rank_distribution(idx, ids)
    kth, order = sort_kth_rank(accu, sample_topk)
    accu_sorted = accu[order]
    del accu
    ranks, cost = find_optimal_sample(idx, idx.n_user, idx.n_item, kth, accu_sorted, n_sample, method)
    return {"ranks": ranks, "train_io_cost": cost, "sample_item_ids": ids[order], "sort_kth_rank": kth,
            "accu_sorted": accu_sorted, "rank_distribution_device_ms": ms}
