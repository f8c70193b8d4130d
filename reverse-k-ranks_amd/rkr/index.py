"""Python mirror of the reference's US index interface over the C ABI.

``build_index(items, users, ...)`` plays the role of US::BuildIndex
(src/impl/US.hpp:246-261) and ``Index.retrieval(query_item, topk, n_execute_query)``
the role of US::Index::Retrieval (src/impl/US.hpp:96-213): it returns, per
query, ``topk`` UserRankElement-like triples (userID, rank, queryIP) plus the
per-query performance records.  Unlike the reference, every query runs (the
reference answers only the first 100 queries when n_sample <= 1000,
US.hpp:112) and errors raise instead of exiting.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


def _check(lib, ctx, rc):
    if rc != L.RKR_OK:
        msg = lib.rkr_last_error(ctx).decode() if ctx else ""
        raise L.RkrError(rc, msg)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class UserRankElement:
    """src/include/struct/UserRankElement.hpp:11-92 (ordering by (rank, userID))."""
    userID: int
    rank: int
    queryIP: float


@dataclass
class QueryResult:
    users: np.ndarray   # (n_query, max_k) int32, -1 past each query's k
    ranks: np.ndarray   # (n_query, max_k) int32
    qips: np.ndarray    # (n_query, max_k) float32
    stats: list         # per-query dicts (SingleQueryPerformance fields)
    timing: dict        # batch timing (device ms per stage)

    def elements(self, qi: int):
        return [UserRankElement(int(u), int(r), float(q))
                for u, r, q in zip(self.users[qi], self.ranks[qi], self.qips[qi]) if u >= 0]


class Index:
    """One device context holding users, items and the sampled score table in HBM."""

    def __init__(self, device: int = 0, mem_budget_bytes: int = 0):
        self.lib = L.load()
        ctx = C.c_void_p()
        rc = self.lib.rkr_create(device, mem_budget_bytes, C.byref(ctx))
        if rc != L.RKR_OK:
            raise L.RkrError(rc, "rkr_create failed (no usable GPU?)")
        self.ctx = ctx
        self.n_user = self.n_item = self.dim = 0

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.rkr_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- options -------------------------------------------------------------
    def set_option(self, opt: int, value: int):
        _check(self.lib, self.ctx, self.lib.rkr_set_option(self.ctx, opt, int(value)))

    def set_form(self, form: str):
        self.set_option(L.RKR_OPT_FORM, {"plain": L.RKR_FORM_PLAIN, "screened": L.RKR_FORM_SCREENED}[form])

    def set_output(self, output: str):
        self.set_option(L.RKR_OPT_OUTPUT, {"exact": L.RKR_OUTPUT_EXACT, "reference": L.RKR_OUTPUT_REFERENCE}[output])

    # ---- data and table --------------------------------------------------------
    def set_data(self, users, items, normalize_users: bool = True):
        users, items = _f32(users), _f32(items)
        if users.ndim != 2 or items.ndim != 2 or users.shape[1] != items.shape[1]:
            raise ValueError("users and items must be 2-D with the same dimension")
        _check(self.lib, self.ctx, self.lib.rkr_set_data(self.ctx, _ptr(users), users.shape[0], _ptr(items),
                                                         items.shape[0], users.shape[1], int(normalize_users)))
        self.n_user, self.n_item, self.dim = users.shape[0], items.shape[0], users.shape[1]

    @staticmethod
    def uniform_sample_ranks(n_item: int, n_sample: int) -> np.ndarray:
        r = np.zeros(max(n_sample, 1), np.int32)
        rc = L.load().rkr_uniform_sample_ranks(n_item, n_sample, _ptr(r))
        if rc != L.RKR_OK:
            raise L.RkrError(rc, "n_sample must satisfy 0 < n_sample < n_item")
        return r[:n_sample]

    def build_table(self, ranks) -> float:
        ranks = np.ascontiguousarray(ranks, np.int32)
        ms = C.c_double(0.0)
        _check(self.lib, self.ctx, self.lib.rkr_build_table(self.ctx, _ptr(ranks), ranks.shape[0], C.byref(ms)))
        return ms.value

    def set_table(self, ranks, table, eps: float = 0.0):
        ranks = np.ascontiguousarray(ranks, np.int32)
        table = _f32(table)
        _check(self.lib, self.ctx, self.lib.rkr_set_table(self.ctx, _ptr(ranks), ranks.shape[0], _ptr(table), eps))

    def get_table(self, user_begin: int = 0, n_rows=None) -> np.ndarray:
        S = C.c_int32()
        _check(self.lib, self.ctx, self.lib.rkr_get_ranks(self.ctx, None, C.byref(S)))
        if n_rows is None:  # the library's count: updates may have come through a server
            self._refresh_counts()
            n_rows = self.n_user - user_begin
        out = np.empty((n_rows, S.value), np.float32)
        _check(self.lib, self.ctx, self.lib.rkr_get_table(self.ctx, user_begin, n_rows, _ptr(out)))
        return out

    def get_ranks(self) -> np.ndarray:
        S = C.c_int32()
        _check(self.lib, self.ctx, self.lib.rkr_get_ranks(self.ctx, None, C.byref(S)))
        r = np.empty(max(S.value, 1), np.int32)
        _check(self.lib, self.ctx, self.lib.rkr_get_ranks(self.ctx, _ptr(r), C.byref(S)))
        return r[:S.value]

    def save_table(self, path: str):
        _check(self.lib, self.ctx, self.lib.rkr_save_table(self.ctx, path.encode()))

    def load_table(self, path: str, eps: float = 0.0):
        _check(self.lib, self.ctx, self.lib.rkr_load_table(self.ctx, path.encode(), eps))

    # ---- full score table (ComputeScoreTableBatch.hpp:79-92, ReadAll.hpp:80-87) ----
    def sorted_scores(self, user_begin: int = 0, n_rows=None) -> np.ndarray:
        """Every item's score for users [user_begin, user_begin + n_rows), each row descending."""
        self._refresh_counts()
        n_rows = self.n_user - user_begin if n_rows is None else n_rows
        out = np.empty((n_rows, self.n_item), np.float32)
        _check(self.lib, self.ctx, self.lib.rkr_sorted_scores(self.ctx, user_begin, n_rows, _ptr(out)))
        return out

    def save_score_table(self, path: str, n_rows: int = -1) -> dict:
        """Writes ReadAll's index file (n_rows x n_item FP32, no header); returns the time split in ms."""
        t = np.zeros(4, np.float64)
        _check(self.lib, self.ctx, self.lib.rkr_save_score_table(self.ctx, path.encode(), n_rows,
                                                                   t.ctypes.data_as(C.c_void_p)))
        return dict(zip(("gemm_ms", "sort_ms", "compute_sort_ms", "save_ms"), t.tolist()))

    # ---- updates (BaseUpdateIndex, src/include/struct/MethodBaseUpdate.hpp:15-40) ----
    def insert_items(self, items):
        items = _f32(items)
        _check(self.lib, self.ctx, self.lib.rkr_insert_items(self.ctx, _ptr(items), items.shape[0]))
        self._refresh_counts()

    def delete_items(self, ids):
        ids = np.ascontiguousarray(ids, np.int64)
        _check(self.lib, self.ctx, self.lib.rkr_delete_items(self.ctx, _ptr(ids), ids.shape[0]))
        self._refresh_counts()

    def insert_users(self, users):
        users = _f32(users)
        _check(self.lib, self.ctx, self.lib.rkr_insert_users(self.ctx, _ptr(users), users.shape[0]))
        self._refresh_counts()

    def delete_users(self, ids):
        ids = np.ascontiguousarray(ids, np.int64)
        _check(self.lib, self.ctx, self.lib.rkr_delete_users(self.ctx, _ptr(ids), ids.shape[0]))
        self._refresh_counts()

    def reserve(self, n_user: int, n_inserted: int):
        """Sizes the update storage for n_user users and n_inserted pending
        inserted items, so updates within those sizes do not allocate."""
        _check(self.lib, self.ctx, self.lib.rkr_reserve(self.ctx, int(n_user), int(n_inserted)))

    def counts(self):
        v = [C.c_int64() for _ in range(4)]
        _check(self.lib, self.ctx, self.lib.rkr_get_counts(self.ctx, *[C.byref(x) for x in v]))
        return dict(zip(("n_user", "n_item", "n_inserted", "n_deleted"), (x.value for x in v)))

    def _refresh_counts(self):
        c = self.counts()
        self.n_user, self.n_item = c["n_user"], c["n_item"]

    def index_size_bytes(self) -> int:
        return int(self.lib.rkr_index_size_bytes(self.ctx))

    # ---- QSRP screen (QSRPNormalLP, src/impl/QSRPNormalLP.hpp:137-253) --------
    def qsrp_build(self, sigma: float = 0.7) -> float:
        """SVD rotation and per-user regression of the current users and table; returns host ms."""
        ms = C.c_double(0.0)
        _check(self.lib, self.ctx, self.lib.rkr_qsrp_build(self.ctx, float(sigma), C.byref(ms)))
        return ms.value

    def qsrp_info(self) -> dict:
        cd, nf = C.c_int32(), C.c_int64()
        _check(self.lib, self.ctx, self.lib.rkr_qsrp_info(self.ctx, C.byref(cd), C.byref(nf)))
        return {"check_dim": cd.value, "n_fallback_fit": nf.value}

    def qsrp_screen(self, queries) -> np.ndarray:
        """(n_query, n_user) uint32 bucket pairs b_lo | b_hi << 16 of the screen alone."""
        q = _f32(queries)
        if q.ndim == 1:
            q = q[None, :]
        self._refresh_counts()
        out = np.empty((q.shape[0], self.n_user), np.uint32)
        _check(self.lib, self.ctx, self.lib.rkr_qsrp_screen(self.ctx, _ptr(q), q.shape[0], out.ctypes.data))
        return out

    def qsrp_phi(self, z) -> np.ndarray:
        z = _f32(z).ravel()
        out = np.empty_like(z)
        _check(self.lib, self.ctx, self.lib.rkr_qsrp_phi(self.ctx, _ptr(z), z.shape[0], _ptr(out)))
        return out

    def query_table_cols(self) -> int:
        """Columns of the query table exact output searches (0: the sampled table)."""
        return int(self.lib.rkr_query_table_cols(self.ctx))

    # ---- queries ----------------------------------------------------------------
    def query(self, queries, topk, max_k=None, out=None) -> QueryResult:
        """Answers a batch.  ``out`` (a QueryResult of an earlier call with the
        same query count and max_k) is written over instead of allocating new
        arrays, which spares a serving loop the first-touch page faults of fresh
        output buffers on every call."""
        queries = _f32(queries)
        if queries.ndim == 1:
            queries = queries[None, :]
        nq = queries.shape[0]
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(topk, np.int32), (nq,)), np.int32)
        mk = int(max_k if max_k is not None else (ks.max() if nq else 1))
        if out is not None and out.users.shape == (nq, mk) and len(out.stats) == nq and nq > 0:
            users, ranks, qips, stats = out.users, out.ranks, out.qips, out.stats
        else:
            users = np.empty((nq, mk), np.int32)
            ranks = np.empty((nq, mk), np.int32)
            qips = np.empty((nq, mk), np.float32)
            # per-query stats land in a structured array with rkr_query_stats'
            # layout (indexable as stats[i]["n_candidate"]; no per-query Python
            # objects)
            stats = np.empty(max(nq, 1), L.QUERY_STATS_DTYPE)
        tm = L.BatchTiming()
        _check(self.lib, self.ctx, self.lib.rkr_query(self.ctx, _ptr(queries), nq, _ptr(ks), mk, _ptr(users),
                                                      _ptr(ranks), _ptr(qips),
                                                      stats.ctypes.data_as(C.POINTER(L.QueryStats)), C.byref(tm)))
        timing = {f: getattr(tm, f) for f, _ in L.BatchTiming._fields_}
        return QueryResult(users, ranks, qips, stats[:nq], timing)

    def retrieval(self, query_item, topk: int, n_execute_query=None):
        """US::Index::Retrieval shape: list (per query) of topk UserRankElement, plus per-query stats."""
        q = _f32(query_item)
        n = q.shape[0] if n_execute_query is None else n_execute_query
        if n > q.shape[0]:
            raise ValueError("n_execute_query larger than n_query_item")
        if topk > self.n_user:
            raise ValueError("top-k is too large")
        res = self.query(q[:n], topk)
        return [res.elements(i) for i in range(n)], res.stats


def build_index(data_item, user, n_sample: int, device: int = 0, table_path=None, normalize_users=True) -> Index:
    """US::BuildIndex counterpart: loads the table from ``table_path`` (the
    reference's memory_index file) when given, else builds it on the device."""
    idx = Index(device)
    idx.set_data(user, data_item, normalize_users=normalize_users)
    if table_path:
        idx.load_table(table_path)
    else:
        idx.build_table(Index.uniform_sample_ranks(idx.n_item, n_sample))
    return idx
