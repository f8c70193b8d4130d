"""Python handle of the serving loop (rkr_server_* in include/rkr.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class Server:
    """Continuous-batching request loop on an Index's device context."""

    def __init__(self, index, max_batch: int = 256, n_slots: int = 2, update_mode: str = "barrier",
                 update_cus: int = 0, rebuild_pending: int = 0):
        self.index = index
        self.lib = index.lib
        mode = {"barrier": L.RKR_SERVE_BARRIER, "snapshot": L.RKR_SERVE_SNAPSHOT}[update_mode]
        cfg = L.ServerConfig(max_batch, n_slots, mode, update_cus, rebuild_pending)
        h = C.c_void_p()
        rc = self.lib.rkr_server_start(index.ctx, C.byref(cfg), C.byref(h))
        if rc != L.RKR_OK:
            raise L.RkrError(rc, self.lib.rkr_last_error(index.ctx).decode())
        self.h = h
        self.dim = index.dim

    def submit_query(self, q, k: int) -> int:
        q = np.ascontiguousarray(q, np.float32)
        t = C.c_uint64()
        rc = self.lib.rkr_server_submit_query(self.h, q.ctypes.data, int(k), C.byref(t))
        if rc != L.RKR_OK:
            raise L.RkrError(rc, "submit_query")
        return t.value

    def submit_update(self, kind: str, vectors=None, ids=None) -> int:
        code = {"insert_users": L.RKR_UPDATE_INSERT_USERS, "insert_items": L.RKR_UPDATE_INSERT_ITEMS,
                "delete_users": L.RKR_UPDATE_DELETE_USERS, "delete_items": L.RKR_UPDATE_DELETE_ITEMS}[kind]
        t = C.c_uint64()
        if vectors is not None:
            v = np.ascontiguousarray(vectors, np.float32)
            rc = self.lib.rkr_server_submit_update(self.h, code, v.ctypes.data, None, v.shape[0], C.byref(t))
        else:
            i = np.ascontiguousarray(ids, np.int64)
            rc = self.lib.rkr_server_submit_update(self.h, code, None, i.ctypes.data, i.shape[0], C.byref(t))
        if rc != L.RKR_OK:
            raise L.RkrError(rc, "submit_update")
        return t.value

    def wait(self, ticket: int, k: int = 0):
        """Returns (users, ranks, qips, info dict); users/ranks are empty for updates."""
        u = np.empty(max(k, 1), np.int32)
        r = np.empty(max(k, 1), np.int32)
        q = np.empty(max(k, 1), np.float32)
        info = L.RequestInfo()
        rc = self.lib.rkr_server_wait(self.h, ticket, u.ctypes.data, r.ctypes.data, q.ctypes.data, C.byref(info))
        d = {"status": info.status, "batch_size": info.batch_size, "t_submit": info.t_submit,
             "t_start": info.t_start, "t_done": info.t_done, "error": info.error.decode(errors="replace"),
             "n_candidate": info.stats.n_candidate, "n_refine_user": info.stats.n_refine_user}
        if rc != L.RKR_OK:
            raise L.RkrError(rc, d["error"])
        return u[:k], r[:k], q[:k], d

    def counters(self):
        v = [C.c_int64() for _ in range(3)]
        self.lib.rkr_server_counters(self.h, *[C.byref(x) for x in v])
        return dict(zip(("n_batches", "n_queries", "n_updates"), (x.value for x in v)))

    def rebuilds(self) -> int:
        """Table rebuilds the loop made when pending item updates reached rebuild_pending."""
        v = C.c_int64()
        self.lib.rkr_server_rebuilds(self.h, C.byref(v))
        return v.value

    def now(self) -> float:
        return float(self.lib.rkr_now())

    def stop(self):
        if self.h:
            self.lib.rkr_server_stop(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()
