"""GPU test of rkr_sorted_scores_noalloc, the entry point that allocates
nothing: its rows equal rkr_sorted_scores' on two calls in a row (the second
smaller), across scratch-sized chunks, and oversize rows are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree through the HIP runtime the library uses."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]

    def malloc(self, n):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), n) == 0
        return p

    def to_host(self, dev, shape):
        out = np.empty(shape, np.float32)
        assert self.h.hipMemcpy(out.ctypes.data, dev, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out

    def free(self, p):
        assert self.h.hipFree(p) == 0


def test_sorted_scores_noalloc_matches_sorted_scores():
    import rkr
    from rkr import _lib as L
    hip = _Hip()
    rng = np.random.default_rng(71)
    U, I, d = 3000, 7000, 40
    users = rng.standard_normal((U, d), dtype=np.float32)
    items = rng.standard_normal((I, d), dtype=np.float32) * rng.lognormal(0, 0.5, (I, 1)).astype(np.float32)
    items[10:20] = items[0]  # tied scores
    users[5] = 0.0           # a row of zeros
    with rkr.Index(0) as idx:
        idx.set_data(users, items)
        lib = idx.lib
        # 1300 rows of 7000 scores: 36 MB, more than the 32 MiB scratch, so two chunks
        for u0, n in ((100, 1300), (2500, 37)):
            _case(f"noalloc rows [{u0}, {u0 + n}) of {I} items")
            buf = hip.malloc(n * I * 4)
            try:
                assert lib.rkr_sorted_scores_noalloc(idx.ctx, u0, n, buf) == L.RKR_OK
                assert lib.rkr_sync(idx.ctx) == L.RKR_OK
                got = hip.to_host(buf, (n, I))
            finally:
                hip.free(buf)
            assert np.array_equal(got, idx.sorted_scores(u0, n)), f"rows [{u0}, {u0 + n}) differ"
        assert lib.rkr_sorted_scores_noalloc(idx.ctx, 0, 0, None) == L.RKR_OK
        assert lib.rkr_sorted_scores_noalloc(idx.ctx, U - 1, 2, C.c_void_p(1)) == L.RKR_ERR_INVALID
    # a second context on the same device shares the array and its address
    with rkr.Index(0) as idx2:
        idx2.set_data(users[:64], items[:900])
        buf = hip.malloc(64 * 900 * 4)
        try:
            assert idx2.lib.rkr_sorted_scores_noalloc(idx2.ctx, 0, 64, buf) == L.RKR_OK
            assert idx2.lib.rkr_sync(idx2.ctx) == L.RKR_OK
            got = hip.to_host(buf, (64, 900))
        finally:
            hip.free(buf)
        assert np.array_equal(got, idx2.sorted_scores(0, 64))


def test_sorted_scores_noalloc_refuses_rows_beyond_scratch():
    import rkr
    from rkr import _lib as L
    _case("noalloc refuses a row wider than its scratch")
    rng = np.random.default_rng(72)
    with rkr.Index(0) as idx:
        idx.set_data(rng.standard_normal((4, 1), dtype=np.float32),
                     rng.standard_normal(((8 << 20) + 1, 1), dtype=np.float32))
        rc = idx.lib.rkr_sorted_scores_noalloc(idx.ctx, 0, 1, C.c_void_p(1))
        assert rc == L.RKR_ERR_INVALID
        assert b"exceeds" in idx.lib.rkr_last_error(idx.ctx)
