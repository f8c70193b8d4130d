"""Writes the golden vectors in tests/golden/ from the reference's own code.

Run in a container where /root/reference exists (oracle/_ref is built from
its headers by oracle/Makefile).  Outputs:

* refine_vectors.npz -- (user, query, <q,u>, rank) for 300 random pairs of the
  fake-normal-30d fixture: <q,u> from the reference's InnerProduct
  (SpaceInnerProduct.hpp:14-20) and rank from RefineCPU::RefineRank
  (RefineCPU.hpp:33-55).
* prune_vectors.npz -- random rank-bound sets and PruneCandidateByBound's
  flags and counters (PruneCandidateByBound.hpp:19-179).
* fake30d_us_top10.npz -- for the 100 fixture queries, k=10, S=128:
  the exact answer ranked with the reference's RefineRank for every user and
  ordered by (rank, userID) (UserRankElement.hpp:59-64); the US-path answer
  of the oracle's restatement (a regression vector: its pieces are pinned by
  the two files above); and a checksum plus the first rows of the table.
"""
import ctypes as C
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "reverse-k-ranks_amd"))

import oracle  # noqa: E402
from rkr import formats  # noqa: E402


def main():
    oracle.build(force=True)
    R = oracle.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref is not built: the reference sources are needed")
    users, items, queries = formats.read_dataset(HERE, "fake-normal-30d")
    un = users.copy()
    R.ref_normalize(un, un.shape[0], un.shape[1])
    U, I, d = un.shape[0], items.shape[0], un.shape[1]
    rng = np.random.default_rng(20240601)

    # refine vectors
    uu = rng.integers(0, U, 300).astype(np.int32)
    qq = rng.integers(0, queries.shape[0], 300).astype(np.int32)
    qip = np.array([R.ref_inner_product(np.ascontiguousarray(queries[q]), np.ascontiguousarray(un[u]), d)
                    for u, q in zip(uu, qq)], np.float32)
    rank = np.array([R.ref_refine_rank(un, items, U, I, d, float(ip), int(u)) for u, ip in zip(uu, qip)], np.int32)
    np.savez(os.path.join(HERE, "refine_vectors.npz"), user=uu, query=qq, qip=qip, rank=rank)

    # prune vectors
    out = {}
    n_cases = 60
    for i in range(n_cases):
        n = int(rng.integers(1, 80))
        k = int(rng.integers(1, n + 1))
        ub = rng.integers(0, 40, n).astype(np.int32)
        lb = (ub + rng.integers(0, 3, n) * rng.integers(0, 15, n)).astype(np.int32)
        p = np.zeros(n, np.int8)
        r = np.zeros(n, np.int8)
        rs, nr, npr = C.c_int(n), C.c_int(0), C.c_int(0)
        R.ref_prune(lb, ub, n, k, C.byref(rs), C.byref(nr), C.byref(npr), p, r)
        out.update({f"lb{i}": lb, f"ub{i}": ub, f"k{i}": np.int32(k), f"prune{i}": p, f"result{i}": r,
                    f"counts{i}": np.array([rs.value, nr.value, npr.value], np.int32)})
    np.savez(os.path.join(HERE, "prune_vectors.npz"), n=np.int32(n_cases), **out)

    # exact answers via the reference's RefineRank for every user
    k = 10
    nq = queries.shape[0]
    ex_u = np.empty((nq, k), np.int32)
    ex_r = np.empty((nq, k), np.int32)
    for qi in range(nq):
        q = np.ascontiguousarray(queries[qi])
        ranks_all = np.empty(U, np.int64)
        for u in range(U):
            ip = R.ref_inner_product(q, np.ascontiguousarray(un[u]), d)
            ranks_all[u] = R.ref_refine_rank(un, items, U, I, d, float(ip), u)
        order = np.lexsort((np.arange(U), ranks_all))[:k]
        ex_u[qi], ex_r[qi] = order, ranks_all[order]
    ranks = oracle.uniform_sample(I, 128)
    T = oracle.build_table(un, items, ranks)
    us_u, us_r, _ = oracle.us_retrieval(un, items, T, ranks, queries, k)
    sha = hashlib.sha256(np.ascontiguousarray(T + np.float32(0)).tobytes()).hexdigest()
    np.savez(os.path.join(HERE, "fake30d_us_top10.npz"), exact_user=ex_u, exact_rank=ex_r, us_user=us_u,
             us_rank=us_r, table_sha256=np.array(sha), table_head=T[:8])
    print("golden vectors written")


if __name__ == "__main__":
    main()
