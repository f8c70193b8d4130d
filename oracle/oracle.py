"""ctypes view of the CPU oracle (liborc.so) and, when built, of the reference
routines compiled into oracle/_ref/libref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_PATH = os.path.join(HERE, "liborc.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_i64 = C.c_int64
_i32 = C.c_int32


class QStat(C.Structure):
    _fields_ = [("n_prune_user", C.c_int32), ("n_result_user", C.c_int32),
                ("n_refine_user", C.c_int32), ("n_candidate", C.c_int32),
                ("ip_cost", C.c_int64)]


def build(force: bool = False) -> None:
    """Compile liborc.so (and oracle/_ref when the reference is present)."""
    if force or not os.path.exists(ORC_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)


_orc = None
_ref = None


def lib():
    global _orc
    if _orc is None:
        build()
        L = C.CDLL(ORC_PATH)
        L.orc_normalize_users.argtypes = [_f32p, _i64, _i32]
        L.orc_ip_f32.argtypes = [_f32p, _f32p, _i32]
        L.orc_ip_f32.restype = C.c_float
        L.orc_ip_f64.argtypes = [_f32p, _f32p, _i32]
        L.orc_ip_f64.restype = C.c_double
        L.orc_uniform_sample.argtypes = [_i64, _i32, _i32p]
        L.orc_uniform_sample.restype = C.c_int
        L.orc_score_rows.argtypes = [_f32p, _i64, _i64, _f32p, _i64, _i32, _f32p]
        L.orc_build_table.argtypes = [_f32p, _i64, _f32p, _i64, _i32, _i32p, _i32, _f32p]
        L.orc_coarse_bsearch.argtypes = [_f32p, _i32, _i32p, _i64, C.c_float,
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.orc_safe_bounds.argtypes = [_f32p, _i32, _i32p, _i64, C.c_float, C.c_float,
                                      C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.orc_prune.argtypes = [_i32p, _i32p, _i64, _i32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                C.POINTER(C.c_int32), _i8p, _i8p]
        L.orc_refine_rank.argtypes = [_f32p, _f32p, _i64, _i32, C.c_float]
        L.orc_refine_rank.restype = C.c_int32
        L.orc_refine_ranks.argtypes = [_f32p, _f32p, _f32p, _i64, _i32, C.c_void_p, C.c_void_p, _i64, _i32p,
                                       C.c_void_p]
        L.orc_us_retrieval.argtypes = [_f32p, _i64, _f32p, _i64, _i32, _f32p, _i32p, _i32, _f32p, _i32, _i32,
                                       _i32p, _i32p, _f32p, C.c_void_p]
        L.orc_us_retrieval.restype = C.c_int
        L.orc_exact_query.argtypes = [_f32p, _i64, _f32p, _i64, _i32, _f32p, _i32p, _i32, C.c_float, _f32p,
                                      _i32, _i32p, _i32p, _f32p, C.c_void_p]
        L.orc_exact_query.restype = C.c_int
        L.orc_bruteforce.argtypes = [_f32p, _i64, _f32p, _i64, _i32, _f32p, _i32, _i32p, _i32p, _f32p]
        L.orc_bruteforce.restype = C.c_int
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        _orc = L
    return _orc


def ref_lib():
    """The reference's own routines (oracle/_ref), or None when not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        L = C.CDLL(REF_PATH)
        L.ref_inner_product.argtypes = [_f32p, _f32p, C.c_int]
        L.ref_inner_product.restype = C.c_float
        L.ref_refine_rank.argtypes = [_f32p, _f32p, _i64, _i64, C.c_int, C.c_float, C.c_int]
        L.ref_refine_rank.restype = C.c_int
        L.ref_prune.argtypes = [_i32p, _i32p, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_int), _i8p, _i8p]
        L.ref_score_list.argtypes = [_f32p, _f32p, _i64, _i64, C.c_int, C.c_int, C.c_int, _f32p]
        L.ref_normalize.argtypes = [_f32p, C.c_int, C.c_int]
        _ref = L
    return _ref


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def normalize_users(users: np.ndarray) -> np.ndarray:
    u = _c32(users).copy()
    lib().orc_normalize_users(u, u.shape[0], u.shape[1])
    return u


def ip_f32(a, b) -> np.float32:
    a, b = _c32(a), _c32(b)
    return np.float32(lib().orc_ip_f32(a, b, a.shape[0]))


def ip_f64(a, b) -> float:
    a, b = _c32(a), _c32(b)
    return lib().orc_ip_f64(a, b, a.shape[0])


def uniform_sample(n_item: int, n_sample: int) -> np.ndarray:
    r = np.zeros(max(n_sample, 1), np.int32)
    if lib().orc_uniform_sample(n_item, n_sample, r) != 0:
        raise ValueError("n_sample must satisfy 0 < n_sample < n_item")
    return r[:n_sample]


def score_rows(users, items, u0=0, nu=None) -> np.ndarray:
    users, items = _c32(users), _c32(items)
    nu = users.shape[0] - u0 if nu is None else nu
    out = np.empty((nu, items.shape[0]), np.float32)
    lib().orc_score_rows(users, u0, nu, items, items.shape[0], users.shape[1], out)
    return out


def sorted_score_rows(users, items, u0=0, nu=None) -> np.ndarray:
    """Each user's scores against every item, sorted in descending order: the
    CPU branch of ComputeScoreTableBatch::ComputeSortItemsBatch
    (src/include/score_computation/ComputeScoreTableBatch.hpp:79-92), the rows
    ReadAll::BuildIndexLoop writes (src/include/alg/DiskIndex/ReadAll.hpp:80-87)."""
    return -np.sort(-score_rows(users, items, u0, nu), axis=1)


def build_table(users, items, ranks) -> np.ndarray:
    users, items = _c32(users), _c32(items)
    ranks = np.ascontiguousarray(ranks, np.int32)
    T = np.empty((users.shape[0], ranks.shape[0]), np.float32)
    lib().orc_build_table(users, users.shape[0], items, items.shape[0], users.shape[1], ranks,
                          ranks.shape[0], T)
    return T


def coarse_bsearch(row, ranks, n_item, qip):
    lb, ub = C.c_int32(), C.c_int32()
    row = _c32(row)
    lib().orc_coarse_bsearch(row, row.shape[0], np.ascontiguousarray(ranks, np.int32), n_item,
                             float(np.float32(qip)), C.byref(lb), C.byref(ub))
    return lb.value, ub.value


def safe_bounds(row, ranks, n_item, qip, eps=0.0):
    lb, ub = C.c_int32(), C.c_int32()
    row = _c32(row)
    lib().orc_safe_bounds(row, row.shape[0], np.ascontiguousarray(ranks, np.int32), n_item,
                          float(np.float32(qip)), float(eps), C.byref(lb), C.byref(ub))
    return lb.value, ub.value


def prune(lb, ub, topk, prune_l=None, result_l=None, n_result=0, n_prune=0, refine=None):
    lb = np.ascontiguousarray(lb, np.int32)
    ub = np.ascontiguousarray(ub, np.int32)
    n = lb.shape[0]
    p = np.zeros(n, np.int8) if prune_l is None else np.ascontiguousarray(prune_l, np.int8).copy()
    r = np.zeros(n, np.int8) if result_l is None else np.ascontiguousarray(result_l, np.int8).copy()
    rs, nr, npr = C.c_int32(n if refine is None else refine), C.c_int32(n_result), C.c_int32(n_prune)
    lib().orc_prune(lb, ub, n, topk, C.byref(rs), C.byref(nr), C.byref(npr), p, r)
    return dict(refine=rs.value, n_result=nr.value, n_prune=npr.value, prune=p, result=r)


def refine_rank(user, items, qip) -> int:
    user, items = _c32(user), _c32(items)
    return int(lib().orc_refine_rank(user, items, items.shape[0], items.shape[1], float(np.float32(qip))))


def refine_ranks(users, queries, items, user_id, q_id):
    """refine_rank of many (user, query) pairs on every oracle thread: returns
    (ranks, <u,q>) with the query's InnerProduct computed as the reference's
    GEMV loop does (US.hpp:124-136)."""
    users, queries, items = _c32(users), _c32(queries), _c32(items)
    uid = np.ascontiguousarray(user_id, np.int64)
    qid = np.ascontiguousarray(q_id, np.int64)
    n = uid.shape[0]
    out = np.empty(n, np.int32)
    qip = np.empty(n, np.float32)
    lib().orc_refine_ranks(users, queries, items, items.shape[0], users.shape[1], uid.ctypes.data,
                           qid.ctypes.data, n, out, qip.ctypes.data)
    return out, qip


def _query_out(nq, k):
    return (np.empty((nq, k), np.int32), np.empty((nq, k), np.int32), np.empty((nq, k), np.float32))


def us_retrieval(users, items, table, ranks, queries, topk, stats=False):
    """The reference US query path (US.hpp:96-213) over every query."""
    users, items, table, queries = _c32(users), _c32(items), _c32(table), _c32(queries)
    ranks = np.ascontiguousarray(ranks, np.int32)
    if table.shape != (users.shape[0], ranks.shape[0]):
        raise ValueError(f"table {table.shape} does not match {users.shape[0]} users x {ranks.shape[0]} ranks")
    nq = queries.shape[0]
    ou, orank, oq = _query_out(nq, topk)
    st = (QStat * max(nq, 1))()
    rc = lib().orc_us_retrieval(users, users.shape[0], items, items.shape[0], users.shape[1], table, ranks,
                                ranks.shape[0], queries, nq, topk, ou, orank, oq, C.cast(st, C.c_void_p))
    if rc != 0:
        raise ValueError("invalid topk")
    if stats:
        return ou, orank, oq, [dict((f, getattr(s, f)) for f, _ in QStat._fields_) for s in st[:nq]]
    return ou, orank, oq


def exact_query(users, items, table, ranks, queries, topk, eps=0.0, stats=False):
    """Exact reverse k-ranks through the sampled table, every query."""
    users, items, table, queries = _c32(users), _c32(items), _c32(table), _c32(queries)
    ranks = np.ascontiguousarray(ranks, np.int32)
    if table.shape != (users.shape[0], ranks.shape[0]):
        raise ValueError(f"table {table.shape} does not match {users.shape[0]} users x {ranks.shape[0]} ranks")
    nq = queries.shape[0]
    ks = np.broadcast_to(np.asarray(topk, np.int32), (nq,))
    kmax = int(ks.max()) if nq else 0
    ou = np.full((nq, kmax), -1, np.int32)
    orank = np.full((nq, kmax), -1, np.int32)
    oq = np.full((nq, kmax), -1, np.float32)
    sts = []
    for i in range(nq):
        k = int(ks[i])
        u_, r_, q_ = (np.empty(k, np.int32), np.empty(k, np.int32), np.empty(k, np.float32))
        st = QStat()
        rc = lib().orc_exact_query(users, users.shape[0], items, items.shape[0], users.shape[1], table, ranks,
                                   ranks.shape[0], float(eps), np.ascontiguousarray(queries[i]), k, u_, r_, q_,
                                   C.cast(C.pointer(st), C.c_void_p))
        if rc != 0:
            raise ValueError("invalid topk")
        ou[i, :k], orank[i, :k], oq[i, :k] = u_, r_, q_
        sts.append(dict((f, getattr(st, f)) for f, _ in QStat._fields_))
    if stats:
        return ou, orank, oq, sts
    return ou, orank, oq


def bruteforce(users, items, queries, topk):
    """Rank every user (strict count), order by (rank, userID), keep topk."""
    users, items, queries = _c32(users), _c32(items), _c32(queries)
    nq = queries.shape[0]
    ou, orank, oq = _query_out(nq, topk)
    for i in range(nq):
        u_, r_, q_ = (np.empty(topk, np.int32), np.empty(topk, np.int32), np.empty(topk, np.float32))
        rc = lib().orc_bruteforce(users, users.shape[0], items, items.shape[0], users.shape[1],
                                  np.ascontiguousarray(queries[i]), topk, u_, r_, q_)
        if rc != 0:
            raise ValueError("invalid topk")
        ou[i], orank[i], oq[i] = u_, r_, q_
    return ou, orank, oq


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ---- QS sample-rank training ---------------------------------------------------

def query_rank_distribution(users, items, sample_ids) -> np.ndarray:
    """accu[s][r] = #{u : rank(u, s) <= r} with rank(u, s) = #{p : <u,p> > <u, item_s>}:
    ComputeQueryRankByComputation (src/app/build-index/QueryRankByComputation.hpp:122-242)
    with QueryRankCPU's lower_bound over the descending row (QueryRankCPU.hpp:101-108):
    #{p : s_p > x} = n_item - #{p : s_p <= x}, searched in the ascending row."""
    S = score_rows(users, items)
    ids = np.asarray(sample_ids, np.int64)
    n_item = items.shape[0]
    ranks = np.empty((S.shape[0], len(ids)), np.int64)
    for u in range(S.shape[0]):
        asc = np.sort(S[u])
        ranks[u] = n_item - np.searchsorted(asc, S[u, ids], side="right")
    accu = np.zeros((len(ids), n_item + 1), np.int64)
    for s in range(len(ids)):
        accu[s] = np.cumsum(np.bincount(ranks[:, s], minlength=n_item + 1))
    return accu.astype(np.int32)


def sort_kth_rank(accu, topk):
    """ComputeSortKthRank (src/app/build-index/QueryRankBySample.hpp:103-134): the first r with
    accu[s][r] >= topk, samples ordered by it (ties in index order here)."""
    kth = np.array([int(np.searchsorted(row, topk, side="left")) for row in np.asarray(accu)], np.int64)
    order = np.argsort(kth, kind="stable")
    return kth[order].astype(np.int32), order.astype(np.int32)


def _fill_ranks(rank_set, n_sample, last_sample, n_item):
    """The fill of FindOptimalSampleFast (OptimalSampleRankFast.hpp:121-141, 245-266)."""
    n_remain = n_sample - len(rank_set)
    if n_remain <= 0:
        return rank_set
    delta = last_sample * 1.0 / n_remain
    x = 0.0
    while len(rank_set) < n_sample and int(x) < last_sample:
        if int(x) not in rank_set:
            rank_set.add(int(x))
            x += delta
        else:
            x += 1
    r = 0
    while len(rank_set) < n_sample and r < n_item:
        rank_set.add(r)
        r += 1
    return rank_set


def candidate_ranks(n_item, sort_kth, method):
    """QueryDistribution's candidate ranks (OptimalSampleRank.hpp:48-91): method 0 is OptimalPart,
    1 OptimalUniform."""
    sort_kth = [int(v) for v in sort_kth]
    if method == 0:
        rs = set()
        for k in sort_kth:
            rs.add(k)
            if k != 0:
                rs.add(k - 1)
        return sorted(rs)
    Q = len(sort_kth)
    delta = (n_item - 1) * 1.0 / Q
    return sorted(int(np.floor(s * delta)) for s in range(Q))


def find_optimal_sample_fast(n_user, n_item, sort_kth, accu, method, n_sample):
    """FindOptimalSampleFast (OptimalSampleRankFast.hpp:102-305) with QueryDistribution's costs
    computed by its own loops over the queries (OptimalSampleRank.hpp:206-318).  Returns
    (ranks, train_io_cost); pure Python, small inputs only."""
    sort_kth = [int(v) for v in sort_kth]
    accu = np.asarray(accu, np.int64)
    Q = len(sort_kth)
    rank = candidate_ranks(n_item, sort_kth, method)
    R = len(rank)
    if R <= n_sample:
        rs = _fill_ranks(set(rank), n_sample, n_item if Q < 500 else rank[-1], n_item)
        return sorted(rs), 0
    import bisect
    last_pos = [bisect.bisect_right(sort_kth, rank[j]) for j in range(R)]      # first q with kth > rank
    next_first = [bisect.bisect_left(sort_kth, rank[j] + 1) for j in range(R)]  # first q with kth >= rank+1

    def begin(j):
        return sum(int(accu[q][rank[j]]) for q in range(last_pos[j]))

    def end(j):
        r1 = rank[j] + 1
        return sum(int(n_user - accu[q][r1 - 1]) for q in range(next_first[j], Q))

    def between(ub, lb):
        r_ub, r_lb = rank[ub] + 1, rank[lb]
        qb, qe = next_first[ub], last_pos[lb]
        if qb >= Q or qb > qe:
            return 0
        return sum(int(accu[q][r_lb] - accu[q][r_ub - 1]) for q in range(qb, qe))

    E = [end(j) for j in range(R)]
    dp = [[0] * R for _ in range(n_sample)]
    pos = [[0] * R for _ in range(n_sample)]
    for r in range(R):
        dp[0][r] = begin(r) + E[r]
        pos[0][r] = r
    T = {}
    for lb in range(1, R):
        for ub in range(lb):
            T[(ub, lb)] = -E[ub] + between(ub, lb) + E[lb]
    for s in range(1, n_sample):
        for r in range(R):
            dp[s][r] = dp[s - 1][r]
            pos[s][r] = pos[s - 1][r]
            for p in range(r):
                c = dp[s - 1][p] + T[(p, r)]
                if dp[s][r] > c:
                    dp[s][r] = c
                    pos[s][r] = p
    best, idx = None, -1
    for r in range(R):
        if best is None or dp[n_sample - 1][r] < best:
            best, idx = dp[n_sample - 1][r], r
    out = [0] * n_sample
    for s in range(n_sample - 1, -1, -1):
        out[s] = rank[idx]
        idx = pos[s][idx]
    if len(set(out)) != n_sample:
        out = sorted(_fill_ranks(set(out), n_sample, rank[-1], n_item))
    return out, int(best)


def sample_set_cost(n_user, sort_kth, accu, ranks):
    """The quantity the DP minimises, from its meaning: for every query, the number of users whose
    rank lies in the bucket of the sorted sample ranks that holds the query's k-th rank
    ([0, r_0], [r_0 + 1, r_1], ..., [r_last + 1, n_item])."""
    accu = np.asarray(accu, np.int64)
    rs = sorted(set(int(r) for r in ranks))
    import bisect
    tot = 0
    for q, k in enumerate(sort_kth):
        i = bisect.bisect_left(rs, int(k))  # first sample rank >= k
        if i == 0:
            tot += int(accu[q][rs[0]])
        elif i == len(rs):
            tot += int(n_user - accu[q][rs[-1]])
        else:
            tot += int(accu[q][rs[i]] - accu[q][rs[i - 1]])
    return tot


# ---- QSRP screen (QSRPNormalLP: SVD + PartDimPartNorm + NormalLinearRegressionLP) ----

def cdf_phi(x: float) -> float:
    """CDFPhi, src/include/alg/RegressionPruning/NormalLinearRegressionLP.hpp:125-147:
    Abramowitz & Stegun 7.1.26 for erf (absolute error <= 1.5e-7), in double."""
    import math
    a1, a2, a3, a4, a5, p = 0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429, 0.3275911
    sign = -1.0 if x < 0 else 1.0
    z = abs(x) / math.sqrt(2.0)
    t = 1.0 / (1.0 + p * z)
    y = 1.0 - (((((a5 * t + a4) * t) + a3) * t + a2) * t + a1) * t * math.exp(-z * z)
    return 0.5 * (1.0 + sign * y)


def regression_points(T, mu, inv_sigma):
    """The constraint points of NormalLinearRegressionLP::AssignCache
    (NormalLinearRegressionLP.hpp:149-181): for sample interval j, sample position j + 1
    against the CDF at both ends of the interval."""
    xs, ys = [], []
    for j in range(len(T) - 1):
        for v in (T[j], T[j + 1]):
            xs.append(cdf_phi(float(np.float32(np.float32(v) - np.float32(mu)) * np.float32(inv_sigma))))
            ys.append(float(j + 1))
    return np.array(xs), np.array(ys)


def lp_min_error_bruteforce(x, y):
    """The LP of NormalLinearRegressionLP::LinearProgram (:183-213): minimise e over (a, b, e)
    subject to |a x_i + b - y_i| <= e, by enumerating every vertex (three tight constraints,
    a 3x3 solve) and keeping the feasible one with the smallest e.  Brute force for small n."""
    import itertools
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    # a constraint is tight when a x_i + b - y_i = s e for s = +1 or -1:
    # the row (x_i, 1, -s) . (a, b, e) = y_i
    rows = [(xi, 1.0, -sgn, yi) for xi, yi in zip(x, y) for sgn in (1.0, -1.0)]
    best = np.inf
    for r1, r2, r3 in itertools.combinations(range(len(rows)), 3):
        A = np.array([rows[r][:3] for r in (r1, r2, r3)])
        rhs = np.array([rows[r][3] for r in (r1, r2, r3)])
        if abs(np.linalg.det(A)) < 1e-12:
            continue
        a, b, e = np.linalg.solve(A, rhs)
        if e < -1e-12:
            continue
        if np.all(np.abs(a * x + b - y) <= e + 1e-9 * (1 + abs(e))):
            best = min(best, e)
    return best
