"""The coarse first pass's error bound, checked on the CPU.

The first pass estimates s(u, q) by one BF16 product per term, u_hi . q_hi with
FP32 sums, and decides buckets from s^ -+ e with

    e = ||u|| * cs(q) + floor,
    cs = (1 + rho) (g1 ||q_hi|| + ||q_r||) + rho (||q_hi|| + ||q_r||) + g_rule ||q||,

rho >= ||u - bf16(u)|| / ||u|| over the users (csrc/k_screen.hip, k_prep_items
and k_user_norms).  Here s is the reference's sequential FP32 score
(src/include/alg/SpaceInnerProduct.hpp:14-20).  The bound must hold for every
pair, whatever order the FP32 sums of the products take, and it is meant to be
well inside the 2^-7 ||u|| ||q|| of the worst-case per-term bound it replaces.
"""
import numpy as np


def bf16_rne(x):
    b = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def rule_scores(U, Q):
    """Sequential FP32 rule: s = s + u_k * q_k, k ascending, each step rounded."""
    s = np.zeros((U.shape[0], Q.shape[0]), dtype=np.float32)
    for k in range(U.shape[1]):
        s = (s + np.outer(U[:, k], Q[:, k]).astype(np.float32)).astype(np.float32)
    return s


def seq_fp32(A, B):
    s = np.zeros((A.shape[0], B.shape[0]), dtype=np.float32)
    for k in range(A.shape[1]):
        s = (s + np.outer(A[:, k], B[:, k]).astype(np.float32)).astype(np.float32)
    return s


def error_scale(U, Q, kp):
    e23, e24 = 2.0 ** -23, 2.0 ** -24
    d = U.shape[1]
    g1 = kp * e23 / (1 - kp * e23)
    g_rule = d * e24 / (1 - d * e24)
    U64, Q64 = U.astype(np.float64), Q.astype(np.float64)
    ur = (U - bf16_rne(U)).astype(np.float64)
    nu = np.sqrt((U64 ** 2).sum(1))
    rho = float(np.max(np.sqrt((ur ** 2).sum(1)) / nu)) * (1 + 2.0 ** -30)
    qh = bf16_rne(Q).astype(np.float64)
    qr = (Q - bf16_rne(Q)).astype(np.float64)
    nh, nr, nq = np.sqrt((qh ** 2).sum(1)), np.sqrt((qr ** 2).sum(1)), np.sqrt((Q64 ** 2).sum(1))
    cs = (1 + rho) * (g1 * nh + nr) + rho * (nh + nr) + g_rule * nq
    return nu, cs, rho


def test_one_product_bound_holds_and_is_tighter():
    rng = np.random.default_rng(5)
    d, kp = 150, 160
    U = (rng.standard_normal((300, d)) * rng.lognormal(0, 0.5, (300, 1))).astype(np.float32)
    Q = (rng.standard_normal((40, d)) / np.sqrt(d) * rng.lognormal(0, 0.5, (40, 1))).astype(np.float32)
    s = rule_scores(U, Q)
    Uh, Qh = bf16_rne(U), bf16_rne(Q)
    nu, cs, rho = error_scale(U, Q, kp)
    e = np.outer(nu, cs)
    # the products' exact sum, and two FP32 summation orders of them
    exact = Uh.astype(np.float64) @ Qh.astype(np.float64).T
    fwd = seq_fp32(Uh, Qh)
    rev = seq_fp32(Uh[:, ::-1], Qh[:, ::-1])
    for est in (exact, fwd, rev):
        assert np.all(np.abs(est - s.astype(np.float64)) <= e)
    # a plausible slip in the formula fails: without the residual terms the
    # bound does not hold for some pair
    g1 = kp * 2.0 ** -23 / (1 - kp * 2.0 ** -23)
    qh_n = np.sqrt((Qh.astype(np.float64) ** 2).sum(1))
    e_bad = np.outer(nu, (1 + rho) * g1 * qh_n)
    assert np.any(np.abs(fwd - s.astype(np.float64)) > e_bad)
    # well inside the worst-case bound 2^-7 ||u|| ||q||
    nq = np.sqrt((Q.astype(np.float64) ** 2).sum(1))
    assert np.median(cs / nq) < 0.5 * 2.0 ** -7
    assert rho < 2.0 ** -8


def test_bound_with_exact_bf16_values():
    """Rows already exact in BF16 have no residual: the scale is only the
    rounding of the sums, and the bound still holds."""
    rng = np.random.default_rng(6)
    d, kp = 64, 64
    U = bf16_rne(rng.standard_normal((50, d)).astype(np.float32))
    Q = bf16_rne(rng.standard_normal((9, d)).astype(np.float32))
    nu, cs, rho = error_scale(U, Q, kp)
    assert rho == 0.0
    s = rule_scores(U, Q)
    est = seq_fp32(U, Q)
    assert np.all(np.abs(est.astype(np.float64) - s.astype(np.float64)) <= np.outer(nu, cs))


def test_upper_count_from_lower_count():
    """The first pass takes c_hi = c_lo - [g[c_lo - 1] <= xh] whenever
    fl(xh - xl) is below half the smallest gap between adjacent guards
    (rounded down), and counts xh in full otherwise; the shortcut must equal
    the full count wherever it is taken, including near-tied guards."""
    rng = np.random.default_rng(8)
    n = 20000
    g = -np.sort(-rng.standard_normal((n, 7)).astype(np.float32), axis=1)
    g[:100, 3] = g[:100, 2]                                         # tied guards
    g[100:200, 4] = np.nextafter(g[100:200, 3], np.float32(-np.inf))  # one ulp apart
    x = rng.standard_normal(n).astype(np.float32)
    e = (np.abs(rng.standard_normal(n)) * rng.choice([1e-4, 1e-2, 0.3], n)).astype(np.float32)
    xl, xh = (x - e).astype(np.float32), (x + e).astype(np.float32)
    c_lo = (g > xl[:, None]).sum(1)
    c_hi = (g > xh[:, None]).sum(1)
    gp = np.concatenate([np.full((n, 1), np.inf, np.float32), g], axis=1)
    last = gp[np.arange(n), c_lo]                                   # g[c_lo - 1], +inf at c_lo = 0
    gaps = (g[:, :-1] - g[:, 1:]).astype(np.float32)
    hg = np.nextafter((0.5 * gaps.min(1).astype(np.float64)).astype(np.float32), np.float32(-np.inf))
    take = (xh - xl).astype(np.float32) < hg
    short = c_lo - (last <= xh)
    assert take.mean() > 0.5 and (~take).sum() > 0
    assert np.array_equal(short[take], c_hi[take])
