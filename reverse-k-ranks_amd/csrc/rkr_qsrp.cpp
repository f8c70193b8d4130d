// QSRP screening (the reference's QSRPNormalLP method, src/impl/QSRPNormalLP.hpp:137-253): an
// inner-product bound from a few leading SVD dimensions (PartDimPartNorm, PartDimPartNorm.hpp:52-95)
// and a per-user regression from score to sample position (NormalLinearRegressionLP,
// NormalLinearRegressionLP.hpp:326-417), both ahead of the exact bounds and the refinement.
//
// Build (host, FP64, once per table): rkr_qsrp_build
//   - the users' Gram matrix and its eigenvectors (Jacobi), which are the left singular vectors
//     of the user matrix that SVD::PerformSVD takes with armadillo's svd_econ (SVD.hpp:62-123):
//     user u becomes u'_r = (w_r . u) / s_r, a query q becomes q'_r = s_r (w_r . q), so that
//     <u', q'> = <u, q>; check_dim is the first index whose cumulative singular-value share
//     reaches SIGMA (SVD.hpp:112-121);
//   - per user, the minimax line fit of NormalLinearRegressionLP::LinearProgram: sample position
//     j + 1 against Phi((T[j] - mu) / sigma) over both ends of each sample interval, solved
//     exactly as the minimal vertical-width strip over the points' convex hull instead of the
//     vendored Seidel LP (same optimum: a 3-variable LP whose solution is that strip), with the
//     reference's fallback line when the fit is not decreasing (:262-283).
// Query (device, exact output only): k_qsrp_bounds turns the left-part score and the
// Cauchy-Schwarz remainder into an interval that provably contains the sequential-rule score,
// and the regression into bucket bounds; candidates of those bounds get the exact score and the
// exact table search (the reference's ComputeRemainDim + SampleSearch::RankBound,
// QSRPNormalLP.hpp:180-200); the exact path's selection then runs unchanged.  Answers equal the
// exact path's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "rkr_internal.hpp"

struct rkr_ctx : public rkr::Ctx {};

namespace rkr {

// CDFPhi of NormalLinearRegressionLP.hpp:125-147: Abramowitz & Stegun 7.1.26 for erf.  Its
// absolute error in erf is at most 1.5e-7, so |cdf_phi - Phi| <= 7.5e-8.
double qsrp_cdf_phi(double x) {
    const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027, a5 = 1.061405429,
                 p = 0.3275911;
    const double sign = x < 0 ? -1.0 : 1.0;
    const double z = std::fabs(x) / std::sqrt(2.0);
    const double t = 1.0 / (1.0 + p * z);
    const double y = 1.0 - (((((a5 * t + a4) * t) + a3) * t + a2) * t + a1) * t * std::exp(-z * z);
    return 0.5 * (1.0 + sign * y);
}

// Eigen-decomposition of the symmetric n x n matrix A (row-major, destroyed) by cyclic Jacobi
// rotations: A = V diag(ev) V^T with the eigenvectors in the columns of V.
static void jacobi_eigen(std::vector<double> &A, int n, std::vector<double> &V, std::vector<double> &ev) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i) {
            diag += A[(size_t)i * n + i] * A[(size_t)i * n + i];
            for (int j = i + 1; j < n; ++j) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
        }
        if (off <= 1e-30 * diag || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    ev.resize(n);
    for (int i = 0; i < n; ++i) ev[i] = A[(size_t)i * n + i];
}

// Minimax (Chebyshev) line through points (x_i, y_i): min over (a, b) of max_i |a x_i + b - y_i|.
// This is NormalLinearRegressionLP's LP (minimise e subject to |a x_i + b - y_i| <= e,
// NormalLinearRegressionLP.hpp:86-99, 149-181) in closed form: the optimal strip between two
// parallel lines has one line on a convex-hull edge and the other through the opposite chain's
// extreme point, so it is the smallest vertical width over all hull-edge slopes (and the
// horizontal strip through the extreme y values when all x are equal).  Returns false when
// there are fewer than two distinct x.
bool qsrp_minimax_line(std::vector<std::pair<double, double>> pts, double &a, double &b, double &e) {
    std::sort(pts.begin(), pts.end());
    // convex hull (monotone chain), lower and upper chains from left to right
    auto cross = [](const std::pair<double, double> &o, const std::pair<double, double> &p,
                    const std::pair<double, double> &q) {
        return (p.first - o.first) * (q.second - o.second) - (p.second - o.second) * (q.first - o.first);
    };
    std::vector<std::pair<double, double>> lo, up;
    for (const auto &p : pts) {
        while (lo.size() >= 2 && cross(lo[lo.size() - 2], lo.back(), p) <= 0) lo.pop_back();
        lo.push_back(p);
    }
    for (auto it = pts.rbegin(); it != pts.rend(); ++it) {
        while (up.size() >= 2 && cross(up[up.size() - 2], up.back(), *it) <= 0) up.pop_back();
        up.push_back(*it);
    }
    std::reverse(up.begin(), up.end());
    if (pts.front().first == pts.back().first) return false;
    // candidate slopes: every edge of both chains, ascending.  The width of the strip of slope m
    // is max_i (y_i - m x_i) - min_i (y_i - m x_i); the maximum is reached on the upper chain at
    // a vertex that moves left as m grows, the minimum on the lower chain at one that moves
    // right, so one sweep with two pointers evaluates every slope (rotating calipers)
    std::vector<double> slopes;
    for (size_t i = 1; i < lo.size(); ++i)
        if (lo[i].first > lo[i - 1].first) slopes.push_back((lo[i].second - lo[i - 1].second) / (lo[i].first - lo[i - 1].first));
    for (size_t i = 1; i < up.size(); ++i)
        if (up[i].first > up[i - 1].first) slopes.push_back((up[i].second - up[i - 1].second) / (up[i].first - up[i - 1].first));
    std::sort(slopes.begin(), slopes.end());
    size_t pu = up.size() - 1, pl = 0;
    double best = INFINITY;
    for (double m : slopes) {
        while (pu > 0 && up[pu - 1].second - m * up[pu - 1].first >= up[pu].second - m * up[pu].first) --pu;
        while (pl + 1 < lo.size() && lo[pl + 1].second - m * lo[pl + 1].first <= lo[pl].second - m * lo[pl].first) ++pl;
        const double hi = up[pu].second - m * up[pu].first, low = lo[pl].second - m * lo[pl].first;
        const double w = (hi - low) / 2.0;
        if (w < best) {
            best = w;
            a = m;
            b = (hi + low) / 2.0;
        }
    }
    e = best;
    return std::isfinite(best);
}

namespace {

// Runs fn(i0, i1) over [0, n) on up to 16 host threads.
template <class F>
void qsrp_threads(int64_t n, F &&fn) {
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>({(int64_t)std::max(1u, std::thread::hardware_concurrency()),
                                                                16, (n + 1023) / 1024}));
    std::vector<std::thread> th;
    int64_t started = 1;
    try {
        for (; started < nt; ++started) th.emplace_back(fn, n * started / nt, n * (started + 1) / nt);
    } catch (...) {
    }
    fn(0, n / nt);
    for (int64_t t = started; t < nt; ++t) fn(n * t / nt, n * (t + 1) / nt);
    for (auto &x : th) x.join();
}

inline float round_up_f(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

}  // namespace

}  // namespace rkr

using namespace rkr;

extern "C" int rkr_qsrp_build(rkr_ctx *c, float sigma, double *build_ms) {
    if (!c) return RKR_ERR_INVALID;
    if (!c->users || !c->table) return set_err(c, RKR_ERR_STATE, "qsrp_build: no data or no table");
    if (!(sigma > 0.0f && sigma <= 1.0f)) return set_err(c, RKR_ERR_INVALID, "qsrp_build: SIGMA must be in (0, 1]");
    if (c->n_ins > 0 || c->n_del > 0)
        return set_err(c, RKR_ERR_STATE, "qsrp_build: item updates are pending; rebuild the table first");
    const double t0 = rkr_now();
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    free_qsrp(c);
    const int64_t U = c->n_user;
    const int32_t d = c->dim, ld = c->ld, S = c->n_sample;
    if (S < 2) return set_err(c, RKR_ERR_INVALID, "qsrp_build: the regression needs n_sample >= 2");

    // users (normalised as stored) and the sampled table, from the device
    std::vector<float> hu((size_t)U * ld), ht((size_t)U * S);
    RKR_HIP(c, hipMemcpy(hu.data(), c->users, sizeof(float) * hu.size(), hipMemcpyDeviceToHost));
    RKR_HIP(c, hipMemcpy(ht.data(), c->table, sizeof(float) * ht.size(), hipMemcpyDeviceToHost));

    // Gram matrix sum_u u u^T, per-thread partial sums
    const int64_t n_part = std::min<int64_t>(16, std::max<int64_t>(1, U / 1024));
    std::vector<std::vector<double>> part((size_t)n_part, std::vector<double>((size_t)d * d, 0.0));
    qsrp_threads(n_part, [&](int64_t p0, int64_t p1) {
        for (int64_t p = p0; p < p1; ++p) {
            std::vector<double> &G = part[(size_t)p];
            std::vector<double> x(d);
            for (int64_t u = U * p / n_part; u < U * (p + 1) / n_part; ++u) {
                for (int k = 0; k < d; ++k) x[k] = hu[(size_t)u * ld + k];
                for (int i = 0; i < d; ++i) {
                    const double xi = x[i];
                    if (xi == 0.0) continue;
                    double *g = &G[(size_t)i * d];
                    for (int j = i; j < d; ++j) g[j] += xi * x[j];
                }
            }
        }
    });
    std::vector<double> G((size_t)d * d, 0.0);
    for (const auto &P : part)
        for (size_t i = 0; i < G.size(); ++i) G[i] += P[i];
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < i; ++j) G[(size_t)i * d + j] = G[(size_t)j * d + i];
    std::vector<double> V, ev;
    jacobi_eigen(G, d, V, ev);
    // singular values s_r = sqrt(lambda_r), descending; w_r = column r of V
    std::vector<int> ord(d);
    for (int i = 0; i < d; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return ev[x] > ev[y]; });
    std::vector<double> W((size_t)d * d), sv(d);  // W row r = w_r
    int32_t rank = 0;
    for (int r = 0; r < d; ++r) {
        sv[r] = std::sqrt(std::max(0.0, ev[ord[r]]));
        for (int k = 0; k < d; ++k) W[(size_t)r * d + k] = V[(size_t)k * d + ord[r]];
        if (sv[r] > 1e-12 * std::max(1.0, sv[0])) rank = r + 1;
    }
    if (rank == 0) return set_err(c, RKR_ERR_STATE, "qsrp_build: the user matrix is zero");
    // directions beyond the rank carry no user energy: keep them with s_r = 1 so the
    // transform stays orthogonal (u'_r is then ~0 for every user)
    for (int r = rank; r < d; ++r) sv[r] = 1.0;
    // check_dim (SVD.hpp:112-121): first index whose cumulative share reaches SIGMA
    double tot = 0.0;
    for (int r = 0; r < rank; ++r) tot += sv[r];
    int32_t cd = 0;
    {
        double cum = 0.0;
        for (int r = 0; r < d; ++r) {
            cum += r < rank ? sv[r] : 0.0;
            if (cum / tot >= (double)sigma) { cd = r; break; }
        }
    }
    // epsilon of orthogonality of W: max |(W W^T - I)_{ij}| * d bounds ||W W^T - I||_2
    double eps_w = 0.0;
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += W[(size_t)i * d + k] * W[(size_t)j * d + k];
            eps_w = std::max(eps_w, std::fabs(s - (i == j ? 1.0 : 0.0)));
        }
    eps_w *= d;

    // rotated users (FP64 -> FP32), norms of the left and right parts, the user norm
    const int32_t ldr = (d + 3) / 4 * 4;
    std::vector<float> ur((size_t)U * ldr, 0.0f);
    std::vector<float> par((size_t)kQsrpPar * U);
    qsrp_threads(U, [&](int64_t u0, int64_t u1) {
        std::vector<double> x(d);
        for (int64_t u = u0; u < u1; ++u) {
            for (int k = 0; k < d; ++k) x[k] = hu[(size_t)u * ld + k];
            double nl = 0.0, nr = 0.0, nu = 0.0;
            for (int r = 0; r < d; ++r) {
                const double *w = &W[(size_t)r * d];
                double s = 0.0;
                for (int k = 0; k < d; ++k) s += w[k] * x[k];
                const float v = (float)(s / sv[r]);
                ur[(size_t)u * ldr + r] = v;
                (r < cd ? nl : nr) += (double)v * v;
            }
            for (int k = 0; k < d; ++k) nu += x[k] * x[k];
            par[(size_t)kQsrpNrmL * U + u] = round_up_f(std::sqrt(nl) * (1 + 1e-15));
            par[(size_t)kQsrpNrmR * U + u] = round_up_f(std::sqrt(nr) * (1 + 1e-15));
            par[(size_t)kQsrpNrmU * U + u] = round_up_f(std::sqrt(nu) * (1 + 1e-15));
        }
    });

    // per-user regression of sample position on the normal CDF of the score
    // (NormalLinearRegressionLP::AssignCache / LinearProgram / CalcError, :149-283)
    const double eps_phi = 7.5e-8 + kQsrpPhiDevErr;  // host formula vs Phi, plus the device's evaluation
    std::atomic<int64_t> n_fallback{0};
    qsrp_threads(U, [&](int64_t u0, int64_t u1) {
        std::vector<std::pair<double, double>> pts;
        for (int64_t u = u0; u < u1; ++u) {
            const float *T = &ht[(size_t)u * S];
            double mu = 0.0;
            for (int j = 0; j < S; ++j) mu += T[j];
            mu /= S;
            double var = 0.0;
            for (int j = 0; j < S; ++j) var += (T[j] - mu) * (T[j] - mu);
            const double sg = std::sqrt(var / S);
            const float mu32 = (float)mu, inv32 = sg > 0 ? (float)(1.0 / sg) : 0.0f;
            // the device evaluates Phi((x - mu32) * inv32); the fit uses the same argument
            auto cdf = [&](float x) { return qsrp_cdf_phi((double)((x - mu32) * inv32)); };
            double a = 0.0, bb = 0.0, e = 0.0;
            bool ok = sg > 0 && inv32 > 0 && std::isfinite(inv32);
            pts.clear();
            if (ok) {
                for (int j = 0; j + 1 < S; ++j) {
                    pts.emplace_back(cdf(T[j]), (double)(j + 1));
                    pts.emplace_back(cdf(T[j + 1]), (double)(j + 1));
                }
                if (!(qsrp_minimax_line(pts, a, bb, e) && a < 0 && bb >= 0 && e > 0 && e < S)) {
                    // the reference's fallback line through the first and last samples (:262-275)
                    const double first = cdf(T[0]), last = cdf(T[S - 1]);
                    ok = first > last;
                    if (ok) {
                        a = (2.0 - S) / (first - last);
                        bb = ((double)S * first - last) / (first - last);
                    }
                    n_fallback.fetch_add(1, std::memory_order_relaxed);
                }
            }
            float *P = &par[u];
            double E = S;
            if (ok) {
                // the error that makes the bound valid for this (a, b): the largest deviation over
                // the constraint points, plus Phi's error times |a|, the FP32 rounding of a and b
                // and of a*Phi+b, and the reference's 0.01 slack
                double emax = 0.0;
                for (const auto &p : pts) emax = std::max(emax, std::fabs(a * p.first + bb - p.second));
                const float a32 = (float)a, b32 = (float)bb;
                E = emax + std::fabs(a) * eps_phi + std::fabs(a - a32) + std::fabs(bb - b32) +
                    std::ldexp(1.0, -21) * (std::fabs(a) + std::fabs(bb) + 1.0) + 0.01;
                P[(size_t)kQsrpA * U] = a32;
                P[(size_t)kQsrpB * U] = b32;
                P[(size_t)kQsrpMu * U] = mu32;
                P[(size_t)kQsrpInvSigma * U] = inv32;
            }
            if (!ok || !(E < S)) {
                // no usable line: the regression says nothing (the span checks still apply)
                P[(size_t)kQsrpA * U] = 0.0f;
                P[(size_t)kQsrpB * U] = 0.0f;
                P[(size_t)kQsrpMu * U] = 0.0f;
                P[(size_t)kQsrpInvSigma * U] = 0.0f;
                E = S;
            }
            P[(size_t)kQsrpErr * U] = E < S ? round_up_f(E) : (float)S;
            P[(size_t)kQsrpTmax * U] = T[0];
            P[(size_t)kQsrpTmin * U] = T[S - 1];
        }
    });

    // query rotation rows P_r = s_r w_r (FP32), the B operand of the query GEMM
    std::vector<float> hp((size_t)d * ld, 0.0f);
    for (int r = 0; r < d; ++r)
        for (int k = 0; k < d; ++k) hp[(size_t)r * ld + k] = (float)(sv[r] * W[(size_t)r * d + k]);

    int rc;
    if ((rc = dev_alloc_t(c, &c->qsrp_users, (size_t)U * ldr, "QSRP rotated users")) ||
        (rc = dev_alloc_t(c, &c->qsrp_rot, (size_t)d * ld, "QSRP query rotation")) ||
        (rc = dev_alloc_t(c, &c->qsrp_par, (size_t)kQsrpPar * U, "QSRP user parameters"))) {
        free_qsrp(c);
        return rc;
    }
    RKR_HIP(c, hipMemcpy(c->qsrp_users, ur.data(), sizeof(float) * ur.size(), hipMemcpyHostToDevice));
    RKR_HIP(c, hipMemcpy(c->qsrp_rot, hp.data(), sizeof(float) * hp.size(), hipMemcpyHostToDevice));
    RKR_HIP(c, hipMemcpy(c->qsrp_par, par.data(), sizeof(float) * par.size(), hipMemcpyHostToDevice));
    // interval margin: |s_fp32 - left_fl| <= ||u'_R|| ||q'_R|| + gamma_cd ||u'_L|| ||q'_L||
    //                                        + C_rot ||u|| ||q||
    const double e24 = std::ldexp(1.0, -24);
    auto gam = [&](double n) { return n * e24 / (1.0 - n * e24); };
    c->qsrp_gcd = round_up_f(1.01 * gam(std::max(cd, 1)));
    c->qsrp_crot = round_up_f(1.01 * (eps_w + std::ldexp(1.0, -23) * (1 + eps_w) +
                                      (gam(d) + std::ldexp(1.0, -23)) * (1 + e24) * d * (1 + eps_w) + gam(d)));
    c->qsrp_cd = cd;
    c->qsrp_ldr = ldr;
    c->qsrp_sigma = sigma;
    c->qsrp_n_user = U;
    c->qsrp_fallback = n_fallback.load();
    if (build_ms) *build_ms = (rkr_now() - t0) * 1e3;
    return RKR_OK;
}

extern "C" int rkr_qsrp_info(const rkr_ctx *c, int32_t *check_dim, int64_t *n_fallback_fit) {
    if (!c) return RKR_ERR_INVALID;
    if (!c->qsrp_users) return RKR_ERR_STATE;
    if (check_dim) *check_dim = c->qsrp_cd;
    if (n_fallback_fit) *n_fallback_fit = c->qsrp_fallback;
    return RKR_OK;
}

namespace {
// A device allocation freed when the scope ends.
struct QsrpDev {
    rkr::Ctx *c;
    void *p = nullptr;
    int alloc(size_t n, const char *what) { return rkr::dev_alloc(c, &p, std::max<size_t>(n, 4), what); }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
    ~QsrpDev() { rkr::dev_free(c, p, 0); }
};
}  // namespace

extern "C" int rkr_qsrp_screen(rkr_ctx *c, const float *queries, int32_t n_query, uint32_t *out) {
    if (!c) return RKR_ERR_INVALID;
    if (n_query < 0 || (n_query > 0 && (!queries || !out))) return set_err(c, RKR_ERR_INVALID, "qsrp_screen: bad arguments");
    if (!c->qsrp_users || c->qsrp_n_user != c->n_user) return set_err(c, RKR_ERR_STATE, "qsrp_screen: no QSRP screen built");
    if (n_query == 0) return RKR_OK;
    RKR_HIP(c, hipSetDevice(c->device));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t U = c->n_user;
    const int32_t d = c->dim, ld = c->ld, ldr = c->qsrp_ldr;
    const bool use_q = c->qtable != nullptr;
    const int32_t S_T = use_q ? c->n_qs : c->n_sample;
    const float eps = use_q ? 0.0f : c->table_eps;
    std::vector<float> hq((size_t)n_query * ld, 0.0f);
    for (int32_t i = 0; i < n_query; ++i) std::memcpy(&hq[(size_t)i * ld], queries + (size_t)i * d, sizeof(float) * d);
    QsrpDev q{c}, qrot{c}, qn{c}, left{c}, bk{c};
    int rc;
    if ((rc = q.alloc(sizeof(float) * hq.size(), "QSRP queries")) ||
        (rc = qrot.alloc(sizeof(float) * (size_t)n_query * ldr, "QSRP query rows")) ||
        (rc = qn.alloc(sizeof(float) * 3 * (size_t)n_query, "QSRP query norms")) ||
        (rc = left.alloc(sizeof(float) * (size_t)n_query * U, "QSRP left scores")) ||
        (rc = bk.alloc(sizeof(uint32_t) * (size_t)n_query * U, "QSRP bounds")))
        return rc;
    RKR_HIP(c, hipMemcpyAsync(q.p, hq.data(), sizeof(float) * hq.size(), hipMemcpyHostToDevice, c->stream));
    RKR_HIP(c, launch_scores_store(c->stream, q.as<float>(), nullptr, n_query, ld, c->qsrp_rot, d, ld, d,
                                   qrot.as<float>(), ldr, true));
    RKR_HIP(c, launch_qsrp_qnorms(c->stream, qrot.as<float>(), ldr, q.as<float>(), ld, d, c->qsrp_cd, n_query,
                                  qn.as<float>()));
    RKR_HIP(c, launch_scores_store(c->stream, c->qsrp_users, nullptr, U, ldr, qrot.as<float>(), n_query, ldr,
                                   c->qsrp_cd, left.as<float>(), U, false));
    const float absf = std::nextafter((float)((double)eps + (double)d * std::ldexp(1.0, -140)), INFINITY);
    RKR_HIP(c, launch_qsrp_bounds(c->stream, left.as<float>(), U, n_query, c->qsrp_par, qn.as<float>(), c->qsrp_gcd,
                                  c->qsrp_crot, absf, c->n_sample, use_q ? c->qcol : nullptr, S_T,
                                  bk.as<uint32_t>()));
    RKR_HIP(c, hipMemcpyAsync(out, bk.p, sizeof(uint32_t) * (size_t)n_query * U, hipMemcpyDeviceToHost, c->stream));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    return RKR_OK;
}

extern "C" int rkr_qsrp_phi(rkr_ctx *c, const float *z, int64_t n, float *out) {
    if (!c) return RKR_ERR_INVALID;
    if (n < 0 || (n > 0 && (!z || !out))) return set_err(c, RKR_ERR_INVALID, "qsrp_phi: bad arguments");
    if (n == 0) return RKR_OK;
    RKR_HIP(c, hipSetDevice(c->device));
    QsrpDev dz{c}, dout{c};
    int rc;
    if ((rc = dz.alloc(sizeof(float) * (size_t)n, "CDF points")) || (rc = dout.alloc(sizeof(float) * (size_t)n, "CDF values")))
        return rc;
    RKR_HIP(c, hipMemcpyAsync(dz.p, z, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    RKR_HIP(c, launch_phi_eval(c->stream, dz.as<float>(), n, dout.as<float>()));
    RKR_HIP(c, hipMemcpyAsync(out, dout.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    RKR_HIP(c, hipStreamSynchronize(c->stream));
    return RKR_OK;
}

extern "C" int rkr_qsrp_fit_line(const double *x, const double *y, int32_t n, double *a, double *b, double *e) {
    if (!x || !y || !a || !b || !e || n < 2) return RKR_ERR_INVALID;
    std::vector<std::pair<double, double>> pts((size_t)n);
    for (int32_t i = 0; i < n; ++i) pts[i] = {x[i], y[i]};
    return rkr::qsrp_minimax_line(std::move(pts), *a, *b, *e) ? RKR_OK : RKR_ERR_INVALID;
}
