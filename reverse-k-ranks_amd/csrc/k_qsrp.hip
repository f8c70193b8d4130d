// k_qsrp.hip -- the QSRP screen on the device (QSRPNormalLP::Retrieval,
// src/impl/QSRPNormalLP.hpp:137-253), exact output.
//
// k_qsrp_qnorms: per query, upper bounds of the norms of the rotated query's
//   left part (dimensions < check_dim), right part, and of the query itself.
// k_qsrp_bounds: per (query, user), PartDimPartNorm::IPBound
//   (PartDimPartNorm.hpp:52-74) -- the left-part score plus and minus the
//   Cauchy-Schwarz bound of the right part -- widened by the rounding margins
//   so that the interval contains the sequential-rule score <u, q>; then
//   NormalLinearRegressionLP::ComputeRankBound (NormalLinearRegressionLP.hpp:
//   326-398) turns the interval's ends into bounds on the sample bucket
//   #{T > <u,q>}, written as the exact path's bucket pair.
// k_qsrp_exact: per (query, user) that those bounds keep as a candidate, the
//   exact score and the exact bucket pair from the table (the reference's
//   ComputeRemainDim + SampleSearch::RankBound, QSRPNormalLP.hpp:180-200).
#include "rkr_internal.hpp"

namespace rkr {
namespace {

// CDFPhi (NormalLinearRegressionLP.hpp:125-147) in FP32.
__device__ __forceinline__ float phi_dev(float x) {
    const float a1 = 0.254829592f, a2 = -0.284496736f, a3 = 1.421413741f, a4 = -1.453152027f, a5 = 1.061405429f,
                p = 0.3275911f;
    const float sign = x < 0.0f ? -1.0f : 1.0f;
    const float z = fabsf(x) * 0.70710678118654752f;
    const float t = 1.0f / (1.0f + p * z);
    const float y = 1.0f - (((((a5 * t + a4) * t) + a3) * t + a2) * t + a1) * t * expf(-z * z);
    return 0.5f * (1.0f + sign * y);
}

__global__ void k_phi_eval(const float *__restrict__ z, int64_t n, float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = phi_dev(z[i]);
}

// Upper bound of sqrt(sum x_k^2) over k in [k0, k1): the FP32 sum of squares is within
// (1 + gamma_n) of the exact one, so (n + 4) * 2^-23 covers it and the square root.
__device__ __forceinline__ float norm_up(const float *x, int32_t k0, int32_t k1) {
    float ss = 0.0f;
    for (int32_t k = k0; k < k1; ++k) ss = ss + x[k] * x[k];
    const double n = (double)(k1 - k0);
    return __double2float_ru((double)sqrtf(ss) * (1.0 + (n + 4.0) * 0x1p-23));
}

__global__ void k_qsrp_qnorms(const float *__restrict__ qrot, int32_t ldr, const float *__restrict__ q, int32_t ld,
                              int32_t d, int32_t cd, int32_t nq, float *__restrict__ qn) {
    const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nq) return;
    const float *r = qrot + (int64_t)b * ldr;
    qn[3 * b + 0] = norm_up(r, 0, cd);
    qn[3 * b + 1] = norm_up(r, cd, d);
    qn[3 * b + 2] = norm_up(q + (int64_t)b * ld, 0, d);
}

__device__ __forceinline__ int32_t bucket_gt(const float *__restrict__ row, int32_t S, float q) {
    int32_t lo = 0, hi = S;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (row[mid] > q) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// grid (user blocks, queries)
__global__ void __launch_bounds__(256) k_qsrp_bounds(const float *__restrict__ left, int64_t n_user,
                                                     const float *__restrict__ par, const float *__restrict__ qn,
                                                     float gcd, float crot, float absf, int32_t S,
                                                     const int32_t *__restrict__ qcol, int32_t S_T,
                                                     uint32_t *__restrict__ bk) {
    const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t b = blockIdx.y;
    if (u >= n_user) return;
    const float L = left[b * n_user + u];
    const float a = par[kQsrpA * n_user + u], bb = par[kQsrpB * n_user + u];
    const float mu = par[kQsrpMu * n_user + u], inv = par[kQsrpInvSigma * n_user + u];
    const float E = par[kQsrpErr * n_user + u];
    const float tmax = par[kQsrpTmax * n_user + u], tmin = par[kQsrpTmin * n_user + u];
    const float nl = par[kQsrpNrmL * n_user + u], nr = par[kQsrpNrmR * n_user + u], nu = par[kQsrpNrmU * n_user + u];
    const float qL = qn[3 * b], qR = qn[3 * b + 1], qq = qn[3 * b + 2];
    // |<u,q> - L| <= ||u'_R|| ||q'_R|| + gcd ||u'_L|| ||q'_L|| + crot ||u|| ||q|| + absf; the
    // FP32 evaluation of that sum (8 roundings at most) is covered by the factor 1 + 2^-20
    const float w = (nr * qR + gcd * (nl * qL) + crot * (nu * qq) + absf) * (1.0f + 0x1p-20f);
    const float x_lb = __double2float_rd((double)L - (double)w);
    const float x_ub = __double2float_ru((double)L + (double)w);
    const float x_pl = nextafterf(x_ub, INFINITY);  // #{T > x_ub} = #{T >= x_pl}
    const bool line = inv > 0.0f;
    // B_up >= #{T >= x_lb} >= #{T > <u,q>}
    int32_t up;
    if (x_lb > tmax) up = 0;
    else if (x_lb <= tmin || !line) up = S;  // no regression line: the span checks only
    else {
        const float p = a * phi_dev((x_lb - mu) * inv) + bb;
        const float c = ceilf(p + E);
        up = c <= 0.0f ? 0 : (c >= (float)S ? S : (int32_t)c);
    }
    // B_dn <= #{T >= x_pl} = #{T > x_ub} <= #{T > <u,q>}
    int32_t dn;
    if (x_pl > tmax) dn = 0;
    else if (x_pl <= tmin) dn = S;
    else if (!line) dn = 0;
    else {
        const float p = a * phi_dev((x_pl - mu) * inv) + bb;
        const float f = floorf(p - E);
        dn = f <= 0.0f ? 0 : (f >= (float)S ? S : (int32_t)f);
    }
    // buckets of the searched table: sampled column j is its column qcol[j]
    uint32_t lo = (uint32_t)up, hi = (uint32_t)dn;
    if (qcol) {
        lo = up == S ? (uint32_t)S_T : (uint32_t)qcol[up];
        hi = dn == 0 ? 0u : (uint32_t)qcol[dn - 1] + 1u;
    }
    bk[b * n_user + u] = lo | (hi << 16);
}

// grid (user blocks, queries): candidates of the screened bounds (b_hi <= beta)
// get the sequential-rule score (SpaceInnerProduct.hpp:14-20) and the exact
// bucket pair; the others keep the screen's bounds, which stay valid.
__global__ void __launch_bounds__(256) k_qsrp_exact(const float *__restrict__ users, const float *__restrict__ q_in,
                                                    int32_t ld, int32_t d, const float *__restrict__ T, int32_t S,
                                                    float eps, int64_t n_user, QueryInfo *__restrict__ qi,
                                                    float *__restrict__ qip, uint32_t *__restrict__ bk) {
    const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t b = blockIdx.y;
    if (u >= n_user) return;
    const uint32_t x = bk[b * n_user + u];
    const bool pass = (int32_t)(x >> 16) <= qi[b].beta;
    const uint64_t m = __ballot(pass);
    if (pass) {
        const float *ur = users + u * (int64_t)ld;
        const float *qr = q_in + b * (int64_t)ld;
        float s = 0.0f;
        for (int32_t k = 0; k < d; ++k) s = s + ur[k] * qr[k];
        const float *row = T + u * (int64_t)S;
        uint32_t lo, hi;
        if (eps > 0.0f) {
            lo = (uint32_t)bucket_gt(row, S, __double2float_rd((double)s - (double)eps));
            hi = (uint32_t)bucket_gt(row, S, __double2float_ru((double)s + (double)eps));
        } else {
            lo = hi = (uint32_t)bucket_gt(row, S, s);
        }
        qip[b * n_user + u] = s;
        bk[b * n_user + u] = lo | (hi << 16);
    }
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&qi[b].n_screen, (int32_t)__popcll(m));
}

}  // namespace

hipError_t launch_phi_eval(hipStream_t s, const float *z, int64_t n, float *out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_phi_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, z, n, out);
    return hipGetLastError();
}

hipError_t launch_qsrp_qnorms(hipStream_t s, const float *qrot, int32_t ldr, const float *q, int32_t ld, int32_t d,
                              int32_t cd, int32_t nq, float *qn) {
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_qsrp_qnorms, dim3((unsigned)((nq + 63) / 64)), dim3(64), 0, s, qrot, ldr, q, ld, d, cd, nq,
                       qn);
    return hipGetLastError();
}

hipError_t launch_qsrp_bounds(hipStream_t s, const float *left, int64_t n_user, int32_t nq, const float *par,
                              const float *qn, float gcd, float crot, float absf, int32_t S, const int32_t *qcol,
                              int32_t S_T, uint32_t *bk) {
    if (n_user <= 0 || nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_qsrp_bounds, dim3((unsigned)((n_user + 255) / 256), (unsigned)nq), dim3(256), 0, s, left,
                       n_user, par, qn, gcd, crot, absf, S, qcol, S_T, bk);
    return hipGetLastError();
}

hipError_t launch_qsrp_exact(hipStream_t s, const float *users, const float *q_in, int32_t ld, int32_t d,
                             const float *T, int32_t S, float eps, int64_t n_user, int32_t nq, QueryInfo *qi,
                             float *qip, uint32_t *bk) {
    if (n_user <= 0 || nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_qsrp_exact, dim3((unsigned)((n_user + 255) / 256), (unsigned)nq), dim3(256), 0, s, users,
                       q_in, ld, d, T, S, eps, n_user, qi, qip, bk);
    return hipGetLastError();
}

}  // namespace rkr
