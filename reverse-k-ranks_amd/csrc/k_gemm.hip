// k_gemm.hip -- score kernels with the reference's sequential FP32 rule.
//
// Every score s(a, b) = sum_{k=0}^{K-1} a_k * b_k is evaluated as the
// reference's InnerProduct (src/include/alg/SpaceInnerProduct.hpp:14-20):
// one accumulator per output, k ascending, each product rounded to FP32
// and then added with FP32 rounding.  The library is compiled with
// -ffp-contract=off, so no product is fused into an FMA; the results are
// bit-identical to a scalar C loop.  That is why these are SIMT register-
// tiled kernels (v_mul_f32 + v_add_f32), not MFMA: matrix cores sum in an
// order of their own and cannot reproduce the rule.
//
// Tiling (gfx950, wave64): 256 threads as a 16 x 16 grid; each thread owns
// an 8 x 8 block of a 128 x 128 tile, split as rows {4ty..4ty+3} and
// {64+4ty..64+4ty+3} (same for columns) so that every ds_read_b128 of a
// wave covers 256 contiguous bytes.  K advances in chunks of 16 staged
// through LDS in k-major layout, double-buffered.  k_scores_store is held to
// 168 VGPRs so that 3 workgroups share a CU: gfx950 issues independent FP32
// VALU work fastest with several waves per SIMD ready (9% over 2 per SIMD).
// A fourth does not fit without spilling the tile.
#include "rkr_internal.hpp"

namespace rkr {
namespace {

constexpr int BM = 128, BN = 128, KC = 16, NT = 256;
constexpr int LDS_PAD = 4;  // k-major rows of BM + pad floats

struct TileSmem {
    float a[2][KC][BM + LDS_PAD];
    float b[2][KC][BN + LDS_PAD];
};

// Loads rows [r0, r0+128) x k-chunk [k0, k0+16) of a row-major matrix
// (leading dimension ld, a multiple of 4) into two float4 registers.
// Rows past n_rows and columns at or past ld load as zero.
__device__ __forceinline__ void load_chunk(const float *__restrict__ X, const int32_t *__restrict__ rows,
                                           int64_t r0, int64_t n_rows, int32_t ld, int32_t k0, float4 (&reg)[2]) {
    const int t = threadIdx.x;
    const int kq = (t & 3) * 4;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int64_t r = r0 + (t >> 2) + h * 64;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (r < n_rows && k0 + kq < ld) {
            const int64_t src = rows ? (int64_t)rows[r] : r;
            v = *reinterpret_cast<const float4 *>(X + src * (int64_t)ld + k0 + kq);
        }
        reg[h] = v;
    }
}

__device__ __forceinline__ void store_chunk(float (*dst)[BM + LDS_PAD], const float4 (&reg)[2]) {
    const int t = threadIdx.x;
    const int kq = (t & 3) * 4;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int m = (t >> 2) + h * 64;
        dst[kq + 0][m] = reg[h].x;
        dst[kq + 1][m] = reg[h].y;
        dst[kq + 2][m] = reg[h].z;
        dst[kq + 3][m] = reg[h].w;
    }
}

// acc[i][j] = score(A row i of the tile, B row j of the tile) over k < K.
__device__ __forceinline__ void tile_scores(TileSmem &sm, const float *__restrict__ A,
                                            const int32_t *__restrict__ a_rows, int64_t m0, int64_t M,
                                            int32_t lda, const float *__restrict__ B, int64_t n0, int64_t N,
                                            int32_t ldb, int32_t K, float (&acc)[8][8]) {
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = 0.0f;

    float4 ra[2], rb[2];
    const int nchunk = (K + KC - 1) / KC;
    load_chunk(A, a_rows, m0, M, lda, 0, ra);
    load_chunk(B, nullptr, n0, N, ldb, 0, rb);
    store_chunk(sm.a[0], ra);
    store_chunk(sm.b[0], rb);
    __syncthreads();
    for (int c = 0; c < nchunk; ++c) {
        const int buf = c & 1;
        const int k0 = c * KC;
        if (c + 1 < nchunk) {
            load_chunk(A, a_rows, m0, M, lda, k0 + KC, ra);
            load_chunk(B, nullptr, n0, N, ldb, k0 + KC, rb);
        }
        const int kend = min(KC, K - k0);
        for (int kk = 0; kk < kend; ++kk) {
            const float4 a0 = *reinterpret_cast<const float4 *>(&sm.a[buf][kk][ty * 4]);
            const float4 a1 = *reinterpret_cast<const float4 *>(&sm.a[buf][kk][64 + ty * 4]);
            const float4 b0 = *reinterpret_cast<const float4 *>(&sm.b[buf][kk][tx * 4]);
            const float4 b1 = *reinterpret_cast<const float4 *>(&sm.b[buf][kk][64 + tx * 4]);
            const float av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
            const float bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const float p = av[i] * bv[j];
                    acc[i][j] = acc[i][j] + p;
                }
        }
        if (c + 1 < nchunk) {
            store_chunk(sm.a[buf ^ 1], ra);
            store_chunk(sm.b[buf ^ 1], rb);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int tile_row(int i) { return (i < 4) ? ((threadIdx.x >> 4) * 4 + i) : (64 + (threadIdx.x >> 4) * 4 + i - 4); }
__device__ __forceinline__ int tile_col(int j) { return (j < 4) ? ((threadIdx.x & 15) * 4 + j) : (64 + (threadIdx.x & 15) * 4 + j - 4); }

__global__ void __launch_bounds__(NT, 3) k_scores_store(const float *__restrict__ A, const int32_t *__restrict__ a_rows,
                                                     int64_t M, int32_t lda, const float *__restrict__ B, int64_t N,
                                                     int32_t ldb, int32_t K, float *__restrict__ C, int64_t ldc,
                                                     int c_row_major_m, int64_t n_tiles_n) {
    __shared__ TileSmem sm;
    const int64_t tile = blockIdx.x;
    const int64_t tm = tile / n_tiles_n, tn = tile % n_tiles_n;
    const int64_t m0 = tm * BM, n0 = tn * BN;
    float acc[8][8];
    tile_scores(sm, A, a_rows, m0, M, lda, B, n0, N, ldb, K, acc);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int64_t m = m0 + tile_row(i);
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t n = n0 + tile_col(j);
            if (n >= N) continue;
            if (c_row_major_m) C[m * ldc + n] = acc[i][j];
            else C[n * ldc + m] = acc[i][j];
        }
    }
}

// Persistent refinement count.  Work unit = (block of 128 refine rows,
// group of kGroup item tiles).  Each unit adds, for every row r of its
// block, #{items p in the group : score(user(r), p) > thr(r)} to cnt[r].
constexpr int kGroup = 8;

__global__ void __launch_bounds__(NT) k_refine_count(const float *__restrict__ users, const float *__restrict__ items,
                                                     int64_t n_item, int32_t ld, int32_t K,
                                                     const int32_t *__restrict__ rows_user,
                                                     const float *__restrict__ rows_thr,
                                                     const int32_t *__restrict__ n_rows_dev, int64_t max_rows,
                                                     int32_t *__restrict__ cnt) {
    __shared__ TileSmem sm;
    __shared__ int32_t red[BM];
    const int64_t n_rows = min((int64_t)*n_rows_dev, max_rows);
    const int64_t n_rb = (n_rows + BM - 1) / BM;
    const int64_t n_it = (n_item + BN - 1) / BN;
    const int64_t n_grp = (n_it + kGroup - 1) / kGroup;
    const int64_t n_units = n_rb * n_grp;
    for (int64_t unit = blockIdx.x; unit < n_units; unit += gridDim.x) {
        const int64_t rb = unit / n_grp, g = unit % n_grp;
        const int64_t m0 = rb * BM;
        float thr[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int64_t m = m0 + tile_row(i);
            thr[i] = m < n_rows ? rows_thr[m] : 0.0f;
        }
        int32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const int64_t t_end = min((g + 1) * kGroup, n_it);
        for (int64_t it = g * kGroup; it < t_end; ++it) {
            const int64_t n0 = it * BN;
            float acc[8][8];
            tile_scores(sm, users, rows_user, m0, n_rows, ld, items, n0, n_item, ld, K, acc);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const bool in = n0 + tile_col(j) < n_item;
#pragma unroll
                for (int i = 0; i < 8; ++i) c[i] += (in && acc[i][j] > thr[i]) ? 1 : 0;
            }
        }
        // reduce over the 16 threads that share each row
        if (threadIdx.x < BM) red[threadIdx.x] = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (c[i]) atomicAdd(&red[tile_row(i)], c[i]);
        __syncthreads();
        if (threadIdx.x < BM) {
            const int64_t m = m0 + threadIdx.x;
            if (m < n_rows && red[threadIdx.x]) atomicAdd(&cnt[m], red[threadIdx.x]);
        }
        __syncthreads();
    }
}

// Small-batch <u, q>: one wave per 64-user panel, lane = user.  Users are
// stored as panels [p][k][lane] so every k step is one coalesced 256-byte
// load; queries sit in LDS and are read as broadcasts.
template <int NQ>
__global__ void __launch_bounds__(NT) k_gemv_panels(const float *__restrict__ usersT, int64_t n_user, int32_t ld,
                                                    int32_t K, const float *__restrict__ q, int32_t nq,
                                                    float *__restrict__ qip) {
    extern __shared__ float qs[];  // nq x K
    for (int i = threadIdx.x; i < nq * K; i += NT) qs[i] = q[(i / K) * (int64_t)ld + (i % K)];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t panel = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    const int64_t u = panel * 64 + lane;
    if (panel * 64 >= n_user) return;
    const float *src = usersT + panel * (int64_t)ld * 64 + lane;
    float acc[NQ];
#pragma unroll
    for (int b = 0; b < NQ; ++b) acc[b] = 0.0f;
    for (int k = 0; k < K; ++k) {
        const float x = src[(int64_t)k * 64];
#pragma unroll
        for (int b = 0; b < NQ; ++b) {
            if (b < nq) {
                const float p = x * qs[b * K + k];
                acc[b] = acc[b] + p;
            }
        }
    }
    if (u < n_user) {
#pragma unroll
        for (int b = 0; b < NQ; ++b)
            if (b < nq) qip[(int64_t)b * n_user + u] = acc[b];
    }
}

__global__ void k_make_panels(const float *__restrict__ users, int64_t n_user, int32_t ld, float *__restrict__ usersT) {
    const int64_t n_panel = (n_user + 63) / 64;
    const int64_t total = n_panel * ld * 64;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t lane = i & 63;
        const int64_t k = (i >> 6) % ld;
        const int64_t p = (i >> 6) / ld;
        const int64_t u = p * 64 + lane;
        usersT[i] = u < n_user ? users[u * ld + k] : 0.0f;
    }
}

}  // namespace

hipError_t launch_scores_store(hipStream_t s, const float *A, const int32_t *a_rows, int64_t M, int32_t lda,
                               const float *B, int64_t N, int32_t ldb, int32_t K, float *C, int64_t ldc,
                               bool c_row_major_m) {
    if (M <= 0 || N <= 0) return hipSuccess;
    const int64_t tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
    hipLaunchKernelGGL(k_scores_store, dim3((unsigned)(tm * tn)), dim3(NT), 0, s, A, a_rows, M, lda, B, N, ldb, K,
                       C, ldc, c_row_major_m ? 1 : 0, tn);
    return hipGetLastError();
}

hipError_t launch_refine_count(hipStream_t s, const float *users, const float *items, int64_t n_item, int32_t ld,
                               int32_t K, const int32_t *rows_user, const float *rows_thr, const int32_t *n_rows_dev,
                               int64_t max_rows, int32_t *cnt) {
    if (max_rows <= 0 || n_item <= 0) return hipSuccess;
    int dev = 0, n_cu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t n_it = (n_item + BN - 1) / BN;
    const int64_t max_units = ((max_rows + BM - 1) / BM) * ((n_it + kGroup - 1) / kGroup);
    const int64_t grid = std::min<int64_t>(max_units, (int64_t)n_cu * 4);
    hipLaunchKernelGGL(k_refine_count, dim3((unsigned)grid), dim3(NT), 0, s, users, items, n_item, ld, K, rows_user,
                       rows_thr, n_rows_dev, max_rows, cnt);
    return hipGetLastError();
}

hipError_t launch_gemv_panels(hipStream_t s, const float *usersT, int64_t n_user, int32_t ld, int32_t K,
                              const float *q, int32_t nq, float *qip) {
    if (n_user <= 0 || nq <= 0) return hipSuccess;
    const int64_t n_panel = (n_user + 63) / 64;
    const unsigned grid = (unsigned)((n_panel + 3) / 4);
    const size_t smem = sizeof(float) * (size_t)nq * K;
    if (nq <= 1) hipLaunchKernelGGL(k_gemv_panels<1>, dim3(grid), dim3(NT), smem, s, usersT, n_user, ld, K, q, nq, qip);
    else if (nq <= 2) hipLaunchKernelGGL(k_gemv_panels<2>, dim3(grid), dim3(NT), smem, s, usersT, n_user, ld, K, q, nq, qip);
    else if (nq <= 4) hipLaunchKernelGGL(k_gemv_panels<4>, dim3(grid), dim3(NT), smem, s, usersT, n_user, ld, K, q, nq, qip);
    else if (nq <= 8) hipLaunchKernelGGL(k_gemv_panels<8>, dim3(grid), dim3(NT), smem, s, usersT, n_user, ld, K, q, nq, qip);
    else hipLaunchKernelGGL(k_gemv_panels<16>, dim3(grid), dim3(NT), smem, s, usersT, n_user, ld, K, q, nq, qip);
    return hipGetLastError();
}

hipError_t launch_make_panels(hipStream_t s, const float *users, int64_t n_user, int32_t ld, float *usersT) {
    if (n_user <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_make_panels, dim3(2048), dim3(256), 0, s, users, n_user, ld, usersT);
    return hipGetLastError();
}

}  // namespace rkr
