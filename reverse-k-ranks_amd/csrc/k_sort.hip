// k_sort.hip -- every user's scores against every item, sorted in descending order.
//
// The full score table (ReadAll's index file, src/include/alg/DiskIndex/ReadAll.hpp:80-87, written by
// src/app/BuildScoreTableBatch.cpp:69-144) holds each user's n_item scores sorted with std::greater
// (src/include/score_computation/ComputeScoreTableBatch.hpp:83-89).  Only the values are kept, so the
// order of equal values is irrelevant and the sort need not be stable.
//
// One 1024-thread workgroup per row.  Rows of up to NMAX values are sorted in LDS (bucket path):
//   1. the row is loaded into registers (coalesced), with its min and max;
//   2. a histogram over NBK value bins, descending, and its exclusive scan.  The bins are
//      apportioned to NB1 linear coarse bins by the coarse bins' counts (a first histogram), so
//      bins are narrow where the row is dense and hold about n / NBK values each even when a
//      heavy tail crowds most of the row into a few linear bins.  FP32 arithmetic: fl(max - x)
//      and fl(t * scale) are monotone in x, so bin order is value order;
//   3. every value is scattered into its bin's segment of an LDS buffer (any order inside a bin);
//   4. each value's final position is its bin's start plus the values of its bin above it (and the
//      equal ones before it): threads walk the buffer in order, so the lanes of a wavefront read the
//      same few bins and write within a few hundred bytes of the output row.
// A row longer than NMAX, or one with a bin over BIN_MAX values (heavy ties, extreme skew), takes
// the merge path instead:
//   1. tiles of TILE values are sorted in LDS by a bitonic network (a partial last tile is padded with
//      -inf, which sorts to its end and is not written back);
//   2. sorted runs are merged pairwise -- run length TILE, 2*TILE, ... -- between the output row and
//      a scratch row.  Each thread finds where its share of a merged run starts by a binary search
//      along the merge path, then merges that share sequentially.  The tiles go to whichever of the
//      two rows makes the last pass land in the output row.
#include "rkr_internal.hpp"

namespace rkr {
namespace {

constexpr int NT = 1024;                 // threads per row
constexpr int NW = NT / 64;              // wavefronts per row
constexpr int TILE = 8192;               // values sorted in LDS at a time (32 KB)
constexpr int PER = TILE / 2 / NT;       // compare-exchange pairs per thread and stage
constexpr int NBK = 8192;                // value bins of the bucket path
constexpr int NMAX = 28672;              // longest row of the bucket path (112 KB of LDS)
constexpr int EPT = NMAX / NT;           // values per thread
constexpr uint32_t BIN_MAX = 192;        // largest bin the bucket path ranks by counting
constexpr int NB1 = 1024;                // linear coarse bins that apportion the NBK bins

struct SortSmem {
    union {
        float tile[TILE];                // merge path
        struct {
            float buf[NMAX];             // the row, bin by bin, then in order
            uint32_t cnt[NBK];           // coarse counts, then bin counts, then bin ends (32 KB)
            uint32_t fseg[NB1];          // bins of coarse bin c: s0 | s1 << 16, bins s0 .. s1 - 1
            float red_f[2 * NW];
            uint32_t red_u[NW + 1];
        } b;
    } u;
};

// Bin of x, descending: its coarse bin c (NB1 linear bins from hi down)
// owns bins s0 .. s1 - 1 (fseg[c] = s0 | s1 << 16: one LDS read), and x takes the one its place
// inside c falls in.  t is clamped to [0, NB1] (NaN from 0 * inf at x == hi
// to 0), so c = NB1 - 1 with frac 1 at and past the range's end.
__device__ __forceinline__ int sort_bin(const uint32_t *fseg, float x, float hi, float scale1) {
    const float t = fminf(fmaxf((hi - x) * scale1, 0.0f), (float)NB1);
    const int c = min((int)t, NB1 - 1);
    const float frac = t - (float)c;
    const uint32_t g = fseg[c];
    const int s0 = (int)(g & 0xFFFFu), nc = (int)(g >> 16) - s0;  // nc >= 1
    return s0 + min(nc - 1, (int)(frac * (float)nc));
}

// Sorts s[0..TILE) in descending order.  Stage (k, j) compares i with i + j for
// the i whose bit j is clear; blocks of k elements alternate direction so that
// each merge step sees bitonic sequences, and the final k = TILE is descending.
__device__ inline void bitonic_desc(float *s) {
    for (int k = 2; k <= TILE; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                const int p = threadIdx.x + e * NT;                   // pair index < TILE / 2
                const int i = ((p & ~(j - 1)) << 1) | (p & (j - 1));  // lower element of the pair
                const float a = s[i], b = s[i + j];
                const bool desc = (i & k) == 0;
                if (desc ? (a < b) : (a > b)) {
                    s[i] = b;
                    s[i + j] = a;
                }
            }
            __syncthreads();
        }
    }
}

// Number of elements of A among the first d outputs of the descending merge of
// A[0..na) and B[0..nb), where A goes first on ties.
__device__ inline int merge_split(const float *A, int na, const float *B, int nb, int d) {
    int lo = max(0, d - nb), hi = min(d, na);
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A[mid] >= B[d - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Bucket path for one row of n <= NMAX values; returns false (having written
// nothing) when a bin holds more than BIN_MAX values.  x and y may be the same
// row: the row is in registers before anything is written.
__device__ __forceinline__ bool bucket_sort_row(SortSmem &sm, const float *x, int32_t n, float *y) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // register slots e < ne hold values (the test is uniform, so the slots
    // past a short row are skipped, not run with no lanes active); ok[e]
    // marks the slots that hold one of the row's values.  Rows of a multiple
    // of 4 values starting 16-byte aligned load 16 bytes per slot group (slot
    // 4k + c holds value 4 (t + k NT) + c), others 4 bytes per slot (value
    // t + e NT); either way every load of a thread is issued before the
    // first is used, from clamped addresses.
    const bool vec = (n & 3) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
    const int ne = vec ? 4 * ((n / 4 + NT - 1) / NT) : (n + NT - 1) / NT;
    float v[EPT];
    bool ok[EPT];
    if (vec) {
        const float4 *x4 = reinterpret_cast<const float4 *>(x);
        const int n4 = n / 4;
#pragma unroll
        for (int k = 0; k < EPT / 4; ++k) {
            const int i = (int)threadIdx.x + k * NT;
            const float4 q = x4[min(i, n4 - 1)];
            v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
#pragma unroll
            for (int c = 0; c < 4; ++c) ok[4 * k + c] = i < n4;
        }
    } else {
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int i = (int)threadIdx.x + e * NT;
            v[e] = x[min(i, n - 1)];
            ok[e] = i < n;
        }
    }
    float mn = INFINITY, mx = -INFINITY;
#pragma unroll
    for (int e = 0; e < EPT; ++e)
        if (ok[e]) {
            mn = fminf(mn, v[e]);
            mx = fmaxf(mx, v[e]);
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, o, 64));
        mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    }
    if (lane == 0) {
        sm.u.b.red_f[w] = mn;
        sm.u.b.red_f[NW + w] = mx;
    }
    for (int b = threadIdx.x; b < NBK; b += NT) sm.u.b.cnt[b] = 0;
    __syncthreads();
    mn = sm.u.b.red_f[0];
    mx = sm.u.b.red_f[NW];
    for (int i = 1; i < NW; ++i) {
        mn = fminf(mn, sm.u.b.red_f[i]);
        mx = fmaxf(mx, sm.u.b.red_f[NW + i]);
    }
    if (!(mx > mn)) {
        // one distinct value: the row is its own sorted order
        for (int i = threadIdx.x; i < n; i += NT) y[i] = x[i];  // each element in place (x may be y)
        return true;
    }
    const float scale = (float)NB1 / (mx - mn);
    // coarse histogram (cnt was zeroed above), then each coarse bin c gets
    // bins s0 .. s1 - 1: one plus its share of the other
    // NBK - NB1 by its count
#pragma unroll
    for (int e = 0; e < EPT; ++e)
        if (e < ne && ok[e])
            atomicAdd(&sm.u.b.cnt[(int)fminf(fmaxf((mx - v[e]) * scale, 0.0f), (float)(NB1 - 1))], 1u);
    __syncthreads();
    {
        static_assert(NB1 == NT, "one coarse bin per thread");
        const uint32_t cc = sm.u.b.cnt[threadIdx.x];
        uint32_t incl = cc;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) sm.u.b.red_u[w] = incl;
        __syncthreads();
        uint32_t below = incl - cc;
        for (int i = 0; i < w; ++i) below += sm.u.b.red_u[i];
        // the last coarse bin ends at bin NBK (below + cc = n there)
        const uint32_t s0 = threadIdx.x + (uint32_t)(((uint64_t)(NBK - NB1) * below) / (uint64_t)n);
        const uint32_t s1 = threadIdx.x + 1u + (uint32_t)(((uint64_t)(NBK - NB1) * (below + cc)) / (uint64_t)n);
        sm.u.b.fseg[threadIdx.x] = s0 | (s1 << 16);
        __syncthreads();  // every thread has read its coarse count
    }
    for (int b = threadIdx.x; b < NBK; b += NT) sm.u.b.cnt[b] = 0;
    __syncthreads();
    // every value's bin, computed once: its LDS reads are issued for all the
    // thread's values before any is used
    int bin[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) bin[e] = e < ne ? sort_bin(sm.u.b.fseg, v[e], mx, scale) : 0;
#pragma unroll
    for (int e = 0; e < EPT; ++e)
        if (e < ne && ok[e]) atomicAdd(&sm.u.b.cnt[bin[e]], 1u);
    __syncthreads();
    // exclusive scan of the bin counts (4 bins per thread) and the largest bin
    constexpr int BPT = NBK / NT;
    uint32_t c[BPT], run = 0, big = 0;
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        c[k] = sm.u.b.cnt[threadIdx.x * BPT + k];
        run += c[k];
        big = max(big, c[k]);
    }
    uint32_t incl = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) big = max(big, __shfl_xor(big, o, 64));
    __syncthreads();  // every thread has read its counts
    if (lane == 63) sm.u.b.red_u[w] = incl;
    if (lane == 0) sm.u.b.red_f[w] = __uint_as_float(big);
    __syncthreads();
    uint32_t base = 0, bmax = 0;
    for (int i = 0; i < NW; ++i) {
        if (i < w) base += sm.u.b.red_u[i];
        bmax = max(bmax, __float_as_uint(sm.u.b.red_f[i]));
    }
    if (bmax > BIN_MAX) return false;  // uniform across the workgroup
    uint32_t start = base + incl - run;
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        sm.u.b.cnt[threadIdx.x * BPT + k] = start;  // scatter cursor, ends as the bin's end
        start += c[k];
    }
    __syncthreads();
    // every slot is claimed before the first value is written
    uint32_t slot[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        slot[e] = 0;
        if (e < ne && ok[e]) slot[e] = atomicAdd(&sm.u.b.cnt[bin[e]], 1u);
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e)
        if (e < ne && ok[e]) sm.u.b.buf[slot[e]] = v[e];
    __syncthreads();
    // final positions, in buffer order (the lanes of a wavefront read the
    // same few bins): the bin's start plus the values of the bin above this
    // one, and the equal ones before it.  The value goes straight to its
    // place in the output row; the row's input is no longer read.  A thread
    // takes RB values at a time, so their bin lookups and the counting reads
    // (four per value and step; a bin holds about n / NBK values) are in
    // flight together; reads past a bin are clamped to the row and masked out.
    constexpr int RB = 4;
    for (int i0 = threadIdx.x; i0 < n; i0 += RB * NT) {
        float x0[RB];
        int b[RB];
        uint32_t b0[RB], b1[RB], r[RB], len = 0;
#pragma unroll
        for (int q = 0; q < RB; ++q) x0[q] = sm.u.b.buf[min(i0 + q * NT, n - 1)];
#pragma unroll
        for (int q = 0; q < RB; ++q) b[q] = sort_bin(sm.u.b.fseg, x0[q], mx, scale);
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            b0[q] = b[q] ? sm.u.b.cnt[b[q] - 1] : 0u;
            b1[q] = sm.u.b.cnt[b[q]];
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            r[q] = b0[q];
            len = max(len, b1[q] - b0[q]);
        }
        for (uint32_t j = 0; j < len; j += 4) {
            float z[RB][4];
#pragma unroll
            for (int q = 0; q < RB; ++q)
#pragma unroll
                for (int k = 0; k < 4; ++k) z[q][k] = sm.u.b.buf[min(b0[q] + j + k, (uint32_t)n - 1u)];
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const uint32_t i = (uint32_t)(i0 + q * NT);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t pos = b0[q] + j + k;
                    r[q] += (pos < b1[q] && (z[q][k] > x0[q] || (z[q][k] == x0[q] && pos < i))) ? 1u : 0u;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < RB; ++q)
            if (i0 + q * NT < n) y[r[q]] = x0[q];
    }
    return true;
}

// Merge path for one row; x and y may be the same row: every value of x is
// read before y is written where it lies.
__device__ void merge_sort_row(float *s, const float *x, int32_t n, float *out_row, float *tmp_row) {
    int passes = 0;
    for (int64_t w = TILE; w < n; w <<= 1) ++passes;
    float *a = (passes & 1) ? tmp_row : out_row;  // sorted tiles
    float *b = (passes & 1) ? out_row : tmp_row;

    for (int t0 = 0; t0 < n; t0 += TILE) {
        const int m = min(TILE, n - t0);
        for (int i = threadIdx.x; i < TILE; i += NT) s[i] = i < m ? x[t0 + i] : -INFINITY;
        __syncthreads();
        bitonic_desc(s);
        for (int i = threadIdx.x; i < m; i += NT) a[t0 + i] = s[i];
        __syncthreads();
    }

    for (int w = TILE; w < n; w <<= 1) {
        __syncthreads();  // the runs written by the previous pass are visible to the workgroup
        for (int s0 = 0; s0 < n; s0 += 2 * w) {
            const int na = min(w, n - s0);
            const int nb = min(w, n - s0 - na);
            const int no = na + nb;
            const int per = (no + NT - 1) / NT;
            const int d = (int)threadIdx.x * per;
            if (d >= no) continue;
            const int cnt = min(per, no - d);
            const float *A = a + s0, *B = a + s0 + na;
            float *D = b + s0 + d;
            int i = merge_split(A, na, B, nb, d), j = d - i;
            float av = i < na ? A[i] : 0.0f, bv = j < nb ? B[j] : 0.0f;
            for (int o = 0; o < cnt; ++o) {
                if (j >= nb || (i < na && av >= bv)) {
                    D[o] = av;
                    if (++i < na) av = A[i];
                } else {
                    D[o] = bv;
                    if (++j < nb) bv = B[j];
                }
            }
        }
        float *t = a;
        a = b;
        b = t;
    }
}

__global__ void __launch_bounds__(NT) k_sort_rows(const float *src, int64_t ld_src, int32_t n, float *out,
                                                  float *tmp, int64_t ld) {
    __shared__ SortSmem sm;
    const int64_t row = blockIdx.x;
    const float *x = src + row * ld_src;
    if (n <= NMAX) {
        if (bucket_sort_row(sm, x, n, out + row * ld)) return;
        __syncthreads();  // the bucket path's LDS is done with before the tiles reuse it
    }
    merge_sort_row(sm.u.tile, x, n, out + row * ld, tmp + row * ld);
}

}  // namespace

hipError_t launch_sort_rows(hipStream_t s, const float *src, int64_t ld_src, int64_t nb, int64_t n_item, float *out,
                            float *tmp, int64_t ld) {
    if (nb <= 0 || n_item <= 0) return hipSuccess;
    if (n_item > (int64_t)1 << 30 || ld < n_item || ld_src < n_item) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sort_rows, dim3((unsigned)nb), dim3(NT), 0, s, src, ld_src, (int32_t)n_item, out, tmp, ld);
    return hipGetLastError();
}

}  // namespace rkr
