// k_sort.hip -- every user's scores against every item, sorted in descending order.
//
// The full score table (ReadAll's index file, src/include/alg/DiskIndex/ReadAll.hpp:80-87, written by
// src/app/BuildScoreTableBatch.cpp:69-144) holds each user's n_item scores sorted with std::greater
// (src/include/score_computation/ComputeScoreTableBatch.hpp:83-89).  Only the values are kept, so the
// order of equal values is irrelevant and the sort need not be stable.
//
// One 1024-thread workgroup per CU, taking rows in turn (rows that load 16 bytes at a time; others
// take one workgroup each).  Rows of up to NMAX values are sorted in LDS (bucket path):
//   1. the row is in registers (coalesced loads, issued while the workgroup's previous row was
//      ranked), and each wavefront's min and max of it were taken before that row's last barrier;
//   2. a histogram over NBK value bins, descending, and its exclusive scan.  The bins are
//      apportioned to NB1 linear coarse bins by the coarse bins' counts (a first histogram, of
//      half the row), so bins are narrow where the row is dense and hold about n / NBK values
//      each even when a heavy tail crowds most of the row into a few linear bins.  FP32
//      arithmetic: fl(max - x) and fl(t * scale) are monotone in x, so bin order is value order;
//   3. every value's order-preserving key is scattered into its bin's segment of an LDS buffer (any
//      order inside a bin);
//   4. each value's final position is its bin's start plus the keys of its bin above it (and the
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
#include <algorithm>

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
constexpr int KPAD = 256;                // zero keys past the row's end, read by the rank step
static_assert(KPAD >= (int)BIN_MAX + 3, "the rank step reads up to BIN_MAX + 2 keys past a bin's start");

struct SortSmem {
    union {
        float tile[TILE];                // merge path
        struct {
            uint32_t key[NMAX + KPAD];   // the row's keys, bin by bin, then KPAD zero keys
            uint32_t cnt[NBK];           // bin counts, then scatter cursors ending as bin ends
            uint32_t fseg[NB1];          // bins of coarse bin c: s0 | s1 << 16, bins s0 .. s1 - 1
            uint32_t coarse[NB1];        // coarse counts (zero between rows)
            float red_f[2 * NW];         // the row's min and max per wavefront, then the largest bin
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

// Order-preserving key of a float: key(a) < key(b) whenever a < b.  Every NaN
// maps to the largest key (the bins put NaN first); -0 keys just below +0.
__device__ __forceinline__ uint32_t sort_key(float f) {
    const uint32_t u = __float_as_uint(f);
    const uint32_t k = u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
    return f == f ? k : 0xFFFFFFFFu;
}

// The float of a key (NaN keys give one quiet NaN).
__device__ __forceinline__ float key_value(uint32_t k) {
    return __uint_as_float(k ^ ((int32_t)k < 0 ? 0x80000000u : 0xFFFFFFFFu));
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

// A row of n <= NMAX values in registers, EPT slots per thread.  Rows of a
// multiple of 4 values whose rows start 16-byte aligned (vec) take 16 bytes
// per slot group (slot 4k + c holds value 4 (t + k NT) + c), others 4 bytes
// per slot (value t + e NT).  Slots e < row_slots<vec>(n) can hold values (the
// test is uniform, so the slots past a short row are skipped, not run with no
// lanes active); in_row says whether this thread's slot e holds one.
template <bool vec>
__device__ __forceinline__ int row_slots(int n) {
    return vec ? 4 * ((n / 4 + NT - 1) / NT) : (n + NT - 1) / NT;
}

template <bool vec>
__device__ __forceinline__ bool in_row(int t, int e, int n) {
    return vec ? t + (e >> 2) * NT < (n >> 2) : t + e * NT < n;
}

// Issues every load of the thread's slots before any is used, from clamped
// addresses; the caller's first use of v waits for them.
template <bool vec>
__device__ __forceinline__ void load_row(int t, float (&v)[EPT], const float *x, int n) {
    if constexpr (vec) {
        const float4 *x4 = reinterpret_cast<const float4 *>(x);
        const int n4 = n / 4;
#pragma unroll
        for (int k = 0; k < EPT / 4; ++k) {
            const float4 q = x4[min(t + k * NT, n4 - 1)];
            v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int e = 0; e < EPT; ++e) v[e] = x[min(t + e * NT, n - 1)];
    }
}

// The inverse of load_row: each slot's value back to its place in y.
template <bool vec>
__device__ __forceinline__ void store_row(int t, const float (&v)[EPT], float *y, int n) {
    if constexpr (vec) {
        float4 *y4 = reinterpret_cast<float4 *>(y);
#pragma unroll
        for (int k = 0; k < EPT / 4; ++k) {
            const int i = t + k * NT;
            if (i < n / 4) y4[i] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
        }
    } else {
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int i = t + e * NT;
            if (i < n) y[i] = v[e];
        }
    }
}

// Each wavefront's min and max of the row held in v, into red_f (read by the
// row's bucket path after the next barrier).
template <bool vec>
__device__ __forceinline__ void row_partials(SortSmem &sm, const float (&v)[EPT], int32_t n, int t) {
    const int lane = t & 63, w = t >> 6;
    float mn = INFINITY, mx = -INFINITY;
#pragma unroll
    for (int e = 0; e < EPT; ++e)
        if (in_row<vec>(t, e, n)) {
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
}

// Bucket path for the row held in v (load_row's layout); returns false, having
// written nothing and with v unchanged, when a bin holds more than BIN_MAX
// values.  Otherwise it writes the row's sorted order to y and (vec: the
// persistent kernel) issues the loads of row xn into v once v's values are in
// LDS, so they are in flight while this row is ranked, and leaves row xn's
// min/max partials in red_f.  On entry red_f holds this row's partials and
// coarse is zero.  y may be the row v came from.
template <bool vec>
__device__ __forceinline__ bool bucket_sort_row(SortSmem &sm, float (&v)[EPT], int32_t n, float *y,
                                                const float *xn, int t) {
    const int lane = t & 63, w = t >> 6;
    const int ne = row_slots<vec>(n);
    float mn = sm.u.b.red_f[0], mx = sm.u.b.red_f[NW];
    for (int i = 1; i < NW; ++i) {
        mn = fminf(mn, sm.u.b.red_f[i]);
        mx = fmaxf(mx, sm.u.b.red_f[NW + i]);
    }
    if (!(mx > mn)) {
        // one distinct value (NaNs aside): the row is its own sorted order
        store_row<vec>(t, v, y, n);
        if constexpr (vec) {
            load_row<vec>(t, v, xn, n);
            __syncthreads();  // every thread has read this row's partials
            row_partials<vec>(sm, v, n, t);
        }
        return true;
    }
    for (int b = t; b < NBK; b += NT) sm.u.b.cnt[b] = 0;  // the bin counts, after the coarse pass
    const float scale = (float)NB1 / (mx - mn);
    // coarse histogram (coarse is zero on entry) of a sample, the values of the
    // slots e < nes (the row's first ns values), then each coarse bin c gets
    // bins s0 .. s1 - 1: one plus its share of the other NBK - NB1 by its
    // sampled count.  The sample only shapes the bins; every value is binned.
    const int nes = vec ? 4 * ((ne / 4 + 1) / 2) : (ne + 1) / 2;
    const uint32_t ns = (uint32_t)min(n, NT * nes);
#pragma unroll
    for (int e = 0; e < EPT; ++e)
        if (e < nes && in_row<vec>(t, e, n))
            atomicAdd(&sm.u.b.coarse[(int)fminf(fmaxf((mx - v[e]) * scale, 0.0f), (float)(NB1 - 1))], 1u);
    __syncthreads();
    {
        static_assert(NB1 == NT, "one coarse bin per thread");
        const uint32_t cc = sm.u.b.coarse[t];
        sm.u.b.coarse[t] = 0u;  // zero again for the next row
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
        // the last coarse bin ends at bin NBK (below + cc = ns there)
        const uint32_t s0 = t + (uint32_t)(((uint64_t)(NBK - NB1) * below) / (uint64_t)ns);
        const uint32_t s1 = t + 1u + (uint32_t)(((uint64_t)(NBK - NB1) * (below + cc)) / (uint64_t)ns);
        sm.u.b.fseg[t] = s0 | (s1 << 16);
        __syncthreads();  // the bins' segments and the zeroed counts are visible
    }
    // every value's bin, computed once: its LDS reads are issued for all the
    // thread's values before any is used
    int bin[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) bin[e] = e < ne ? sort_bin(sm.u.b.fseg, v[e], mx, scale) : 0;
#pragma unroll
    for (int e = 0; e < EPT; ++e)
        if (e < ne && in_row<vec>(t, e, n)) atomicAdd(&sm.u.b.cnt[bin[e]], 1u);
    __syncthreads();
    // exclusive scan of the bin counts (BPT bins per thread) and the largest bin
    constexpr int BPT = NBK / NT;
    uint32_t c[BPT], run = 0, big = 0;
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        c[k] = sm.u.b.cnt[t * BPT + k];
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
    // red_u and red_f were last read before the barriers of the coarse pass and
    // the bin histogram, and each thread overwrites only its own counts below
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
        sm.u.b.cnt[t * BPT + k] = start;  // scatter cursor, ends as the bin's end
        start += c[k];
    }
    __syncthreads();
    // every slot is claimed before the first value is written
    uint32_t slot[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        slot[e] = 0;
        if (e < ne && in_row<vec>(t, e, n)) slot[e] = atomicAdd(&sm.u.b.cnt[bin[e]], 1u);
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e)
        if (e < ne && in_row<vec>(t, e, n)) sm.u.b.key[slot[e]] = sort_key(v[e]);
    if (t < KPAD) sm.u.b.key[n + t] = 0u;  // below every value's key
    __syncthreads();
    if constexpr (vec) load_row<vec>(t, v, xn, n);  // v's values are in LDS
    // final positions, in buffer order (the lanes of a wavefront read the
    // same few bins): the bin's start plus the number of the bin's keys that
    // order before this one.  (key, position) pairs order by key, larger
    // first, then by position, lower first, so equal values take consecutive
    // places; one 64-bit compare of (key, ~position) does both.  Each value
    // reads a window from its bin's start, as long as the longest bin of the
    // thread's RB values rounded up to 4: the keys past its bin belong to later
    // bins (smaller values, smaller keys) or to the zero padding, and never
    // order before it, so the window needs no bounds.  The value goes straight
    // to its place in the output row.  A thread takes RB values at a time, so
    // their bin lookups and the window reads are in flight together.
    constexpr int RB = 4;
    for (int i0 = t; i0 < n; i0 += RB * NT) {
        uint32_t kx[RB], b0[RB], b1[RB], r[RB], len = 0;
        float x0[RB];
        uint64_t me[RB];
#pragma unroll
        for (int q = 0; q < RB; ++q) kx[q] = sm.u.b.key[min(i0 + q * NT, n - 1)];
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            x0[q] = key_value(kx[q]);
            const int b = sort_bin(sm.u.b.fseg, x0[q], mx, scale);
            b0[q] = b ? sm.u.b.cnt[b - 1] : 0u;
            b1[q] = sm.u.b.cnt[b];
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            r[q] = 0;
            len = max(len, b1[q] - b0[q]);
            me[q] = ((uint64_t)kx[q] << 32) | (uint32_t)~(uint32_t)(i0 + q * NT);
        }
        for (uint32_t j = 0; j < len; j += 4) {
            uint32_t z[RB][4];
#pragma unroll
            for (int q = 0; q < RB; ++q)
#pragma unroll
                for (int k = 0; k < 4; ++k) z[q][k] = sm.u.b.key[b0[q] + j + k];
#pragma unroll
            for (int q = 0; q < RB; ++q)
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    r[q] += (((uint64_t)z[q][k] << 32) | (uint32_t)~(b0[q] + j + k)) > me[q] ? 1u : 0u;
        }
#pragma unroll
        for (int q = 0; q < RB; ++q)
            if (i0 + q * NT < n) y[b0[q] + r[q]] = x0[q];
    }
    // the largest-bin reads of red_f ended before the scatter's barrier
    if constexpr (vec) row_partials<vec>(sm, v, n, t);
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

// vec (every row starts 16-byte aligned and n is a multiple of 4): persistent,
// workgroup p sorts rows p, p + G, p + 2G, ... (G = gridDim.x, one workgroup
// per CU), and on the bucket path the next row's loads are in flight while the
// current row is ranked.  Otherwise one workgroup per row (G = n_rows).
template <bool vec>
__global__ void __launch_bounds__(NT) k_sort_rows(const float *src, int64_t ld_src, int64_t n_rows, int32_t n,
                                                  float *out, float *tmp, int64_t ld) {
    __shared__ SortSmem sm;
    int64_t row = blockIdx.x;
    if (n > NMAX) {
        for (; row < n_rows; row += gridDim.x) {
            merge_sort_row(sm.u.tile, src + row * ld_src, n, out + row * ld, tmp + row * ld);
            __syncthreads();  // the tiles are free before the next row's first tile
        }
        return;
    }
    float v[EPT];
    if (row >= n_rows) return;
    for (int i = threadIdx.x; i < NB1; i += NT) sm.u.b.coarse[i] = 0u;
    if constexpr (!vec) {
        const float *x = src + row * ld_src;
        load_row<false>(threadIdx.x, v, x, n);
        row_partials<false>(sm, v, n, threadIdx.x);
        __syncthreads();
        if (bucket_sort_row<false>(sm, v, n, out + row * ld, nullptr, threadIdx.x)) return;
        __syncthreads();  // the bucket path's LDS is done with before the tiles reuse it
        merge_sort_row(sm.u.tile, x, n, out + row * ld, tmp + row * ld);
        return;
    }
    load_row<vec>(threadIdx.x, v, src + row * ld_src, n);
    row_partials<vec>(sm, v, n, threadIdx.x);
    __syncthreads();
    for (; row < n_rows; row += gridDim.x) {
        // the next row's loads are issued unconditionally (so v is dead across
        // the merge path); a workgroup's last row loads its own row again, which
        // it has just read, and the values are not used.  n and the thread index
        // pass through an empty asm statement each iteration, so the per-slot
        // indexes and LDS addresses derived from them are computed where they
        // are used instead of being hoisted out of the loop and held in
        // registers (which spilled)
        int32_t nn = n;
        int t = threadIdx.x;
        asm volatile("" : "+s"(nn), "+v"(t));
        const int64_t next = row + gridDim.x;
        const float *xn = src + (next < n_rows ? next : row) * ld_src;
        if (!bucket_sort_row<vec>(sm, v, nn, out + row * ld, xn, t)) {
            __syncthreads();  // the bucket path's LDS is done with before the tiles reuse it
            merge_sort_row(sm.u.tile, src + row * ld_src, n, out + row * ld, tmp + row * ld);
            load_row<vec>(t, v, xn, nn);
            row_partials<vec>(sm, v, nn, t);
        }
        __syncthreads();  // this row's LDS is done with, the next row's partials are visible
    }
}

}  // namespace

hipError_t launch_sort_rows(hipStream_t s, const float *src, int64_t ld_src, int64_t nb, int64_t n_item, float *out,
                            float *tmp, int64_t ld, int grid_cap) {
    if (nb <= 0 || n_item <= 0) return hipSuccess;
    if (n_item > (int64_t)1 << 30 || ld < n_item || ld_src < n_item || grid_cap <= 0) return hipErrorInvalidValue;
    const bool vec = (n_item & 3) == 0 && (ld_src & 3) == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0;
    const unsigned grid = (unsigned)(vec ? std::min<int64_t>(nb, grid_cap) : nb);
    if (vec)
        hipLaunchKernelGGL(k_sort_rows<true>, dim3(grid), dim3(NT), 0, s, src, ld_src, nb, (int32_t)n_item, out, tmp, ld);
    else
        hipLaunchKernelGGL(k_sort_rows<false>, dim3(grid), dim3(NT), 0, s, src, ld_src, nb, (int32_t)n_item, out, tmp, ld);
    return hipGetLastError();
}

}  // namespace rkr
