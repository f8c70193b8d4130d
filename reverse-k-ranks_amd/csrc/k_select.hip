// k_select.hip -- per-user order statistics at the sample ranks.
//
// For one row x[0..I) of exact scores and strictly increasing sample ranks
// r_0 < ... < r_{S-1} < I, writes T[j] = the r_j-th largest value of x
// (0-based, descending).  This is the value SampleSearch::LoopPreprocess
// reads from a row sorted with std::greater
// (src/include/alg/RankBoundRefinement/SampleSearch.hpp:80-85;
//  src/include/score_computation/ComputeScoreTableBatch.hpp:83-89): the
// r-th order statistic is unique as a value, so no full sort is needed.
//
// One 1024-thread workgroup per row, all state in LDS (up to 160 KB on
// gfx950).  Rows of up to 53,248 values (a multiple of 4) are read from
// memory once, into registers; every later pass runs from there.
//   1. min/max of the row;
//   2. a coarse histogram of 1024 linear bins apportions NB level-0 bins to
//      the coarse bins by their element counts (narrow bins where the row is
//      dense), then a histogram over the level-0 bins (FP32 arithmetic:
//      fl(x - lo) and fl(t * scale) are monotone in x, so the bin is too); a
//      reverse prefix sum tells, for every r_j, its bin and the elements
//      above that bin;
//   3. one more pass over the row gathers the elements of the bins holding
//      targets into LDS (the histogram's storage), with a bitmap of target
//      bins deciding membership -- several passes only when they exceed it;
//   4. bins of up to 64 elements are ranked element-parallel (each element
//      counts the elements of its bin that order before it), larger ones by
//      one wavefront (register bitonic sort up to 256, counting above) or
//      the whole workgroup (wg_select, which narrows the bin by repeated
//      histograms -- this also makes ties and skewed rows exact).
#include "rkr_internal.hpp"

namespace rkr {
namespace {

constexpr int NT = 1024;          // threads per row
constexpr int NW = NT / 64;       // wavefronts per row
constexpr int NB = 16384;         // level-0 bins
constexpr int NB2 = 1024;         // bins of the narrowing loop
constexpr int NB1 = 1024;         // coarse bins that apportion the level-0 bins
constexpr int SMAX = 4096;        // largest n_sample handled
constexpr int GCAP = NB;          // floats gathered per pass (the histogram's 64 KB)
constexpr int EMAX = 64;            // bins up to this size are ranked element-parallel
constexpr int WAVE_SEL_MAX = 2048;  // bins larger than this use wg_select
constexpr int MIDCAP = GCAP / (EMAX + 1) + 1;          // wavefront bins of one pass, at most
constexpr int COOPCAP = GCAP / (WAVE_SEL_MAX + 1) + 1; // workgroup bins of one pass, at most
constexpr uint16_t DEEP = 0x8000;

// The target ranks: from LDS (16-bit copies, rows of up to 65,536 items) or
// from memory.  The choice is uniform over the workgroup.
struct Ranks {
    const int32_t *g;
    const uint16_t *l;  // the LDS copy
    bool lds;
    __device__ __forceinline__ uint32_t operator[](int j) const { return lds ? (uint32_t)l[j] : (uint32_t)g[j]; }
};

struct SelSmem {
    union {
        uint32_t incl[NB + 1];    // incl[b] = #elements in bins >= b (steps 2-3)
        float gbuf[GCAP];         // gathered elements of the target bins (step 4)
    } u;
    uint32_t tbits[NB / 32];      // target bins of the current pass
    uint16_t tpre[NB / 32];       // target bins of the pass below each bitmap word
    uint16_t tb[SMAX];            // bin of target j
    uint16_t fb[SMAX];            // first target index of each distinct bin (bins descending)
    uint32_t fh[SMAX];            // elements in distinct bin f
    uint32_t fabove[SMAX];        // elements in the bins above distinct bin f
    uint32_t fcur[SMAX];          // gather cursor of distinct bin f (ends at its segment's end)
    uint16_t grp[SMAX + 1];       // group start (index into fb) | DEEP
    uint16_t mid[MIDCAP];         // distinct bins of a pass ranked by one wavefront
    uint16_t coop[COOPCAP];       // distinct bins of a pass needing wg_select
    uint32_t sbits[GCAP / 32];    // first gathered element of each bin of the pass
    uint16_t spre[GCAP / 32];     // bins starting below each sbits word
    uint32_t hist2[NB2 + 1];
    uint32_t fseg[NB1];           // level-0 bins of coarse bin c: s0 | s1 << 16, bins s0 .. s1 - 1
    uint16_t srk[SMAX];           // the target ranks, when n_item <= 65536
    float red_f[2 * NW];
    uint32_t red_u[NW + 1];
    int32_t nfb, ngrp, ncoop, nmid, sel_b;
};

// Monotone in x for x >= lo with FP32 arithmetic (every rounding step is
// monotone); NaN (only from 0 * inf at x == lo) maps to bin 0 (fmaxf returns
// the number), overflow to the last bin, so an LDS index is always in range.
// Clamped with min/max rather than branches: this runs once per element.
__device__ __forceinline__ int bin_f32(float x, float lo, float scale, int nb) {
    const float t = (x - lo) * scale;
    return (int)fminf(fmaxf(t, 0.0f), (float)(nb - 1));
}

// Level-0 bins of N values.  The coarse bin c of x (NB1 linear bins over
// [lo, hi]) owns level-0 bins s0 .. s1 - 1 (fseg[c] = s0 | s1 << 16: one LDS
// read), in proportion
// to the row's elements in c, and x takes the one its place inside c falls
// in.  Monotone in x: t is (see bin_f32), so are c and, within one c, t - c.
// t is clamped to [0, NB1] (NaN to 0): below the range c = 0 with frac 0, at
// or past its end c = NB1 - 1 with frac 1.  The LDS reads of all N values are
// issued together.  (The two 16-bit reads of a value fall in one or two
// adjacent banks; one 8-byte entry per coarse bin was slower for its bank
// conflicts.)
template <int N>
__device__ __forceinline__ void fine_bins(const uint32_t *fseg, const float (&v)[N], float lo, float scale1,
                                          int (&b)[N]) {
    int c[N];
    float frac[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const float t = fminf(fmaxf((v[q] - lo) * scale1, 0.0f), (float)NB1);
        c[q] = min((int)t, NB1 - 1);
        frac[q] = t - (float)c[q];
    }
    int s0[N], s1[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const uint32_t w = fseg[c[q]];
        s0[q] = (int)(w & 0xFFFFu);
        s1[q] = (int)(w >> 16);
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const int n = s1[q] - s0[q];  // >= 1
        b[q] = s0[q] + min(n - 1, (int)(frac[q] * (float)n));
    }
}

// A row held in registers: thread t owns float4 t + k*NT for k < C4 (rows of
// up to 16*C4*NT values whose length is a multiple of 4 and whose start is
// 16-byte aligned).  The row is then read from memory once; every later pass
// over it runs from registers.
template <int C4>
struct RowRegs {
    float4 r[C4];
    int nv;  // float4s this thread owns
    __device__ __forceinline__ void load(const float *__restrict__ x, int64_t n) {
        const float4 *x4 = reinterpret_cast<const float4 *>(x);
        const int n4 = (int)(n >> 2), t = threadIdx.x;
        nv = n4 > t ? (n4 - t + NT - 1) / NT : 0;
        // unconditional loads of a clamped index (every thread owns at least
        // one float4 whenever the row has NT of them; others read x4[0]), so
        // all C4 loads are in flight together
#pragma unroll
        for (int k = 0; k < C4; ++k) r[k] = x4[min(t + k * NT, max(n4 - 1, 0))];
    }
    // f(const float (&v)[4]) for every float4 the thread owns
    template <class F>
    __device__ __forceinline__ void for_each(F &&f) const {
#pragma unroll
        for (int k = 0; k < C4; ++k)
            if (k < nv) {
                const float v[4] = {r[k].x, r[k].y, r[k].z, r[k].w};
                f(v);
            }
    }
};

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Visits every element of row x[0..n), in no particular order, in small
// batches: f(const float (&v)[N]) with N = 4 (aligned rows, 16-byte loads,
// four in flight per thread) or N = 1.  A batch lets f issue the LDS reads
// of all its values before it waits for any of them.
template <class F>
__device__ __forceinline__ void row_for_each(const float *__restrict__ x, int64_t n, F &&f) {
    if ((n & 3) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0) {
        const float4 *x4 = reinterpret_cast<const float4 *>(x);
        const int64_t n4 = n >> 2;
        int64_t i = threadIdx.x;
        for (; i + 3 * NT < n4; i += 4 * NT) {
            const float4 a = x4[i], b = x4[i + NT], c = x4[i + 2 * NT], d = x4[i + 3 * NT];
            const float va[4] = {a.x, a.y, a.z, a.w}, vb[4] = {b.x, b.y, b.z, b.w};
            const float vc[4] = {c.x, c.y, c.z, c.w}, vd[4] = {d.x, d.y, d.z, d.w};
            f(va); f(vb); f(vc); f(vd);
        }
        for (; i < n4; i += NT) {
            const float4 a = x4[i];
            const float va[4] = {a.x, a.y, a.z, a.w};
            f(va);
        }
    } else {
        for (int64_t i = threadIdx.x; i < n; i += NT) {
            const float v[1] = {x[i]};
            f(v);
        }
    }
}

// Sorts the 64*E values a wavefront holds (value v[e] of lane l is element
// e*64 + l) in descending order with a bitonic network: partners in other
// lanes are reached by shuffles, partners in the same lane directly.
template <int E>
__device__ __forceinline__ void wave_bitonic_desc(float (&v)[E]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int size = 2; size <= 64 * E; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            float nv[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int idx = e * 64 + lane;
                const float other = stride < 64 ? __shfl_xor(v[e], stride, 64) : v[e ^ (stride >> 6)];
                // blocks with (idx & size) == 0 run descending; the lower index of a
                // pair keeps the larger value in a descending block
                const bool desc = (idx & size) == 0;
                const bool lower = (idx & stride) == 0;
                nv[e] = (lower == desc) ? fmaxf(v[e], other) : fminf(v[e], other);
            }
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = nv[e];
        }
    }
}

// Writes T[j] for the targets j0 <= j < j1 of one bin whose h elements sit at
// gbuf[seg..seg+h) (h <= 64*E): sorts them in registers, then reads the
// element at position ranks[j] - above.
template <int E>
__device__ __forceinline__ void wave_sort_select(const float *gbuf, uint32_t seg, uint32_t h, uint32_t above,
                                                 const Ranks &ranks, int j0, int j1, float *T) {
    const int lane = threadIdx.x & 63;
    float v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = (uint32_t)(e * 64 + lane);
        v[e] = i < h ? gbuf[seg + i] : -INFINITY;  // padding sorts last (all scores are finite)
    }
    wave_bitonic_desc<E>(v);
    for (int j = j0; j < j1; ++j) {
        const uint32_t pos = (uint32_t)ranks[j] - above;  // < h
        float x = 0.0f;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const float y = __shfl(v[e], (int)(pos & 63u), 64);
            if ((int)(pos >> 6) == e) x = y;
        }
        if (lane == 0) T[j] = x;
    }
}

// Workgroup min/max; every thread gets the result.  One barrier: every call
// site reaches another barrier before red_f is written again.
__device__ __forceinline__ void wg_minmax(SelSmem &sm, float &mn, float &mx) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    mn = wave_min(mn);
    mx = wave_max(mx);
    if (lane == 0) { sm.red_f[w] = mn; sm.red_f[NW + w] = mx; }
    __syncthreads();
    mn = sm.red_f[0];
    mx = sm.red_f[NW];
#pragma unroll
    for (int i = 1; i < NW; ++i) { mn = fminf(mn, sm.red_f[i]); mx = fmaxf(mx, sm.red_f[NW + i]); }
}

// Exclusive scan of one value per thread; returns the thread's prefix and
// writes the total to *total.  One barrier: each thread adds the wave totals
// itself, and every call site reaches another barrier before red_u is written
// again.
__device__ __forceinline__ uint32_t wg_excl_scan(SelSmem &sm, uint32_t v, uint32_t *total) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sm.red_u[w] = x;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const uint32_t t = sm.red_u[i];
        before += i < w ? t : 0u;
        tot += t;
    }
    *total = tot;
    return before + x - v;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Turns counts h[0..nb) into incl[b] = sum_{b' >= b} h[b'] in place
// (incl[nb] = 0).  nb is a multiple of NT; thread t owns h[t*per, t*per+per).
// With per = 16 a thread's bins are four 16-byte granules: lane t reads (and
// writes) granule (i + t/2) mod 4 in step i, so the lanes of one step cover
// all 32 LDS banks instead of the 8 that a stride of 16 words reaches.
template <int nb>
__device__ __forceinline__ void wg_reverse_incl(SelSmem &sm, uint32_t *h) {
    constexpr int per = nb / NT;
    static_assert(per >= 1 && per * NT == nb, "bins must be a multiple of the workgroup");
    const int base = threadIdx.x * per;
    if constexpr (per == 16) {
        const int rot = (threadIdx.x >> 1) & 3;
        u32x4 q[4], nat[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) q[i] = *reinterpret_cast<const u32x4 *>(&h[base + 4 * ((i + rot) & 3)]);
#pragma unroll
        for (int g = 0; g < 4; ++g) {  // granule g was read in step (g - rot) mod 4
            const int i = (g - rot) & 3;
            u32x4 v = q[0];
#pragma unroll
            for (int k = 1; k < 4; ++k) v = __builtin_unpredictable(i == k) ? q[k] : v;
            nat[g] = v;
        }
        uint32_t s = 0;
#pragma unroll
        for (int g = 0; g < 4; ++g) s += nat[g].x + nat[g].y + nat[g].z + nat[g].w;
        uint32_t total;
        uint32_t e = wg_excl_scan(sm, s, &total);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            u32x4 o;
            o.x = total - e; e += nat[g].x;
            o.y = total - e; e += nat[g].y;
            o.z = total - e; e += nat[g].z;
            o.w = total - e; e += nat[g].w;
            nat[g] = o;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {  // step i writes granule (i + rot) mod 4
            const int g = (i + rot) & 3;
            u32x4 v = nat[0];
#pragma unroll
            for (int k = 1; k < 4; ++k) v = __builtin_unpredictable(g == k) ? nat[k] : v;
            *reinterpret_cast<u32x4 *>(&h[base + 4 * g]) = v;
        }
    } else {
        uint32_t loc[per];
        uint32_t s = 0;
#pragma unroll
        for (int i = 0; i < per; ++i) { loc[i] = h[base + i]; s += loc[i]; }
        uint32_t total;
        uint32_t e = wg_excl_scan(sm, s, &total);
#pragma unroll
        for (int i = 0; i < per; ++i) { h[base + i] = total - e; e += loc[i]; }
    }
    if (threadIdx.x == 0) h[nb] = 0;
    __syncthreads();
}

// Largest b in [0, nb) with incl[b] > t (incl non-increasing, incl[0] > t).
__device__ __forceinline__ int find_bin(const uint32_t *incl, int nb, uint32_t t) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (incl[mid] > t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// t-th largest (0-based) among the elements of src[0..n) that lie in the
// value range [lo, hi], which must hold more than t of them.  All threads
// call it and get the value.  Each round keeps the one bin of NB2 that
// holds the answer, so the range shrinks until it is a single value.
__device__ float wg_select(SelSmem &sm, const float *src, int64_t n, float lo, float hi, uint32_t t) {
    // Every round drops at least one distinct value from the range, so the
    // loop ends for finite inputs; the round cap only guards non-finite data
    // (rejected by the host before any kernel runs).
    for (int round = 0; round < 256 && lo < hi; ++round) {
        const float scale = (float)NB2 / (hi - lo);
        for (int i = threadIdx.x; i < NB2; i += NT) sm.hist2[i] = 0;
        __syncthreads();
        for (int64_t i = threadIdx.x; i < n; i += NT) {
            const float x = src[i];
            if (x >= lo && x <= hi) atomicAdd(&sm.hist2[bin_f32(x, lo, scale, NB2)], 1u);
        }
        __syncthreads();
        wg_reverse_incl<NB2>(sm, sm.hist2);
        if (threadIdx.x == 0) sm.sel_b = find_bin(sm.hist2, NB2, t);
        __syncthreads();
        const int b = sm.sel_b;
        t -= sm.hist2[b + 1];
        float mn = INFINITY, mx = -INFINITY;
        for (int64_t i = threadIdx.x; i < n; i += NT) {
            const float x = src[i];
            if (x >= lo && x <= hi && bin_f32(x, lo, scale, NB2) == b) { mn = fminf(mn, x); mx = fmaxf(mx, x); }
        }
        wg_minmax(sm, mn, mx);
        lo = mn;
        hi = mx;
    }
    return lo;
}

// One group of target bins of the row (a pass of the gather over the row, or
// one bin larger than LDS narrowed on its own): `each` visits the row's values.
template <class Each>
__device__ __forceinline__ void select_group(SelSmem &sm, int g, int ngrp, int nfb, const float *__restrict__ x,
                                             int64_t n_item, const Ranks &ranks, int32_t S, float *T,
                                             float lo0, float scale1, Each &&each) {
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const bool deep = (sm.grp[g] & DEEP) != 0;
    const int f0 = sm.grp[g] & ~DEEP;
    const int f1 = sm.grp[g + 1] & ~DEEP;
    if (deep) {
        // one bin larger than LDS: narrow it from the row in global memory
        const int j0 = sm.fb[f0];
        const int j1 = (f0 + 1 < nfb) ? sm.fb[f0 + 1] : S;
        const int b = sm.tb[j0];
        float bmn = INFINITY, bmx = -INFINITY;
        each([&](const auto &v) {
            constexpr int N = sizeof(v) / sizeof(float);
            int bv[N];
            fine_bins<N>(sm.fseg, v, lo0, scale1, bv);
#pragma unroll
            for (int q = 0; q < N; ++q)
                if (bv[q] == b) { bmn = fminf(bmn, v[q]); bmx = fmaxf(bmx, v[q]); }
        });
        wg_minmax(sm, bmn, bmx);
        for (int j = j0; j < j1; ++j) {
            const uint32_t t = (uint32_t)ranks[j] - sm.fabove[f0];
            const float v = wg_select(sm, x, n_item, bmn, bmx, t);
            if (tid == 0) T[j] = v;
        }
        return;
    }
    // member bins: bitmap bits, then segment starts (bins descending in f)
    for (int f = f0 + tid; f < f1; f += NT) {
        const int b = sm.tb[sm.fb[f]];
        atomicOr(&sm.tbits[b >> 5], 1u << (b & 31));
    }
    if (ngrp > 1) {
        // several passes: offsets restart at each group's first bin
        __syncthreads();
        const uint32_t base = sm.fcur[f0];
        __syncthreads();
        for (int f = f0 + tid; f < f1; f += NT) sm.fcur[f] -= base;
    }
    for (int i = tid; i < GCAP / 32; i += NT) sm.sbits[i] = 0;
    if (tid == 0) { sm.ncoop = 0; sm.nmid = 0; }
    __syncthreads();
    {
        // member bins below each bitmap word (ascending bins)
        const uint32_t cnt = tid < NB / 32 ? (uint32_t)__popc(sm.tbits[tid]) : 0u;
        uint32_t tot;
        const uint32_t pre = wg_excl_scan(sm, cnt, &tot);
        if (tid < NB / 32) sm.tpre[tid] = (uint16_t)pre;
    }
    __syncthreads();
    // gather: the member bin of rank m in ascending bin order is f1 - 1 - m
    // A batch's bins, bitmap words and prefixes are read together, and its
    // members' slots are all claimed before the first is written.
    each([&](const auto &v) {
        constexpr int N = sizeof(v) / sizeof(float);
        int b[N];
        fine_bins<N>(sm.fseg, v, lo0, scale1, b);
        uint32_t word[N], pre[N];
#pragma unroll
        for (int q = 0; q < N; ++q) { word[q] = sm.tbits[b[q] >> 5]; pre[q] = sm.tpre[b[q] >> 5]; }
        uint32_t slot[N];
        bool mem[N];
#pragma unroll
        for (int q = 0; q < N; ++q) {
            const uint32_t bit = 1u << (b[q] & 31);
            mem[q] = (word[q] & bit) != 0;
            slot[q] = 0;
            if (mem[q]) slot[q] = atomicAdd(&sm.fcur[f1 - 1 - ((int)pre[q] + __popc(word[q] & (bit - 1u)))], 1u);
        }
#pragma unroll
        for (int q = 0; q < N; ++q)
            if (mem[q]) sm.u.gbuf[slot[q]] = v[q];
    });
    __syncthreads();
    // Bins of up to EMAX elements are ranked element-parallel: each gathered
    // element finds its bin from the segment starts (a bitmap over the gather
    // buffer and its popcount prefix), counts the elements of its segment
    // that order before it (larger, or equal at a lower index) and, if that
    // rank is a target, writes it.  Larger bins are listed for a wavefront or
    // the whole workgroup.
    for (int f = f0 + tid; f < f1; f += NT) {
        const uint32_t h = sm.fh[f];
        const uint32_t st = sm.fcur[f] - h;
        atomicOr(&sm.sbits[st >> 5], 1u << (st & 31));
        if (h > WAVE_SEL_MAX) sm.coop[atomicAdd(&sm.ncoop, 1)] = (uint16_t)f;
        else if (h > EMAX) sm.mid[atomicAdd(&sm.nmid, 1)] = (uint16_t)f;
    }
    __syncthreads();
    {
        const uint32_t cnt = tid < GCAP / 32 ? (uint32_t)__popc(sm.sbits[tid]) : 0u;
        uint32_t tot;
        const uint32_t pre = wg_excl_scan(sm, cnt, &tot);
        if (tid < GCAP / 32) sm.spre[tid] = (uint16_t)pre;
    }
    __syncthreads();
    const uint32_t gtot = sm.fcur[f1 - 1];  // elements gathered in this pass
    for (uint32_t i = tid; i < gtot; i += NT) {
        const uint32_t word = sm.sbits[i >> 5];
        const int f = f0 + (int)sm.spre[i >> 5] + __popc(word & (0xffffffffu >> (31 - (i & 31)))) - 1;
        const uint32_t h = sm.fh[f];
        if (h > EMAX) continue;
        const uint32_t seg = sm.fcur[f] - h;
        const float v = sm.u.gbuf[i];
        uint32_t pos = 0;
        for (uint32_t q = seg; q < seg + h; ++q) {
            const float o = sm.u.gbuf[q];
            pos += (o > v || (o == v && q < i)) ? 1u : 0u;
        }
        const uint32_t r = sm.fabove[f] + pos;
        const int j0 = sm.fb[f], j1 = (f + 1 < nfb) ? sm.fb[f + 1] : S;
        const uint32_t r0 = (uint32_t)ranks[j0];
        if (r < r0) continue;
        // ranks increase strictly, so rank r can only sit at index <= j0 + (r - r0)
        int hi = min(j0 + (int)(r - r0), j1 - 1), lo = j0;
        if ((uint32_t)ranks[hi] <= r) {
            lo = hi;
        } else {
            --hi;
            while (lo < hi) {
                const int m = (lo + hi + 1) >> 1;
                if ((uint32_t)ranks[m] <= r) lo = m; else hi = m - 1;
            }
        }
        if ((uint32_t)ranks[lo] == r) T[lo] = v;
    }
    // mid-sized bins: one wavefront each
    const int nmid = sm.nmid;
    for (int m = w; m < nmid; m += NW) {
        const int f = sm.mid[m];
        const uint32_t h = sm.fh[f];
        const int j0 = sm.fb[f];
        const int j1 = (f + 1 < nfb) ? sm.fb[f + 1] : S;
        const uint32_t seg = sm.fcur[f] - h, above = sm.fabove[f];
        if (h <= 128) { wave_sort_select<2>(sm.u.gbuf, seg, h, above, ranks, j0, j1, T); continue; }
        if (h <= 256) { wave_sort_select<4>(sm.u.gbuf, seg, h, above, ranks, j0, j1, T); continue; }
        // larger bins (rare): rank each element by counting
        for (uint32_t i = lane; i < h; i += 64) {
            const float e = sm.u.gbuf[seg + i];
            uint32_t pos = 0;
            for (uint32_t k = 0; k < h; ++k) {
                const float o = sm.u.gbuf[seg + k];
                pos += (o > e || (o == e && k < i)) ? 1u : 0u;
            }
            const uint32_t r = above + pos;
            for (int j = j0; j < j1; ++j)
                if ((uint32_t)ranks[j] == r) T[j] = e;
        }
    }
    __syncthreads();
    // bins too large for a wavefront: the whole workgroup
    const int ncoop = sm.ncoop;
    for (int c = 0; c < ncoop; ++c) {
        const int f = sm.coop[c];
        const int j0 = sm.fb[f];
        const int j1 = (f + 1 < nfb) ? sm.fb[f + 1] : S;
        const uint32_t h = sm.fh[f];
        const uint32_t seg = sm.fcur[f] - h;
        float bmn = INFINITY, bmx = -INFINITY;
        for (uint32_t i = tid; i < h; i += NT) { const float v = sm.u.gbuf[seg + i]; bmn = fminf(bmn, v); bmx = fmaxf(bmx, v); }
        wg_minmax(sm, bmn, bmx);
        for (int j = j0; j < j1; ++j) {
            const uint32_t t = (uint32_t)ranks[j] - sm.fabove[f];
            const float v = wg_select(sm, sm.u.gbuf + seg, h, bmn, bmx, t);
            if (tid == 0) T[j] = v;
        }
    }
    // release the member bins
    for (int f = f0 + tid; f < f1; f += NT) {
        const int b = sm.tb[sm.fb[f]];
        sm.tbits[b >> 5] = 0;
    }
    __syncthreads();
}

// C4 > 0: the row is held in registers (RowRegs<C4>); C4 == 0: every pass
// reads it from memory.
template <int C4>
__global__ void __launch_bounds__(NT) k_select_rows(const float *__restrict__ scores, int64_t n_item,
                                                   const int32_t *__restrict__ granks, int32_t S,
                                                   float *__restrict__ table_out) {
    __shared__ SelSmem sm;
    const int64_t row = blockIdx.x;
    const float *x = scores + row * n_item;
    float *T = table_out + row * (int64_t)S;
    const int tid = threadIdx.x;
    // the target ranks in LDS when they fit 16 bits (first read after the
    // barriers of the passes below)
    const bool lds_ranks = n_item <= 65536;
    if (lds_ranks)
        for (int j = tid; j < S; j += NT) sm.srk[j] = (uint16_t)granks[j];
    const Ranks ranks{granks, sm.srk, lds_ranks};
    RowRegs<(C4 > 0 ? C4 : 1)> rr;
    if constexpr (C4 > 0) rr.load(x, n_item);
    auto each = [&](auto &&f) {
        if constexpr (C4 > 0) rr.for_each(f);
        else row_for_each(x, n_item, f);
    };

    // 1. range
    float mn = INFINITY, mx = -INFINITY;
    each([&](const auto &v) {
        for (float e : v) { mn = fminf(mn, e); mx = fmaxf(mx, e); }
    });
    wg_minmax(sm, mn, mx);
    const float lo0 = mn, hi0 = mx;
    if (!(lo0 < hi0) || isinf(lo0) || isinf(hi0)) {
        for (int j = tid; j < S; j += NT) T[j] = lo0;
        return;
    }
    // 2a. coarse histogram: the level-0 bins are apportioned to NB1 linear
    // coarse bins by their element counts (each gets at least one), so
    // level-0 bins are narrow where the row is dense (a row with a heavy tail
    // otherwise puts hundreds of elements in each central linear bin)
    const float scale1 = (float)NB1 / (hi0 - lo0);
    for (int i = tid; i < NB1; i += NT) sm.hist2[i] = 0;
    __syncthreads();
    each([&](const auto &v) {
        for (float e : v) atomicAdd(&sm.hist2[bin_f32(e, lo0, scale1, NB1)], 1u);
    });
    __syncthreads();
    {
        constexpr int PER1 = NB1 / NT > 0 ? NB1 / NT : 1;  // coarse bins per thread
        uint32_t c_loc[PER1], sum = 0;
        for (int i = 0; i < PER1; ++i) {
            const int c = tid * PER1 + i;
            c_loc[i] = c < NB1 ? sm.hist2[c] : 0u;
            sum += c_loc[i];
        }
        uint32_t total;
        uint32_t cum = wg_excl_scan(sm, sum, &total);  // elements below the thread's first coarse bin
        // coarse bin c starts at level-0 bin c + floor((NB - NB1) * below(c) / total); the last ends at NB
        const uint64_t den = max(total, 1u);
        for (int i = 0; i < PER1; ++i) {
            const int c = tid * PER1 + i;
            const uint32_t s0 = (uint32_t)c + (uint32_t)(((uint64_t)(NB - NB1) * cum) / den);
            cum += c_loc[i];
            const uint32_t s1 = (uint32_t)c + 1u + (uint32_t)(((uint64_t)(NB - NB1) * cum) / den);
            if (c < NB1) sm.fseg[c] = s0 | (s1 << 16);
        }
    }

    // 2b. histogram + reverse prefix sums
    for (int i = tid; i < NB; i += NT) sm.u.incl[i] = 0;
    for (int i = tid; i < NB / 32; i += NT) sm.tbits[i] = 0;
    __syncthreads();
    each([&](const auto &v) {
        constexpr int N = sizeof(v) / sizeof(float);
        int b[N];
        fine_bins<N>(sm.fseg, v, lo0, scale1, b);
#pragma unroll
        for (int q = 0; q < N; ++q) atomicAdd(&sm.u.incl[b[q]], 1u);
    });
    __syncthreads();
    wg_reverse_incl<NB>(sm, sm.u.incl);

    // bin of every target: the largest b with incl[b] > r, by a fixed-length
    // search (incl[0] > r for every target, incl[NB] = 0), the thread's
    // targets searched side by side so their LDS reads overlap
    {
        constexpr int PER = SMAX / NT;
        uint32_t r[PER];
        int b[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int j = tid + i * NT;
            r[i] = j < S ? (uint32_t)ranks[j] : 0u;
            b[i] = 0;
        }
#pragma unroll
        for (int step = NB / 2; step >= 1; step >>= 1)
#pragma unroll
            for (int i = 0; i < PER; ++i) b[i] += sm.u.incl[b[i] + step] > r[i] ? step : 0;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int j = tid + i * NT;
            if (j < S) sm.tb[j] = (uint16_t)b[i];
        }
    }
    __syncthreads();

    // 3. distinct target bins (parallel scans): fb[f] is the first target of
    // distinct bin f, fh[f] its element count, fabove[f] the elements above it
    {
        constexpr int PER = SMAX / NT;  // targets per thread (S <= SMAX)
        uint32_t nf = 0;
        bool fl[PER];
        for (int i = 0; i < PER; ++i) {
            const int j = tid * PER + i;
            fl[i] = j < S && (j == 0 || sm.tb[j] != sm.tb[j - 1]);
            nf += fl[i] ? 1u : 0u;
        }
        uint32_t nf_tot;
        uint32_t fpos = wg_excl_scan(sm, nf, &nf_tot);
        for (int i = 0; i < PER; ++i) {
            if (!fl[i]) continue;
            const int j = tid * PER + i;
            const int b = sm.tb[j];
            sm.fb[fpos] = (uint16_t)j;
            sm.fh[fpos] = sm.u.incl[b] - sm.u.incl[b + 1];
            sm.fabove[fpos] = sm.u.incl[b + 1];
            ++fpos;
        }
        if (tid == 0) sm.nfb = (int)nf_tot;
    }
    __syncthreads();
    // the histogram is no longer needed: its storage becomes the gather buffer.
    // fcur[f] = elements of the distinct bins before f (a parallel scan); when
    // they all fit, one pass gathers them at those offsets.  Otherwise bins are
    // packed greedily into passes of <= GCAP floats, and a bin larger than
    // GCAP is narrowed from global memory on its own.
    const int nfb = sm.nfb;
    {
        constexpr int PER = SMAX / NT;
        uint32_t h[PER], sum = 0;
        for (int i = 0; i < PER; ++i) {
            const int f = tid * PER + i;
            h[i] = f < nfb ? sm.fh[f] : 0u;
            sum += h[i];
        }
        uint32_t total;
        uint32_t off = wg_excl_scan(sm, sum, &total);
        for (int i = 0; i < PER; ++i) {
            const int f = tid * PER + i;
            if (f < nfb) sm.fcur[f] = off;
            off += h[i];
        }
        if (tid == 0) {
            if (total <= GCAP) {
                sm.grp[0] = 0;
                sm.grp[1] = (uint16_t)nfb;
                sm.ngrp = 1;
            } else {
                int ngrp = 0;
                uint32_t fill = GCAP;
                for (int f = 0; f < nfb; ++f) {
                    const uint32_t hf = sm.fh[f];
                    if (hf > GCAP) {
                        sm.grp[ngrp++] = (uint16_t)(f | DEEP);
                        fill = GCAP;  // next bin starts a new group
                    } else {
                        if (fill + hf > GCAP) { sm.grp[ngrp++] = (uint16_t)f; fill = 0; }
                        fill += hf;
                    }
                }
                sm.grp[ngrp] = (uint16_t)nfb;
                sm.ngrp = ngrp;
            }
        }
    }
    __syncthreads();
    const int ngrp = sm.ngrp;

    // the first group reads the row from wherever the kernel holds it; any
    // later one (rows whose target bins exceed LDS) reads it from memory, so
    // the registers of a register-held row are free after the first gather
    if (ngrp > 0) select_group(sm, 0, ngrp, nfb, x, n_item, ranks, S, T, lo0, scale1, each);
    for (int g = 1; g < ngrp; ++g)
        select_group(sm, g, ngrp, nfb, x, n_item, ranks, S, T, lo0, scale1,
                     [&](auto &&f) { row_for_each(x, n_item, f); });
}

}  // namespace

hipError_t launch_select_rows(hipStream_t s, const float *scores, int64_t nb, int64_t n_item, const int32_t *ranks,
                              int32_t S, float *table_out) {
    if (nb <= 0) return hipSuccess;
    if (S > SMAX || S <= 0) return hipErrorInvalidValue;
    // rows of a multiple of 4 values (every row then starts 16-byte aligned)
    // up to 16 * C4 * NT values go through registers
    const bool regs = (n_item & 3) == 0 && (reinterpret_cast<uintptr_t>(scores) & 15) == 0;
    const int64_t per = (n_item / 4 + NT - 1) / NT;  // float4s per thread
    const dim3 grid((unsigned)nb), block(NT);
    if (regs && per <= 4) hipLaunchKernelGGL(k_select_rows<4>, grid, block, 0, s, scores, n_item, ranks, S, table_out);
    else if (regs && per <= 8) hipLaunchKernelGGL(k_select_rows<8>, grid, block, 0, s, scores, n_item, ranks, S, table_out);
    else if (regs && per <= 13) hipLaunchKernelGGL(k_select_rows<13>, grid, block, 0, s, scores, n_item, ranks, S, table_out);
    else hipLaunchKernelGGL(k_select_rows<0>, grid, block, 0, s, scores, n_item, ranks, S, table_out);
    return hipGetLastError();
}

static_assert(sizeof(SelSmem) <= 163840, "selection state must fit the 160 KB LDS of a gfx950 CU");

}  // namespace rkr
