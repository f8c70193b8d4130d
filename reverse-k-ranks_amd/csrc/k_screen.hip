// k_screen.hip -- screened exact-rank refinement (RKR_FORM_SCREENED).
//
// For each refine row r (a (query, user) pair with threshold thr_r = <u,q>)
// the rank is #{items p : s(u,p) > thr_r}, with s the sequential FP32 score
// rule.  The screened form computes an approximation s^ of every s(u,p) on
// the matrix cores and decides most items from it:
//
//   x = x_hi + x_lo + x_res with x_hi = bf16(x), x_lo = bf16(x - x_hi)
//   s^ = sum_k u_hi p_hi + u_hi p_lo + u_lo p_hi     (v_mfma_f32_16x16x32_bf16,
//                                                      FP32 accumulation)
//   |s^ - s| <= e(u,p) = C * ||u|| * ||p|| + abs_floor
//
// C covers the dropped terms (u_lo p_lo and the residues, <= 3.2*2^-16 of
// sum|u_k p_k|), the FP32 accumulation of the 3*Kp exact products (any
// order, <= 2^-23 per addition), the rounding of the sequential rule itself
// (gamma_d), and Cauchy-Schwarz sum|u_k p_k| <= ||u|| ||p||; abs_floor covers
// denormals a matrix core may flush.  Then
//   s^ - e > thr  => counted (s > thr);   s^ + e <= thr => not counted;
// anything in between is appended to a recheck list and decided by the exact
// sequential FP32 score (k_recheck), so every rank stays bit-exact.
//
// Work unit = 128 refine rows (4 waves x 2 blocks of 16 rows; A fragments
// held in VGPRs) x a chunk of items streamed through LDS 64 items at a time.
// Units are ordered chunk-major, so the workgroups resident at one time all
// stream the same item chunk, which stays in each XCD's L2.  The next stage
// is loaded into registers while the current one is computed from LDS.
#include "rkr_internal.hpp"

namespace rkr {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;
constexpr int ITEMS_STAGE = 64;
constexpr int KPAD_LDS = 8;  // bf16 pad per LDS row

// Round-to-nearest-even FP32 -> BF16 (finite inputs).
__device__ __forceinline__ uint16_t bf16_bits(float x) {
    const uint32_t b = __float_as_uint(x);
    return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

union Frag {
    bf16x8 v;
    uint16_t h[8];
    uint4 q;
};

// Splits every item row into hi/lo BF16 rows (Kp columns, zero padded) and
// stores an upper bound of each item's L2 norm (FP64 sum, rounded up).
// With cs (queries of the coarse first pass) it also stores each row q's
// error scale of the one-product estimate u_hi . q_hi,
//   cs = (1 + rho) (g1 ||q_hi|| + ||q_r||) + rho (||q_hi|| + ||q_r||) + g_rule ||q||,
// rounded up, where q_r = q - q_hi (exact in FP32) and rho >= ||u_r|| / ||u||
// over the users: |u_hi . q_hi (FP32 sums) - s(u, q)| <= ||u|| cs + the
// denormal floor.  (u_hi . q_r + u_r . q_hi + u_r . q_r by Cauchy-Schwarz with
// ||u_hi|| <= (1 + rho) ||u||; g1 the FP32 accumulation of the Kp exact
// products, g_rule the sequential rule's own rounding.)
__global__ void k_prep_items(const float *__restrict__ items, int64_t n_item, int32_t ld, int32_t K, int32_t Kp,
                             uint16_t *__restrict__ hi, uint16_t *__restrict__ lo, float *__restrict__ norm,
                             float *__restrict__ cs, const float *__restrict__ rho, double g1, double g_rule) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.y + threadIdx.y;
    if (p >= n_item) return;
    const float *x = items + p * ld;
    double ss = 0.0, sh = 0.0, sr = 0.0;
    for (int k = threadIdx.x; k < Kp; k += blockDim.x) {
        const float v = k < K ? x[k] : 0.0f;
        const uint16_t h = bf16_bits(v);
        const float r = v - bf16_to_f32(h);  // exact: bf16(v) is within a factor 2 of v
        const uint16_t l = bf16_bits(r);
        hi[p * Kp + k] = h;
        lo[p * Kp + k] = l;
        ss += (double)v * (double)v;
        sh += (double)bf16_to_f32(h) * (double)bf16_to_f32(h);
        sr += (double)r * (double)r;
    }
    for (int o = 16; o > 0; o >>= 1) {
        ss += __shfl_xor(ss, o, 32);
        sh += __shfl_xor(sh, o, 32);
        sr += __shfl_xor(sr, o, 32);
    }
    if (threadIdx.x == 0 && norm) norm[p] = __double2float_ru(sqrt(ss));
    if (threadIdx.x == 0 && cs) {
        // the FP64 sums' own rounding (< 2^-40 relative) is covered by the
        // 2^-30 margin
        const double rh = (double)*rho, nh = sqrt(sh), nr = sqrt(sr);
        const double c = (1.0 + rh) * (g1 * nh + nr) + rh * (nh + nr) + g_rule * sqrt(ss);
        cs[p] = __double2float_ru(c * (1.0 + 0x1p-30));
    }
}

// Upper bound of each user's L2 norm; with rho, also raises *rho to at least
// ||u - bf16(u)|| / ||u|| of every user (rounded up; non-negative floats
// order as their bits).
__global__ void k_user_norms(const float *__restrict__ users, int64_t n_user, int32_t ld, int32_t K,
                             float *__restrict__ norm, float *__restrict__ rho) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n_user) return;
    double ss = 0.0, sr = 0.0;
    for (int k = 0; k < K; ++k) {
        const float f = users[u * ld + k];
        const double v = f, r = f - bf16_to_f32(bf16_bits(f));
        ss += v * v;
        sr += r * r;
    }
    norm[u] = __double2float_ru(sqrt(ss));
    if (rho && ss > 0.0) atomicMax(reinterpret_cast<int *>(rho), __float_as_int(__double2float_ru(sqrt(sr / ss) * (1.0 + 0x1p-30))));
}

struct ScreenParams {
    const float *users;        // n_user x ld FP32
    const float *user_norm;    // n_user
    const uint16_t *user_hi, *user_lo;  // n_user x Kp BF16 split
    const uint16_t *item_hi;   // n_item x Kp
    const uint16_t *item_lo;
    const float *item_norm;    // n_item
    int64_t n_item;
    int32_t ld, K, Kp;
    const int32_t *rows_user;  // refine rows
    const float *rows_thr;
    const int32_t *n_rows_dev;
    int64_t max_rows;
    int64_t target_units;      // work units wanted to fill the device
    int32_t chunk_items;       // most items per work-unit chunk
    float C;                   // relative error factor
    float abs_floor;           // absolute error floor
    int32_t *cnt;              // per refine row
    int32_t *recheck_n;        // appended rechecks (may exceed capacity)
    int4 *recheck;             // (row, item, s^ bits, 0)
    int64_t recheck_cap;
};

// RB row blocks of 16 per wave: 2 (32 rows, two waves per SIMD) or 4 (64
// rows, one wave per SIMD, half the LDS reads per matrix-core operation).
template <int KS, int RB>
__global__ void __launch_bounds__(NT, RB == 2 ? 2 : 1) k_refine_screen(ScreenParams P) {
    constexpr int RW = 4 * 16 * RB;  // rows per workgroup
    // items per LDS stage: 32 keeps the two prefetch register sets at 8 KS
    // VGPRs
    constexpr int IST = 32;
    // two stage buffers: stage m of a unit is computed from buffer m & 1
    // while stage m + 1 is written to the other and stage m + 3 is loaded
    // into registers, with one barrier per stage
    __shared__ uint4 sB[2][2][IST * (KS * 32 + KPAD_LDS) / 8];  // [buf][hi|lo][item][k] as 16-byte chunks
    __shared__ float sN[2][IST];                                // item norm bounds of the stage
    // undecided (row, item) pairs of the unit, appended with LDS atomics and
    // flushed to the recheck list when the unit ends (a global atomic that
    // returns a slot would make the stage loop wait for its prefetch loads)
    constexpr int QCAP = 256;
    __shared__ int4 sQ[QCAP];
    __shared__ int32_t sQn, sQbase;
    constexpr int ROWQ = (KS * 32 + KPAD_LDS) / 8;                   // uint4 per LDS item row
    constexpr int KQ = KS * 4;                                       // uint4 per global item row (Kp = 32 KS)
    constexpr int PER_T = 2 * IST * KQ / NT;                 // uint4 each thread stages per stage
    static_assert(2 * IST * KQ % NT == 0, "stage must split evenly over the workgroup");
    // Persistent grid: the work split is sized here from the refine count the
    // compaction produced, so the host launches without reading it back.
    const int64_t n_rows = min((int64_t)*P.n_rows_dev, P.max_rows);
    const int64_t n_rb = (n_rows + RW - 1) / RW;
    const int64_t n_stage_total = (P.n_item + IST - 1) / IST;
    const int64_t fill_chunks = n_rb > 0 ? (P.target_units + n_rb - 1) / n_rb : 1;
    const int64_t st_per_chunk =
        max((int64_t)1, min((int64_t)max(1, P.chunk_items / IST), (n_stage_total + fill_chunks - 1) / fill_chunks));
    const int64_t n_chunks = (n_stage_total + st_per_chunk - 1) / st_per_chunk;
    const int64_t n_units = n_rb * n_chunks;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    // Register sets for the stages loaded ahead: with two (where the A
    // fragments leave room), stage m of a unit lives in set A (m even) or B
    // (m odd) from its load, three stages ahead, until it is written to LDS;
    // with one, two stages ahead.  Every lane issues the same
    // loads in every call, with no branch around them and no use of their
    // values until the set is written to LDS, so the memory counter then
    // waits for exactly the older set, never for the loads just issued.  Rows
    // past n_item load the last item; compute() gives them a norm bound of
    // -inf, so they never count.
    constexpr int NSET = (KS <= 4 || (RB == 4 && KS <= 5)) ? 2 : 1;
    // native vectors, not HIP's uint4 struct: a struct copy is a memcpy that
    // keeps the sets in scratch memory
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    u32x4 preA[PER_T], preB[NSET == 2 ? PER_T : 1];
    float preA_n = -INFINITY, preB_n = -INFINITY;
#define RKR_LOAD_STAGE(st, dst, dst_n)                                                                        \
    do {                                                                                                    \
        const int64_t p0_ = (st) * IST;                                                                     \
        _Pragma("unroll") for (int j = 0; j < PER_T; ++j) {                                                 \
            const int c = threadIdx.x + j * NT;                                                             \
            const int half = c / (IST * KQ);                                                                \
            const int rem = c % (IST * KQ);                                                                 \
            const int64_t p = min(p0_ + rem / KQ, P.n_item - 1);                                            \
            dst[j] = reinterpret_cast<const u32x4 *>(half ? P.item_lo : P.item_hi)[p * KQ + rem % KQ];     \
        }                                                                                                   \
        dst_n = P.item_norm[min(p0_ + (int64_t)(threadIdx.x % IST), P.n_item - 1)];                         \
    } while (0)
#define RKR_STORE_STAGE(src, src_n, buf)                                                                      \
    do {                                                                                                    \
        _Pragma("unroll") for (int j = 0; j < PER_T; ++j) {                                                 \
            const int c = threadIdx.x + j * NT;                                                             \
            const int half = c / (IST * KQ);                                                                \
            const int rem = c % (IST * KQ);                                                                 \
            *reinterpret_cast<u32x4 *>(&sB[buf][half][(rem / KQ) * ROWQ + rem % KQ]) = src[j];              \
        }                                                                                                   \
        if (threadIdx.x < IST) sN[buf][threadIdx.x] = src_n;                                                \
    } while (0)

    if (threadIdx.x == 0) sQn = 0;
    __syncthreads();
    for (int64_t unit = blockIdx.x; unit < n_units; unit += gridDim.x) {
        const int64_t chunk = unit / n_rb;
        const int64_t rb = unit % n_rb;
        const int64_t row_base = rb * RW + wave * 16 * RB;

        // A fragments: row (row_base + b*16 + lr), k = ks*32 + 8*lq + i
        Frag ah[RB][KS], al[RB][KS];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int64_t r = row_base + b * 16 + lr;
            // A fragments straight from the users' BF16 split (k = ks*32 + 8*lq .. +7)
            const int64_t ur = r < n_rows ? (int64_t)P.rows_user[r] : -1;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const uint4 z = make_uint4(0, 0, 0, 0);
                ah[b][ks].q = ur >= 0 ? reinterpret_cast<const uint4 *>(P.user_hi + ur * P.Kp)[ks * 4 + lq] : z;
                al[b][ks].q = ur >= 0 ? reinterpret_cast<const uint4 *>(P.user_lo + ur * P.Kp)[ks * 4 + lq] : z;
            }
        }
        // epilogue constants of the 4 rows this lane's accumulator covers:
        // D[4*lq + i][lr] of each 16x16 block
        float thr_hi[RB][4], thr_lo[RB][4], Ar[RB][4];
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t r = row_base + b * 16 + 4 * lq + i;
                if (r < n_rows) {
                    const float t = P.rows_thr[r];
                    thr_hi[b][i] = t - P.abs_floor;  // s^ + e > thr - floor  <=>  may count
                    thr_lo[b][i] = t + P.abs_floor;  // s^ - e > thr + floor  =>  counts
                    Ar[b][i] = P.C * P.user_norm[P.rows_user[r]];
                } else {
                    thr_hi[b][i] = INFINITY;
                    thr_lo[b][i] = INFINITY;
                    Ar[b][i] = 0.0f;
                }
            }
        int32_t cnt[RB][4];
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) cnt[b][i] = 0;

        // Wait here for the unit's A fragments and epilogue constants: the
        // stage loop's prefetch loads are issued after them, and a use of an
        // older load inside the loop would otherwise wait for every load in
        // flight (the memory counter is in order), the prefetch included.
        __builtin_amdgcn_s_waitcnt(0);
        const int64_t st0 = chunk * st_per_chunk;
        const int64_t n_st = max((int64_t)0, min(n_stage_total, st0 + st_per_chunk) - st0);
        // Stage m (of this unit) is computed from LDS buffer m & 1.  Before it,
        // stage m + 1 goes from register set (m + 1) & 1 to the other buffer
        // (its last readers, stage m - 1, passed the barrier that ended it) and
        // stage m + 3 starts loading into that register set; one barrier per
        // stage then orders the buffers.
        auto compute = [&](int buf, int64_t st) {
            const int64_t p0 = st * IST;
#pragma unroll 1
            for (int g = 0; g < IST / 16; ++g) {
                const int it = g * 16 + lr;
                const int64_t p = p0 + it;
                // all 3*Kp products of a 16x16 block into one FP32 accumulator
                // (the bound holds for any order of the additions); B
                // fragments are read from LDS one k-step at a time
                f32x4 accs[RB];
#pragma unroll
                for (int b = 0; b < RB; ++b) accs[b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    Frag bh, bl;
                    bh.q = sB[buf][0][it * ROWQ + ks * 4 + lq];
                    bl.q = sB[buf][1][it * ROWQ + ks * 4 + lq];
#pragma unroll
                    for (int b = 0; b < RB; ++b) {
                        accs[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[b][ks].v, bh.v, accs[b], 0, 0, 0);
                        accs[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[b][ks].v, bl.v, accs[b], 0, 0, 0);
                        accs[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[b][ks].v, bh.v, accs[b], 0, 0, 0);
                    }
                }
                // decide the 8 outputs without branches; only undecided ones
                // (rare) take the recheck path
                const float np = p < P.n_item ? sN[buf][it] : -INFINITY;
                uint32_t undecided = 0;
#pragma unroll
                for (int b = 0; b < RB; ++b)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float sc = accs[b][i];
                        const bool may = __builtin_fmaf(Ar[b][i], np, sc) > thr_hi[b][i];
                        const bool sure = __builtin_fmaf(-Ar[b][i], np, sc) > thr_lo[b][i];
                        cnt[b][i] += (may && sure) ? 1 : 0;
                        undecided |= (may && !sure) ? (1u << (b * 4 + i)) : 0u;
                    }
                if (undecided) {
#pragma unroll
                    for (int b = 0; b < RB; ++b)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            if (undecided & (1u << (b * 4 + i))) {
                                const int64_t r = row_base + b * 16 + 4 * lq + i;
                                const int4 e = make_int4((int32_t)r, (int32_t)p, __float_as_int(accs[b][i]), 0);
                                const int32_t q = atomicAdd(&sQn, 1);
                                if (q < QCAP) {
                                    sQ[q] = e;
                                } else {
                                    // the queue is full: straight to the list
                                    const int32_t slot = atomicAdd(P.recheck_n, 1);
                                    if (slot < P.recheck_cap) P.recheck[slot] = e;
                                }
                            }
                }
            }
        };
        // stages past the unit's last load its last again (never computed),
        // so the loads stay unconditional
        const int64_t st_last = st0 + n_st - 1;
        auto ahead = [&](int64_t m) { return min(st0 + m, st_last); };
        if constexpr (NSET == 2) {
            if (n_st > 0) {
                RKR_LOAD_STAGE(ahead(0), preA, preA_n);
                RKR_LOAD_STAGE(ahead(1), preB, preB_n);
                RKR_STORE_STAGE(preA, preA_n, 0);  // the previous unit ended with a barrier
                RKR_LOAD_STAGE(ahead(2), preA, preA_n);
                __syncthreads();
            }
            for (int64_t m = 0; m < n_st; m += 2) {
                // even stage: buffer 0; stage m + 1 is in set B
                RKR_STORE_STAGE(preB, preB_n, 1);
                RKR_LOAD_STAGE(ahead(m + 3), preB, preB_n);
                compute(0, st0 + m);
                __syncthreads();
                if (m + 1 >= n_st) break;
                // odd stage: buffer 1; stage m + 2 is in set A
                RKR_STORE_STAGE(preA, preA_n, 0);
                RKR_LOAD_STAGE(ahead(m + 4), preA, preA_n);
                compute(1, st0 + m + 1);
                __syncthreads();
            }
        } else {
            if (n_st > 0) {
                RKR_LOAD_STAGE(ahead(0), preA, preA_n);
                RKR_STORE_STAGE(preA, preA_n, 0);  // the previous unit ended with a barrier
                RKR_LOAD_STAGE(ahead(1), preA, preA_n);
                __syncthreads();
            }
            for (int64_t m = 0; m < n_st; ++m) {
                const int buf = (int)(m & 1);
                RKR_STORE_STAGE(preA, preA_n, buf ^ 1);
                RKR_LOAD_STAGE(ahead(m + 2), preA, preA_n);
                compute(buf, st0 + m);
                __syncthreads();
            }
        }
        // the unit's queued rechecks go to the list (the stage loop ended
        // with a barrier, so the queue is complete)
        {
            const int32_t nqd = min(sQn, QCAP);
            if (nqd > 0) {
                if (threadIdx.x == 0) sQbase = atomicAdd(P.recheck_n, nqd);
                __syncthreads();
                const int32_t base = sQbase;
                for (int t = threadIdx.x; t < nqd; t += NT)
                    if ((int64_t)base + t < P.recheck_cap) P.recheck[base + t] = sQ[t];
                __syncthreads();  // every thread has read sQbase and the queue
                if (threadIdx.x == 0) sQn = 0;
                __syncthreads();
            }
        }
        // reduce each row's count over the 16 lanes (columns) that share it
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                int32_t v = cnt[b][i];
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
                const int64_t r = row_base + b * 16 + 4 * lq + i;
                if (lr == 0 && r < n_rows && v) atomicAdd(&P.cnt[r], v);
            }
    }
}

#undef RKR_LOAD_STAGE
#undef RKR_STORE_STAGE

// Exact decision of the rechecked (row, item) pairs with the sequential FP32
// rule; also records the largest |s^ - s_fp64| seen (as float bits).
__global__ void k_recheck(const float *__restrict__ users, const float *__restrict__ items, int32_t ld, int32_t K,
                          const int32_t *__restrict__ rows_user, const float *__restrict__ rows_thr,
                          const int32_t *__restrict__ rows_query, const int32_t *__restrict__ recheck_n,
                          const int4 *__restrict__ recheck, int64_t cap, int32_t *__restrict__ cnt,
                          int32_t *__restrict__ q_recheck, uint32_t *__restrict__ max_err_bits) {
    const int64_t n = min((int64_t)*recheck_n, cap);
    float worst = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int4 rc = recheck[i];
        const float *u = users + (int64_t)rows_user[rc.x] * ld;
        const float *p = items + (int64_t)rc.y * ld;
        float s = 0.0f;
        double s64 = 0.0;
        int k = 0;
        // 32 terms per load burst: each lane reads its own rows, so a whole
        // 128-byte stretch is taken before its cache lines can be evicted
        for (; k + 32 <= K; k += 32) {
            float4 a[8], b[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                a[t] = *reinterpret_cast<const float4 *>(u + k + 4 * t);
                b[t] = *reinterpret_cast<const float4 *>(p + k + 4 * t);
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const float av[4] = {a[t].x, a[t].y, a[t].z, a[t].w}, bv[4] = {b[t].x, b[t].y, b[t].z, b[t].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float prod = av[j] * bv[j];
                    s = s + prod;
                    s64 += (double)av[j] * (double)bv[j];
                }
            }
        }
        for (; k < K; k += 4) {
            const float4 a = *reinterpret_cast<const float4 *>(u + k);
            const float4 b = *reinterpret_cast<const float4 *>(p + k);
            const float av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (k + j < K) {
                    const float prod = av[j] * bv[j];
                    s = s + prod;
                    s64 += (double)av[j] * (double)bv[j];
                }
            }
        }
        if (s > rows_thr[rc.x]) atomicAdd(&cnt[rc.x], 1);
        atomicAdd(&q_recheck[rows_query[rc.x]], 1);
        const float err = (float)fabs((double)__int_as_float(rc.z) - s64);
        worst = fmaxf(worst, err);
    }
    if (worst > 0.0f) atomicMax(max_err_bits, __float_as_uint(worst));
}


// Coarse buckets of a query batch.  For every (query b, user u) a screened
// estimate s^ of <u,q_b> with |s^ - s| <= e = C ||u|| ||q|| + abs_floor is
// compared with G <= 7 guard values of u's table row, T[u][col_0] >=
// T[u][col_1] >= ... (col ascending):
//   c_lo = #{g : guard_g > s^ - e},  c_hi = #{g : guard_g > s^ + e}
// and cb[row(b) * ldc + u] = c_lo | c_hi << 4.  abs_floor also carries the
// table's own eps.  The first pass (one guard set) takes the single product
// s^ = sum_k u_hi q_hi, so C there covers the BF16 rounding of both factors
// (2^-7 + 2^-16 of sum|u_k q_k|): a wider interval only lists a few more
// users, and the pass does a third of the matrix work.  The second pass
// (guard sets inside one bin) keeps the three-product form above.  Queries
// are the A operand (staged through LDS, 64 per stage in the first pass, 32
// with guard sets), users the B operand held in VGPRs: lane (lr, lq) of a
// 16x16 tile owns user lr and queries 4 lq .. 4 lq + 3, so its guards stay in
// registers.  qmap (optional) maps the launch's queries to rows of q_hi /
// q_lo / q_norm and cb.
struct IpCoarseParams {
    const float *users;           // n_user x ld
    const float *user_norm;       // n_user
    const uint16_t *user_hi, *user_lo;  // n_user x Kp BF16 split
    const float *table;           // n_user x S
    const float *gvals;           // n_user x kGuardRow guard values (all sets), or null
    int64_t n_user;
    int32_t S;
    GuardCols gc;
    const uint16_t *q_hi, *q_lo;  // rows of Kp
    const float *q_norm;
    const int32_t *qmap;          // nq entries, or null
    int32_t nq;
    int32_t ld, K, Kp;
    float C, abs_floor;
    uint8_t *cb;                  // rows of ldc bytes (ldc % 128 == 0)
    int64_t ldc;
    int64_t spare_row;            // row written for the stage's padding queries
    // per-query guard sets (second pass): query qq uses sets[qset[qq]]
    const GuardCols *sets;
    const int32_t *qset;
    int32_t n_sets;
    int32_t set_id[7];            // guard-row set of each launch set (with gvals)
};

constexpr int kMaxGuardSets = 7;

// t + t + b for a lane predicate b: one v_addc_co_u32 with b's lane mask as
// the carry-in (the ballot of a compare is that compare's mask, so no
// 0/1 value is materialised first).
__device__ __forceinline__ uint32_t twice_plus(uint32_t t, bool b) {
    uint32_t r;
    uint64_t carry_out;
    asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry_out) : "v"(t), "s"(__builtin_amdgcn_ballot_w64(b)));
    return r;
}

// c_lo | c_hi << 4 with c = #{j < 7 : g[j] > x} for x = xl and x = xh, over
// g[0] >= g[1] >= ... >= g[6] (-inf padded past G).  The predicate g[j] > x
// is monotone in j, so three tests of a branchless binary search decide each
// count, c = 4 b2 + 2 b1 + b0: g[3] splits {0..3} from {4..7}, then g[5] or
// g[1], then one of g[6], g[4], g[2], g[0].  The byte is assembled from the
// six test bits, high bit first, by doubling with carry-in (bit 3 stays 0).
__device__ __forceinline__ uint32_t guard_byte(const float (&g)[8], float xl, float xh) {
    const bool l2 = g[3] > xl, h2 = g[3] > xh;
    const float lp1 = __builtin_unpredictable(l2) ? g[5] : g[1];
    const float hp1 = __builtin_unpredictable(h2) ? g[5] : g[1];
    const bool l1 = lp1 > xl, h1 = hp1 > xh;
    const float lpa = __builtin_unpredictable(l1) ? g[6] : g[4], lpb = __builtin_unpredictable(l1) ? g[2] : g[0];
    const float hpa = __builtin_unpredictable(h1) ? g[6] : g[4], hpb = __builtin_unpredictable(h1) ? g[2] : g[0];
    const bool l0 = (__builtin_unpredictable(l2) ? lpa : lpb) > xl;
    const bool h0 = (__builtin_unpredictable(h2) ? hpa : hpb) > xh;
    uint32_t t = h2 ? 1u : 0u;
    t = twice_plus(t, h1);
    t = twice_plus(t, h0);
    t = t + t;
    t = twice_plus(t, l2);
    t = twice_plus(t, l1);
    return twice_plus(t, l0);
}

// c = #{j < 7 : g[j] > x} by the same three-step search as guard_byte, and
// last = g[c - 1], the smallest guard above x (+inf when c = 0), which is
// always the last of the three tested values that passed.
__device__ __forceinline__ uint32_t guard_count_last(const float (&g)[8], float x, float &last) {
    const bool b2 = g[3] > x;
    const float p1 = __builtin_unpredictable(b2) ? g[5] : g[1];
    const bool b1 = p1 > x;
    const float pa = __builtin_unpredictable(b1) ? g[6] : g[4], pb = __builtin_unpredictable(b1) ? g[2] : g[0];
    const float p0 = __builtin_unpredictable(b2) ? pa : pb;
    const bool b0 = p0 > x;
    const float l1 = __builtin_unpredictable(b2) ? g[3] : INFINITY;
    const float l2 = __builtin_unpredictable(b1) ? p1 : l1;
    last = __builtin_unpredictable(b0) ? p0 : l2;
    uint32_t t = b2 ? 1u : 0u;
    t = twice_plus(t, b1);
    return twice_plus(t, b0);
}

template <int KS, bool SETS>
__global__ void __launch_bounds__(NT, SETS ? 2 : coarse_first_pass_wgs(KS)) k_ip_coarse(IpCoarseParams P) {
    // first pass: one product, so only the queries' hi halves are staged and
    // only the users' hi halves are held (4 workgroups per CU fit)
    constexpr bool ONE = !SETS;
    constexpr int NH = ONE ? 1 : 2;
    // queries per LDS stage: 64 halves the stage barriers and query refills per
    // user block; the second pass keeps 32 so its guard sets also fit 3
    // workgroups per CU in LDS
    constexpr int QST = SETS ? 32 : 64;
    constexpr int ROWQ = (KS * 32 + KPAD_LDS) / 8;
    constexpr int KQ = KS * 4;
    __shared__ uint4 sA[NH][QST * ROWQ];
    __shared__ float sQn[QST];
    __shared__ int64_t sQo[QST];  // byte offset of the query's cb row
    __shared__ int32_t sQs[QST];
    __shared__ float4 sG[SETS ? kMaxGuardSets * 128 * 2 : 1];      // [set][user][8 guards]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const int64_t n_rb = (P.n_user + 127) / 128;
    for (int64_t rb = blockIdx.x; rb < n_rb; rb += gridDim.x) {
        const int64_t row_base = rb * 128 + wave * 32;
        Frag uh[2][KS], ul[2][ONE ? 1 : KS];
        float gd[2][8], ar[2], hg[2];
        bool live[2];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int64_t r = row_base + b * 16 + lr;
            live[b] = r < P.n_user;
            // rows past n_user load row 0 and are zeroed after the load (a
            // select between a load and a constant can become a load through
            // a selected pointer, flat and from scratch)
            const int64_t rr = live[b] ? r : 0;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const uint4 z = make_uint4(0, 0, 0, 0);
                const uint4 h = reinterpret_cast<const uint4 *>(P.user_hi + rr * P.Kp)[ks * 4 + lq];
                uh[b][ks].q = live[b] ? h : z;
                if (!ONE) {
                    const uint4 l = reinterpret_cast<const uint4 *>(P.user_lo + rr * P.Kp)[ks * 4 + lq];
                    ul[b][ONE ? 0 : ks].q = live[b] ? l : z;
                }
            }
            if (!SETS) {
                if (P.gvals) {
                    // the user's guard values, one 32-byte row
                    const float4 *g4 = reinterpret_cast<const float4 *>(P.gvals + (live[b] ? r : 0) * kGuardRow);
                    const float4 x0 = g4[0], x1 = g4[1];
                    const float ninf = -INFINITY;
                    gd[b][0] = live[b] ? x0.x : ninf; gd[b][1] = live[b] ? x0.y : ninf;
                    gd[b][2] = live[b] ? x0.z : ninf; gd[b][3] = live[b] ? x0.w : ninf;
                    gd[b][4] = live[b] ? x1.x : ninf; gd[b][5] = live[b] ? x1.y : ninf;
                    gd[b][6] = live[b] ? x1.z : ninf; gd[b][7] = ninf;
                } else {
                    const float *trow = P.table + (live[b] ? r : 0) * (int64_t)P.S;
#pragma unroll
                    for (int j = 0; j < 8; ++j) gd[b][j] = (live[b] && j < P.gc.n) ? trow[P.gc.col[j]] : -INFINITY;
                }
            }
            ar[b] = live[b] ? P.C * P.user_norm[r] : 0.0f;
            if (!SETS) {
                // half the smallest gap between adjacent guards, rounded down
                // (differences with -inf padding are +inf or NaN, which
                // fminf passes over)
                float m = INFINITY;
#pragma unroll
                for (int j = 0; j < 6; ++j) m = fminf(m, gd[b][j] - gd[b][j + 1]);
                hg[b] = __double2float_rd(0.5 * (double)m);
            }
        }
        if (SETS) {
            __syncthreads();  // the previous row block is done with sG
            float *g = reinterpret_cast<float *>(sG);
            for (int t = threadIdx.x; t < P.n_sets * 128 * 8; t += NT) {
                const int set = t / (128 * 8), ul = (t / 8) % 128, j = t % 8;
                const int64_t r = rb * 128 + ul;
                float v = -INFINITY;
                if (r < P.n_user) {
                    if (P.gvals) {
                        // 32 contiguous bytes of the user's guard row
                        v = P.gvals[r * kGuardRow + P.set_id[set] * 8 + j];
                    } else {
                        const GuardCols &gc = P.sets[set];
                        if (j < gc.n) v = P.table[r * P.S + gc.col[j]];
                    }
                }
                g[t] = v;
            }
        }
        for (int32_t q0 = 0; q0 < P.nq; q0 += QST) {
            __syncthreads();  // every wave is done with the previous stage
            {
                // the stage's query rows: KS 16-byte pieces per thread, all
                // loads in flight before the first LDS write (rows past nq
                // load the last query and are zeroed after the load).  The
                // query map's entries are all read before the first row load,
                // so the row loads do not wait for one another.
                static_assert(NH * QST * KQ == KS * NT, "one piece per thread and k-step");
                uint4 v[KS];
                int dst[KS], qr[KS];
                bool ok[KS];
#pragma unroll
                for (int k = 0; k < KS; ++k) {
                    const int c = (int)threadIdx.x + k * NT;
                    const int rem = c % (QST * KQ);
                    const int qq = q0 + rem / KQ;
                    ok[k] = qq < P.nq;
                    qr[k] = min(qq, P.nq - 1);
                }
                if (P.qmap) {
#pragma unroll
                    for (int k = 0; k < KS; ++k) qr[k] = P.qmap[qr[k]];
                }
#pragma unroll
                for (int k = 0; k < KS; ++k) {
                    const int c = (int)threadIdx.x + k * NT;
                    const int half = c / (QST * KQ);
                    const int rem = c % (QST * KQ);
                    v[k] = reinterpret_cast<const uint4 *>(half ? P.q_lo : P.q_hi)[(int64_t)qr[k] * KQ + rem % KQ];
                    dst[k] = half * (QST * ROWQ) + (rem / KQ) * ROWQ + rem % KQ;
                }
#pragma unroll
                for (int k = 0; k < KS; ++k) (&sA[0][0])[dst[k]] = ok[k] ? v[k] : make_uint4(0, 0, 0, 0);
            }
            if (threadIdx.x < QST) {
                const int qq = q0 + (int)threadIdx.x;
                const int32_t qr = qq < P.nq ? (P.qmap ? P.qmap[qq] : qq) : 0;
                sQo[threadIdx.x] = (qq < P.nq ? (int64_t)qr : P.spare_row) * P.ldc;
                sQn[threadIdx.x] = qq < P.nq ? P.q_norm[qr] : 0.f;
                if (SETS) sQs[threadIdx.x] = qq < P.nq ? P.qset[qq] : 0;
            }
            __syncthreads();
#pragma unroll
            for (int g = 0; g < QST / 16; ++g) {
                const int it = g * 16 + lr;
                f32x4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    Frag qh, ql;
                    qh.q = sA[0][it * ROWQ + ks * 4 + lq];
                    if (!ONE) ql.q = sA[NH - 1][it * ROWQ + ks * 4 + lq];
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(qh.v, uh[b][ks].v, acc[b], 0, 0, 0);
                        if (!ONE) {
                            acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(qh.v, ul[b][ONE ? 0 : ks].v, acc[b], 0,
                                                                              0, 0);
                            acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ql.v, uh[b][ks].v, acc[b], 0, 0, 0);
                        }
                    }
                }
                // D[4 lq + i][lr]: query q0 + g*16 + 4 lq + i, user row_base + b*16 + lr
                if (SETS) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        // rows are padded to whole row blocks and padding queries
                        // write the spare row, so every lane stores
                        const int ql_ = g * 16 + 4 * lq + i;
                        const float qn = sQn[ql_];
                        uint8_t *out = P.cb + sQo[ql_] + row_base + lr;
#pragma unroll
                        for (int b = 0; b < 2; ++b) {
                            const float4 *gp = sG + (sQs[ql_] * 128 + wave * 32 + b * 16 + lr) * 2;
                            const float4 x0 = gp[0], x1 = gp[1];
                            gd[b][0] = x0.x; gd[b][1] = x0.y; gd[b][2] = x0.z; gd[b][3] = x0.w;
                            gd[b][4] = x1.x; gd[b][5] = x1.y; gd[b][6] = x1.z; gd[b][7] = x1.w;
                            const float e = __fmaf_rn(ar[b], qn, P.abs_floor);
                            out[b * 16] = (uint8_t)guard_byte(gd[b], acc[b][i] - e, acc[b][i] + e);
                        }
                    }
                } else {
                    // c_hi = #{g > xh} <= c_lo.  When at most one guard can lie
                    // in (xl, xh] -- the interval is narrower than half the
                    // smallest gap between adjacent guards -- c_hi is c_lo less
                    // [g[c_lo - 1] <= xh].  Two guards inside would be closer
                    // than xh - xl (< 2 fl(xh - xl) in FP32), so a wavefront
                    // with any pair failing the test counts xh for all its
                    // pairs.
                    uint32_t byte[4][2];
                    bool redo = false;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float qn = sQn[g * 16 + 4 * lq + i];
#pragma unroll
                        for (int b = 0; b < 2; ++b) {
                            const float e = __fmaf_rn(ar[b], qn, P.abs_floor);
                            const float xl = acc[b][i] - e, xh = acc[b][i] + e;
                            float last;
                            const uint32_t c = guard_count_last(gd[b], xl, last);
                            redo |= !(xh - xl < hg[b]);
                            byte[i][b] = c | (c - (last <= xh ? 1u : 0u)) << 4;
                        }
                    }
                    if (__builtin_amdgcn_ballot_w64(redo) != 0) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const float qn = sQn[g * 16 + 4 * lq + i];
#pragma unroll
                            for (int b = 0; b < 2; ++b) {
                                const float e = __fmaf_rn(ar[b], qn, P.abs_floor);
                                byte[i][b] = guard_byte(gd[b], acc[b][i] - e, acc[b][i] + e);
                            }
                        }
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        // rows are padded to whole row blocks and padding
                        // queries write the spare row, so every lane stores
                        uint8_t *out = P.cb + sQo[g * 16 + 4 * lq + i] + row_base + lr;
#pragma unroll
                        for (int b = 0; b < 2; ++b) out[b * 16] = (uint8_t)byte[i][b];
                    }
                }
            }
        }
    }
}

// gv[u * kGuardRow + 8 s + j] = T[u][set s column j], -inf past the set's n:
// the coarse passes' guard values of every user in one 256-byte row (the
// columns sit in different cache lines of a table row), gathered once per
// table.
__global__ void __launch_bounds__(256) k_gather_guards(const float *__restrict__ T, int64_t S, GuardSets gs,
                                                       int64_t u0, int64_t n_rows, float *__restrict__ gv) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n_rows * kGuardRow) return;
    const int64_t u = u0 + i / kGuardRow;
    const int sj = (int)(i % kGuardRow), j = sj & 7;
    const GuardCols &g = gs.set[sj >> 3];
    gv[u * kGuardRow + sj] = j < g.n ? T[u * S + g.col[j]] : -INFINITY;
}

}  // namespace

hipError_t launch_gather_guards(hipStream_t s, const float *T, int64_t S, const GuardSets &gs, int64_t u0,
                                int64_t n_rows, float *gv) {
    if (n_rows <= 0) return hipSuccess;
    for (const GuardCols &g : gs.set) {
        if (g.n < 0 || g.n > 7) return hipErrorInvalidValue;
        for (int j = 0; j < g.n; ++j)
            if (g.col[j] < 0 || g.col[j] >= S) return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_gather_guards, dim3((unsigned)((n_rows * kGuardRow + 255) / 256)), dim3(256), 0, s, T, S, gs,
                       u0, n_rows, gv);
    return hipGetLastError();
}

hipError_t launch_prep_items(hipStream_t s, const float *items, int64_t n_item, int32_t ld, int32_t K, int32_t Kp,
                             uint16_t *hi, uint16_t *lo, float *norm, float *cs, const float *rho, double g1,
                             double g_rule) {
    if (n_item <= 0) return hipSuccess;
    if (cs && !rho) return hipErrorInvalidValue;
    dim3 block(32, 8);
    hipLaunchKernelGGL(k_prep_items, dim3((unsigned)((n_item + 7) / 8)), block, 0, s, items, n_item, ld, K, Kp, hi, lo,
                       norm, cs, rho, g1, g_rule);
    return hipGetLastError();
}

hipError_t launch_user_norms(hipStream_t s, const float *users, int64_t n_user, int32_t ld, int32_t K, float *norm,
                             float *rho) {
    if (n_user <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_user_norms, dim3((unsigned)((n_user + 255) / 256)), dim3(256), 0, s, users, n_user, ld, K,
                       norm, rho);
    return hipGetLastError();
}

int screen_max_kp() { return 256; }

hipError_t launch_refine_screen(hipStream_t s, const ScreenLaunch &L) {
    if (L.max_rows <= 0 || L.n_item <= 0) return hipSuccess;
    ScreenParams P;
    if (!L.user_hi || !L.user_lo) return hipErrorInvalidValue;
    P.users = L.users;
    P.user_norm = L.user_norm;
    P.user_hi = L.user_hi;
    P.user_lo = L.user_lo;
    P.item_hi = L.item_hi;
    P.item_lo = L.item_lo;
    P.item_norm = L.item_norm;
    P.n_item = L.n_item;
    P.ld = L.ld;
    P.K = L.K;
    P.Kp = L.Kp;
    P.rows_user = L.rows_user;
    P.rows_thr = L.rows_thr;
    P.n_rows_dev = L.n_rows_dev;
    P.max_rows = L.max_rows;
    P.C = L.C;
    P.abs_floor = L.abs_floor;
    P.cnt = L.cnt;
    P.recheck_n = L.recheck_n;
    P.recheck = reinterpret_cast<int4 *>(L.recheck);
    P.recheck_cap = L.recheck_cap;
    P.target_units = std::max<int64_t>(1, L.target_units);
    P.chunk_items = std::max(1, L.chunk_items);
    // persistent grid: enough workgroups to fill the device, never more than
    // the largest possible number of units
    const int rb = L.rows_per_wave == 64 ? 4 : 2;
    const int64_t n_rb_max = (L.max_rows + 64 * rb - 1) / (64 * rb);
    const int64_t n_stage = (L.n_item + ITEMS_STAGE - 1) / ITEMS_STAGE;
    const int64_t max_units = std::max<int64_t>(n_rb_max, std::min<int64_t>(P.target_units, n_rb_max * n_stage));
    const int64_t cap = rb == 4 ? L.grid_cap / 2 : L.grid_cap;  // one workgroup per CU at 64 rows per wave
    const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>(max_units, cap)));
#define RKR_SCREEN_CASE(ks)                                                                              \
    case ks:                                                                                           \
        if (rb == 4) hipLaunchKernelGGL((k_refine_screen<ks, 4>), grid, dim3(NT), 0, s, P);             \
        else hipLaunchKernelGGL((k_refine_screen<ks, 2>), grid, dim3(NT), 0, s, P);                     \
        break;
    switch (L.Kp / 32) {
        RKR_SCREEN_CASE(1)
        RKR_SCREEN_CASE(2)
        RKR_SCREEN_CASE(3)
        RKR_SCREEN_CASE(4)
        RKR_SCREEN_CASE(5)
        RKR_SCREEN_CASE(6)
        RKR_SCREEN_CASE(7)
        RKR_SCREEN_CASE(8)
        default: return hipErrorInvalidValue;
    }
#undef RKR_SCREEN_CASE
    return hipGetLastError();
}

hipError_t launch_recheck(hipStream_t s, const float *users, const float *items, int32_t ld, int32_t K,
                          const int32_t *rows_user, const float *rows_thr, const int32_t *rows_query,
                          const int32_t *recheck_n, const void *recheck, int64_t cap, int32_t *cnt,
                          int32_t *q_recheck, uint32_t *max_err_bits) {
    if (cap <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_recheck, dim3(2048), dim3(256), 0, s, users, items, ld, K, rows_user, rows_thr, rows_query,
                       recheck_n, reinterpret_cast<const int4 *>(recheck), cap, cnt, q_recheck, max_err_bits);
    return hipGetLastError();
}

hipError_t launch_ip_coarse(hipStream_t s, const CoarseLaunch &L) {
    if (L.n_user <= 0 || L.nq <= 0) return hipSuccess;
    const bool sets = L.sets != nullptr;
    if (sets ? (L.n_sets < 1 || L.n_sets > kMaxGuardSets || !L.qset) : (L.gc.n < 1 || L.gc.n > 7))
        return hipErrorInvalidValue;
    if (L.ldc % 128 || L.ldc < L.n_user) return hipErrorInvalidValue;
    if (!L.user_hi || !L.user_lo) return hipErrorInvalidValue;
    IpCoarseParams P{L.users, L.user_norm, L.user_hi, L.user_lo, L.table, L.gvals, L.n_user, L.S, L.gc, L.q_hi,
                     L.q_lo, L.q_norm, L.qmap, L.nq, L.ld, L.K, L.Kp, L.C, L.abs_floor, L.cb, L.ldc, L.spare_row,
                     L.sets, L.qset, L.n_sets, {}};
    if (sets && L.gvals)
        for (int t = 0; t < L.n_sets; ++t) {
            if (L.set_id[t] < 1 || L.set_id[t] >= kGuardSets) return hipErrorInvalidValue;
            P.set_id[t] = L.set_id[t];
        }
    const int64_t n_rb = (L.n_user + 127) / 128;
    const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>(n_rb, L.grid_cap)));
#define RKR_COARSE_CASE(ks)                                                                                   \
    case ks:                                                                                                  \
        if (sets) hipLaunchKernelGGL((k_ip_coarse<ks, true>), grid, dim3(NT), 0, s, P);                      \
        else hipLaunchKernelGGL((k_ip_coarse<ks, false>), grid, dim3(NT), 0, s, P);                          \
        break;
    switch (L.Kp / 32) {
        RKR_COARSE_CASE(1)
        RKR_COARSE_CASE(2)
        RKR_COARSE_CASE(3)
        RKR_COARSE_CASE(4)
        RKR_COARSE_CASE(5)
        RKR_COARSE_CASE(6)
        RKR_COARSE_CASE(7)
        RKR_COARSE_CASE(8)
        default: return hipErrorInvalidValue;
    }
#undef RKR_COARSE_CASE
    return hipGetLastError();
}

}  // namespace rkr
