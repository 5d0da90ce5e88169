// Kernels of the join path (gfx950, wave64). Included by tdp_device.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tdp_geom.hpp"

namespace tdp {

constexpr int kWave = 64;
constexpr int kBlock = 256;                 // 4 waves
constexpr int kWavesPerBlock = kBlock / kWave;

enum : uint8_t { ST_DEAD = 0, ST_ALIVE = 1, ST_REPORTED = 2 };
enum { MODE_REF32 = 0, MODE_FP64 = 1, MODE_MIXED = 2 };

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ size_t wave_global() {
    return (size_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
}

template <typename R>
__device__ __forceinline__ R rmin2(R a, R b) { return b < a ? b : a; }

template <typename R>
__device__ __forceinline__ R wave_min(R v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = rmin2<R>(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// order-preserving unsigned keys for atomic min over floating values
__device__ __forceinline__ uint32_t fkey(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funkey(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}
__device__ __forceinline__ unsigned long long dkey(double f) {
    unsigned long long u = (unsigned long long)__double_as_longlong(f);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dunkey(unsigned long long k) {
    unsigned long long u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}
template <typename R> struct Key;
template <> struct Key<float> {
    using type = uint32_t;
    static __device__ __forceinline__ uint32_t enc(float v) { return fkey(v); }
    static __device__ __forceinline__ float dec(uint32_t k) { return funkey(k); }
};
template <> struct Key<double> {
    using type = unsigned long long;
    static __device__ __forceinline__ unsigned long long enc(double v) { return dkey(v); }
    static __device__ __forceinline__ double dec(unsigned long long k) { return dunkey(k); }
};

// ----------------------------------------------------------------- filter

// object-level [min boxmin, min coredist] over all voxel pairs of each
// candidate pair (voxel_pair_dist_kernel_for_all_aggregate_distance,
// computing_gpu.cu:1984-2043); one wave per object pair
template <typename R>
__global__ void __launch_bounds__(kBlock) k_aggregate(
    size_t P, const uint32_t *__restrict__ pair_target, const uint32_t *__restrict__ pair_obj2,
    const uint32_t *__restrict__ vo1, const float *__restrict__ vb1,
    const uint32_t *__restrict__ vo2, const float *__restrict__ vb2, R *__restrict__ rmin,
    R *__restrict__ rmax) {
    const size_t p = wave_global();
    if (p >= P) return;
    const int lane = lane_id();
    const uint32_t a0 = vo1[pair_target[p]], n1 = vo1[pair_target[p] + 1] - a0;
    const uint32_t b0 = vo2[pair_obj2[p]], n2 = vo2[pair_obj2[p] + 1] - b0;
    const uint32_t total = n1 * n2;
    R mn = (R)__builtin_inff(), mx = (R)__builtin_inff();
    for (uint32_t t = lane; t < total; t += kWave) {
        const uint32_t i = t / n2, j = t - i * n2;
        R a, b;
        voxel_bound<R>(vb1 + 9 * (size_t)(a0 + i), vb2 + 9 * (size_t)(b0 + j), a, b);
        mn = rmin2(mn, a);
        mx = rmin2(mx, b);
    }
    mn = wave_min(mn);
    mx = wave_min(mx);
    if (lane == 0) {
        rmin[p] = mn;
        rmax[p] = mx;
    }
}

// kNN status of candidate i among the live candidates of its target
// (evaluate_knn_kernel, computing_gpu.cu:105-151): 1 CONFIRMED, 2 REMOVE, 0 KEEP
template <typename R>
__device__ __forceinline__ int knn_status_one(const R *rmin, const R *rmax, const uint8_t *state,
                                              uint32_t b, uint32_t e, uint32_t i, int cand_left,
                                              bool use_state) {
    int sure = 0, maybe = 0;
    const R mi = rmin[i], xi = rmax[i];
    for (uint32_t j = b; j < e; j++) {
        if (j == i) continue;
        if (use_state && state[j] != ST_ALIVE) continue;
        if (mi >= rmax[j]) sure++;
        if (!(xi <= rmin[j])) maybe++;
    }
    if (maybe < cand_left) return 1;
    if (sure >= cand_left) return 2;
    return 0;
}

// filter-time kNN decisions per target (mbb_knn_op_gpu_one_kernel,
// KNNJoin.cpp:643-754, on the status of computing_gpu.cu:4572-4608):
//  - exactly K candidates: all reported, no distance check;
//  - K of them CONFIRMED (counted in list order): every CONFIRMED reported;
//  - otherwise CONFIRMED reported, KEEP refined, REMOVE dropped.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_filter_knn(size_t N, const uint32_t *__restrict__ prefix,
                                                       const R *__restrict__ rmin,
                                                       const R *__restrict__ rmax, int K,
                                                       uint8_t *__restrict__ state,
                                                       uint8_t *__restrict__ round,
                                                       int32_t *__restrict__ confirmed,
                                                       uint8_t *__restrict__ scratch) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const uint32_t b = prefix[t], e = prefix[t + 1];
    const int m = (int)(e - b);
    int conf = 0;
    for (uint32_t i = b; i < e; i++) {
        int s = knn_status_one<R>(rmin, rmax, state, b, e, i, K, false);
        scratch[i] = (uint8_t)s;
    }
    for (uint32_t i = b; i < e; i++) {
        if (scratch[i] == 1) conf++;
        if (conf == K) break;
    }
    int c = 0;
    for (uint32_t i = b; i < e; i++) {
        uint8_t st = ST_DEAD, rd = 255;
        if (m == K) {
            st = ST_REPORTED; rd = 0;
        } else if (conf == K) {
            if (scratch[i] == 1) { st = ST_REPORTED; rd = 0; }
        } else {
            if (scratch[i] == 1) { st = ST_REPORTED; rd = 0; c++; }
            else if (scratch[i] == 0) st = ST_ALIVE;
        }
        state[i] = st;
        round[i] = rd;
    }
    confirmed[t] = c;
}

// filter-time within decisions per object pair (WithinJoin.cpp:178-187):
// min > tau dropped, max <= tau reported, else refined
template <typename R>
__global__ void __launch_bounds__(kBlock) k_filter_within(size_t P, const R *__restrict__ rmin,
                                                          const R *__restrict__ rmax, R tau,
                                                          uint8_t *__restrict__ state,
                                                          uint8_t *__restrict__ round) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    uint8_t st = ST_DEAD, rd = 255;
    if (rmin[p] > tau) st = ST_DEAD;
    else if (rmax[p] <= tau) { st = ST_REPORTED; rd = 0; }
    else st = ST_ALIVE;
    state[p] = st;
    round[p] = rd;
}

// voxel pairs kept for refinement: boxmin <= object max (kNN), and also
// boxmin <= tau (within) (computing_gpu.cu:2604-2812)
template <typename R>
__device__ __forceinline__ bool survives(R mn, R objmax, bool within, R tau) {
    return within ? (mn <= objmax && mn <= tau) : (mn <= objmax);
}

// per live pair: number of surviving voxel pairs; for within, a surviving
// pair with max <= tau settles the object pair at once (WithinJoin.cpp:190-225)
template <typename R>
__global__ void __launch_bounds__(kBlock) k_survivor_count(
    size_t P, const uint32_t *__restrict__ pair_target, const uint32_t *__restrict__ pair_obj2,
    const uint32_t *__restrict__ vo1, const float *__restrict__ vb1,
    const uint32_t *__restrict__ vo2, const float *__restrict__ vb2, const R *__restrict__ rmax,
    bool within, R tau, uint8_t *__restrict__ state, uint8_t *__restrict__ round,
    uint32_t *__restrict__ count) {
    const size_t p = wave_global();
    if (p >= P) return;
    const int lane = lane_id();
    if (state[p] != ST_ALIVE) {
        if (lane == 0) count[p] = 0;
        return;
    }
    const uint32_t a0 = vo1[pair_target[p]], n1 = vo1[pair_target[p] + 1] - a0;
    const uint32_t b0 = vo2[pair_obj2[p]], n2 = vo2[pair_obj2[p] + 1] - b0;
    const uint32_t total = n1 * n2;
    const R objmax = rmax[p];
    uint32_t cnt = 0;
    bool settled = false;
    for (uint32_t t = lane; t < total; t += kWave) {
        const uint32_t i = t / n2, j = t - i * n2;
        R mn, mx;
        voxel_bound<R>(vb1 + 9 * (size_t)(a0 + i), vb2 + 9 * (size_t)(b0 + j), mn, mx);
        if (survives<R>(mn, objmax, within, tau)) {
            cnt++;
            if (within && mx <= tau) settled = true;
        }
    }
    cnt = (uint32_t)wave_sum_u64(cnt);
    const bool any_settled = __any(settled);
    if (lane == 0) {
        if (within && any_settled) {
            state[p] = ST_REPORTED;
            round[p] = 0;
            cnt = 0;
        } else if (cnt == 0) {
            state[p] = ST_DEAD;
        }
        count[p] = cnt;
    }
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_survivor_write(
    size_t P, const uint32_t *__restrict__ pair_target, const uint32_t *__restrict__ pair_obj2,
    const uint32_t *__restrict__ vo1, const float *__restrict__ vb1,
    const uint32_t *__restrict__ vo2, const float *__restrict__ vb2, const R *__restrict__ rmax,
    bool within, R tau, const uint8_t *__restrict__ state, const uint64_t *__restrict__ offs,
    uint32_t *__restrict__ vp_v1, uint32_t *__restrict__ vp_v2, uint32_t *__restrict__ vp_pair,
    R *__restrict__ vp_min, R *__restrict__ vp_max) {
    const size_t p = wave_global();
    if (p >= P) return;
    if (state[p] != ST_ALIVE) return;
    const int lane = lane_id();
    const uint32_t a0 = vo1[pair_target[p]], n1 = vo1[pair_target[p] + 1] - a0;
    const uint32_t b0 = vo2[pair_obj2[p]], n2 = vo2[pair_obj2[p] + 1] - b0;
    const uint32_t total = n1 * n2;
    const R objmax = rmax[p];
    uint64_t pos = offs[p];
    for (uint32_t base = 0; base < total; base += kWave) {
        const uint32_t t = base + lane;
        bool keep = false;
        R mn = 0, mx = 0;
        uint32_t i = 0, j = 0;
        if (t < total) {
            i = t / n2;
            j = t - i * n2;
            voxel_bound<R>(vb1 + 9 * (size_t)(a0 + i), vb2 + 9 * (size_t)(b0 + j), mn, mx);
            keep = survives<R>(mn, objmax, within, tau);
        }
        const uint64_t mask = __ballot(keep);
        if (keep) {
            const uint64_t below = mask & ((1ull << lane) - 1ull);
            const uint64_t w = pos + (uint64_t)__popcll(below);
            vp_v1[w] = a0 + i;
            vp_v2[w] = b0 + j;
            vp_pair[w] = (uint32_t)p;
            vp_min[w] = mn;
            vp_max[w] = mx;
        }
        pos += (uint64_t)__popcll(mask);
    }
}

// --------------------------------------------------------------- refinement

template <typename R, typename K>
__global__ void __launch_bounds__(kBlock) k_reset_acc(size_t P, const uint8_t *__restrict__ state,
                                                      K *__restrict__ acc_min,
                                                      K *__restrict__ acc_max) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    if (state[p] != ST_ALIVE) return;
    acc_min[p] = Key<R>::enc((R)__builtin_inff());
    acc_max[p] = Key<R>::enc((R)__builtin_inff());
}

struct TriRec {
    float v0x, v0y, v0z, ph, v1x, v1y, v1z, h, v2x, v2y, v2z, pad;
};

template <typename R>
__device__ __forceinline__ void load_tri(const float4 *__restrict__ rec, uint64_t idx, R t[9],
                                         float &ph, float &h) {
    const float4 a = rec[3 * idx], b = rec[3 * idx + 1], c = rec[3 * idx + 2];
    t[0] = (R)a.x; t[1] = (R)a.y; t[2] = (R)a.z; ph = a.w;
    t[3] = (R)b.x; t[4] = (R)b.y; t[5] = (R)b.z; h = b.w;
    t[6] = (R)c.x; t[7] = (R)c.y; t[8] = (R)c.z;
}

// exact per-pair values (d, max(d-ph1-ph2,0), d+h1+h2) in precision R
template <typename R>
__device__ __forceinline__ void pair_values(const float4 *rec1, uint64_t i1, const float4 *rec2,
                                            uint64_t i2, R &d, R &lo, R &hi) {
    R S[9], T[9];
    float ph1, h1, ph2, h2;
    load_tri<R>(rec1, i1, S, ph1, h1);
    load_tri<R>(rec2, i2, T, ph2, h2);
    d = tri_dist<R>(S, T);
    R l = d - (R)ph1 - (R)ph2;
    lo = l < (R)0 ? (R)0 : l;
    hi = d + (R)h1 + (R)h2;
}

// FP32 estimate for the mixed form: TriDist on coordinates translated to
// the first vertex of S (each translated coordinate carries one rounding),
// with an absolute bound e on |d32 - d64|. A result without a certificate
// (see tri_dist) or from a near-degenerate triangle is "fragile": it gets the
// interval [0, inf) and is always re-evaluated in FP64.
struct F32Est {
    float d, e;
    bool fragile;
};

constexpr float kU = 5.9604645e-08f;    // 2^-24
constexpr float kBoundK = 1024.f;       // e = kBoundK * u * (L + d)

__device__ __forceinline__ F32Est f32_estimate(const float4 *rec1, uint64_t i1,
                                               const float4 *rec2, uint64_t i2, float &ph1,
                                               float &h1, float &ph2, float &h2) {
    float S[9], T[9];
    load_tri<float>(rec1, i1, S, ph1, h1);
    load_tri<float>(rec2, i2, T, ph2, h2);
    const float ox = S[0], oy = S[1], oz = S[2];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        S[3 * k] -= ox; S[3 * k + 1] -= oy; S[3 * k + 2] -= oz;
        T[3 * k] -= ox; T[3 * k + 1] -= oy; T[3 * k + 2] -= oz;
    }
    float L = 0.f;
#pragma unroll
    for (int k = 0; k < 9; k++) L = fmaxf(L, fmaxf(fabsf(S[k]), fabsf(T[k])));
    F32Est r;
    bool cert = false;
    r.d = tri_dist<float>(S, T, &cert);
    float e1[3], e2[3], n[3];
    TDP_SUB(e1, (S + 3), S); TDP_SUB(e2, (S + 6), (S + 3)); TDP_CROSS(n, e1, e2);
    const float sn = TDP_DOT(n, n);
    TDP_SUB(e1, (T + 3), T); TDP_SUB(e2, (T + 6), (T + 3)); TDP_CROSS(n, e1, e2);
    const float tn = TDP_DOT(n, n);
    const float L4 = (L * L) * (L * L);
    r.fragile = !cert || !(sn > 1e-12f * L4) || !(tn > 1e-12f * L4);
    r.e = kBoundK * kU * (L + r.d);
    return r;
}

// intervals for (d, lo, hi) of one triangle pair in FP64 from its FP32 estimate
__device__ __forceinline__ void f32_intervals(const F32Est &es, float ph1, float h1, float ph2,
                                              float h2, float lw[3], float uw[3]) {
    if (es.fragile) {
        lw[0] = lw[1] = lw[2] = 0.f;
        uw[0] = uw[1] = uw[2] = INFINITY;
        return;
    }
    const float s = es.e + 8.f * kU * (es.d + ph1 + ph2 + h1 + h2);
    lw[0] = fmaxf(es.d - s, 0.f);
    uw[0] = es.d + s;
    lw[1] = fmaxf(es.d - ph1 - ph2 - s, 0.f);
    uw[1] = fmaxf(es.d - ph1 - ph2 + s, 0.f);
    lw[2] = es.d + h1 + h2 - s;
    uw[2] = es.d + h1 + h2 + s;
}

// Exact FP64 minima of (d, lo, hi) over the n1 x n2 triangle pairs of a
// voxel pair, evaluating FP64 only where an FP32 interval reaches the
// smallest FP32 upper bound. Every pair whose FP64 value could be the
// minimum is re-evaluated, so the result equals the all-FP64 result bit for
// bit. Per lane, the best and second-best lower bound are tracked for each
// quantity; a lane whose second-best also qualifies re-scans its pairs.
// Called by all 64 lanes of a wave; results are per lane (wave-min after).
__device__ __forceinline__ void mixed_minima(const float4 *rec1, uint64_t o1,
                                             const float4 *rec2, uint64_t o2, uint64_t n2,
                                             uint64_t total, int lane, double &d, double &lo,
                                             double &hi, uint64_t &rechecks) {
    float up[3] = {INFINITY, INFINITY, INFINITY};
    float b1[3] = {INFINITY, INFINITY, INFINITY}, b2[3] = {INFINITY, INFINITY, INFINITY};
    uint64_t i1[3] = {0, 0, 0};
    for (uint64_t q = lane; q < total; q += kWave) {
        const uint64_t i = q / n2, j = q - i * n2;
        float ph1, h1, ph2, h2, lw[3], uw[3];
        const F32Est es = f32_estimate(rec1, o1 + i, rec2, o2 + j, ph1, h1, ph2, h2);
        f32_intervals(es, ph1, h1, ph2, h2, lw, uw);
#pragma unroll
        for (int f = 0; f < 3; f++) {
            up[f] = fminf(up[f], uw[f]);
            if (lw[f] < b1[f]) { b2[f] = b1[f]; b1[f] = lw[f]; i1[f] = q; }
            else if (lw[f] < b2[f]) b2[f] = lw[f];
        }
    }
    float U[3];
#pragma unroll
    for (int f = 0; f < 3; f++) U[f] = wave_min(up[f]);
    const bool rescan = (b2[0] <= U[0]) || (b2[1] <= U[1]) || (b2[2] <= U[2]);
    if (!rescan) {
#pragma unroll
        for (int f = 0; f < 3; f++) {
            if (!(b1[f] <= U[f])) continue;
            bool seen = false;
            for (int g = 0; g < f; g++) seen |= (b1[g] <= U[g]) && (i1[g] == i1[f]);
            if (seen) continue;
            const uint64_t i = i1[f] / n2, j = i1[f] - i * n2;
            double a, b, c;
            pair_values<double>(rec1, o1 + i, rec2, o2 + j, a, b, c);
            d = rmin2(d, a); lo = rmin2(lo, b); hi = rmin2(hi, c);
            rechecks++;
        }
    } else {
        for (uint64_t q = lane; q < total; q += kWave) {
            const uint64_t i = q / n2, j = q - i * n2;
            float ph1, h1, ph2, h2, lw[3], uw[3];
            const F32Est es = f32_estimate(rec1, o1 + i, rec2, o2 + j, ph1, h1, ph2, h2);
            f32_intervals(es, ph1, h1, ph2, h2, lw, uw);
            if (lw[0] <= U[0] || lw[1] <= U[1] || lw[2] <= U[2]) {
                double a, b, c;
                pair_values<double>(rec1, o1 + i, rec2, o2 + j, a, b, c);
                d = rmin2(d, a); lo = rmin2(lo, b); hi = rmin2(hi, c);
                rechecks++;
            }
        }
    }
}

// Per voxel pair: (d, lo, hi) = min over the n1 x n2 triangle pairs, folded
// into the voxel pair's range and min-accumulated per object pair.
// Replaces TriDist_cuda_op_shm_flat_fused (computing_gpu.cu:1069-1152) plus
// the host fold (KNNJoin.cpp:1663-1700, WithinJoin.cpp:596-632):
//   last LOD        -> [d, d]
//   hausdorff lvl 2 -> [min(min, lo), min(max, hi)]
//   hausdorff lvl 1 -> [min(min, d - H1 - H2), min(max, d + P1 + P2)] (object-level)
//   hausdorff lvl 0 -> [min, min(max, d)]
// Empty voxels at this LOD keep their range and do not enter the object max.
template <typename R, int MODE>
__global__ void __launch_bounds__(kBlock) k_refine(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint32_t *__restrict__ vp_pair, R *__restrict__ vp_min, R *__restrict__ vp_max,
    const uint64_t *__restrict__ off1, const uint32_t *__restrict__ cnt1,
    const float4 *__restrict__ rec1, const uint64_t *__restrict__ off2,
    const uint32_t *__restrict__ cnt2, const float4 *__restrict__ rec2,
    const uint32_t *__restrict__ vobj1, const uint32_t *__restrict__ vobj2,
    const float *__restrict__ ohd1, const float *__restrict__ ohd2, int slot, int is_last,
    int level, typename Key<R>::type *__restrict__ acc_min,
    typename Key<R>::type *__restrict__ acc_max, unsigned long long *__restrict__ stats) {
    const size_t w = wave_global();
    if (w >= n_vp) return;
    const int lane = lane_id();
    const uint32_t v1 = vp_v1[w], v2 = vp_v2[w], p = vp_pair[w];
    uint32_t n1 = 0, n2 = 0;
    uint64_t o1 = 0, o2 = 0;
    if (slot >= 0) {
        n1 = cnt1[5 * (size_t)v1 + slot];
        n2 = cnt2[5 * (size_t)v2 + slot];
        o1 = off1[5 * (size_t)v1 + slot];
        o2 = off2[5 * (size_t)v2 + slot];
    }
    R omin = vp_min[w], omax = vp_max[w];
    using KT = typename Key<R>::type;
    if (n1 == 0 || n2 == 0) {
        if (lane == 0) atomicMin(&acc_min[p], Key<R>::enc(omin));
        return;
    }
    const uint32_t total = n1 * n2;
    R d = (R)__builtin_inff(), lo = (R)__builtin_inff(), hi = (R)__builtin_inff();
    uint64_t rechecks = 0;
    if (MODE != MODE_MIXED) {
        for (uint32_t q = lane; q < total; q += kWave) {
            const uint32_t i = q / n2, j = q - i * n2;
            R a, b, c;
            pair_values<R>(rec1, o1 + i, rec2, o2 + j, a, b, c);
            d = rmin2(d, a);
            lo = rmin2(lo, b);
            hi = rmin2(hi, c);
        }
    } else {
        double dd = INFINITY, dl = INFINITY, dh = INFINITY;
        mixed_minima(rec1, o1, rec2, o2, n2, total, lane, dd, dl, dh, rechecks);
        d = (R)dd; lo = (R)dl; hi = (R)dh;
    }
    d = wave_min(d);
    lo = wave_min(lo);
    hi = wave_min(hi);
    if (MODE == MODE_MIXED) rechecks = wave_sum_u64(rechecks);
    if (lane == 0) {
        R nmin = omin, nmax = omax;
        if (is_last) {
            nmin = d;
            nmax = d;
        } else if (level == 2) {
            if (lo < nmin) nmin = lo;
            if (hi < nmax) nmax = hi;
        } else if (level == 1) {
            const R a = d - (R)ohd1[10 * (size_t)vobj1[v1] + 2 * slot] -
                        (R)ohd2[10 * (size_t)vobj2[v2] + 2 * slot];
            const R b = d + (R)ohd1[10 * (size_t)vobj1[v1] + 2 * slot + 1] +
                        (R)ohd2[10 * (size_t)vobj2[v2] + 2 * slot + 1];
            if (a < nmin) nmin = a;
            if (b < nmax) nmax = b;
        } else {
            if (d < nmax) nmax = d;
        }
        vp_min[w] = nmin;
        vp_max[w] = nmax;
        atomicMin(&acc_min[p], Key<R>::enc(nmin));
        atomicMin(&acc_max[p], Key<R>::enc(nmax));
        atomicAdd(&stats[0], (unsigned long long)total);
        atomicAdd(&stats[1], 1ull);
        if (MODE == MODE_MIXED) atomicAdd(&stats[2], (unsigned long long)rechecks);
    }
}

// object range after the round: min over all its voxel pairs, max = min over
// the non-empty ones (SpatialJoin::update_voxel_pair_list, SpatialJoin.cpp:48-69)
template <typename R>
__global__ void __launch_bounds__(kBlock) k_pair_finish(
    size_t P, const uint8_t *__restrict__ state, const typename Key<R>::type *__restrict__ acc_min,
    const typename Key<R>::type *__restrict__ acc_max, R *__restrict__ rmin, R *__restrict__ rmax) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    if (state[p] != ST_ALIVE) return;
    rmin[p] = Key<R>::dec(acc_min[p]);
    rmax[p] = Key<R>::dec(acc_max[p]);
}

// per-LOD kNN walk (evaluate_candidate_lists_gpu, KNNJoin.cpp:133-202): status
// over the live candidates with K - confirmed left; in list order CONFIRMED is
// reported (stopping at K, the rest dropped), KEEP stays, REMOVE is dropped
template <typename R>
__global__ void __launch_bounds__(kBlock) k_decide_knn(size_t N, const uint32_t *__restrict__ prefix,
                                                       const R *__restrict__ rmin,
                                                       const R *__restrict__ rmax, int K,
                                                       uint8_t *__restrict__ state,
                                                       uint8_t *__restrict__ round,
                                                       int32_t *__restrict__ confirmed,
                                                       uint8_t *__restrict__ scratch, uint8_t rnd) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const uint32_t b = prefix[t], e = prefix[t + 1];
    bool any = false;
    for (uint32_t i = b; i < e; i++) any |= state[i] == ST_ALIVE;
    if (!any) return;
    int c = confirmed[t];
    const int left = K - c;
    for (uint32_t i = b; i < e; i++)
        scratch[i] = state[i] == ST_ALIVE
                         ? (uint8_t)knn_status_one<R>(rmin, rmax, state, b, e, i, left, true)
                         : (uint8_t)255;
    bool done = false;
    for (uint32_t i = b; i < e; i++) {
        if (scratch[i] == 255) continue;
        if (done) { state[i] = ST_DEAD; continue; }
        if (scratch[i] == 1) {
            state[i] = ST_REPORTED;
            round[i] = rnd;
            c++;
            if (c == K) done = true;
        } else if (scratch[i] == 2) {
            state[i] = ST_DEAD;
        }
    }
    confirmed[t] = c;
}

// per-LOD within decision (WithinJoin.cpp:769-792)
template <typename R>
__global__ void __launch_bounds__(kBlock) k_decide_within(size_t P, const R *__restrict__ rmin,
                                                          const R *__restrict__ rmax, R tau,
                                                          uint8_t *__restrict__ state,
                                                          uint8_t *__restrict__ round, uint8_t rnd) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    if (state[p] != ST_ALIVE) return;
    if (rmax[p] <= tau) { state[p] = ST_REPORTED; round[p] = rnd; }
    else if (rmin[p] > tau) state[p] = ST_DEAD;
}

// a voxel pair survives the round if its object pair is still refined and
// it is not evicted (mindist > the object's min of maxima)
template <typename R>
__global__ void __launch_bounds__(kBlock) k_vp_keep(uint64_t n_vp, const uint32_t *__restrict__ vp_pair,
                                                    const R *__restrict__ vp_min,
                                                    const uint8_t *__restrict__ state,
                                                    const R *__restrict__ rmax,
                                                    uint32_t *__restrict__ keep) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_vp) return;
    const uint32_t p = vp_pair[w];
    keep[w] = (state[p] == ST_ALIVE && !(vp_min[w] > rmax[p])) ? 1u : 0u;
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_vp_scatter(
    uint64_t n_vp, const uint32_t *__restrict__ keep, const uint64_t *__restrict__ pos,
    const uint32_t *__restrict__ a1, const uint32_t *__restrict__ a2,
    const uint32_t *__restrict__ ap, const R *__restrict__ amin, const R *__restrict__ amax,
    uint32_t *__restrict__ b1, uint32_t *__restrict__ b2, uint32_t *__restrict__ bp,
    R *__restrict__ bmin, R *__restrict__ bmax) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_vp) return;
    if (!keep[w]) return;
    const uint64_t o = pos[w];
    b1[o] = a1[w];
    b2[o] = a2[w];
    bp[o] = ap[w];
    bmin[o] = amin[w];
    bmax[o] = amax[w];
}

// ------------------------------------------------------------------ scan

// exclusive scan of n values (u32 or u64) into u64, 1024 per block
template <typename In>
__global__ void __launch_bounds__(kBlock) k_scan_local(const In *__restrict__ in, size_t n,
                                                       uint64_t *__restrict__ out,
                                                       uint64_t *__restrict__ block_sum) {
    __shared__ uint64_t wsum[kWavesPerBlock];
    const size_t base = (size_t)blockIdx.x * 1024 + 4 * threadIdx.x;
    uint64_t v[4];
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = (base + k < n) ? (uint64_t)in[base + k] : 0;
        s += v[k];
    }
    // inclusive scan of s across the wave
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    uint64_t x = s;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        uint64_t y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) wsum[wv] = x;
    __syncthreads();
    uint64_t woff = 0;
    for (int k = 0; k < wv; k++) woff += wsum[k];
    uint64_t run = woff + x - s;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kBlock - 1) block_sum[blockIdx.x] = woff + x;
}

__global__ void __launch_bounds__(kBlock) k_scan_add(uint64_t *__restrict__ out, size_t n,
                                                     const uint64_t *__restrict__ block_off) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] += block_off[i / 1024];
}

// single-thread tail: total = last offset + last value
template <typename In>
__global__ void k_scan_total(const In *__restrict__ in, const uint64_t *__restrict__ out, size_t n,
                             uint64_t *__restrict__ total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *total = n ? out[n - 1] + (uint64_t)in[n - 1] : 0;
}

// ------------------------------------------------------- batch parity ops

template <typename R>
__global__ void __launch_bounds__(kBlock) k_tridist_batch(size_t n, const float *__restrict__ S,
                                                          const float *__restrict__ T,
                                                          R *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    R s[9], t[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        s[k] = (R)S[9 * i + k];
        t[k] = (R)T[9 * i + k];
    }
    out[i] = tri_dist<R>(s, t);
}

// per voxel pair (d, lo, hi) over 12-float records; one wave per pair
template <typename R, int MODE>
__global__ void __launch_bounds__(kBlock) k_meshdist_batch(size_t P, const uint64_t *__restrict__ os,
                                                           const float4 *__restrict__ rec,
                                                           R *__restrict__ out,
                                                           unsigned long long *__restrict__ stats) {
    const size_t w = wave_global();
    if (w >= P) return;
    const int lane = lane_id();
    const uint64_t o1 = os[4 * w], n1 = os[4 * w + 1], o2 = os[4 * w + 2], n2 = os[4 * w + 3];
    const uint64_t total = n1 * n2;
    R d = (R)__builtin_inff(), lo = (R)__builtin_inff(), hi = (R)__builtin_inff();
    if (MODE != MODE_MIXED) {
        for (uint64_t q = lane; q < total; q += kWave) {
            const uint64_t i = q / n2, j = q - i * n2;
            R a, b, c;
            pair_values<R>(rec, o1 + i, rec, o2 + j, a, b, c);
            d = rmin2(d, a); lo = rmin2(lo, b); hi = rmin2(hi, c);
        }
    } else {
        double dd = INFINITY, dl = INFINITY, dh = INFINITY;
        uint64_t rechecks = 0;
        mixed_minima(rec, o1, rec, o2, n2, total, lane, dd, dl, dh, rechecks);
        d = (R)dd; lo = (R)dl; hi = (R)dh;
        rechecks = wave_sum_u64(rechecks);
        if (lane == 0) atomicAdd(&stats[2], (unsigned long long)rechecks);
    }
    d = wave_min(d);
    lo = wave_min(lo);
    hi = wave_min(hi);
    if (lane == 0) {
        out[3 * w] = d;
        out[3 * w + 1] = lo;
        out[3 * w + 2] = hi;
    }
}

// kNN status for host-given lists (one thread per target)
__global__ void __launch_bounds__(kBlock) k_knn_status(size_t N, const uint32_t *__restrict__ prefix,
                                                       const int32_t *__restrict__ confirmed,
                                                       const double *__restrict__ rmin,
                                                       const double *__restrict__ rmax, int K,
                                                       int32_t *__restrict__ status) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const uint32_t b = prefix[t], e = prefix[t + 1];
    for (uint32_t i = b; i < e; i++)
        status[i] = knn_status_one<double>(rmin, rmax, nullptr, b, e, i, K - confirmed[t], false);
}

}  // namespace tdp
