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

// filter predicate of a join: kNN, within, intersect
enum : int { FILTER_KNN = 0, FILTER_WITHIN = 1, FILTER_INTERSECT = 2 };
// or-ed into the Hausdorff level given to the folds: the intersection join's
// per-round rule instead of the kNN / within range fold
constexpr int kFoldIntersect = 8;
enum { MODE_REF32 = 0, MODE_FP64 = 1, MODE_MIXED = 2 };

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
// orders a wave's LDS writes before its later reads (compiler and memory)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// one wave per item; the grid may be 2-D (x fastest) when the item count
// exceeds what a 1-D launch can address (2^32 work-items per dimension)
__device__ __forceinline__ size_t wave_global() {
    const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
    return b * kWavesPerBlock + (threadIdx.x >> 6);
}

template <typename R>
__device__ __forceinline__ R rmin2(R a, R b) { return b < a ? b : a; }

template <typename R>
__device__ __forceinline__ R wave_min(R v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = rmin2<R>(v, __shfl_xor(v, o, kWave));
    return v;
}
// Wave reductions on the DPP lane network instead of LDS permutes: an
// inclusive scan inside each row of 16 lanes (shifts right by 1, 2, 4, 8),
// then row broadcasts 15 and 31 carry it across the rows, so lane 63 holds
// the reduction of all 64 lanes, read into a scalar register. Lanes whose
// source lies outside their row take the identity (bound control off). Every
// lane of the wave must be active.
template <typename Op>
__device__ __forceinline__ uint32_t wave_reduce_u32(uint32_t v, uint32_t id, Op op) {
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x111, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x112, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x114, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x118, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x142, 0xa, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x143, 0xc, 0xf, false));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
// inclusive prefix sum across the wave (the scan above, lane i holding the
// sum of lanes 0..i); every lane of the wave must be active
__device__ __forceinline__ uint32_t wave_incl_sum_u32(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
    return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    return wave_reduce_u32(v, 0xffffffffu, [](uint32_t a, uint32_t b) { return min(a, b); });
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    return wave_reduce_u32(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// A join's statistics counters: kStatN of them, each spread over kStatSlots
// copies picked by block, so that the one-per-block atomics of a launch with
// tens of thousands of blocks do not queue on one address (they took most of
// the time of the short kernels that count); readers add up the copies.
constexpr int kStatN = 8, kStatSlots = 64;
__device__ __forceinline__ void stat_add(unsigned long long *stats, int k, unsigned long long v) {
    atomicAdd(&stats[kStatN * (blockIdx.x & (kStatSlots - 1)) + k], v);
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
// minimum of a float over the wave, as the minimum of its order-preserving
// keys (no NaN among the values)
__device__ __forceinline__ float wave_min_f32(float v) { return funkey(wave_min_u32(fkey(v))); }
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
// computing_gpu.cu:1984-2043); one wave per object pair. The wave walks the
// n1 x n2 voxel grid in tiles of R rows x cw columns (R * cw <= 64, the fewest
// iterations over 1-3 column chunks): lanes keep their column voxel (of the
// object with more voxels) in registers and load their row voxel per
// iteration; squared distances are min-reduced and the square roots taken once
// (sqrt is monotone, and the min does not depend on the order).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_aggregate(
    size_t P, const uint32_t *__restrict__ pair_target, const uint32_t *__restrict__ pair_obj2,
    const uint32_t *__restrict__ vo1, const float *__restrict__ vb1,
    const uint32_t *__restrict__ vo2, const float *__restrict__ vb2, R *__restrict__ rmin,
    R *__restrict__ rmax) {
    const uint32_t p = __builtin_amdgcn_readfirstlane((uint32_t)wave_global());
    if (p >= P) return;
    const uint32_t lane = (uint32_t)lane_id();
    const uint32_t a0 = vo1[pair_target[p]], n1 = vo1[pair_target[p] + 1] - a0;
    const uint32_t b0 = vo2[pair_obj2[p]], n2 = vo2[pair_obj2[p] + 1] - b0;
    R smin = (R)__builtin_inff(), cmin = (R)__builtin_inff();
    const bool sw = n1 > n2;  // columns: tile-1 voxels
    const uint32_t nl = sw ? n1 : n2, ns = sw ? n2 : n1;
    if (ns > 0) {
        uint32_t cw = kWave, rows = 1, nch = 0, ngr = 0, best = 0xffffffffu;
        const uint32_t m0 = (nl + kWave - 1) / kWave;
        for (uint32_t m = m0; m < m0 + 3; m++) {
            const uint32_t w = (nl + m - 1) / m;
            const uint32_t r = min((uint32_t)kWave / w, ns);
            const uint32_t c = (nl + w - 1) / w, g = (ns + r - 1) / r;
            if (c * g < best) {
                best = c * g;
                cw = w;
                rows = r;
                nch = c;
                ngr = g;
            }
        }
        const float *__restrict__ lb = sw ? vb1 + 9 * (size_t)a0 : vb2 + 9 * (size_t)b0;
        const float *__restrict__ sb = sw ? vb2 + 9 * (size_t)b0 : vb1 + 9 * (size_t)a0;
        const uint32_t r = lane / cw, c = lane - r * cw;
        const bool in_tile = r < rows;
        // the rows' voxel records staged in LDS in R when they fit (one load
        // and one conversion per voxel instead of one per voxel and
        // iteration); the column's converted once per chunk
        __shared__ R stage_all[kWavesPerBlock][kWave * 9];
        R *const stage = stage_all[threadIdx.x / kWave];
        const bool staged = ns <= (uint32_t)kWave;
        if (staged) {
            if (lane < ns) {
#pragma unroll
                for (int k = 0; k < 9; k++) stage[9 * lane + k] = (R)sb[9 * (size_t)lane + k];
            }
            wave_sync();
        }
        // rows from LDS or from memory: two copies of the walk, so that each
        // reads through its own address space
        auto walk = [&](auto row_ptr) {
            for (uint32_t ch = 0; ch < nch; ch++) {
                const uint32_t col = ch * cw + c;
                const bool cv = in_tile && col < nl;
                R lv[9];
                const float *src = lb + 9 * (size_t)(cv ? col : nl - 1);
#pragma unroll
                for (int k = 0; k < 9; k++) lv[k] = (R)src[k];
                for (uint32_t gr = 0; gr < ngr; gr++) {
                    const uint32_t row = gr * rows + r;
                    const auto *rv = row_ptr(row < ns ? row : ns - 1);
                    R sv[9];
#pragma unroll
                    for (int k = 0; k < 9; k++) sv[k] = (R)rv[k];
                    R s, cd;
                    // the tile-1 voxel is always the first operand
                    if (sw)
                        voxel_sq<R>(lv, sv, s, cd);
                    else
                        voxel_sq<R>(sv, lv, s, cd);
                    // cells outside the grid count as +inf (a select)
                    const bool v = cv && row < ns;
                    smin = rmin2(smin, v ? s : (R)__builtin_inff());
                    cmin = rmin2(cmin, v ? cd : (R)__builtin_inff());
                }
            }
        };
        if (staged)
            walk([&](uint32_t rw) { return (const R *)(stage + 9 * rw); });
        else
            walk([&](uint32_t rw) { return sb + 9 * (size_t)rw; });
    }
    smin = wave_min(smin);
    cmin = wave_min(cmin);
    if (lane == 0) {
        rmin[p] = dsqrt<R>(smin);
        rmax[p] = dsqrt<R>(cmin);
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
                                                       const R *__restrict__ rmax,
                                                       const int32_t *__restrict__ k_t,
                                                       uint8_t *__restrict__ state,
                                                       uint8_t *__restrict__ round,
                                                       int32_t *__restrict__ confirmed,
                                                       uint8_t *__restrict__ scratch,
                                                       const uint8_t *__restrict__ within_t) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    if (within_t && within_t[t]) return;  // a within target of a mixed join
    const int K = k_t[t];
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
                                                          const R *__restrict__ rmax,
                                                          const uint32_t *__restrict__ pair_target,
                                                          const float *__restrict__ tau_t,
                                                          uint8_t *__restrict__ state,
                                                          uint8_t *__restrict__ round,
                                                          const uint8_t *__restrict__ within_t) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    if (within_t && !within_t[pair_target[p]]) return;  // a kNN target of a mixed join
    const R tau = (R)tau_t[pair_target[p]];
    uint8_t st = ST_DEAD, rd = 255;
    if (rmin[p] > tau) st = ST_DEAD;
    else if (rmax[p] <= tau) { st = ST_REPORTED; rd = 0; }
    else st = ST_ALIVE;
    state[p] = st;
    round[p] = rd;
}

// voxel pairs kept for refinement: boxmin <= object max (kNN), and also
// boxmin <= tau (within) (computing_gpu.cu:2604-2812); for the intersection
// join, voxel boxes that overlap with positive volume (aab::intersect,
// himesh_box.cpp:95-99, as mbb_intersect uses it, IntersectionJoin.cpp:36)
__device__ __forceinline__ bool boxes_overlap(const float *a, const float *b) {
    return !(b[0] >= a[3] || b[3] <= a[0] || b[1] >= a[4] || b[4] <= a[1] || b[2] >= a[5] ||
             b[5] <= a[2]);
}

template <typename R>
__device__ __forceinline__ bool survives(const float *a, const float *b, R mn, R objmax, int mode,
                                         R tau) {
    if (mode == FILTER_INTERSECT) return boxes_overlap(a, b);
    return mode == FILTER_WITHIN ? (mn <= objmax && mn <= tau) : (mn <= objmax);
}

// per live pair: number of surviving voxel pairs; for within, a surviving
// pair with max <= tau settles the object pair at once (WithinJoin.cpp:190-225)
template <typename R>
__global__ void __launch_bounds__(kBlock) k_survivor_count(
    size_t P, const uint32_t *__restrict__ pair_target, const uint32_t *__restrict__ pair_obj2,
    const uint32_t *__restrict__ vo1, const float *__restrict__ vb1,
    const uint32_t *__restrict__ vo2, const float *__restrict__ vb2, const R *__restrict__ rmax,
    int mode, const float *__restrict__ tau_t, const uint8_t *__restrict__ within_t,
    uint8_t *__restrict__ state, uint8_t *__restrict__ round, uint32_t *__restrict__ count) {
    const size_t p = wave_global();
    if (p >= P) return;
    const int lane = lane_id();
    if (state[p] != ST_ALIVE) {
        if (lane == 0) count[p] = 0;
        return;
    }
    if (within_t) mode = within_t[pair_target[p]] ? FILTER_WITHIN : FILTER_KNN;
    const uint32_t a0 = vo1[pair_target[p]], n1 = vo1[pair_target[p] + 1] - a0;
    const uint32_t b0 = vo2[pair_obj2[p]], n2 = vo2[pair_obj2[p] + 1] - b0;
    const uint32_t total = n1 * n2;
    const bool within = mode == FILTER_WITHIN;
    const R objmax = mode == FILTER_INTERSECT ? (R)0 : rmax[p];
    const R tau = within ? (R)tau_t[pair_target[p]] : (R)0;
    uint32_t cnt = 0;
    bool settled = false;
    for (uint32_t t = lane; t < total; t += kWave) {
        const uint32_t i = t / n2, j = t - i * n2;
        R mn, mx;
        const float *va = vb1 + 9 * (size_t)(a0 + i), *vb = vb2 + 9 * (size_t)(b0 + j);
        voxel_bound<R>(va, vb, mn, mx);
        if (survives<R>(va, vb, mn, objmax, mode, tau)) {
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
    int mode, const float *__restrict__ tau_t, const uint8_t *__restrict__ within_t,
    const uint8_t *__restrict__ state, const uint64_t *__restrict__ offs,
    uint32_t *__restrict__ vp_v1, uint32_t *__restrict__ vp_v2, uint32_t *__restrict__ vp_pair,
    R *__restrict__ vp_min, R *__restrict__ vp_max) {
    const size_t p = wave_global();
    if (p >= P) return;
    if (state[p] != ST_ALIVE) return;
    if (within_t) mode = within_t[pair_target[p]] ? FILTER_WITHIN : FILTER_KNN;
    const int lane = lane_id();
    const uint32_t a0 = vo1[pair_target[p]], n1 = vo1[pair_target[p] + 1] - a0;
    const uint32_t b0 = vo2[pair_obj2[p]], n2 = vo2[pair_obj2[p] + 1] - b0;
    const uint32_t total = n1 * n2;
    const R objmax = mode == FILTER_INTERSECT ? (R)0 : rmax[p];
    const R tau = mode == FILTER_WITHIN ? (R)tau_t[pair_target[p]] : (R)0;
    uint64_t pos = offs[p];
    for (uint32_t base = 0; base < total; base += kWave) {
        const uint32_t t = base + lane;
        bool keep = false;
        R mn = 0, mx = 0;
        uint32_t i = 0, j = 0;
        if (t < total) {
            i = t / n2;
            j = t - i * n2;
            const float *va = vb1 + 9 * (size_t)(a0 + i), *vb = vb2 + 9 * (size_t)(b0 + j);
            voxel_bound<R>(va, vb, mn, mx);
            keep = survives<R>(va, vb, mn, objmax, mode, tau);
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

// ------------------------------------------------------- mixed precision
//
// The mixed form returns, per voxel pair, exactly the FP64 minima of
// (d, lo, hi) = (TriDist, max(d - ph1 - ph2, 0), d + h1 + h2) over its
// n1 x n2 triangle pairs while evaluating FP64 TriDist on few of them:
//   1. FP32, every pair: a rigorous interval for its FP64 TriDist value from
//      per-triangle bounding spheres (about 20 flops), giving the smallest
//      upper bound U of each quantity;
//   2. FP64, one pair per lane (the lane's smallest lower bound): exact
//      values tighten U;
//   3. FP64 on every remaining pair whose interval still reaches U,
//      compacted through an LDS queue so that all 64 lanes work.
// A pair that is not evaluated has a lower bound above U, so it cannot hold
// a minimum; the result equals the all-FP64 evaluation bit for bit.
//
// Why the bounds hold for the reference's TriDist: every value it returns is
// the distance between two points of the triangles, hence at most the
// sphere upper bound and at least the true distance -- except a 0 returned
// for separated triangles when none of its tests showed them disjoint. A
// positive lower bound is therefore used only when the other triangle's
// sphere lies strictly on one side of a triangle's plane, which makes
// TriDist's face test flag the pair as disjoint (no 0 return). (An FP32
// evaluation of TriDist itself is not used: its t = (A.T|B|^2 - B.T A.B)/denom
// loses accuracy like 1/sin^2 of the edge angle and its branch structure is
// discontinuous, so no forward error bound for it is rigorous.)

constexpr float kU = 5.9604645e-08f;    // 2^-24
constexpr uint32_t kMixedMinPairs = 2 * kWave;  // below this, evaluate every pair in FP64

// Per-triangle bound record, kBnd float4 per triangle:
//   {centre.xyz, radius}      sphere around the triangle (centre rounded to
//                             FP32, radius measured from the rounded centre
//                             with a relative margin);
//   {unit normal.xyz, r + max |centre coord|}   normal 0 when near-degenerate
//                             (the margins of the FP32 tests take r and the
//                             centre's magnitude together, rounded upwards);
//   {proxy hausdorff ph, hausdorff h, alo, ahi}: alo = r - ph + 8u t and
//                             ahi = r + h + 8u t (t = r + ph + h, u = 2^-24),
//                             the triangle's parts of the upper bounds of lo
//                             and hi in k_mx_bounds.
// The normal is formed in FP64 from the (exact) FP64 edge differences and
// rounded once, so its direction is within a few FP32 ulps of the true
// plane's even for thin triangles (the FP64 cross product's error relative
// to |n| stays below 1e-9 under the 1e-6 aspect cut-off).
constexpr int kBnd = 3;
// zeroed triangles (records and bounds) after the last one of every device
// tile: a tiled walk reads up to kWave - 1 rows past a voxel's last triangle
// without clamping its index
constexpr size_t kTriPad = kWave;
// most triangles of one voxel at one LOD: the pair walks number a voxel
// pair's n1 x n2 triangle pairs in 32 bits (tiles are checked at upload)
constexpr uint32_t kMaxVoxTris = 65535;

// owning object of every voxel, from the per-object voxel offsets
__global__ void __launch_bounds__(kBlock) k_vox_owner(size_t n_obj,
                                                      const uint32_t *__restrict__ obj_vox_off,
                                                      uint32_t *__restrict__ vox_obj) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_obj) return;
    for (uint32_t v = obj_vox_off[o]; v < obj_vox_off[o + 1]; v++) vox_obj[v] = (uint32_t)o;
}

// the bound record of the triangle (a, b, c) (records carry ph in a.w and h in b.w)
__device__ __forceinline__ void tri_bound(const float4 a, const float4 b, const float4 c,
                                          float4 *__restrict__ out) {
    const float cx = ((a.x + b.x) + c.x) * (1.f / 3.f);
    const float cy = ((a.y + b.y) + c.y) * (1.f / 3.f);
    const float cz = ((a.z + b.z) + c.z) * (1.f / 3.f);
    auto d = [&](float x, float y, float z) {
        return __builtin_sqrtf((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
    };
    float r = fmaxf(d(a.x, a.y, a.z), fmaxf(d(b.x, b.y, b.z), d(c.x, c.y, c.z)));
    r = r * (1.f + 64.f * kU) + 1e-30f;
    const double e1x = (double)b.x - a.x, e1y = (double)b.y - a.y, e1z = (double)b.z - a.z;
    const double e2x = (double)c.x - b.x, e2y = (double)c.y - b.y, e2z = (double)c.z - b.z;
    const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z,
                 nz = e1x * e2y - e1y * e2x;
    const double nl = sqrt(nx * nx + ny * ny + nz * nz);
    const double e = fmax(e1x * e1x + e1y * e1y + e1z * e1z, e2x * e2x + e2y * e2y + e2z * e2z);
    float fx = 0.f, fy = 0.f, fz = 0.f;  // near-degenerate: no plane test
    if (nl > 1e-6 * e && nl * nl > 1e-12) {
        fx = (float)(nx / nl);
        fy = (float)(ny / nl);
        fz = (float)(nz / nl);
    }
    const float cm = (r + fmaxf(fabsf(cx), fmaxf(fabsf(cy), fabsf(cz)))) * (1.f + 2.f * kU);
    // upper-bound constants of lo and hi, rounded upwards by 8u (r + ph + h)
    const float ph = a.w, h = b.w, tt = (r + ph) + h;
    const float alo = (r - ph) + 8.f * kU * tt, ahi = (r + h) + 8.f * kU * tt;
    out[0] = make_float4(cx, cy, cz, r);
    out[1] = make_float4(fx, fy, fz, cm);
    out[2] = make_float4(ph, h, alo, ahi);
}

__global__ void __launch_bounds__(kBlock) k_tri_bounds(size_t n, const float4 *__restrict__ rec,
                                                       float4 *__restrict__ bnd) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tri_bound(rec[3 * i], rec[3 * i + 1], rec[3 * i + 2], bnd + kBnd * i);
}

// ---- records supplied per LOD round (tiles filled without their records)
//
// have[v] bit s: the records of voxel v at LOD slot s are in the device
// arrays (or listed to be put there before the round runs).

// flag[v] = 1 for every voxel a live voxel pair names (v1 when side bit 0,
// v2 when side bit 1) whose records at `slot` the tile does not hold
__global__ void __launch_bounds__(kBlock) k_lazy_flag(size_t n, const uint32_t *__restrict__ v1,
                                                      const uint32_t *__restrict__ v2, int side,
                                                      int slot, const uint32_t *__restrict__ have,
                                                      uint32_t *__restrict__ flag) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t bit = 1u << slot;
    if (side & 1) {
        const uint32_t v = v1[i];
        if (!(have[v] & bit)) flag[v] = 1u;
    }
    if (side & 2) {
        const uint32_t v = v2[i];
        if (!(have[v] & bit)) flag[v] = 1u;
    }
}

// the flagged voxels in voxel order (pos: exclusive scan of flag) with their
// record counts at `slot`
__global__ void __launch_bounds__(kBlock) k_lazy_list(size_t n_vox, int slot,
                                                      const uint32_t *__restrict__ flag,
                                                      const uint64_t *__restrict__ pos,
                                                      const uint32_t *__restrict__ vox_tri_cnt,
                                                      uint32_t *__restrict__ list,
                                                      uint32_t *__restrict__ cnt) {
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vox || !flag[v]) return;
    const uint64_t k = pos[v];
    list[k] = (uint32_t)v;
    cnt[k] = vox_tri_cnt[5 * v + slot];
}

// Listed voxels' records into the tile's record array, with their bound
// records: the packed records (each voxel's records at `slot`, voxel after
// voxel, voxel k's first at off[k]) numbered [p0, p1) are in `piece`. The
// last piece (last != 0) marks the listed voxels held. One wave per voxel,
// grid-stride.
__global__ void __launch_bounds__(kBlock) k_lazy_put(size_t n_list, int slot,
                                                     const uint32_t *__restrict__ list,
                                                     const uint64_t *__restrict__ off,
                                                     const uint64_t *__restrict__ vox_tri_off,
                                                     const uint32_t *__restrict__ vox_tri_cnt,
                                                     const float4 *__restrict__ piece, uint64_t p0,
                                                     uint64_t p1, int last,
                                                     float4 *__restrict__ tri,
                                                     float4 *__restrict__ bnd,
                                                     uint32_t *__restrict__ have) {
    const unsigned lane = threadIdx.x % kWave;
    const size_t waves = (size_t)gridDim.x * (kBlock / kWave);
    for (size_t k = (size_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; k < n_list;
         k += waves) {
        const uint32_t v = list[k];
        const uint64_t dst = vox_tri_off[5 * (size_t)v + slot], src = off[k];
        const uint32_t c = vox_tri_cnt[5 * (size_t)v + slot];
        if (last && lane == 0) have[v] |= 1u << slot;
        if (src >= p1 || src + c <= p0) continue;
        // this voxel's records inside the piece
        const uint32_t t0 = src < p0 ? (uint32_t)(p0 - src) : 0u;
        const uint32_t t1 = src + c > p1 ? (uint32_t)(p1 - src) : c;
        for (uint32_t t = t0 + lane; t < t1; t += kWave) {
            const uint64_t q = src + t - p0;
            const float4 a = piece[3 * q], b = piece[3 * q + 1], cc = piece[3 * q + 2];
            tri[3 * (dst + t)] = a;
            tri[3 * (dst + t) + 1] = b;
            tri[3 * (dst + t) + 2] = cc;
            tri_bound(a, b, cc, bnd + kBnd * (dst + t));
        }
    }
}

// ---- voxel spheres: per voxel and LOD slot, a ball holding every vertex of
// the voxel's triangles at that LOD, for the row test of the mixed mark pass
// (k_mx_mark). Two records per (voxel, slot), at sph[kSph * (5 v + slot)]:
//   {C.xyz, R}   every vertex lies within R of C (C is the middle of the
//                vertices' box; any centre would do, R is what the test
//                relies on: its square is summed in FP64 from exact vertex
//                differences, and the root is rounded upwards to FP32);
//   {R + max |C coord| (rounded upwards), the largest proxy Hausdorff ph,
//    the smallest Hausdorff h, the largest cm of the triangles' bound
//    records}   the margin term of the sphere, the extreme per-triangle
//                Hausdorff values that bound the threshold of any pair with
//                one of these triangles, and the largest per-triangle margin
//                term (r + max |centre coord| as k_tri_bounds stored it).
// A voxel without triangles at the slot gets zeros (its voxel pairs are
// EMPTY and never read it).
constexpr int kSph = 2;

__device__ __forceinline__ float wave_max_f(float v) { return -wave_min<float>(-v); }

// one wave: the sphere records of the cnt triangles at off
__device__ __forceinline__ void vox_sphere(const float4 *__restrict__ tri,
                                           const float4 *__restrict__ bnd, uint64_t off,
                                           uint32_t cnt, float4 *__restrict__ out) {
    const uint32_t lane = (uint32_t)lane_id();
    if (cnt == 0) {
        if (lane < kSph) out[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    float lx = INFINITY, ly = INFINITY, lz = INFINITY;
    float hx = -INFINITY, hy = -INFINITY, hz = -INFINITY, mph = -INFINITY, mh = INFINITY;
    float mcm = 0.f;
    for (uint32_t t = lane; t < cnt; t += kWave) {
        const float4 a = tri[3 * (off + t)], b = tri[3 * (off + t) + 1], c = tri[3 * (off + t) + 2];
        mcm = fmaxf(mcm, bnd[kBnd * (off + t) + 1].w);
        lx = fminf(lx, fminf(a.x, fminf(b.x, c.x)));
        ly = fminf(ly, fminf(a.y, fminf(b.y, c.y)));
        lz = fminf(lz, fminf(a.z, fminf(b.z, c.z)));
        hx = fmaxf(hx, fmaxf(a.x, fmaxf(b.x, c.x)));
        hy = fmaxf(hy, fmaxf(a.y, fmaxf(b.y, c.y)));
        hz = fmaxf(hz, fmaxf(a.z, fmaxf(b.z, c.z)));
        mph = fmaxf(mph, a.w);  // records carry ph in a.w and h in b.w
        mh = fminf(mh, b.w);
    }
    lx = wave_min<float>(lx);
    ly = wave_min<float>(ly);
    lz = wave_min<float>(lz);
    hx = wave_max_f(hx);
    hy = wave_max_f(hy);
    hz = wave_max_f(hz);
    mph = wave_max_f(mph);
    mh = wave_min<float>(mh);
    mcm = wave_max_f(mcm);
    const float cx = 0.5f * (lx + hx), cy = 0.5f * (ly + hy), cz = 0.5f * (lz + hz);
    // FP32 differences are exact in FP64 and so are their squares; the sum
    // and the root add a few 2^-53, which the factor below covers
    double r2 = 0.0;
    for (uint32_t t = lane; t < cnt; t += kWave) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float4 v = tri[3 * (off + t) + k];
            const double dx = (double)v.x - cx, dy = (double)v.y - cy, dz = (double)v.z - cz;
            r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
        }
    }
    float r = (float)sqrt(r2);  // nearest: at most u below
    r = wave_max_f(r) * (1.f + 4.f * kU) + 1e-30f;
    const float cm = (r + fmaxf(fabsf(cx), fmaxf(fabsf(cy), fabsf(cz)))) * (1.f + 2.f * kU);
    if (lane == 0) {
        out[0] = make_float4(cx, cy, cz, r);
        out[1] = make_float4(cm, mph, mh, mcm);
    }
}

// every (voxel, slot) of a tile, one wave each
__global__ void __launch_bounds__(kBlock) k_vox_sph(size_t n_vox,
                                                    const uint64_t *__restrict__ vox_tri_off,
                                                    const uint32_t *__restrict__ vox_tri_cnt,
                                                    const float4 *__restrict__ tri,
                                                    const float4 *__restrict__ bnd,
                                                    float4 *__restrict__ sph) {
    const size_t waves = (size_t)gridDim.x * kWavesPerBlock;
    for (size_t k = (size_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; k < 5 * n_vox;
         k += waves)
        vox_sphere(tri, bnd, vox_tri_off[k], vox_tri_cnt[k], sph + kSph * k);
}

// the listed voxels at one slot (records put for a round), one wave each
__global__ void __launch_bounds__(kBlock) k_vox_sph_list(size_t n_list, int slot,
                                                         const uint32_t *__restrict__ list,
                                                         const uint64_t *__restrict__ vox_tri_off,
                                                         const uint32_t *__restrict__ vox_tri_cnt,
                                                         const float4 *__restrict__ tri,
                                                         const float4 *__restrict__ bnd,
                                                         float4 *__restrict__ sph) {
    const size_t waves = (size_t)gridDim.x * kWavesPerBlock;
    for (size_t k = (size_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; k < n_list;
         k += waves) {
        const size_t vs = 5 * (size_t)list[k] + slot;
        vox_sphere(tri, bnd, vox_tri_off[vs], vox_tri_cnt[vs], sph + kSph * vs);
    }
}

// (d, lo, hi) intervals from an interval [dl, du] of the FP64 TriDist value
__device__ __forceinline__ void intervals(float dl, float du, float ph1, float h1, float ph2,
                                          float h2, float lw[3], float uw[3]) {
    const float s = 8.f * kU * (du + ph1 + ph2 + h1 + h2);
    lw[0] = dl;
    uw[0] = du;
    lw[1] = fmaxf(dl - ph1 - ph2 - s, 0.f);
    uw[1] = fmaxf(du - ph1 - ph2 + s, 0.f);
    lw[2] = dl + h1 + h2 - s;
    uw[2] = du + h1 + h2 + s;
}

// Interval of one triangle pair's FP64 TriDist value.
//  upper: centre distance + both radii (every point pair is closer);
//  lower: the gap between the two triangles' projections on the axis through
//         their centres (no point pair is closer than a separating gap),
//         used only when one triangle's vertices lie strictly on one side of
//         the other's plane (see above), else 0.
// All terms carry margins for FP32 rounding, including the rounded centroid's
// distance from its triangle's plane.
__device__ __forceinline__ void pair_interval(const float4 *r1, const float4 *b1, uint64_t i1,
                                              const float4 *r2, const float4 *b2, uint64_t i2,
                                              float lw[3], float uw[3]) {
    const float4 s1 = b1[kBnd * i1], p1 = b1[kBnd * i1 + 1];
    const float4 s2 = b2[kBnd * i2], p2 = b2[kBnd * i2 + 1];
    const float4 a0 = r1[3 * i1], a1 = r1[3 * i1 + 1], a2 = r1[3 * i1 + 2];
    const float4 c0 = r2[3 * i2], c1 = r2[3 * i2 + 1], c2 = r2[3 * i2 + 2];
    const float dx = s2.x - s1.x, dy = s2.y - s1.y, dz = s2.z - s1.z;
    const float dc = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
    const float rs = s1.w + s2.w;
    const float m = 32.f * kU * (dc + p1.w + p2.w) + 1e-6f;  // p.w = r + max |centre coord|
    const float du = dc + rs + m;
    float dl = 0.f;
    // vertices relative to the first centre
    const float ax0 = a0.x - s1.x, ay0 = a0.y - s1.y, az0 = a0.z - s1.z;
    const float ax1 = a1.x - s1.x, ay1 = a1.y - s1.y, az1 = a1.z - s1.z;
    const float ax2 = a2.x - s1.x, ay2 = a2.y - s1.y, az2 = a2.z - s1.z;
    const float cx0 = c0.x - s1.x, cy0 = c0.y - s1.y, cz0 = c0.z - s1.z;
    const float cx1 = c1.x - s1.x, cy1 = c1.y - s1.y, cz1 = c1.z - s1.z;
    const float cx2 = c2.x - s1.x, cy2 = c2.y - s1.y, cz2 = c2.z - s1.z;
    // plane of triangle 1 against the vertices of triangle 2 (and vice versa)
    const float t0 = p1.x * cx0 + p1.y * cy0 + p1.z * cz0;
    const float t1 = p1.x * cx1 + p1.y * cy1 + p1.z * cz1;
    const float t2 = p1.x * cx2 + p1.y * cy2 + p1.z * cz2;
    const float w0 = p2.x * (ax0 - dx) + p2.y * (ay0 - dy) + p2.z * (az0 - dz);
    const float w1 = p2.x * (ax1 - dx) + p2.y * (ay1 - dy) + p2.z * (az1 - dz);
    const float w2 = p2.x * (ax2 - dx) + p2.y * (ay2 - dy) + p2.z * (az2 - dz);
    const bool sep1 = fminf(t0, fminf(t1, t2)) > m || fmaxf(t0, fmaxf(t1, t2)) < -m;
    const bool sep2 = fminf(w0, fminf(w1, w2)) > m || fmaxf(w0, fmaxf(w1, w2)) < -m;
    if ((sep1 || sep2) && dc > 0.f) {
        const float inv = __builtin_amdgcn_rcpf(dc);
        const float ux = dx * inv, uy = dy * inv, uz = dz * inv;
        const float smax = fmaxf(ux * ax0 + uy * ay0 + uz * az0,
                                 fmaxf(ux * ax1 + uy * ay1 + uz * az1, ux * ax2 + uy * ay2 + uz * az2));
        const float tmin = fminf(ux * cx0 + uy * cy0 + uz * cz0,
                                 fminf(ux * cx1 + uy * cy1 + uz * cz1, ux * cx2 + uy * cy2 + uz * cz2));
        dl = fmaxf(tmin - smax - m, 0.f);
    }
    intervals(dl, du, a0.w, a1.w, c0.w, c1.w, lw, uw);
}


// the next float above a non-negative x (x itself for inf / nan)
__device__ __forceinline__ float ulp_up(float x) {
    // selects, not branches: +inf and NaN stay as they are
    const uint32_t b = __float_as_uint(x);
    const float up = __uint_as_float(x >= 0.f ? b + 1u : b - 1u);
    return x < INFINITY ? up : x;
}

__device__ __forceinline__ bool reaches(const float lw[3], const float U[3]) {
    return lw[0] <= U[0] || lw[1] <= U[1] || lw[2] <= U[2];
}

// Exact FP64 minima of (d, lo, hi) for one voxel pair; called by all 64 lanes
// of a wave (per-lane partial minima, wave-min after). `queue` is this
// wave's 128-entry LDS buffer.
__device__ __forceinline__ void mixed_minima(const float4 *rec1, const float4 *bnd1, uint64_t o1,
                                             const float4 *rec2, const float4 *bnd2, uint64_t o2,
                                             uint32_t n2, uint32_t total, int lane,
                                             uint32_t *queue, double &d, double &lo, double &hi,
                                             uint64_t &evals64, uint64_t &culled) {
    float lw[3], uw[3];
    // 1. sphere intervals: smallest upper bounds, and each lane's most
    //    promising pair (smallest lower bound of the distance)
    float up[3] = {INFINITY, INFINITY, INFINITY};
    float best = INFINITY;
    uint32_t best_q = 0xffffffffu;
    for (uint32_t q = lane; q < total; q += kWave) {
        const uint32_t i = q / n2, j = q - i * n2;
        pair_interval(rec1, bnd1, o1 + i, rec2, bnd2, o2 + j, lw, uw);
#pragma unroll
        for (int f = 0; f < 3; f++) up[f] = fminf(up[f], uw[f]);
        if (lw[0] < best) { best = lw[0]; best_q = q; }
    }
    float U[3];
#pragma unroll
    for (int f = 0; f < 3; f++) U[f] = wave_min(up[f]);
    // lo >= 0: if some pair certainly has lo = 0, the minimum is exactly +0
    const bool lo_zero = (U[1] == 0.f);
    if (lo_zero) { lo = 0.0; U[1] = -1.f; }
    // 2. FP64 on each lane's most promising pair
    if (best_q != 0xffffffffu) {
        const uint32_t i = best_q / n2, j = best_q - i * n2;
        double a, b, c;
        pair_values<double>(rec1, o1 + i, rec2, o2 + j, a, b, c);
        d = rmin2(d, a); lo = rmin2(lo, b); hi = rmin2(hi, c);
        evals64++;
    }
    {
        const float Ud = wave_min((float)d), Uh = wave_min((float)hi);
        // the FP64 values rounded to FP32 may round down; step up one ulp
        U[0] = fminf(U[0], ulp_up(Ud));
        U[2] = fminf(U[2], ulp_up(Uh));
        if (!lo_zero) {
            const float Ul = wave_min((float)lo);
            U[1] = fminf(U[1], ulp_up(Ul));
        }
    }
    // 3. FP64 on every other pair whose interval reaches U
    uint32_t cnt = 0;
    for (uint32_t base = 0; base < total; base += kWave) {
        const uint32_t q = base + lane;
        bool cand = false;
        if (q < total && q != best_q) {
            const uint32_t i = q / n2, j = q - i * n2;
            pair_interval(rec1, bnd1, o1 + i, rec2, bnd2, o2 + j, lw, uw);
            cand = reaches(lw, U);
            culled += cand ? 0 : 1;
        }
        const uint64_t mask = __ballot(cand);
        if (cand) queue[cnt + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = q;
        cnt += (uint32_t)__popcll(mask);
        wave_sync();
        if (cnt >= kWave) {
            const uint32_t qq = queue[lane];
            const uint32_t i = qq / n2, j = qq - i * n2;
            double a, b, c;
            pair_values<double>(rec1, o1 + i, rec2, o2 + j, a, b, c);
            d = rmin2(d, a); lo = rmin2(lo, b); hi = rmin2(hi, c);
            evals64++;
            const uint32_t rest = cnt - kWave;
            const uint32_t mv = lane < (int)rest ? queue[kWave + lane] : 0u;
            wave_sync();
            if (lane < (int)rest) queue[lane] = mv;
            cnt = rest;
            wave_sync();
        }
    }
    if (lane < (int)cnt) {
        const uint32_t qq = queue[lane];
        const uint32_t i = qq / n2, j = qq - i * n2;
        double a, b, c;
        pair_values<double>(rec1, o1 + i, rec2, o2 + j, a, b, c);
        d = rmin2(d, a); lo = rmin2(lo, b); hi = rmin2(hi, c);
        evals64++;
    }
    wave_sync();
    if (lo_zero) lo = 0.0;
}

// The range a voxel pair carries after a round, from this round's minima
// (d, lo, hi) and its previous range:
//   last LOD        -> [d, d]
//   hausdorff lvl 2 -> [min(min, lo), min(max, hi)]
//   hausdorff lvl 1 -> [min(min, d - H1 - H2), min(max, d + P1 + P2)] (object-level H, P)
//   hausdorff lvl 0 -> [min, min(max, d)]
// (KNNJoin.cpp:1663-1700, WithinJoin.cpp:596-632). With kFoldIntersect the
// intersection join's per-round values instead (IntersectionJoin.cpp:100-115,
// computing_gpu.cu:1441-1512): max = hi of this round (0 exactly when a
// triangle pair at distance 0 has no Hausdorff slack: an intersection), min =
// the value whose positivity rules the pair out -- lo (level 2),
// d - P1 - P2 (level 1), +inf (level 0: never ruled out).
// object-level Hausdorff (hd) and proxy (pr) values of a voxel pair's two
// objects at a LOD slot, subtracted / added one at a time as the reference
// does (KNNJoin.cpp:1681-1683, IntersectionJoin.cpp:104-105)
template <typename R>
struct ObjSlack {
    R hd1 = 0, hd2 = 0, pr1 = 0, pr2 = 0;
};

template <typename R>
__device__ __forceinline__ ObjSlack<R> obj_slack(const float *__restrict__ ohd1,
                                                 const float *__restrict__ ohd2, uint32_t o1,
                                                 uint32_t o2, int slot) {
    ObjSlack<R> x;
    x.hd1 = (R)ohd1[10 * (size_t)o1 + 2 * slot];
    x.hd2 = (R)ohd2[10 * (size_t)o2 + 2 * slot];
    x.pr1 = (R)ohd1[10 * (size_t)o1 + 2 * slot + 1];
    x.pr2 = (R)ohd2[10 * (size_t)o2 + 2 * slot + 1];
    return x;
}

template <typename R>
__device__ __forceinline__ void fold_range(int is_last, int level, R d, R lo, R hi,
                                           const ObjSlack<R> &x, R &nmin, R &nmax) {
    if (level & kFoldIntersect) {
        const int lv = level & 3;
        nmax = hi;
        nmin = lv == 2 ? lo : lv == 1 ? d - x.pr1 - x.pr2 : (R)__builtin_inff();
        return;
    }
    if (is_last) {
        nmin = d;
        nmax = d;
    } else if (level == 2) {
        if (lo < nmin) nmin = lo;
        if (hi < nmax) nmax = hi;
    } else if (level == 1) {
        const R a = d - x.hd1 - x.hd2, b = d + x.pr1 + x.pr2;
        if (a < nmin) nmin = a;
        if (b < nmax) nmax = b;
    } else {
        if (d < nmax) nmax = d;
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
    const float4 *__restrict__ bnd1, const float4 *__restrict__ bnd2,
    const uint32_t *__restrict__ vobj1, const uint32_t *__restrict__ vobj2,
    const float *__restrict__ ohd1, const float *__restrict__ ohd2, int slot, int is_last,
    int level, typename Key<R>::type *__restrict__ acc_min,
    typename Key<R>::type *__restrict__ acc_max, unsigned long long *__restrict__ stats) {
    static_assert(MODE != MODE_MIXED, "the mixed form runs as the k_mx_* pipeline");
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
        // an empty voxel keeps its range (the intersection join: no values)
        if (lane == 0 && !(level & kFoldIntersect)) atomicMin(&acc_min[p], Key<R>::enc(omin));
        return;
    }
    const uint32_t total = n1 * n2;
    R d = (R)__builtin_inff(), lo = (R)__builtin_inff(), hi = (R)__builtin_inff();
    for (uint32_t q = lane; q < total; q += kWave) {
        const uint32_t i = q / n2, j = q - i * n2;
        R a, b, c;
        pair_values<R>(rec1, o1 + i, rec2, o2 + j, a, b, c);
        d = rmin2(d, a);
        lo = rmin2(lo, b);
        hi = rmin2(hi, c);
    }
    d = wave_min(d);
    lo = wave_min(lo);
    hi = wave_min(hi);
    if (lane == 0) {
        R nmin = omin, nmax = omax;
        ObjSlack<R> x;
        if ((level & 3) == 1) x = obj_slack<R>(ohd1, ohd2, vobj1[v1], vobj2[v2], slot);
        fold_range<R>(is_last, level, d, lo, hi, x, nmin, nmax);
        vp_min[w] = nmin;
        vp_max[w] = nmax;
        atomicMin(&acc_min[p], Key<R>::enc(nmin));
        atomicMin(&acc_max[p], Key<R>::enc(nmax));
    }
}

// per round: triangle pairs covered (stats[0]) and non-empty voxel pairs
// evaluated (stats[1]); one thread per voxel pair, one atomic per block
__global__ void __launch_bounds__(kBlock) k_round_stats(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint32_t *__restrict__ cnt1, const uint32_t *__restrict__ cnt2, int slot,
    unsigned long long *__restrict__ stats) {
    __shared__ unsigned long long part[2][kWavesPerBlock];
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t pairs = 0, vps = 0;
    if (w < n_vp && slot >= 0) {
        const uint64_t t = (uint64_t)cnt1[5 * (size_t)vp_v1[w] + slot] * cnt2[5 * (size_t)vp_v2[w] + slot];
        pairs = t;
        vps = t ? 1 : 0;
    }
    pairs = wave_sum_u64(pairs);
    vps = wave_sum_u64(vps);
    if (lane_id() == 0) {
        part[0][threadIdx.x / kWave] = pairs;
        part[1][threadIdx.x / kWave] = vps;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int k = 0; k < kWavesPerBlock; k++) {
            a += part[0][k];
            b += part[1][k];
        }
        if (a) stat_add(stats, 0, a);
        if (b) stat_add(stats, 1, b);
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
                                                       const R *__restrict__ rmax,
                                                       const int32_t *__restrict__ k_t,
                                                       uint8_t *__restrict__ state,
                                                       uint8_t *__restrict__ round,
                                                       int32_t *__restrict__ confirmed,
                                                       uint8_t *__restrict__ scratch, uint8_t rnd,
                                                       const uint8_t *__restrict__ within_t) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    if (within_t && within_t[t]) return;  // a within target of a mixed join
    const int K = k_t[t];
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

// per-LOD intersection decision (IntersectionJoin.cpp:96-130): a voxel pair
// with hi = 0 reports the object pair (its distance range [0, 0]); else, at
// Hausdorff levels 1 and 2, the pair is ruled out when every voxel pair's
// lower value is positive; otherwise it stays for the next LOD
template <typename R>
__global__ void __launch_bounds__(kBlock) k_decide_intersect(size_t P, R *__restrict__ rmin,
                                                             R *__restrict__ rmax, int level,
                                                             uint8_t *__restrict__ state,
                                                             uint8_t *__restrict__ round,
                                                             uint8_t rnd) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    if (state[p] != ST_ALIVE) return;
    if (rmax[p] == (R)0) {
        state[p] = ST_REPORTED;
        round[p] = rnd;
        rmin[p] = 0;
        rmax[p] = 0;
    } else if (level >= 1 && rmin[p] > (R)0) {
        state[p] = ST_DEAD;
    }
}

// the target of every candidate pair, from the per-target list offsets
__global__ void __launch_bounds__(kBlock) k_pair_targets(size_t N, const uint32_t *__restrict__ prefix,
                                                         uint32_t *__restrict__ pair_target) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    for (uint32_t i = prefix[t]; i < prefix[t + 1]; i++) pair_target[i] = (uint32_t)t;
}

// every object pair of the intersection join enters the filter undecided
__global__ void __launch_bounds__(kBlock) k_state_alive(size_t P, uint8_t *__restrict__ state,
                                                        uint8_t *__restrict__ round) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    state[p] = ST_ALIVE;
    round[p] = 255;
}

// per-LOD within decision (WithinJoin.cpp:769-792)
template <typename R>
__global__ void __launch_bounds__(kBlock) k_decide_within(size_t P, const R *__restrict__ rmin,
                                                          const R *__restrict__ rmax,
                                                          const uint32_t *__restrict__ pair_target,
                                                          const float *__restrict__ tau_t,
                                                          uint8_t *__restrict__ state,
                                                          uint8_t *__restrict__ round, uint8_t rnd,
                                                          const uint8_t *__restrict__ within_t) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    if (state[p] != ST_ALIVE) return;
    if (within_t && !within_t[pair_target[p]]) return;  // a kNN target of a mixed join
    const R tau = (R)tau_t[pair_target[p]];
    if (rmax[p] <= tau) { state[p] = ST_REPORTED; round[p] = rnd; }
    else if (rmin[p] > tau) state[p] = ST_DEAD;
}

// a voxel pair survives the round if its object pair is still refined and
// (evict) it is not evicted (mindist > the object's min of maxima); the
// intersection join evicts no voxel pair of an undecided object pair
template <typename R>
__global__ void __launch_bounds__(kBlock) k_vp_keep(uint64_t n_vp, const uint32_t *__restrict__ vp_pair,
                                                    const R *__restrict__ vp_min,
                                                    const uint8_t *__restrict__ state,
                                                    const R *__restrict__ rmax, int evict,
                                                    uint32_t *__restrict__ keep) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_vp) return;
    const uint32_t p = vp_pair[w];
    keep[w] = (state[p] == ST_ALIVE && !(evict && vp_min[w] > rmax[p])) ? 1u : 0u;
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
                                                           const float4 *__restrict__ bnd,
                                                           R *__restrict__ out,
                                                           unsigned long long *__restrict__ stats) {
    __shared__ uint32_t queue[kWavesPerBlock][2 * kWave];
    const size_t w = wave_global();
    if (w >= P) return;
    const int lane = lane_id();
    const uint64_t o1 = os[4 * w], o2 = os[4 * w + 2];
    const uint32_t n1 = (uint32_t)os[4 * w + 1], n2 = (uint32_t)os[4 * w + 3];
    const uint32_t total = n1 * n2;
    R d = (R)__builtin_inff(), lo = (R)__builtin_inff(), hi = (R)__builtin_inff();
    if (MODE != MODE_MIXED) {
        for (uint32_t q = lane; q < total; q += kWave) {
            const uint32_t i = q / n2, j = q - i * n2;
            R a, b, c;
            pair_values<R>(rec, o1 + i, rec, o2 + j, a, b, c);
            d = rmin2(d, a); lo = rmin2(lo, b); hi = rmin2(hi, c);
        }
    } else {
        double dd = INFINITY, dl = INFINITY, dh = INFINITY;
        uint64_t evals64 = 0, culled = 0;
        mixed_minima(rec, bnd, o1, rec, bnd, o2, n2, total, lane, queue[threadIdx.x >> 6], dd,
                     dl, dh, evals64, culled);
        d = (R)dd; lo = (R)dl; hi = (R)dh;
        evals64 = wave_sum_u64(evals64);
        culled = wave_sum_u64(culled);
        if (lane == 0) {
            atomicAdd(&stats[2], (unsigned long long)evals64);
            atomicAdd(&stats[3], (unsigned long long)culled);
        }
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

// n 64-bit words set to v (a kernel rather than a memset: usable wherever a
// caller's stream is captured)
__global__ void k_fill_u64(unsigned long long *__restrict__ p, uint32_t n, unsigned long long v) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// *bad = 1 when a pair {offset1, n1, offset2, n2} reaches past n_tri records
// or has more than 2^31 triangle pairs (the kernels' 32-bit pair index)
__global__ void __launch_bounds__(kBlock) k_check_pairs(size_t P, const uint64_t *__restrict__ os,
                                                        uint64_t n_tri,
                                                        unsigned long long *__restrict__ bad) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const uint64_t o1 = os[4 * p], n1 = os[4 * p + 1], o2 = os[4 * p + 2], n2 = os[4 * p + 3];
    const bool out = o1 > n_tri || n1 > n_tri - o1 || o2 > n_tri || n2 > n_tri - o2 ||
                     (n1 && n2 > 0x7fffffffull / n1);
    if (out) atomicOr(bad, 1ull);
}

// n 32-bit words set to v
__global__ void __launch_bounds__(kBlock) k_fill_u32(size_t n, uint32_t *__restrict__ p, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
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
