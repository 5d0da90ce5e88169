// kNN and within candidate lists on the device (included after
// tdp_kernels.hpp).
//
// The octree's best-first query (octree.cpp:215-291) with add_candidate
// (:120-154) returns, for a target with k wanted: every object whose box
// distance is below M, the k-th smallest transform distance, together with
// the k smallest, ordered by transform distance; equal distances keep the
// order in which the octree reached the objects. Here one wave per target
// computes every object's two distances with the host's float arithmetic,
// finds M by k rounds of wave-wide extraction of the smallest (distance,
// index), and ranks the chosen objects by distance. Where equal distances
// decide the order or the k-th place, the target is flagged for the host
// octree instead.
#pragma once

#include "tdp_kernels.hpp"

namespace tdp {

constexpr int kCandMaxObj = 1024;              // tile-2 objects on the device path
constexpr int kCandPerLane = kCandMaxObj / kWave;
constexpr uint32_t kCandMaxCap = 256;          // longest list ranked on the device

// aab::distance min part (himesh_box.cpp:149-168) with a = object, b = query
__device__ __forceinline__ float cand_box_min(const float *a, const float *b) {
    float mn = 0.f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float t1 = a[i] - b[3 + i];
        const float t2 = a[3 + i] - b[i];
        if (t2 < 0) mn += t2 * t2;
        else if (t1 > 0) mn += t1 * t1;
    }
    return __builtin_sqrtf(mn);
}

__device__ __forceinline__ float cand_point(const float *a, const float *b) {
    return __builtin_sqrtf((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                           (a[2] - b[2]) * (a[2] - b[2]));
}

// (m, o) < (bm, bo) lexicographically
__device__ __forceinline__ bool key_less(float m, uint32_t o, float bm, uint32_t bo) {
    return m < bm || (m == bm && o < bo);
}

__global__ void __launch_bounds__(kBlock) k_knn_candidates(
    uint64_t N, const float *__restrict__ box1, const float *__restrict__ tr1, uint32_t n2,
    const float *__restrict__ box2, const float *__restrict__ tr2, int self,
    const int32_t *__restrict__ k_t, uint32_t cap, uint32_t *__restrict__ count,
    uint32_t *__restrict__ ids, uint8_t *__restrict__ fallback) {
    __shared__ float lm[kWavesPerBlock][kCandMaxCap];
    __shared__ uint32_t lo[kWavesPerBlock][kCandMaxCap];
    const uint32_t t = wave_item();
    if (t >= N) return;
    const int lane = lane_id();
    const int wv = threadIdx.x / kWave;
    float qb[6], qt[3];
#pragma unroll
    for (int i = 0; i < 6; i++) qb[i] = box1[6 * (size_t)t + i];
#pragma unroll
    for (int i = 0; i < 3; i++) qt[i] = tr1[3 * (size_t)t + i];
    const int K = k_t[t];
    float mx[kCandPerLane], mn[kCandPerLane];
    uint32_t nv = 0;
#pragma unroll
    for (int j = 0; j < kCandPerLane; j++) {
        const uint32_t o = (uint32_t)lane + kWave * j;
        const bool v = o < n2 && !(self && o == t);
        mx[j] = v ? cand_point(qt, tr2 + 3 * (size_t)o) : INFINITY;
        mn[j] = v ? cand_box_min(box2 + 6 * (size_t)o, qb) : INFINITY;
        nv += (uint32_t)__popcll(__ballot(v));
    }
    // the K-th smallest (transform distance, index)
    float M = INFINITY;
    uint32_t oK = 0xffffffffu;
    bool tie = false;
    if ((uint32_t)K <= nv) {
        float pm = -INFINITY;
        uint32_t po = 0;
        for (int r = 0; r < K; r++) {
            float bm = INFINITY;
            uint32_t bo = 0xffffffffu;
#pragma unroll
            for (int j = 0; j < kCandPerLane; j++) {
                const uint32_t o = (uint32_t)lane + kWave * j;
                if (mx[j] < INFINITY && key_less(pm, po, mx[j], o) && key_less(mx[j], o, bm, bo)) {
                    bm = mx[j];
                    bo = o;
                }
            }
            const float wm = wave_min(bm);
            uint32_t c = bm == wm ? bo : 0xffffffffu;
#pragma unroll
            for (int s = 32; s > 0; s >>= 1) c = min(c, (uint32_t)__shfl_xor((int)c, s, kWave));
            pm = wm;
            po = c;
        }
        M = pm;
        oK = po;
        // an object outside the first K at distance M: the octree's order decides
        bool outside = false;
#pragma unroll
        for (int j = 0; j < kCandPerLane; j++) {
            const uint32_t o = (uint32_t)lane + kWave * j;
            outside |= mx[j] == M && o > oK;
        }
        tie = __any(outside);
    }
    // chosen objects, compacted into LDS
    uint32_t c = 0;
    bool over = false;
#pragma unroll
    for (int j = 0; j < kCandPerLane; j++) {
        const uint32_t o = (uint32_t)lane + kWave * j;
        const bool take = mx[j] < INFINITY && (mn[j] < M || key_less(mx[j], o, M, oK) ||
                                               (mx[j] == M && o == oK));
        const uint64_t m = __ballot(take);
        const uint32_t at = c + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (take && at < cap && at < kCandMaxCap) {
            lm[wv][at] = mx[j];
            lo[wv][at] = o;
        }
        c += (uint32_t)__popcll(m);
    }
    over = c > cap || c > kCandMaxCap;
    wave_sync();
    // rank by distance; equal distances among the chosen: the octree's order decides
    bool dup = false;
    if (!tie && !over) {
        for (uint32_t i = lane; i < c; i += kWave) {
            const float m = lm[wv][i];
            uint32_t rank = 0, eq = 0;
            for (uint32_t q = 0; q < c; q++) {
                rank += lm[wv][q] < m;
                eq += lm[wv][q] == m;
            }
            dup |= eq > 1;
            ids[(size_t)t * cap + rank] = lo[wv][i];
        }
    }
    dup = __any(dup);
    if (lane == 0) {
        const bool fb = tie || over || dup;
        fallback[t] = fb ? 1 : 0;
        count[t] = fb ? 0 : c;
    }
}

// Within candidate lists on the device (octree.cpp:293-330, query_within):
// the octree's depth-first walk visits a node when its box distance to the
// target's box is at most tau and, in a leaf, takes every object (not the
// target itself, not taken before) whose box distance is at most tau, in the
// leaf's order. A child's box lies inside its parent's, and the FP32 box
// distance is monotone in the box, so a leaf is reached exactly when its own
// box passes: the walk is the leaves in depth-first order, each tested on its
// own. One wave per target tests 64 leaves at a time, then walks the passing
// ones in order, 64 objects at a time, with an LDS bit set of the objects
// taken. Lists longer than cap are flagged for the host octree.
__global__ void __launch_bounds__(kBlock) k_within_candidates(
    uint64_t N, const float *__restrict__ box1, uint32_t n2, const float *__restrict__ box2,
    int self, const float *__restrict__ tau_t, uint32_t n_leaves,
    const float *__restrict__ leaf_box, const uint32_t *__restrict__ leaf_off,
    const uint32_t *__restrict__ leaf_obj, uint32_t cap, uint32_t *__restrict__ count,
    uint32_t *__restrict__ ids, uint8_t *__restrict__ fallback) {
    __shared__ uint32_t seen_all[kWavesPerBlock][kCandMaxObj / 32];
    const uint32_t t = wave_item();
    if (t >= N) return;
    const int lane = lane_id();
    uint32_t *const seen = seen_all[threadIdx.x / kWave];
    for (uint32_t i = (uint32_t)lane; i < kCandMaxObj / 32; i += kWave) seen[i] = 0u;
    wave_sync();
    float qb[6];
#pragma unroll
    for (int i = 0; i < 6; i++) qb[i] = box1[6 * (size_t)t + i];
    const float tau = tau_t[t];
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t c = 0;
    for (uint32_t l0 = 0; l0 < n_leaves; l0 += kWave) {
        const uint32_t l = l0 + (uint32_t)lane;
        const bool visit = l < n_leaves && cand_box_min(leaf_box + 6 * (size_t)l, qb) <= tau;
        unsigned long long vm = __ballot(visit);
        while (vm) {
            const uint32_t leaf = l0 + (uint32_t)(__ffsll((long long)vm) - 1);
            vm &= vm - 1ull;
            const uint32_t k1 = leaf_off[leaf + 1];
            for (uint32_t k0 = leaf_off[leaf]; k0 < k1; k0 += kWave) {
                const uint32_t k = k0 + (uint32_t)lane;
                uint32_t o = 0;
                bool take = false;
                if (k < k1) {
                    o = leaf_obj[k];
                    // an object appears once per leaf: no two lanes share o
                    take = o < n2 && !(self && o == t) && !((seen[o >> 5] >> (o & 31u)) & 1u) &&
                           cand_box_min(box2 + 6 * (size_t)o, qb) <= tau;
                }
                const unsigned long long tm = __ballot(take);
                if (take) {
                    const uint32_t at = c + (uint32_t)__popcll(tm & below);
                    if (at < cap) ids[(size_t)t * cap + at] = o;
                    atomicOr(&seen[o >> 5], 1u << (o & 31u));
                }
                c += (uint32_t)__popcll(tm);
                wave_sync();
            }
        }
    }
    if (lane == 0) {
        fallback[t] = c > cap ? 1 : 0;
        count[t] = c > cap ? 0 : c;
    }
}

// per-target lists (stride cap) to one array at offs[t]
__global__ void __launch_bounds__(kBlock) k_cand_compact(uint64_t N, uint32_t cap,
                                                        const uint32_t *__restrict__ count,
                                                        const uint64_t *__restrict__ offs,
                                                        const uint32_t *__restrict__ lists,
                                                        uint32_t *__restrict__ out) {
    const uint32_t t = wave_item();
    if (t >= N) return;
    const uint32_t n = count[t];
    for (uint32_t i = lane_id(); i < n; i += kWave) out[offs[t] + i] = lists[(size_t)t * cap + i];
}

}  // namespace tdp
