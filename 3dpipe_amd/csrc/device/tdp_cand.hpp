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
constexpr int kWithinMaxObj = 16384;           // tile-2 objects of the within lists

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
// taken (kWithinMaxObj bits per wave). Lists longer than cap are flagged for
// the host octree.
__global__ void __launch_bounds__(kBlock) k_within_candidates(
    uint64_t N, const float *__restrict__ box1, uint32_t n2, const float *__restrict__ box2,
    int self, const float *__restrict__ tau_t, uint32_t n_leaves,
    const float *__restrict__ leaf_box, const uint32_t *__restrict__ leaf_off,
    const uint32_t *__restrict__ leaf_obj, uint32_t cap, uint32_t *__restrict__ count,
    uint32_t *__restrict__ ids, uint8_t *__restrict__ fallback) {
    __shared__ uint32_t seen_all[kWavesPerBlock][kWithinMaxObj / 32];
    const uint32_t t = wave_item();
    if (t >= N) return;
    const int lane = lane_id();
    uint32_t *const seen = seen_all[threadIdx.x / kWave];
    const uint32_t words = (n2 + 31) / 32;
    for (uint32_t i = (uint32_t)lane; i < words; i += kWave) seen[i] = 0u;
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

// ------------------------------------------------------- kNN lists on a grid
//
// For tile 2 of more than kCandMaxObj objects: the same lists as
// k_knn_candidates (every object whose box distance is below M, the k-th
// smallest transform distance, together with the k smallest, ordered by
// transform distance; ties to the host octree), with tile 2's objects binned
// by transform point into a grid of cubic cells (x fastest, one contiguous run
// of objects per cell), so that a target reads only the cells near it:
//  1. the cube of cells around the target's cell grows until it holds k
//     objects besides the target; the k-th smallest transform distance among
//     them, U, bounds M from above;
//  2. every object within U of the target's transform point lies in the cells
//     that q +- U covers: their (distance, index) keys give M and the k-th key
//     exactly, and every object at distance M;
//  3. an object with box distance below M has its transform point within
//     M + E of the target's box per axis, E bounding the distance from any
//     object's transform point to the far side of its own box: the cells
//     covering the target's box and point grown by M + E hold every object of
//     the list.
// Windows are grown by a relative margin and an absolute one for the rounding
// of the float distances and bounds; a coordinate's cell, floor((x - lo) / h)
// clamped to the grid, is monotone in x, so a window's end cells bound the
// cells of every point inside it.

constexpr uint32_t kGridSel = 1024;  // keys a wave holds in steps 1-2

struct CandGrid {
    float lo[3];
    float inv_h;
    float emax;  // E above
    float tol;   // absolute margin for coordinate rounding
    int n[3];
};

__device__ __forceinline__ int grid_cell(const CandGrid &g, int a, float x) {
    float f = (x - g.lo[a]) * g.inv_h;
    f = fminf(fmaxf(f, 0.f), (float)(g.n[a] - 1));
    return (int)f;
}

// box distance min part with the object's box as two records {low, -}, {high, -}
__device__ __forceinline__ float cand_box_min4(float4 l, float4 h, const float *b) {
    const float a[6] = {l.x, l.y, l.z, h.x, h.y, h.z};
    return cand_box_min(a, b);
}

// bounds of tile 2's transform points and E: out = {lo.xyz, hi.xyz, E}
// (one block)
__global__ void __launch_bounds__(1024) k_grid_bounds(uint32_t n, const float *__restrict__ box,
                                                     const float *__restrict__ tr,
                                                     float *__restrict__ out) {
    __shared__ float red[7][1024 / kWave];
    float v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.f};
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        float e = 0.f;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float t = tr[3 * (size_t)i + a];
            v[a] = fminf(v[a], t);
            v[3 + a] = fmaxf(v[3 + a], t);
            const float d = fmaxf(fabsf(t - box[6 * (size_t)i + a]), fabsf(box[6 * (size_t)i + 3 + a] - t));
            e += d * d;
        }
        v[6] = fmaxf(v[6], e);
    }
#pragma unroll
    for (int a = 0; a < 7; a++) {
        float x = v[a];
        for (int s = 32; s > 0; s >>= 1) {
            const float y = __shfl_xor(x, s, kWave);
            x = a < 3 ? fminf(x, y) : fmaxf(x, y);
        }
        if (lane_id() == 0) red[a][threadIdx.x / kWave] = x;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        const int a = threadIdx.x;
        float x = red[a][0];
        for (uint32_t w = 1; w < blockDim.x / kWave; w++) x = a < 3 ? fminf(x, red[a][w]) : fmaxf(x, red[a][w]);
        out[a] = a == 6 ? __builtin_sqrtf(x) : x;
    }
}

__device__ __forceinline__ uint32_t grid_cell_of(const CandGrid &g, const float *t) {
    return (uint32_t)grid_cell(g, 0, t[0]) +
           (uint32_t)g.n[0] * ((uint32_t)grid_cell(g, 1, t[1]) + (uint32_t)g.n[1] * (uint32_t)grid_cell(g, 2, t[2]));
}

__global__ void __launch_bounds__(kBlock) k_grid_count(uint32_t n, const float *__restrict__ tr,
                                                      CandGrid g, uint32_t *__restrict__ cell_of,
                                                      uint32_t *__restrict__ cnt) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = grid_cell_of(g, tr + 3 * (size_t)i);
    cell_of[i] = c;
    atomicAdd(&cnt[c], 1u);
}

// objects in cell order: {transform.xyz, index bits}, {low.xyz, 0}, {high.xyz, 0}
// (the order inside a cell is arbitrary: nothing downstream depends on it)
__global__ void __launch_bounds__(kBlock) k_grid_fill(uint32_t n, const float *__restrict__ box,
                                                     const float *__restrict__ tr,
                                                     const uint32_t *__restrict__ cell_of,
                                                     const uint64_t *__restrict__ off,
                                                     uint32_t *__restrict__ cursor,
                                                     float4 *__restrict__ gobj) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = cell_of[i];
    const uint64_t p = off[c] + atomicAdd(&cursor[c], 1u);
    const float *t = tr + 3 * (size_t)i, *b = box + 6 * (size_t)i;
    gobj[3 * p] = make_float4(t[0], t[1], t[2], __uint_as_float(i));
    gobj[3 * p + 1] = make_float4(b[0], b[1], b[2], 0.f);
    gobj[3 * p + 2] = make_float4(b[3], b[4], b[5], 0.f);
}

// the k-th smallest key (transform distance, index) of the c keys in LDS, by
// k rounds of wave-wide extraction; keys as 64-bit words (distance bits high:
// the distances are non-negative, so their bits order as the values)
__device__ __forceinline__ uint64_t grid_kth(const float *lm, const uint32_t *lo, uint32_t c, int K) {
    uint64_t prev = 0;
    bool first = true;
    for (int r = 0; r < K; r++) {
        uint64_t best = ~0ull;
        for (uint32_t i = lane_id(); i < c; i += kWave) {
            const uint64_t key = ((uint64_t)__float_as_uint(lm[i]) << 32) | lo[i];
            if ((first || key > prev) && key < best) best = key;
        }
#pragma unroll
        for (int s = 32; s > 0; s >>= 1) {
            const uint64_t o = __shfl_xor(best, s, kWave);
            best = o < best ? o : best;
        }
        prev = best;
        first = false;
    }
    return prev;
}

// One wave per target t of the slice [t0, t_end). Lists go to
// ids[(t - t0) * cap ...], counts and flags to count[t - t0], fallback[t - t0].
__global__ void __launch_bounds__(kBlock) k_knn_candidates_grid(
    uint64_t t_end, uint64_t t0, const float *__restrict__ box1, const float *__restrict__ tr1,
    CandGrid g, const uint64_t *__restrict__ off, const float4 *__restrict__ gobj, int self,
    const int32_t *__restrict__ k_t, uint32_t cap, uint32_t *__restrict__ count,
    uint32_t *__restrict__ ids, uint8_t *__restrict__ fallback) {
    __shared__ float lm_all[kWavesPerBlock][kGridSel];
    __shared__ uint32_t lo_all[kWavesPerBlock][kGridSel];
    __shared__ uint32_t lc_all[kWavesPerBlock];
    const uint32_t w = wave_item();
    const uint64_t t = t0 + w;
    if (t >= t_end) return;
    const int lane = lane_id();
    const int wv = threadIdx.x / kWave;
    float *const lm = lm_all[wv];
    uint32_t *const lo = lo_all[wv];
    uint32_t *const lc = &lc_all[wv];
    float qb[6], qt[3];
#pragma unroll
    for (int i = 0; i < 6; i++) qb[i] = box1[6 * t + i];
#pragma unroll
    for (int i = 0; i < 3; i++) qt[i] = tr1[3 * t + i];
    const int K = k_t[t];
    const uint32_t skip = self ? (uint32_t)t : 0xffffffffu;
    bool ok = K >= 1;
#pragma unroll
    for (int i = 0; i < 6; i++) ok &= __builtin_isfinite(qb[i]);
#pragma unroll
    for (int i = 0; i < 3; i++) ok &= __builtin_isfinite(qt[i]);

    // the objects of the cells [x0, x1] x [y0, y1] x [z0, z1] that pass
    // pred(record, &key) go to LDS (the first kGridSel of them); returns how
    // many passed. Lanes take rows (y, z) of the window; a row's cells hold
    // one contiguous run of objects.
    auto gather = [&](const int *c0, const int *c1, auto pred) -> uint32_t {
        if (lane == 0) *lc = 0u;
        wave_sync();
        const uint32_t ny = (uint32_t)(c1[1] - c0[1] + 1);
        const uint32_t rows = ny * (uint32_t)(c1[2] - c0[2] + 1);
        for (uint32_t row = (uint32_t)lane; row < rows; row += kWave) {
            const uint32_t yy = (uint32_t)c0[1] + row % ny, zz = (uint32_t)c0[2] + row / ny;
            const size_t base = ((size_t)zz * (uint32_t)g.n[1] + yy) * (uint32_t)g.n[0];
            const uint64_t e = off[base + (uint32_t)c1[0] + 1];
            for (uint64_t p = off[base + (uint32_t)c0[0]]; p < e; p++) {
                float key;
                const uint32_t o = __float_as_uint(gobj[3 * p].w);
                if (o != skip && pred(p, key)) {
                    const uint32_t at = atomicAdd(lc, 1u);
                    if (at < kGridSel) {
                        lm[at] = key;
                        lo[at] = o;
                    }
                }
            }
        }
        wave_sync();
        return __builtin_amdgcn_readfirstlane(*lc);
    };
    auto point_key = [&](uint64_t p, float &key) {
        const float4 r = gobj[3 * p];
        const float b[3] = {r.x, r.y, r.z};
        key = cand_point(qt, b);
        return true;
    };

    // 1. a cube around the target's cell with k objects besides the target
    int c[3], a0[3], a1[3];
#pragma unroll
    for (int a = 0; a < 3; a++) c[a] = grid_cell(g, a, qt[a]);
    const uint32_t need = (uint32_t)K + (self ? 1u : 0u);
    uint32_t have = 0;
    for (int r = 0; ok;) {
        bool full = true;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            a0[a] = max(c[a] - r, 0);
            a1[a] = min(c[a] + r, g.n[a] - 1);
            full &= a0[a] == 0 && a1[a] == g.n[a] - 1;
        }
        const uint32_t ny = (uint32_t)(a1[1] - a0[1] + 1);
        const uint32_t rows = ny * (uint32_t)(a1[2] - a0[2] + 1);
        uint32_t s = 0;
        for (uint32_t row = (uint32_t)lane; row < rows; row += kWave) {
            const uint32_t yy = (uint32_t)a0[1] + row % ny, zz = (uint32_t)a0[2] + row / ny;
            const size_t base = ((size_t)zz * (uint32_t)g.n[1] + yy) * (uint32_t)g.n[0];
            s += (uint32_t)(off[base + (uint32_t)a1[0] + 1] - off[base + (uint32_t)a0[0]]);
        }
        have = wave_sum_u32(s);
        if (have >= need || full) break;
        r = r < 4 ? r + 1 : 2 * r;
    }
    ok &= have >= need;
    float U = 0.f;
    if (ok) {
        const uint32_t n1 = gather(a0, a1, point_key);
        ok = n1 <= kGridSel && n1 >= (uint32_t)K;
        if (ok) U = __uint_as_float((uint32_t)(grid_kth(lm, lo, n1, K) >> 32));
    }
    // 2. every object within U: M and the k-th key
    float M = 0.f;
    uint32_t oK = 0;
    if (ok) {
        int w0[3], w1[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float W = U * (1.f + 0x1p-12f) + 0x1p-18f * fabsf(qt[a]) + g.tol;
            w0[a] = grid_cell(g, a, qt[a] - W);
            w1[a] = grid_cell(g, a, qt[a] + W);
        }
        const uint32_t n2 = gather(w0, w1, [&](uint64_t p, float &key) {
            point_key(p, key);
            return key <= U;
        });
        ok = n2 <= kGridSel && n2 >= (uint32_t)K;
        if (ok) {
            const uint64_t kk = grid_kth(lm, lo, n2, K);
            M = __uint_as_float((uint32_t)(kk >> 32));
            oK = (uint32_t)kk;
            // an object past the k-th key at distance M: the octree's order decides
            bool tie = false;
            for (uint32_t i = (uint32_t)lane; i < n2; i += kWave) tie |= lm[i] == M && lo[i] > oK;
            ok = !__any(tie);
        }
        wave_sync();
    }
    // 3. the list: box distance below M, or a key up to the k-th
    uint32_t cn = 0;
    if (ok) {
        int w0[3], w1[3];
        const float W = (M + g.emax) * (1.f + 0x1p-12f);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float m = 0x1p-18f * fmaxf(fabsf(qt[a]), fmaxf(fabsf(qb[a]), fabsf(qb[3 + a]))) + g.tol;
            w0[a] = grid_cell(g, a, fminf(qb[a], qt[a]) - (W + m));
            w1[a] = grid_cell(g, a, fmaxf(qb[3 + a], qt[a]) + (W + m));
        }
        cn = gather(w0, w1, [&](uint64_t p, float &key) {
            point_key(p, key);
            const uint32_t o = __float_as_uint(gobj[3 * p].w);
            const float mn = cand_box_min4(gobj[3 * p + 1], gobj[3 * p + 2], qb);
            return mn < M || key < M || (key == M && o <= oK);
        });
        ok = cn <= cap && cn <= kGridSel;
    }
    // rank by distance; equal distances among the list: the octree's order decides
    bool dup = false;
    if (ok) {
        for (uint32_t i = (uint32_t)lane; i < cn; i += kWave) {
            const float m = lm[i];
            uint32_t rank = 0, eq = 0;
            for (uint32_t q = 0; q < cn; q++) {
                rank += lm[q] < m;
                eq += lm[q] == m;
            }
            dup |= eq > 1;
            ids[(size_t)w * cap + rank] = lo[i];
        }
    }
    dup = __any(dup);
    if (lane == 0) {
        const bool fb = !ok || dup;
        fallback[w] = fb ? 1 : 0;
        count[w] = fb ? 0 : cn;
    }
}

// ----------------------------------------------------- within lists on a grid
//
// For tile 2 of more than kWithinMaxObj objects: the lists of
// k_within_candidates (objects, not the target, whose box distance is at
// most tau, in the order of the octree's depth-first walk, each at the first
// leaf that takes it), from the grid of k_knn_candidates_grid. An object with
// box distance at most tau has its transform point within tau + E of the
// target's box per axis, so the cells of the target's box grown by tau + E
// hold every such object. Each object's leaves (depth-first rank and its
// place in the leaf, in rank order: mem[mem_off[o] ..]) give its place in
// the walk: the first of them whose box passes (a leaf is reached exactly
// when its own box passes). An object in no leaf that passes (none: a box
// without volume overlaps no leaf) is not listed, as the walk never meets it.
// The list is ranked by (leaf rank, place); longer than cap: to the host.
__global__ void __launch_bounds__(kBlock) k_within_candidates_grid(
    uint64_t t_end, uint64_t t0, const float *__restrict__ box1, CandGrid g,
    const uint64_t *__restrict__ off, const float4 *__restrict__ gobj, int self,
    const float *__restrict__ tau_t, const uint32_t *__restrict__ mem_off,
    const uint2 *__restrict__ mem, const float *__restrict__ leaf_box, uint32_t cap,
    uint32_t *__restrict__ count, uint32_t *__restrict__ ids, uint8_t *__restrict__ fallback) {
    __shared__ uint64_t lk_all[kWavesPerBlock][kCandMaxCap];
    __shared__ uint32_t lo_all[kWavesPerBlock][kCandMaxCap];
    __shared__ uint32_t lc_all[kWavesPerBlock];
    const uint32_t w = wave_item();
    const uint64_t t = t0 + w;
    if (t >= t_end) return;
    const int lane = lane_id();
    const int wv = threadIdx.x / kWave;
    uint64_t *const lk = lk_all[wv];
    uint32_t *const lo = lo_all[wv];
    uint32_t *const lc = &lc_all[wv];
    float qb[6];
#pragma unroll
    for (int i = 0; i < 6; i++) qb[i] = box1[6 * t + i];
    const float tau = tau_t[t];
    const uint32_t skip = self ? (uint32_t)t : 0xffffffffu;
    bool ok = tau >= 0.f && __builtin_isfinite(tau);
#pragma unroll
    for (int i = 0; i < 6; i++) ok &= __builtin_isfinite(qb[i]);
    uint32_t cn = 0;
    if (ok) {
        int c0[3], c1[3];
        const float W = (tau + g.emax) * (1.f + 0x1p-12f);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float m = 0x1p-18f * fmaxf(fabsf(qb[a]), fabsf(qb[3 + a])) + g.tol;
            c0[a] = grid_cell(g, a, qb[a] - (W + m));
            c1[a] = grid_cell(g, a, qb[3 + a] + (W + m));
        }
        if (lane == 0) *lc = 0u;
        wave_sync();
        const uint32_t ny = (uint32_t)(c1[1] - c0[1] + 1);
        const uint32_t rows = ny * (uint32_t)(c1[2] - c0[2] + 1);
        for (uint32_t row = (uint32_t)lane; row < rows; row += kWave) {
            const uint32_t yy = (uint32_t)c0[1] + row % ny, zz = (uint32_t)c0[2] + row / ny;
            const size_t base = ((size_t)zz * (uint32_t)g.n[1] + yy) * (uint32_t)g.n[0];
            const uint64_t e = off[base + (uint32_t)c1[0] + 1];
            for (uint64_t p = off[base + (uint32_t)c0[0]]; p < e; p++) {
                const uint32_t o = __float_as_uint(gobj[3 * p].w);
                if (o == skip || !(cand_box_min4(gobj[3 * p + 1], gobj[3 * p + 2], qb) <= tau))
                    continue;
                // its first leaf on the walk
                uint64_t key = ~0ull;
                for (uint32_t k = mem_off[o]; k < mem_off[o + 1]; k++) {
                    const uint2 mk = mem[k];
                    if (cand_box_min(leaf_box + 6 * (size_t)mk.x, qb) <= tau) {
                        key = ((uint64_t)mk.x << 32) | mk.y;
                        break;
                    }
                }
                if (key == ~0ull) continue;
                const uint32_t at = atomicAdd(lc, 1u);
                if (at < kCandMaxCap) {
                    lk[at] = key;
                    lo[at] = o;
                }
            }
        }
        wave_sync();
        cn = __builtin_amdgcn_readfirstlane(*lc);
        ok = cn <= cap && cn <= kCandMaxCap;
    }
    // rank by place on the walk (keys differ: one object per place)
    if (ok) {
        for (uint32_t i = (uint32_t)lane; i < cn; i += kWave) {
            const uint64_t k = lk[i];
            uint32_t rank = 0;
            for (uint32_t q = 0; q < cn; q++) rank += lk[q] < k;
            ids[(size_t)w * cap + rank] = lo[i];
        }
    }
    if (lane == 0) {
        fallback[w] = ok ? 0 : 1;
        count[w] = ok ? cn : 0;
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
