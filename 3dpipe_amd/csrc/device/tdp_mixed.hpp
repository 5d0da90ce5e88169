// Mixed-precision refinement as a pipeline of kernels (included after
// tdp_kernels.hpp). Per live voxel pair the FP64 minima of (d, lo, hi) are
// formed exactly -- see the comment above pair_interval() for why the FP32
// intervals are rigorous -- but the FP64 TriDist runs only on the pairs those
// intervals cannot exclude:
//   k_mx_bounds1  FP32 sphere intervals of every triangle pair: the smallest
//                 upper bounds U and up to three probe pairs (the pairs most
//                 likely to hold the minima of d, lo and hi);
//   k_mx_eval     FP64 TriDist on the probes, atomically min-folded into the
//                 voxel pair's minima (order-independent, hence exact);
//   k_mx_mark     FP32 intervals again against U tightened by the probes:
//                 a bit mask of the pairs that can still hold a minimum,
//                 and their number;
//   (scan)
//   k_mx_expand   the bit masks to a flat work list;
//   k_mx_eval     FP64 TriDist on the work list, one pair per lane;
//   k_mx_fold     the voxel pair's new range and the object-pair minima.
// Splitting the phases keeps the FP32 passes free of the FP64 TriDist's
// register footprint, and the FP64 work runs as dense, balanced lists however
// uneven the voxel pairs are.
//
// FP32 passes: one wave per voxel pair. The lanes hold the triangles of the
// larger voxel (64 at a time, bound records in registers); the wave walks the
// smaller voxel's triangles in a uniform loop, so their records come through
// scalar loads and every pair costs arithmetic only. A pair is first tested
// with its sphere interval (centre distance +- radii, about 30 flops); only
// pairs that pass go on to the vertex-level projection interval.
#pragma once

#include "tdp_kernels.hpp"

namespace tdp {

enum : uint32_t { MX_EMPTY = 1u, MX_SMALL = 2u, MX_LOZERO = 4u };
constexpr uint32_t kNoProbe = 0xffffffffu;
// voxel pairs of at most this many triangle pairs go straight to FP64
constexpr uint32_t kMxSmallPairs = 8;

struct VpLod {
    uint64_t o1, o2;
    uint32_t n1, n2;
};

__device__ __forceinline__ VpLod vp_lod(uint32_t v1, uint32_t v2, const uint64_t *off1,
                                        const uint32_t *cnt1, const uint64_t *off2,
                                        const uint32_t *cnt2, int slot) {
    VpLod g{0, 0, 0, 0};
    if (slot >= 0) {
        g.n1 = cnt1[5 * (size_t)v1 + slot];
        g.n2 = cnt2[5 * (size_t)v2 + slot];
        g.o1 = off1[5 * (size_t)v1 + slot];
        g.o2 = off2[5 * (size_t)v2 + slot];
    }
    return g;
}

// the wave's voxel pair index, in a scalar register
__device__ __forceinline__ uint32_t wave_item() {
    return __builtin_amdgcn_readfirstlane((uint32_t)wave_global());
}

// ---------------------------------------------------------------- intervals

// sphere-level data of one triangle (first and third bound records)
struct TriS {
    float cx, cy, cz, r, nx, ny, nz, cm, ph, h, dg;  // dg = 1: no plane
};

__device__ __forceinline__ TriS load_s(const float4 *__restrict__ bnd, uint64_t t) {
    const float4 a = bnd[kBnd * t], b = bnd[kBnd * t + 1], c = bnd[kBnd * t + 2];
    return TriS{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z};
}

// vertices of one triangle
struct TriV {
    float x0, y0, z0, x1, y1, z1, x2, y2, z2;
};

__device__ __forceinline__ TriV load_v(const float4 *__restrict__ rec, uint64_t t) {
    const float4 a = rec[3 * t], b = rec[3 * t + 1], c = rec[3 * t + 2];
    return TriV{a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
}

// Disc interval of the FP64 TriDist value of (A, B), A from tile 1, from
// per-triangle data only. Each triangle lies in the disc of its plane around
// its centre with its sphere's radius, so:
//  upper: centre distance + radii;
//  lower: the gap between the discs' projections on the centre axis u,
//         dc - rA sin(u, nA) - rB sin(u, nB), used only when one disc lies
//         on one side of the other's plane by more than twice the margin:
//         B's vertices are within rB sin(nA, nB) of the plane parallel to A's
//         through B's centre. Then every vertex passes the vertex-level plane
//         test of pair_interval(), which is what guarantees that TriDist's
//         face test shows the triangles disjoint.
// Sines are rounded up (16 ulp inside the root for the cosine's error); a
// triangle without a plane counts as a ball (sine 1).
// Operation count on the separated path, as used by k_mx_mark (lower bounds
// and the three comparisons; add/sub/mul/sqrt/rcp/min/max/compare = 1, abs
// free): centre difference 3, dc 6, rs 1, m 5, e1 and e2 10, cross product 9,
// st 10, guard 7, disc gap 21, du 2, s 5, lower bounds of lo and hi 7,
// comparisons with U 3 -- 89 per triangle pair.
__device__ __forceinline__ void disc_interval(const TriS &A, const TriS &B, float lw[3],
                                              float uw[3]) {
    const float dx = B.cx - A.cx, dy = B.cy - A.cy, dz = B.cz - A.cz;
    const float dc = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
    const float rs = A.r + B.r;
    const float m = 32.f * kU * (dc + rs + A.cm + B.cm) + 1e-6f;
    const float e1 = A.nx * dx + A.ny * dy + A.nz * dz;
    const float e2 = B.nx * dx + B.ny * dy + B.nz * dz;
    const float xx = A.ny * B.nz - A.nz * B.ny, xy = A.nz * B.nx - A.nx * B.nz,
                xz = A.nx * B.ny - A.ny * B.nx;
    const float st = fminf(__builtin_amdgcn_sqrtf(xx * xx + xy * xy + xz * xz) + 8.f * kU + A.dg + B.dg,
                           1.f);
    const bool g = fabsf(e1) > B.r * st + 2.f * m || fabsf(e2) > A.r * st + 2.f * m;
    float dl = 0.f;
    if (g) {  // dc > 0 here
        const float inv = __builtin_amdgcn_rcpf(dc);
        const float c1 = e1 * inv, c2 = e2 * inv;
        const float sa = __builtin_amdgcn_sqrtf(fmaxf(1.f - c1 * c1, 0.f) + 16.f * kU);
        const float sb = __builtin_amdgcn_sqrtf(fmaxf(1.f - c2 * c2, 0.f) + 16.f * kU);
        dl = fmaxf(dc - A.r * fminf(sa, 1.f) - B.r * fminf(sb, 1.f) - m, 0.f);
    }
    intervals(dl, dc + rs + m, A.ph, A.h, B.ph, B.h, lw, uw);
}

// Vertex-level interval (the arithmetic of pair_interval on register data).
__device__ __forceinline__ void vertex_interval(const TriS &A, const TriV &a, const TriS &B,
                                                const TriV &c, float lw[3], float uw[3]) {
    const float dx = B.cx - A.cx, dy = B.cy - A.cy, dz = B.cz - A.cz;
    const float dc = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
    const float rs = A.r + B.r;
    const float m = 32.f * kU * (dc + rs + A.cm + B.cm) + 1e-6f;
    const float du = dc + rs + m;
    float dl = 0.f;
    const float ax0 = a.x0 - A.cx, ay0 = a.y0 - A.cy, az0 = a.z0 - A.cz;
    const float ax1 = a.x1 - A.cx, ay1 = a.y1 - A.cy, az1 = a.z1 - A.cz;
    const float ax2 = a.x2 - A.cx, ay2 = a.y2 - A.cy, az2 = a.z2 - A.cz;
    const float cx0 = c.x0 - A.cx, cy0 = c.y0 - A.cy, cz0 = c.z0 - A.cz;
    const float cx1 = c.x1 - A.cx, cy1 = c.y1 - A.cy, cz1 = c.z1 - A.cz;
    const float cx2 = c.x2 - A.cx, cy2 = c.y2 - A.cy, cz2 = c.z2 - A.cz;
    const float t0 = A.nx * cx0 + A.ny * cy0 + A.nz * cz0;
    const float t1 = A.nx * cx1 + A.ny * cy1 + A.nz * cz1;
    const float t2 = A.nx * cx2 + A.ny * cy2 + A.nz * cz2;
    const float w0 = B.nx * (ax0 - dx) + B.ny * (ay0 - dy) + B.nz * (az0 - dz);
    const float w1 = B.nx * (ax1 - dx) + B.ny * (ay1 - dy) + B.nz * (az1 - dz);
    const float w2 = B.nx * (ax2 - dx) + B.ny * (ay2 - dy) + B.nz * (az2 - dz);
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
    intervals(dl, du, A.ph, A.h, B.ph, B.h, lw, uw);
}

// ------------------------------------------------------------ pair walking
//
// Calls f(A, B, p, va, vc) for every triangle pair of the voxel pair: A/B
// the tile-1/tile-2 sphere data, p the pair's place, va()/vc() the vertices
// of A/B. The lanes hold the larger side, 64 triangles at a time (sphere data
// and vertices in registers); the smaller side is walked in a wave-uniform
// loop, so its records come through scalar loads. SWAP: the lanes hold
// tile-1 triangles. Idle lanes (p.valid = false) take part in every call, so
// f may use wave-wide operations. p.word numbers the (lane chunk, uniform
// step) iterations 0, 1, ... in order; the pair is bit (lane) of that word.

struct PairAt {
    uint32_t i, j;     // triangle indices within the two voxels
    uint32_t word;     // iteration number: chunk * n_small + step
    bool valid;
};

__device__ __forceinline__ bool mx_swap(const VpLod &g) { return g.n1 > g.n2; }
__device__ __forceinline__ uint32_t mx_words(const VpLod &g) {
    const uint32_t nl = mx_swap(g) ? g.n1 : g.n2, ns = mx_swap(g) ? g.n2 : g.n1;
    return ns * ((nl + kWave - 1) / kWave);
}

template <bool SWAP, typename F>
__device__ __forceinline__ void walk_pairs(const VpLod &g, int lane, const float4 *__restrict__ rec1,
                                           const float4 *__restrict__ bnd1,
                                           const float4 *__restrict__ rec2,
                                           const float4 *__restrict__ bnd2, F &&f) {
    const uint32_t nl = SWAP ? g.n1 : g.n2, ns = SWAP ? g.n2 : g.n1;
    const uint64_t ol = SWAP ? g.o1 : g.o2, os = SWAP ? g.o2 : g.o1;
    const float4 *__restrict__ bl = SWAP ? bnd1 : bnd2;
    const float4 *__restrict__ bs = SWAP ? bnd2 : bnd1;
    const float4 *__restrict__ rl = SWAP ? rec1 : rec2;
    const float4 *__restrict__ rs = SWAP ? rec2 : rec1;
    uint32_t word = 0;
    for (uint32_t lb = 0; lb < nl; lb += kWave) {
        const uint32_t li = lb + lane;
        const bool valid = li < nl;
        const uint32_t lc = valid ? li : nl - 1;  // in bounds for idle lanes
        const TriS L = load_s(bl, ol + lc);
        const TriV LV = load_v(rl, ol + lc);
        for (uint32_t si = 0; si < ns; si++, word++) {
            const TriS S = load_s(bs, os + si);
            auto lane_v = [&]() { return LV; };
            auto step_v = [&]() { return load_v(rs, os + si); };
            if (SWAP)
                f(L, S, PairAt{lc, si, word, valid}, lane_v, step_v);
            else
                f(S, L, PairAt{si, lc, word, valid}, step_v, lane_v);
        }
    }
}

template <typename F>
__device__ __forceinline__ void for_pairs(const VpLod &g, int lane, const float4 *__restrict__ rec1,
                                          const float4 *__restrict__ bnd1,
                                          const float4 *__restrict__ rec2,
                                          const float4 *__restrict__ bnd2, F &&f) {
    if (mx_swap(g))
        walk_pairs<true>(g, lane, rec1, bnd1, rec2, bnd2, f);
    else
        walk_pairs<false>(g, lane, rec1, bnd1, rec2, bnd2, f);
}

// lowest q among the lanes holding the wave minimum of v (kNoProbe if none)
__device__ __forceinline__ uint32_t wave_argmin_q(float v, uint32_t q) {
    const float vmin = wave_min(v);
    uint32_t c = (v == vmin) ? q : kNoProbe;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c = min(c, (uint32_t)__shfl_xor((int)c, o, kWave));
    return c;
}

// ------------------------------------------------------------------ kernels
//
// `need` selects the quantities the round folds (bit 0: d, bit 1: lo, bit 2:
// hi); the others get no probes and no candidates, and their minima are left
// unformed.

__global__ void __launch_bounds__(kBlock) k_mx_bounds1(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint64_t *__restrict__ off1, const uint32_t *__restrict__ cnt1,
    const float4 *__restrict__ rec1, const float4 *__restrict__ bnd1,
    const uint64_t *__restrict__ off2, const uint32_t *__restrict__ cnt2,
    const float4 *__restrict__ rec2, const float4 *__restrict__ bnd2, int slot, uint32_t need,
    float *__restrict__ vpU, uint32_t *__restrict__ probe, uint32_t *__restrict__ flags,
    unsigned long long *__restrict__ mkey, uint32_t *__restrict__ nwords) {
    const uint32_t w = wave_item();
    if (w >= n_vp) return;
    const int lane = lane_id();
    const VpLod g = vp_lod(vp_v1[w], vp_v2[w], off1, cnt1, off2, cnt2, slot);
    const unsigned long long inf_key = dkey(INFINITY);
    if (lane < 3) {
        mkey[3 * (size_t)w + lane] = inf_key;
        probe[3 * (size_t)w + lane] = kNoProbe;
        vpU[3 * (size_t)w + lane] = -INFINITY;
    }
    if (g.n1 == 0 || g.n2 == 0) {
        if (lane == 0) {
            flags[w] = MX_EMPTY;
            nwords[w] = 0;
        }
        return;
    }
    const uint32_t n2 = g.n2;
    if (g.n1 * n2 <= kMxSmallPairs) {
        if (lane == 0) {
            flags[w] = MX_SMALL;
            nwords[w] = 0;
        }
        return;
    }
    float up[3] = {INFINITY, INFINITY, INFINITY};
    float best[3] = {INFINITY, INFINITY, INFINITY};
    uint32_t bq[3] = {kNoProbe, kNoProbe, kNoProbe};
    for_pairs(g, lane, rec1, bnd1, rec2, bnd2,
              [&](const TriS &A, const TriS &B, const PairAt &p, auto &&, auto &&) {
                  if (!p.valid) return;
                  float lw[3], uw[3];
                  disc_interval(A, B, lw, uw);
                  // probe score: centre distance less the radii (not a
                  // bound, only a guess at the closest pair)
                  const float dx = B.cx - A.cx, dy = B.cy - A.cy, dz = B.cz - A.cz;
                  const float sc = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz) - A.r - B.r;
                  const float key[3] = {sc, sc - A.ph - B.ph, sc + A.h + B.h};
                  const uint32_t q = p.i * n2 + p.j;
#pragma unroll
                  for (int f = 0; f < 3; f++) {
                      up[f] = fminf(up[f], uw[f]);
                      if (key[f] < best[f]) {
                          best[f] = key[f];
                          bq[f] = q;
                      }
                  }
              });
    float U[3];
    uint32_t pq[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        U[f] = wave_min(up[f]);
        pq[f] = wave_argmin_q(best[f], bq[f]);
        if (!((need >> f) & 1u)) {
            U[f] = -INFINITY;
            pq[f] = kNoProbe;
        }
    }
    uint32_t fl = 0;
    if (U[1] == 0.f) {  // some pair certainly has lo = +0: the minimum is +0
        fl |= MX_LOZERO;
        pq[1] = kNoProbe;
    }
    if (pq[1] == pq[0]) pq[1] = kNoProbe;
    if (pq[2] == pq[0] || pq[2] == pq[1]) pq[2] = kNoProbe;
    if (lane < 3) {
        vpU[3 * (size_t)w + lane] = lane == 0 ? U[0] : (lane == 1 ? U[1] : U[2]);
        probe[3 * (size_t)w + lane] = lane == 0 ? pq[0] : (lane == 1 ? pq[1] : pq[2]);
    }
    if (lane == 0) {
        flags[w] = fl;
        nwords[w] = mx_words(g);
        if (fl & MX_LOZERO) mkey[3 * (size_t)w + 1] = dkey(0.0);
    }
}

// Minima over the lanes of one voxel pair w (lanes are sorted by w; kNoProbe
// marks an empty lane), folded with one atomic per voxel pair and quantity.
__device__ __forceinline__ void seg_min_fold(uint32_t w, unsigned long long a,
                                             unsigned long long b, unsigned long long c,
                                             unsigned long long *__restrict__ mkey) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t ow = (uint32_t)__shfl_down((int)w, o, kWave);
        const unsigned long long oa = __shfl_down(a, o, kWave);
        const unsigned long long ob = __shfl_down(b, o, kWave);
        const unsigned long long oc = __shfl_down(c, o, kWave);
        if (lane + o < kWave && ow == w) {
            a = min(a, oa);
            b = min(b, ob);
            c = min(c, oc);
        }
    }
    const uint32_t pw = (uint32_t)__shfl_up((int)w, 1, kWave);
    const bool head = lane == 0 || pw != w;
    if (head && w != kNoProbe) {
        const unsigned long long inf_key = dkey(INFINITY);
        if (a < inf_key) atomicMin(&mkey[3 * (size_t)w], a);
        if (b < inf_key) atomicMin(&mkey[3 * (size_t)w + 1], b);
        if (c < inf_key) atomicMin(&mkey[3 * (size_t)w + 2], c);
    }
}

// FP64 TriDist of work items (voxel pair w, triangle pair q), min-folded into
// the voxel pair's minima keys; items with q == kNoProbe are skipped. Items
// arrive sorted by voxel pair (probe lists: w = k / 3; work lists: written
// in voxel-pair order).
__global__ void __launch_bounds__(kBlock) k_mx_eval(
    uint64_t n_items, const uint32_t *__restrict__ item_vp, const uint32_t *__restrict__ item_q,
    int vp_from_index3, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint64_t *__restrict__ off1, const uint32_t *__restrict__ cnt1,
    const float4 *__restrict__ rec1, const uint64_t *__restrict__ off2,
    const uint32_t *__restrict__ cnt2, const float4 *__restrict__ rec2, int slot,
    unsigned long long *__restrict__ mkey, unsigned long long *__restrict__ stats) {
    // grid-stride: the work list can exceed one launch's thread count
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const unsigned long long inf_key = dkey(INFINITY);
    uint64_t n = 0;
    for (size_t k0 = (size_t)blockIdx.x * blockDim.x; k0 < n_items; k0 += stride) {
        const size_t k = k0 + threadIdx.x;
        uint32_t w = kNoProbe;
        unsigned long long a = inf_key, b = inf_key, c = inf_key;
        bool act = false;
        if (k < n_items) {
            const uint32_t q = item_q[k];
            w = vp_from_index3 ? (uint32_t)(k / 3) : item_vp[k];
            if (q != kNoProbe) {
                act = true;
                const VpLod g = vp_lod(vp_v1[w], vp_v2[w], off1, cnt1, off2, cnt2, slot);
                const uint32_t i = q / g.n2, j = q - i * g.n2;
                double d, lo, hi;
                pair_values<double>(rec1, g.o1 + i, rec2, g.o2 + j, d, lo, hi);
                a = dkey(d);
                b = dkey(lo);
                c = dkey(hi);
            }
        }
        seg_min_fold(w, a, b, c, mkey);
        n += (uint64_t)__popcll(__ballot(act));
    }
    if (lane_id() == 0 && n) atomicAdd(&stats[2], (unsigned long long)n);
}

// U tightened by the exact values found so far (rounded up to FP32)
__device__ __forceinline__ void tightened_U(const float *vpU, const unsigned long long *mkey,
                                            uint32_t fl, size_t w, float U[3]) {
#pragma unroll
    for (int f = 0; f < 3; f++) U[f] = fminf(vpU[3 * w + f], ulp_up((float)dunkey(mkey[3 * w + f])));
    if (fl & MX_LOZERO) U[1] = -INFINITY;
}

// Candidate test: the sphere interval first, the vertex interval for pairs
// that pass it; probes (already evaluated) are never candidates.
template <typename VA, typename VC>
__device__ __forceinline__ bool mx_candidate(const TriS &A, const TriS &B, const PairAt &p,
                                             VA &&va, VC &&vc, const float U[3], uint32_t q,
                                             uint32_t p0, uint32_t p1, uint32_t p2) {
    if (!p.valid) return false;
    float lw[3], uw[3];
    disc_interval(A, B, lw, uw);
    if (!reaches(lw, U)) return false;
    if (q == p0 || q == p1 || q == p2) return false;
    vertex_interval(A, va(), B, vc(), lw, uw);
    return reaches(lw, U);
}

// Candidate bit masks: word k of voxel pair w (at mask[woff[w] + k]) holds
// the ballot of iteration k of the pair walk; count[w] = number of bits.
// SMALL voxel pairs: every pair, no mask.
__global__ void __launch_bounds__(kBlock) k_mx_mark(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint64_t *__restrict__ off1, const uint32_t *__restrict__ cnt1,
    const float4 *__restrict__ rec1, const float4 *__restrict__ bnd1,
    const uint64_t *__restrict__ off2, const uint32_t *__restrict__ cnt2,
    const float4 *__restrict__ rec2, const float4 *__restrict__ bnd2, int slot,
    const float *__restrict__ vpU, const uint32_t *__restrict__ probe,
    const uint32_t *__restrict__ flags, const unsigned long long *__restrict__ mkey,
    const uint64_t *__restrict__ woff, unsigned long long *__restrict__ mask,
    uint32_t *__restrict__ count) {
    const uint32_t w = wave_item();
    if (w >= n_vp) return;
    const int lane = lane_id();
    const uint32_t fl = flags[w];
    if (fl & MX_EMPTY) {
        if (lane == 0) count[w] = 0;
        return;
    }
    const VpLod g = vp_lod(vp_v1[w], vp_v2[w], off1, cnt1, off2, cnt2, slot);
    if (fl & MX_SMALL) {
        if (lane == 0) count[w] = g.n1 * g.n2;
        return;
    }
    float U[3];
    tightened_U(vpU, mkey, fl, w, U);
    const uint32_t p0 = probe[3 * (size_t)w], p1 = probe[3 * (size_t)w + 1],
                   p2 = probe[3 * (size_t)w + 2];
    const uint32_t n2 = g.n2, nw = mx_words(g);
    unsigned long long *__restrict__ out = mask + woff[w];
    uint32_t c = 0;
    unsigned long long mine = 0;  // lane k keeps word (k mod 64) of the current group
    for_pairs(g, lane, rec1, bnd1, rec2, bnd2,
              [&](const TriS &A, const TriS &B, const PairAt &p, auto &&va, auto &&vc) {
                  const bool take = mx_candidate(A, B, p, va, vc, U, p.i * n2 + p.j, p0, p1, p2);
                  const unsigned long long m = __ballot(take);
                  c += (uint32_t)__popcll(m);
                  const uint32_t k = p.word & (kWave - 1);
                  if ((uint32_t)lane == k) mine = m;
                  if (k == kWave - 1 || p.word + 1 == nw) {
                      if ((uint32_t)lane <= k) out[(p.word & ~(uint32_t)(kWave - 1)) + lane] = mine;
                  }
              });
    if (lane == 0) count[w] = c;
}

// Bit masks (and SMALL voxel pairs) to the flat work list at offs[w].
__global__ void __launch_bounds__(kBlock) k_mx_expand(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint64_t *__restrict__ off1, const uint32_t *__restrict__ cnt1,
    const uint64_t *__restrict__ off2, const uint32_t *__restrict__ cnt2, int slot,
    const uint32_t *__restrict__ flags, const uint64_t *__restrict__ woff,
    const unsigned long long *__restrict__ mask, const uint64_t *__restrict__ offs,
    uint32_t *__restrict__ item_vp, uint32_t *__restrict__ item_q) {
    const uint32_t w = wave_item();
    if (w >= n_vp) return;
    const int lane = lane_id();
    const uint32_t fl = flags[w];
    if (fl & MX_EMPTY) return;
    const VpLod g = vp_lod(vp_v1[w], vp_v2[w], off1, cnt1, off2, cnt2, slot);
    uint64_t pos = offs[w];
    if (fl & MX_SMALL) {
        const uint32_t total = g.n1 * g.n2;
        for (uint32_t q = lane; q < total; q += kWave) {
            item_vp[pos + q] = w;
            item_q[pos + q] = q;
        }
        return;
    }
    const bool swap = mx_swap(g);
    const uint32_t ns = swap ? g.n2 : g.n1, n2 = g.n2;
    const uint32_t nw = mx_words(g);
    const unsigned long long *__restrict__ in = mask + woff[w];
    for (uint32_t wb = 0; wb < nw; wb += kWave) {
        const uint32_t k = wb + lane;
        unsigned long long m = k < nw ? in[k] : 0ull;
        const uint32_t cnt = (uint32_t)__popcll(m);
        // exclusive prefix of cnt over the lanes
        uint32_t incl = cnt;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t v = (uint32_t)__shfl_up((int)incl, o, kWave);
            if (lane >= o) incl += v;
        }
        const uint32_t total = (uint32_t)__shfl((int)incl, kWave - 1, kWave);
        uint64_t at = pos + (incl - cnt);
        const uint32_t chunk = k / (ns ? ns : 1), si = k - chunk * ns;
        while (m) {
            const uint32_t b = (uint32_t)__ffsll((long long)m) - 1u;
            m &= m - 1ull;
            const uint32_t li = chunk * kWave + b;
            const uint32_t i = swap ? li : si, j = swap ? si : li;
            item_vp[at] = w;
            item_q[at] = i * n2 + j;
            at++;
        }
        pos += total;
    }
}

// fold the exact minima into the voxel pair's range (same rules as k_refine)
template <typename R>
__global__ void __launch_bounds__(kBlock) k_mx_fold(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint32_t *__restrict__ vp_pair, R *__restrict__ vp_min, R *__restrict__ vp_max,
    const uint32_t *__restrict__ cnt1, const uint32_t *__restrict__ cnt2,
    const uint32_t *__restrict__ vobj1, const uint32_t *__restrict__ vobj2,
    const float *__restrict__ ohd1, const float *__restrict__ ohd2, int slot, int is_last,
    int level, const uint32_t *__restrict__ flags, const unsigned long long *__restrict__ mkey,
    typename Key<R>::type *__restrict__ acc_min, typename Key<R>::type *__restrict__ acc_max,
    unsigned long long *__restrict__ stats) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool counted = false;
    uint64_t pairs = 0;
    if (w < n_vp) {
        const uint32_t v1 = vp_v1[w], v2 = vp_v2[w], p = vp_pair[w];
        const R omin = vp_min[w], omax = vp_max[w];
        if (flags[w] & MX_EMPTY) {
            atomicMin(&acc_min[p], Key<R>::enc(omin));
        } else {
            const R d = (R)dunkey(mkey[3 * w]), lo = (R)dunkey(mkey[3 * w + 1]),
                    hi = (R)dunkey(mkey[3 * w + 2]);
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
            counted = true;
            pairs = (uint64_t)cnt1[5 * (size_t)v1 + slot] * cnt2[5 * (size_t)v2 + slot];
        }
    }
    const uint64_t np = wave_sum_u64(pairs);
    const uint64_t nv = (uint64_t)__popcll(__ballot(counted));
    if (lane_id() == 0 && nv) {
        atomicAdd(&stats[0], (unsigned long long)np);
        atomicAdd(&stats[1], (unsigned long long)nv);
    }
}

}  // namespace tdp
