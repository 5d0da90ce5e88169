// Mixed-precision refinement as a pipeline of kernels (included after
// tdp_kernels.hpp). Per live voxel pair the FP64 minima of (d, lo, hi) are
// formed exactly -- see the comment above pair_interval() for why the FP32
// intervals are rigorous -- but the FP64 TriDist runs only on the pairs those
// intervals cannot exclude:
//   k_mx_bounds1  FP32 intervals of every triangle pair: the smallest upper
//                 bounds U and up to three probe pairs (the smallest lower
//                 bound of d, lo and hi);
//   k_mx_eval     FP64 TriDist on the probes, atomically min-folded into the
//                 voxel pair's minima (order-independent, hence exact);
//   k_mx_count    FP32 intervals again against U tightened by the probes:
//                 the number of pairs that can still hold a minimum;
//   (scan)
//   k_mx_write    the same test, writing those pairs to a flat work list;
//   k_mx_eval     FP64 TriDist on the work list, one pair per lane;
//   k_mx_fold     the voxel pair's new range and the object-pair minima.
// Splitting the phases keeps the FP32 passes free of the FP64 TriDist's
// register footprint, and the FP64 work runs as dense, balanced lists however
// uneven the voxel pairs are.
#pragma once

#include "tdp_kernels.hpp"

namespace tdp {

enum : uint32_t { MX_EMPTY = 1u, MX_SMALL = 2u, MX_LOZERO = 4u };
constexpr uint32_t kNoProbe = 0xffffffffu;

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

// q -> (i, j) with q = i * n2 + j, for q < 2^24 and n2 < 2^16 (one
// correction step after the FP32 estimate)
__device__ __forceinline__ void split_q(uint32_t q, uint32_t n2, float rn2, uint32_t &i,
                                        uint32_t &j) {
    int32_t ii = (int32_t)((float)q * rn2);
    int32_t jj = (int32_t)q - ii * (int32_t)n2;
    if (jj < 0) { ii--; jj += (int32_t)n2; }
    else if (jj >= (int32_t)n2) { ii++; jj -= (int32_t)n2; }
    i = (uint32_t)ii;
    j = (uint32_t)jj;
}

// lowest lane holding the wave minimum of v; returns that lane's q
__device__ __forceinline__ uint32_t wave_argmin_q(float v, uint32_t q, float &vmin) {
    vmin = wave_min(v);
    const uint64_t m = __ballot(v == vmin && q != kNoProbe);
    if (m == 0) return kNoProbe;
    const int src = __ffsll((long long)m) - 1;
    return (uint32_t)__shfl((int)q, src, kWave);
}

__global__ void __launch_bounds__(kBlock) k_mx_bounds1(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint64_t *__restrict__ off1, const uint32_t *__restrict__ cnt1,
    const float4 *__restrict__ rec1, const float4 *__restrict__ bnd1,
    const uint64_t *__restrict__ off2, const uint32_t *__restrict__ cnt2,
    const float4 *__restrict__ rec2, const float4 *__restrict__ bnd2, int slot,
    float *__restrict__ vpU, uint32_t *__restrict__ probe, uint32_t *__restrict__ flags,
    unsigned long long *__restrict__ mkey) {
    const size_t w = wave_global();
    if (w >= n_vp) return;
    const int lane = lane_id();
    const VpLod g = vp_lod(vp_v1[w], vp_v2[w], off1, cnt1, off2, cnt2, slot);
    const unsigned long long inf_key = dkey(INFINITY);
    if (lane < 3) {
        mkey[3 * w + lane] = inf_key;
        probe[3 * w + lane] = kNoProbe;
        vpU[3 * w + lane] = INFINITY;
    }
    if (g.n1 == 0 || g.n2 == 0) {
        if (lane == 0) flags[w] = MX_EMPTY;
        return;
    }
    const uint32_t total = g.n1 * g.n2;
    if (total <= kMixedMinPairs) {
        if (lane == 0) flags[w] = MX_SMALL;
        return;
    }
    const float rn2 = 1.f / (float)g.n2;
    float lw[3], uw[3];
    float up[3] = {INFINITY, INFINITY, INFINITY};
    float best[3] = {INFINITY, INFINITY, INFINITY};
    uint32_t bq[3] = {kNoProbe, kNoProbe, kNoProbe};
    for (uint32_t q = lane; q < total; q += kWave) {
        uint32_t i, j;
        split_q(q, g.n2, rn2, i, j);
        pair_interval(rec1, bnd1, g.o1 + i, rec2, bnd2, g.o2 + j, lw, uw);
#pragma unroll
        for (int f = 0; f < 3; f++) {
            up[f] = fminf(up[f], uw[f]);
            if (lw[f] < best[f]) { best[f] = lw[f]; bq[f] = q; }
        }
    }
    float U[3];
    uint32_t pq[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        U[f] = wave_min(up[f]);
        float m;
        pq[f] = wave_argmin_q(best[f], bq[f], m);
    }
    uint32_t fl = 0;
    if (U[1] == 0.f) {  // some pair certainly has lo = +0: the minimum is +0
        fl |= MX_LOZERO;
        pq[1] = kNoProbe;
    }
    if (pq[1] == pq[0]) pq[1] = kNoProbe;
    if (pq[2] == pq[0] || pq[2] == pq[1]) pq[2] = kNoProbe;
    if (lane < 3) {
        vpU[3 * w + lane] = lane == 0 ? U[0] : (lane == 1 ? U[1] : U[2]);
        probe[3 * w + lane] = lane == 0 ? pq[0] : (lane == 1 ? pq[1] : pq[2]);
    }
    if (lane == 0) {
        flags[w] = fl;
        if (fl & MX_LOZERO) mkey[3 * w + 1] = dkey(0.0);
    }
}

// FP64 TriDist of work items (voxel pair w, triangle pair q), min-folded into
// the voxel pair's minima keys; items with q == kNoProbe are skipped
__global__ void __launch_bounds__(kBlock) k_mx_eval(
    uint64_t n_items, const uint32_t *__restrict__ item_vp, const uint32_t *__restrict__ item_q,
    int vp_from_index3, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint64_t *__restrict__ off1, const uint32_t *__restrict__ cnt1,
    const float4 *__restrict__ rec1, const uint64_t *__restrict__ off2,
    const uint32_t *__restrict__ cnt2, const float4 *__restrict__ rec2, int slot,
    unsigned long long *__restrict__ mkey, unsigned long long *__restrict__ stats) {
    // grid-stride: the work list can exceed one launch's thread count
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    uint64_t n = 0;
    for (size_t k0 = (size_t)blockIdx.x * blockDim.x; k0 < n_items; k0 += stride) {
        const size_t k = k0 + threadIdx.x;
        bool act = false;
        if (k < n_items) {
            const uint32_t q = item_q[k];
            const size_t w = vp_from_index3 ? k / 3 : item_vp[k];
            if (q != kNoProbe) {
                act = true;
                const VpLod g = vp_lod(vp_v1[w], vp_v2[w], off1, cnt1, off2, cnt2, slot);
                const uint32_t i = q / g.n2, j = q - i * g.n2;
                double a, b, c;
                pair_values<double>(rec1, g.o1 + i, rec2, g.o2 + j, a, b, c);
                atomicMin(&mkey[3 * w], dkey(a));
                atomicMin(&mkey[3 * w + 1], dkey(b));
                atomicMin(&mkey[3 * w + 2], dkey(c));
            }
        }
        n += (uint64_t)__popcll(__ballot(act));
    }
    if (lane_id() == 0 && n) atomicAdd(&stats[2], (unsigned long long)n);
}

// U tightened by the exact values found so far (rounded up to FP32)
__device__ __forceinline__ void tightened_U(const float *vpU, const unsigned long long *mkey,
                                            uint32_t fl, size_t w, float U[3]) {
#pragma unroll
    for (int f = 0; f < 3; f++) U[f] = fminf(vpU[3 * w + f], ulp_up((float)dunkey(mkey[3 * w + f])));
    if (fl & MX_LOZERO) U[1] = -1.f;
}

__device__ __forceinline__ bool is_probe(uint32_t q, const uint32_t *probe, size_t w) {
    return q == probe[3 * w] || q == probe[3 * w + 1] || q == probe[3 * w + 2];
}

__global__ void __launch_bounds__(kBlock) k_mx_count(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint64_t *__restrict__ off1, const uint32_t *__restrict__ cnt1,
    const float4 *__restrict__ rec1, const float4 *__restrict__ bnd1,
    const uint64_t *__restrict__ off2, const uint32_t *__restrict__ cnt2,
    const float4 *__restrict__ rec2, const float4 *__restrict__ bnd2, int slot,
    const float *__restrict__ vpU, const uint32_t *__restrict__ probe,
    const uint32_t *__restrict__ flags, const unsigned long long *__restrict__ mkey,
    uint32_t *__restrict__ count, unsigned long long *__restrict__ stats) {
    const size_t w = wave_global();
    if (w >= n_vp) return;
    const int lane = lane_id();
    const uint32_t fl = flags[w];
    if (fl & MX_EMPTY) {
        if (lane == 0) count[w] = 0;
        return;
    }
    const VpLod g = vp_lod(vp_v1[w], vp_v2[w], off1, cnt1, off2, cnt2, slot);
    const uint32_t total = g.n1 * g.n2;
    if (fl & MX_SMALL) {
        if (lane == 0) count[w] = total;
        return;
    }
    float U[3];
    tightened_U(vpU, mkey, fl, w, U);
    const float rn2 = 1.f / (float)g.n2;
    float lw[3], uw[3];
    uint32_t c = 0;
    for (uint32_t q = lane; q < total; q += kWave) {
        uint32_t i, j;
        split_q(q, g.n2, rn2, i, j);
        pair_interval(rec1, bnd1, g.o1 + i, rec2, bnd2, g.o2 + j, lw, uw);
        if (reaches(lw, U) && !is_probe(q, probe, w)) c++;
    }
    c = (uint32_t)wave_sum_u64(c);
    if (lane == 0) {
        count[w] = c;
        uint32_t np = 0;
        for (int f = 0; f < 3; f++) np += probe[3 * w + f] != kNoProbe;
        atomicAdd(&stats[3], (unsigned long long)(total - c - np));
    }
}

__global__ void __launch_bounds__(kBlock) k_mx_write(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint64_t *__restrict__ off1, const uint32_t *__restrict__ cnt1,
    const float4 *__restrict__ rec1, const float4 *__restrict__ bnd1,
    const uint64_t *__restrict__ off2, const uint32_t *__restrict__ cnt2,
    const float4 *__restrict__ rec2, const float4 *__restrict__ bnd2, int slot,
    const float *__restrict__ vpU, const uint32_t *__restrict__ probe,
    const uint32_t *__restrict__ flags, const unsigned long long *__restrict__ mkey,
    const uint64_t *__restrict__ offs, uint32_t *__restrict__ item_vp,
    uint32_t *__restrict__ item_q) {
    const size_t w = wave_global();
    if (w >= n_vp) return;
    const int lane = lane_id();
    const uint32_t fl = flags[w];
    if (fl & MX_EMPTY) return;
    const VpLod g = vp_lod(vp_v1[w], vp_v2[w], off1, cnt1, off2, cnt2, slot);
    const uint32_t total = g.n1 * g.n2;
    uint64_t pos = offs[w];
    if (fl & MX_SMALL) {
        for (uint32_t q = lane; q < total; q += kWave) {
            item_vp[pos + q] = (uint32_t)w;
            item_q[pos + q] = q;
        }
        return;
    }
    float U[3];
    tightened_U(vpU, mkey, fl, w, U);
    const float rn2 = 1.f / (float)g.n2;
    float lw[3], uw[3];
    for (uint32_t base = 0; base < total; base += kWave) {
        const uint32_t q = base + lane;
        bool cand = false;
        if (q < total) {
            uint32_t i, j;
            split_q(q, g.n2, rn2, i, j);
            pair_interval(rec1, bnd1, g.o1 + i, rec2, bnd2, g.o2 + j, lw, uw);
            cand = reaches(lw, U) && !is_probe(q, probe, w);
        }
        const uint64_t mask = __ballot(cand);
        if (cand) {
            const uint64_t at = pos + (uint64_t)__popcll(mask & ((1ull << lane) - 1ull));
            item_vp[at] = (uint32_t)w;
            item_q[at] = q;
        }
        pos += (uint64_t)__popcll(mask);
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
