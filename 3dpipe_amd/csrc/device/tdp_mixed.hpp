// Mixed-precision refinement as a pipeline of kernels (included after
// tdp_kernels.hpp). Per live voxel pair the FP64 minima of (d, lo, hi) are
// formed exactly -- see the comment above pair_interval() for why the FP32
// intervals are rigorous -- but the FP64 TriDist runs only on the pairs those
// intervals cannot exclude:
//   k_mx_bounds   FP32 upper bounds U of every triangle pair's quantities and
//                 up to three probe pairs (the pairs most likely to hold the
//                 minima of d, lo and hi);
//   k_mx_eval     FP64 TriDist on the probes, atomically min-folded into the
//                 voxel pair's minima (order-independent, hence exact); their
//                 values tighten U;
//   k_mx_mark     FP32 intervals of every triangle pair against U: a bit mask
//                 of the pairs that can still hold a minimum, and their number;
//   (scan)
//   k_mx_expand   the bit masks to a flat work list;
//   k_mx_eval     FP64 TriDist on the work list, one pair per lane;
//   k_mx_fold     the voxel pair's new range and the object-pair minima.
// Splitting the phases keeps the FP32 passes free of the FP64 TriDist's
// register footprint, and the FP64 work runs as dense, balanced lists however
// uneven the voxel pairs are.
//
// FP32 passes: one wave per voxel pair, walking its n1 x n2 grid in tiles of
// R rows x cw columns (R * cw <= 64 lanes, chosen per voxel pair for the
// fewest iterations). The larger voxel's triangles are the columns: a lane
// keeps its column's records in registers for the whole walk. The rows'
// records are staged in LDS a block at a time and each lane reads its row
// from there. A pair is first tested with its disc bound (per-triangle data
// only); only pairs that pass go on to the vertex-level bound.
#pragma once

#include <type_traits>

#include "tdp_kernels.hpp"

namespace tdp {

enum : uint32_t { MX_EMPTY = 1u, MX_SMALL = 2u };
constexpr uint32_t kNoProbe = 0xffffffffu;
// voxel pairs of at most this many triangle pairs go straight to FP64 (default)
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
    float cx, cy, cz, r, nx, ny, nz, cm, ph, h;
    float alo, ahi;  // upper-bound constants of lo and hi (k_tri_bounds)
};

__device__ __forceinline__ TriS load_s_at(const float4 *__restrict__ p) {
    const float4 a = p[0], b = p[1], c = p[2];
    return TriS{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
}
__device__ __forceinline__ TriS load_s(const float4 *__restrict__ bnd, uint64_t t) {
    return load_s_at(bnd + kBnd * t);
}

// vertices of one triangle
struct TriV {
    float x0, y0, z0, x1, y1, z1, x2, y2, z2;
};

__device__ __forceinline__ TriV load_v_at(const float4 *__restrict__ p) {
    const float4 a = p[0], b = p[1], c = p[2];
    return TriV{a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
}
__device__ __forceinline__ TriV load_v(const float4 *__restrict__ rec, uint64_t t) {
    return load_v_at(rec + 3 * t);
}

// The disc bound of a pair's FP64 TriDist value, from per-triangle data
// only. Each triangle lies in the disc of its plane around its centre with
// its sphere's radius, so the gap between the discs' projections on the
// centre axis u, dc - rA sin(u, nA) - rB sin(u, nB), bounds the distance of
// the triangles from below. It is used only when one triangle lies on one
// side of the other's plane by more than twice the margin m: B lies within
// rB of its centre, so |nA . (cB - cA)| > rB + 2m puts every vertex of B on
// one side of A's plane (and the same with A and B exchanged). Then every
// vertex passes the vertex-level plane test of pair_interval(), which is what
// guarantees that TriDist's face test shows the triangles disjoint (no false
// 0 return), so its value is a point-pair distance, at least the distance of
// the triangles. (A tighter guard, B within rB sin(nA, nB) of the plane
// parallel to A's through B's centre, keeps 0.5% fewer pairs out and costs
// 19 more operations.) Sines are rounded up (16 ulp inside the root for the
// cosine's error); a triangle without a plane has a zero normal, so only the
// other triangle's plane can pass the guard. m = 32u (dc + rA + rB + cmA +
// cmB) + 1e-6 covers the FP32 rounding of these operations, including the
// rounded centroid's distance from its triangle's plane (u = 2^-24). The
// bound is evaluated by mark_pair(), with one margin M for all pairs of a
// voxel pair (mark_margin): a larger margin only narrows the guard and lowers
// the bound.

// ------------------------------------------------------------ pair walking

// Tiling of a voxel pair's grid: the larger side (nl triangles) is split
// into nch column chunks of width cw, the smaller side (ns) into ngr groups
// of R rows; iteration (chunk k, group g) is word k * ngr + g of the voxel
// pair's candidate mask, and lane r * cw + c is its bit for row g * R + r,
// column k * cw + c.
struct Tiling {
    uint32_t cw, R, nch, ngr;
    bool swap;  // the columns are tile-1 triangles
};

// four floats in LDS (32-bit addresses)
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) f4v lds_f4;

__device__ __forceinline__ TriS load_s_lds(const lds_f4 *p) {
    const f4v a = p[0], b = p[1], c = p[2];
    return TriS{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
}
__device__ __forceinline__ TriV load_v_lds(const lds_f4 *p) {
    const f4v a = p[0], b = p[1], c = p[2];
    return TriV{a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
}

// rows of a staged block of the walk (the most rows a tile may have)
constexpr uint32_t kStageRows = 32;

__device__ __forceinline__ Tiling mx_tiling(const VpLod &g, uint32_t max_rows) {
    Tiling t;
    t.swap = g.n1 > g.n2;
    const uint32_t nl = t.swap ? g.n1 : g.n2, ns = t.swap ? g.n2 : g.n1;
    const uint32_t m0 = (nl + kWave - 1) / kWave;
    uint32_t best = 0xffffffffu;
    for (uint32_t m = m0; m < m0 + 3; m++) {
        const uint32_t cw = (nl + m - 1) / m;
        const uint32_t R = min(min(min((uint32_t)kWave / cw, ns), max(max_rows, 1u)), kStageRows);
        const uint32_t nch = (nl + cw - 1) / cw, ngr = (ns + R - 1) / R;
        if (nch * ngr < best) {
            best = nch * ngr;
            t.cw = cw;
            t.R = R;
            t.nch = nch;
            t.ngr = ngr;
        }
    }
    return t;
}

__device__ __forceinline__ uint32_t mx_words(const Tiling &t) { return t.nch * t.ngr; }

// A voxel pair's figures for the round, worked out once (k_mx_plan, one
// thread per voxel pair) and read by every pass in one load: where its
// triangles are, how its grid is tiled, whether it is EMPTY or SMALL. (Each
// wave working them out again cost a chain of dependent loads and the
// tiling's integer divisions per pass.)
struct VpPlan {
    uint64_t o1, o2;    // first triangle of each voxel at the round's LOD
    uint32_t n1, n2;    // triangles of each voxel
    uint32_t nch, ngr;  // tiling (zero for EMPTY / SMALL)
    uint32_t shape;     // cw | R << 8 | swap << 16 | flags << 24
    uint32_t nw;        // mask words (zero for EMPTY / SMALL)
    float M;            // the mark pass's margin (mark_margin; tiled only)
    uint32_t pad;
};
static_assert(sizeof(VpPlan) == 48, "VpPlan layout");
static_assert(kBnd == 3, "a bound record and a vertex record stage alike (3 float4)");

__device__ __forceinline__ VpLod plan_lod(const VpPlan &p) { return VpLod{p.o1, p.o2, p.n1, p.n2}; }
__device__ __forceinline__ uint32_t plan_flags(const VpPlan &p) { return p.shape >> 24; }
__device__ __forceinline__ Tiling plan_tiling(const VpPlan &p) {
    Tiling t;
    t.cw = p.shape & 255u;
    t.R = (p.shape >> 8) & 255u;
    t.swap = ((p.shape >> 16) & 1u) != 0;
    t.nch = p.nch;
    t.ngr = p.ngr;
    return t;
}

// The margin of mark_pair for every pair of a voxel pair. A pair's margin
// m = fmaf(32u, dc + cm_S, fmaf(32u, cm_L, 1e-6)) grows with dc, cm_S and
// cm_L (each rounded operation is monotone), so the same operations on upper
// bounds of the three give an M at least every pair's m:
//   cm_S, cm_L  the largest cm of each voxel's triangles (sphere record
//               field w, k_vox_sph);
//   dc          Dmax = ((1 + 16u) D + R_S + R_L + 16u (cmV_S + cmV_L))
//               (1 + 16u), D the FP32 distance of the spheres' centres, R
//               their radii, cmV their R + max |C coord|.
// A triangle's exact centroid lies in its voxel's sphere and its FP32
// centroid within 5u of a coordinate bound (cmV) of it on each axis, so the
// centroids of a pair lie at most |C_S - C_L| + R_S + R_L + 9u (cmV_S +
// cmV_L) apart; D is at most 4u below |C_S - C_L|, and a pair's FP32 dc at
// most 4u above the distance of its centroids, which the factors cover
// with the rounding of Dmax's own three operations. (sa, sb): the column
// voxel's sphere records, (ra, rb) the row voxel's. (Formed per voxel pair
// by k_mx_plan: in the mark pass the dependent sphere loads lengthened every
// wave's set-up.)
__device__ __forceinline__ float mark_margin(f4v sa, f4v sb, f4v ra, f4v rb) {
    const float ex = sa.x - ra.x, ey = sa.y - ra.y, ez = sa.z - ra.z;
    const float D = __builtin_amdgcn_sqrtf(fmaf(ex, ex, fmaf(ey, ey, ez * ez)));
    const float dmax =
        (fmaf(D, 1.f + 16.f * kU, sa.w + ra.w) + 16.f * kU * (sb.x + rb.x)) * (1.f + 16.f * kU);
    return fmaf(32.f * kU, dmax + rb.w, fmaf(32.f * kU, sb.w, 1e-6f));
}

// Plans of the round's voxel pairs, with their flags and mask word counts:
// EMPTY (a voxel without triangles at this LOD), SMALL (at most small_pairs
// triangle pairs: all go to FP64, no mask), else tiled for the FP32 passes.
// Also the round's statistics as k_round_stats counts them (triangle pairs
// covered, non-empty voxel pairs; one atomic per block).
__global__ void __launch_bounds__(kBlock) k_mx_plan(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint64_t *__restrict__ off1, const uint32_t *__restrict__ cnt1,
    const uint64_t *__restrict__ off2, const uint32_t *__restrict__ cnt2, int slot,
    const float4 *__restrict__ sph1, const float4 *__restrict__ sph2,
    uint32_t small_pairs, uint32_t max_rows, VpPlan *__restrict__ plan,
    uint32_t *__restrict__ flags, uint32_t *__restrict__ nwords,
    unsigned long long *__restrict__ stats) {
    __shared__ unsigned long long part[2][kWavesPerBlock];
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t pairs = 0;
    if (w < n_vp) {
        const uint32_t v1 = vp_v1[w], v2 = vp_v2[w];
        const VpLod g = vp_lod(v1, v2, off1, cnt1, off2, cnt2, slot);
        VpPlan p{g.o1, g.o2, g.n1, g.n2, 0u, 0u, 0u, 0u, 0.f, 0u};
        uint32_t fl = 0;
        pairs = (uint64_t)g.n1 * g.n2;
        if (pairs == 0) {
            fl = MX_EMPTY;
        } else if (pairs <= small_pairs) {
            fl = MX_SMALL;
        } else {
            const Tiling t = mx_tiling(g, max_rows);
            p.nch = t.nch;
            p.ngr = t.ngr;
            p.shape = t.cw | (t.R << 8) | (t.swap ? 1u << 16 : 0u);
            p.nw = mx_words(t);
            const float4 *c = (t.swap ? sph1 : sph2) + kSph * (5 * (size_t)(t.swap ? v1 : v2) + slot);
            const float4 *r = (t.swap ? sph2 : sph1) + kSph * (5 * (size_t)(t.swap ? v2 : v1) + slot);
            const float4 c0 = c[0], c1 = c[1], r0 = r[0], r1 = r[1];
            p.M = mark_margin(f4v{c0.x, c0.y, c0.z, c0.w}, f4v{c1.x, c1.y, c1.z, c1.w},
                              f4v{r0.x, r0.y, r0.z, r0.w}, f4v{r1.x, r1.y, r1.z, r1.w});
        }
        p.shape |= fl << 24;
        plan[w] = p;
        flags[w] = fl;
        nwords[w] = p.nw;
    }
    const uint64_t sp = wave_sum_u64(pairs), sv = wave_sum_u64(pairs ? 1u : 0u);
    if (lane_id() == 0) {
        part[0][threadIdx.x / kWave] = sp;
        part[1][threadIdx.x / kWave] = sv;
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

struct PairAt {
    uint32_t i, j;     // triangle indices within the two voxels
    uint32_t word;     // iteration number
    bool valid;
    uint32_t k;        // iteration number within the staged block
    bool last;         // the iteration is the chunk's last row group
    uint64_t cvm, cvl; // the valid lanes of the chunk's row groups: all but
                       // the last, the last
};

// lane masks straight from a compare (one v_cmp into a scalar register
// pair; a ballot of a combined condition goes through a select and a
// compare). All 64 lanes must be active.
constexpr int kCmpOGT = 2, kCmpOLE = 5;  // ordered >, <= (LLVM predicates)
__device__ __forceinline__ uint64_t lanes_gt(float a, float b) {
    return __builtin_amdgcn_fcmpf(a, b, kCmpOGT);
}
__device__ __forceinline__ uint64_t lanes_le(float a, float b) {
    return __builtin_amdgcn_fcmpf(a, b, kCmpOLE);
}

// Row records of the tiled walk, staged in LDS per wave: kStageRows rows of
// bound records and of triangle records. mx_tiling never gives more than 10
// rows per tile (R <= 64 / cw, R <= the smaller voxel, and cw is at least the
// larger voxel over its chunk count, checked for all voxel sizes up to 700)
// and clamps R to kStageRows, so a staged block covers whole row groups.
struct RowStage {
    float4 s[kStageRows][kBnd];
    float4 v[kStageRows][3];
};

// Calls f(A, B, p, va, vc, swap) for every cell of the tiling: A/B the
// tile-1 / tile-2 sphere data, p the place, va()/vc() the vertices of A/B,
// swap an std::integral_constant<bool> that is true when the column triangle
// (the lane's, fixed over a column chunk) is A. All 64 lanes take part in
// every call (p.valid = false for cells outside the grid), so f may use
// wave-wide operations. fc(L, LV, cv) is called by every lane at the start of
// each column chunk with the lane's column triangle (sphere data, vertices;
// cv false for lanes outside the grid, which hold the last column's).
// Each lane reads its row from `stage`, which the wave fills with the records
// of the next R * floor(32 / R) rows whenever it starts a block of row groups
// (the row vertices too when VERTS, less their triangle's centre): one
// coalesced load per row and chunk
// instead of one per row and iteration, and no memory latency inside the
// iterations. (With one row per iteration the row is the same in every lane
// and its LDS reads are broadcasts.)
//
// With CULL, each staging lane also applies rt(row's bound record) to its row
// (rows past the smaller voxel count as passed over), and an iteration all of
// whose rows rt passes over is not evaluated: f gets skip = true (wave-
// uniform) and an empty row record. Without CULL, skip is always false.
//
// fb(w0, nb, vmask) is called by every lane after each staged block, whose
// iterations were w0 .. w0 + nb - 1 (nb <= 32), vmask the valid lanes of
// iteration w0 + lane. sx(record, beyond) may rewrite a row's bound record
// (after rt has seen it) before it is staged; beyond: the row lies past the
// smaller voxel.
template <bool SWAP, bool VERTS, bool CULL, typename FC, typename F, typename RT, typename FB,
          typename SX>
__device__ __forceinline__ void walk_tiles(const VpLod &g, const Tiling &t, int lane,
                                           const float4 *__restrict__ rec1,
                                           const float4 *__restrict__ bnd1,
                                           const float4 *__restrict__ rec2,
                                           const float4 *__restrict__ bnd2, RowStage *stage,
                                           FC &&fc, F &&f, RT &&rt, FB &&fb, SX &&sx) {
    const uint32_t nl = SWAP ? g.n1 : g.n2, ns = SWAP ? g.n2 : g.n1;
    const uint64_t ol = SWAP ? g.o1 : g.o2, os = SWAP ? g.o2 : g.o1;
    const float4 *__restrict__ bl = SWAP ? bnd1 : bnd2;
    const float4 *__restrict__ bs = SWAP ? bnd2 : bnd1;
    const float4 *__restrict__ rl = SWAP ? rec1 : rec2;
    const float4 *__restrict__ rs = SWAP ? rec2 : rec1;
    const uint32_t r = (uint32_t)lane / t.cw;
    const uint32_t c = (uint32_t)lane - r * t.cw;
    const bool in_tile = r < t.R;
    // Blocks of gpb row groups. Rows staged past the smaller voxel (at most
    // 31 past its last triangle) read the next voxel's records or the tile's
    // padding (kTriPad); their cells are not valid. Lanes outside the tile
    // read row R - 1 of a group.
    const uint32_t gpb = kStageRows / t.R, rpb = gpb * t.R;
    const uint32_t rr = min(r, t.R - 1u);
    const lds_f4 *const s_base = (const lds_f4 *)&stage->s[0][0] + kBnd * rr;
    const lds_f4 *const v_base = (const lds_f4 *)&stage->v[0][0] + 3 * rr;
    const uint64_t gmask = (t.R >= 64u) ? ~0ull : ((1ull << t.R) - 1ull);
    uint32_t word = 0;
    for (uint32_t ch = 0; ch < t.nch; ch++) {
        const uint32_t col = ch * t.cw + c;
        const bool cv = in_tile && col < nl;
        const uint64_t lt = ol + (cv ? col : nl - 1);  // in bounds for idle lanes
        const TriS L = load_s(bl, lt);
        const TriV LV = load_v(rl, lt);
        // the chunk's valid lanes (cvm), and those of the last row group:
        // its first ns - (ngr - 1) R rows, held by the lanes below that
        // many times cw
        const uint64_t cvm = __ballot(cv);
        const uint32_t nlast = (ns - (t.ngr - 1u) * t.R) * t.cw;
        const uint64_t cvl = cvm & (nlast >= 64u ? ~0ull : (1ull << nlast) - 1ull);
        fc(L, LV, cv);
        auto lane_v = [&]() { return LV; };
        // staged blocks of gpb row groups, then the block's iterations in a
        // loop of their own (no per-iteration block test)
        for (uint32_t gb = 0; gb < t.ngr; gb += gpb) {
            const uint64_t row0 = os + (uint64_t)gb * t.R;
            wave_sync();  // the previous block's reads are done
            // the block's records: bound records to lanes 0..rpb-1, vertex
            // records (VERTS) to lanes 32..32+rpb-1
            const bool is_s = lane < 32;
            const uint32_t q = is_s ? (uint32_t)lane : (uint32_t)lane - 32u;
            const bool ld = q < rpb && (is_s || VERTS);
            bool over = false;
            if (ld) {
                const float4 *p = is_s ? bs + kBnd * (row0 + q) : rs + 3 * (row0 + q);
                float4 rq[3];
#pragma unroll
                for (int i = 0; i < 3; i++) rq[i] = p[i];
                if (VERTS && !is_s) {
                    // vertices staged less their triangle's centre, formed as
                    // the bound record forms it (tri_bound)
                    const float cx = ((rq[0].x + rq[1].x) + rq[2].x) * (1.f / 3.f);
                    const float cy = ((rq[0].y + rq[1].y) + rq[2].y) * (1.f / 3.f);
                    const float cz = ((rq[0].z + rq[1].z) + rq[2].z) * (1.f / 3.f);
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        rq[i].x -= cx;
                        rq[i].y -= cy;
                        rq[i].z -= cz;
                    }
                }
                if constexpr (CULL) over = is_s && (gb * t.R + q >= ns || rt(rq));
                if (is_s) sx(rq, gb * t.R + q >= ns);
                float4 *dst = is_s ? &stage->s[q][0] : &stage->v[q][0];
#pragma unroll
                for (int i = 0; i < 3; i++) dst[i] = rq[i];
            }
            // rows of the block rt does not pass over (bit per row)
            const uint64_t kept = CULL ? ~__ballot(over) : ~0ull;
            wave_sync();
            const uint32_t nb = min(gpb, t.ngr - gb);
            uint32_t row = gb * t.R + r;
            uint64_t gm = gmask;  // the rows of iteration k
            for (uint32_t k = 0, k0 = 0; k < nb; k++, k0 += t.R, row += t.R, word++) {
                const bool valid = cv && row < ns;
                // the lane's row slot: a 32-bit LDS address, the lane's part
                // formed once
                const bool skip = CULL && (kept & gm) == 0ull;
                gm <<= t.R;
                TriS S{};
                if (!skip) S = load_s_lds(s_base + kBnd * k0);
                auto row_v = [&]() { return load_v_lds(v_base + 3 * k0); };
                const bool last = gb + k == t.ngr - 1u;
                if (SWAP)
                    f(L, S, PairAt{cv ? col : 0u, row, word, valid, k, last, cvm, cvl}, lane_v,
                      row_v, std::integral_constant<bool, true>{}, skip);
                else
                    f(S, L, PairAt{row, cv ? col : 0u, word, valid, k, last, cvm, cvl}, row_v,
                      lane_v, std::integral_constant<bool, false>{}, skip);
            }
            fb(word - nb, nb, gb + (uint32_t)lane == t.ngr - 1u ? cvl : cvm);
        }
    }
}

template <bool VERTS, bool CULL, typename FC, typename F, typename RT, typename FB, typename SX>
__device__ __forceinline__ void for_tiles(const VpLod &g, const Tiling &t, int lane,
                                          const float4 *__restrict__ rec1,
                                          const float4 *__restrict__ bnd1,
                                          const float4 *__restrict__ rec2,
                                          const float4 *__restrict__ bnd2, RowStage *stage,
                                          FC &&fc, F &&f, RT &&rt, FB &&fb, SX &&sx) {
    if (t.swap)
        walk_tiles<true, VERTS, CULL>(g, t, lane, rec1, bnd1, rec2, bnd2, stage, fc, f, rt, fb,
                                      sx);
    else
        walk_tiles<false, VERTS, CULL>(g, t, lane, rec1, bnd1, rec2, bnd2, stage, fc, f, rt, fb,
                                       sx);
}
// nothing at the end of a staged block
struct NoBlockEnd {
    __device__ void operator()(uint32_t, uint32_t, uint64_t) const {}
};
// no row test
struct NoRowTest {
    __device__ bool operator()(const float4 *) const { return false; }
};
// row bound records staged as loaded
struct NoStageEdit {
    __device__ void operator()(float4 *, bool) const {}
};

// ------------------------------------------------------------------ bounds

// v_min_f32 as written: fminf in a loop re-canonicalises its running value
// (one more instruction per call); the operands here are never signalling
// NaNs
__device__ __forceinline__ float min_f32(float a, float b) {
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// lowest q among the lanes holding the wave minimum of v (kNoProbe if none)
__device__ __forceinline__ uint32_t wave_argmin_q(float v, uint32_t q) {
    const float vmin = wave_min_f32(v);
    return wave_min_u32(v == vmin ? q : kNoProbe);
}

// Upper bounds for k_mx_bounds. Every value TriDist returns is the distance
// between two points of the triangles, so the FP64 d of a pair is at most
// dc* + r_A + r_B, dc* the exact distance of the (FP32) centres: with dc the
// FP32 centre distance (dc* <= dc (1 + 5u), u = 2^-24) and F = 1 + 20u,
//   d  <= fl(dc F + fl(r_A G + fl(r_B G)))              (G = 1 + 4u),
//   lo <= max(fl(dc F + fl(alo_A + alo_B)), 0),  alo = r - ph + 8u t,
//   hi <=     fl(dc F + fl(ahi_A + ahi_B)),      ahi = r + h + 8u t,
// t = r + ph + h (alo and ahi are formed once per triangle by k_tri_bounds):
// each rounding is at most u of its result, which the factors and the 8u t
// terms cover with room for the FP64 rounding of lo and hi themselves. The
// clamp of lo at 0 is applied to the minimum.
constexpr float kUF = 1.f + 20.f * kU;
constexpr float kUG = 1.f + 4.f * kU;

// Per voxel pair: the smallest FP32 upper bound of each quantity the round
// folds (vpU; -inf for the others) and up to three probe pairs, the pairs of
// smallest score (centre distance less the radii, with the quantity's
// Hausdorff terms), whose FP64 values usually are the minima. NEED: the
// quantities the round folds (bit 0 d, 1 lo, 2 hi), a launch constant so that
// the arithmetic of the others drops out (and `need` the same bits at run
// time). The column triangle's parts of the bounds and scores are per-lane
// constants of a column chunk; a pair adds the row triangle's.
//
// Rows may be passed over freely: the minimum over any subset of the pairs
// is still an upper bound of the minimum, and any pair a valid probe. With
// CULL the walk passes over the rows whose pairs cannot reach below what the
// round's fold keeps anyway (the caps of k_mx_mark): in a round that folds
// lo and hi, the voxel pair's range (vp_min, vp_max); in the last round, the
// object pair's range max so far (pair_max; an estimate of its final
// distance). The row test is the sphere gap to the column voxel's sphere
// (k_vox_sph), with the quantities' Hausdorff terms; it needs no margins
// (a wrong call costs a looser U, never a wrong result).
template <uint32_t NEED, bool CULL>
__global__ void __launch_bounds__(kBlock) k_mx_bounds(
    uint64_t n_vp, const VpPlan *__restrict__ plan, const float4 *__restrict__ rec1,
    const float4 *__restrict__ bnd1, const float4 *__restrict__ rec2,
    const float4 *__restrict__ bnd2, uint32_t need, float *__restrict__ vpU,
    uint32_t *__restrict__ probe, unsigned long long *__restrict__ mkey,
    const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const float4 *__restrict__ sph1, const float4 *__restrict__ sph2, int slot,
    const double *__restrict__ vp_min, const double *__restrict__ vp_max,
    const uint32_t *__restrict__ vp_pair, const double *__restrict__ pair_max, int caps) {
    const uint32_t w = wave_item();
    if (w >= n_vp) return;
    const int lane = lane_id();
    const VpPlan pl = plan[w];
    // the cull thresholds (requested with the plan; +inf where none applies)
    float Td = INFINITY, Tlo = INFINITY, Thi = INFINITY;
    uint32_t vw1 = 0, vw2 = 0;
    if constexpr (CULL) {
        vw1 = vp_v1[w];
        vw2 = vp_v2[w];
        const double mn = vp_min[w], mx = vp_max[w], pm = pair_max[vp_pair[w]];
        if (caps & 1) {
            Tlo = (float)mn;
            Thi = (float)mx;
        }
        if (caps & 2) Td = (float)pm;
    }
    if (lane < 3) {
        mkey[3 * (size_t)w + lane] = dkey(INFINITY);
        probe[3 * (size_t)w + lane] = kNoProbe;
        vpU[3 * (size_t)w + lane] = -INFINITY;
    }
    if (plan_flags(pl)) return;  // EMPTY or SMALL
    const VpLod g = plan_lod(pl);
    const uint32_t n2 = g.n2;
    const Tiling t = plan_tiling(pl);
    float up[3] = {INFINITY, INFINITY, INFINITY};
    float best[3] = {INFINITY, INFINITY, INFINITY};
    uint32_t bq[3] = {kNoProbe, kNoProbe, kNoProbe};
    // column constants: r G, alo, ahi; the scores' -r, -(r + ph), h - r
    float crg = 0.f, calo = 0.f, cahi = 0.f, ck0 = 0.f, ck1 = 0.f, ck2 = 0.f;
    __shared__ RowStage stage[kWavesPerBlock];
    // the column voxel's sphere (scalar loads) and the row test
    typedef __attribute__((address_space(4))) const f4v cst_f4;
    f4v sa = {0.f, 0.f, 0.f, 0.f}, sb = {0.f, 0.f, 0.f, 0.f};
    if constexpr (CULL) {
        const cst_f4 *const sp = (const cst_f4 *)((t.swap ? sph1 : sph2) +
                                                  kSph * (5 * (size_t)(t.swap ? vw1 : vw2) + slot));
        sa = sp[0];
        sb = sp[1];
    }
    auto row_test = [&](const float4 *pr) -> bool {
        const float dx = pr[0].x - sa.x, dy = pr[0].y - sa.y, dz = pr[0].z - sa.z;
        const float lb = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz) - pr[0].w - sa.w;
        bool over = true;
        if ((NEED & need) & 1u) over = over && lb > Td;
        if ((NEED & need) & 2u) over = over && lb - pr[2].x - sb.y > Tlo;
        if ((NEED & need) & 4u) over = over && lb + pr[2].y + sb.z > Thi;
        return over;
    };
    // Cells outside the grid get +inf bounds and scores, so every lane
    // computes every cell with no select and no lane mask: a lane outside
    // the grid has +inf column constants, and a row past the smaller voxel is
    // staged with +inf row terms. The staged row record carries the scores'
    // row terms (-r, -(r + ph), h - r) in place of its normal, which this
    // pass does not read.
    for_tiles<false, CULL>(g, t, lane, rec1, bnd1, rec2, bnd2, &stage[threadIdx.x / kWave],
              [&](const TriS &L, const TriV &, bool cv) {
                  crg = cv ? L.r * kUG : INFINITY;
                  calo = cv ? L.alo : INFINITY;
                  cahi = cv ? L.ahi : INFINITY;
                  ck0 = cv ? -L.r : INFINITY;
                  ck1 = cv ? -(L.r + L.ph) : INFINITY;
                  ck2 = cv ? L.h - L.r : INFINITY;
              },
              [&](const TriS &A, const TriS &B, const PairAt &p, auto &&, auto &&, auto swap,
                  bool skip) {
                  // an iteration whose rows the row test passed over adds
                  // nothing (a wave-uniform branch)
                  if (skip) return;
                  constexpr bool col_is_a = decltype(swap)::value;
                  const TriS &S = col_is_a ? B : A;
                  const TriS &L = col_is_a ? A : B;
                  const float dx = S.cx - L.cx, dy = S.cy - L.cy, dz = S.cz - L.cz;
                  float dc;
                  {
#pragma clang fp contract(fast)
                      dc = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
                  }
                  float x[3], key[3];
                  // S.nx, S.ny, S.nz: the staged score terms (see above)
                  if (NEED & 1u) {
                      x[0] = fmaf(dc, kUF, fmaf(S.r, kUG, crg));
                      key[0] = dc + (ck0 + S.nx);
                  }
                  if (NEED & 2u) {
                      x[1] = fmaf(dc, kUF, S.alo + calo);
                      key[1] = dc + (ck1 + S.ny);
                  }
                  if (NEED & 4u) {
                      x[2] = fmaf(dc, kUF, S.ahi + cahi);
                      key[2] = dc + (ck2 + S.nz);
                  }
                  const uint32_t q = p.i * n2 + p.j;
#pragma unroll
                  for (int f = 0; f < 3; f++) {
                      if (!((NEED >> f) & 1u)) continue;
                      up[f] = min_f32(up[f], x[f]);
                      bq[f] = key[f] < best[f] ? q : bq[f];
                      best[f] = min_f32(best[f], key[f]);
                  }
              },
              row_test, NoBlockEnd{},
              [](float4 *r, bool beyond) {
                  // the scores' row terms in place of the normal; +inf row
                  // terms for a row past the smaller voxel
                  const float k0 = -r[0].w, k1 = -(r[0].w + r[2].x), k2 = r[2].y - r[0].w;
                  r[1].x = beyond ? INFINITY : k0;
                  r[1].y = beyond ? INFINITY : k1;
                  r[1].z = beyond ? INFINITY : k2;
                  if (beyond) {
                      r[0].w = INFINITY;
                      r[2].z = INFINITY;
                      r[2].w = INFINITY;
                  }
              });
    up[1] = fmaxf(up[1], 0.f);
    float U[3];
    uint32_t pq[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        if (((NEED & need) >> f) & 1u) {
            U[f] = wave_min_f32(up[f]);
            pq[f] = wave_argmin_q(best[f], bq[f]);
        } else {
            U[f] = -INFINITY;
            pq[f] = kNoProbe;
        }
    }
    // Slots of quantities the round does not fold take further probes: the
    // pairs next to the last folded quantity's probe in the voxel pair's
    // order (neighbouring triangles of the smaller voxel's row, where the
    // minimum often lies when the probe misses it).
    {
        constexpr int mq = (NEED & 4u) ? 2 : ((NEED & 2u) ? 1 : 0);
        const uint32_t q0 = pq[mq], nq = g.n1 * n2;
        int step = 1;
#pragma unroll
        for (int f = 0; f < 3; f++) {
            if ((((NEED & need) >> f) & 1u) || !(((NEED & need) >> mq) & 1u) || q0 == kNoProbe)
                continue;
            const int64_t qq = (int64_t)q0 + step;
            pq[f] = (qq >= 0 && qq < (int64_t)nq) ? (uint32_t)qq : kNoProbe;
            step = -step;
        }
    }
    const bool lo_zero = U[1] == 0.f;  // some pair certainly has lo = +0
    if (lo_zero) pq[1] = kNoProbe;
    if (pq[1] == pq[0]) pq[1] = kNoProbe;
    if (pq[2] == pq[0] || pq[2] == pq[1]) pq[2] = kNoProbe;
    if (lane < 3) {
        vpU[3 * (size_t)w + lane] = lane == 0 ? U[0] : (lane == 1 ? U[1] : U[2]);
        probe[3 * (size_t)w + lane] = lane == 0 ? pq[0] : (lane == 1 ? pq[1] : pq[2]);
    }
    if (lane == 0 && lo_zero) mkey[3 * (size_t)w + 1] = dkey(0.0);  // the minimum is +0
}

// ------------------------------------------------------------- FP64 lists

// Minima over the lanes of one voxel pair w (lanes are sorted by w; kNoProbe
// marks an empty lane), folded with one atomic per voxel pair and quantity.
// NEED: the quantities folded (bits 0-2: a, b, c); the others are left alone.
template <uint32_t NEED = 7u>
__device__ __forceinline__ void seg_min_fold(uint32_t w, unsigned long long a,
                                             unsigned long long b, unsigned long long c,
                                             unsigned long long *__restrict__ mkey) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t ow = (uint32_t)__shfl_down((int)w, o, kWave);
        const bool same = lane + o < kWave && ow == w;
        if constexpr ((NEED & 1u) != 0) {
            const unsigned long long oa = __shfl_down(a, o, kWave);
            if (same) a = min(a, oa);
        }
        if constexpr ((NEED & 2u) != 0) {
            const unsigned long long ob = __shfl_down(b, o, kWave);
            if (same) b = min(b, ob);
        }
        if constexpr ((NEED & 4u) != 0) {
            const unsigned long long oc = __shfl_down(c, o, kWave);
            if (same) c = min(c, oc);
        }
    }
    const uint32_t pw = (uint32_t)__shfl_up((int)w, 1, kWave);
    const bool head = lane == 0 || pw != w;
    if (head && w != kNoProbe) {
        const unsigned long long inf_key = dkey(INFINITY);
        if ((NEED & 1u) && a < inf_key) atomicMin(&mkey[3 * (size_t)w], a);
        if ((NEED & 2u) && b < inf_key) atomicMin(&mkey[3 * (size_t)w + 1], b);
        if ((NEED & 4u) && c < inf_key) atomicMin(&mkey[3 * (size_t)w + 2], c);
    }
}

// FP64 TriDist of work items (voxel pair w, triangle pair q), min-folded into
// the voxel pair's minima keys; items with q == kNoProbe are skipped. Items
// arrive sorted by voxel pair (probe lists: w = k / 3; work lists: written
// in voxel-pair order). NEED: the quantities the round folds (the others'
// keys are not touched: the round does not use them).
//
// With vp_from_index3 (the probes: item k is probe k % 3 of voxel pair k / 3)
// it also records where each probe lies in its voxel pair's candidate masks,
// pwb[k] = word << 6 | bit (~0 for none), so that k_mx_mark can clear the
// probes' bits without dividing per voxel pair.
template <uint32_t NEED>
__global__ void __launch_bounds__(kBlock) k_mx_eval(
    uint64_t n_items, const uint32_t *__restrict__ item_vp, const uint32_t *__restrict__ item_q,
    int vp_from_index3, const VpPlan *__restrict__ plan, const float4 *__restrict__ rec1,
    const float4 *__restrict__ rec2, unsigned long long *__restrict__ mkey,
    unsigned long long *__restrict__ stats, unsigned long long *__restrict__ pwb) {
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
            unsigned long long at = ~0ull;
            if (q != kNoProbe) {
                act = true;
                const VpPlan &pl = plan[w];
                const uint64_t o1 = pl.o1, o2 = pl.o2;
                const uint32_t n2 = pl.n2;
                const VpLod g{o1, o2, 0u, n2};
                const uint32_t i = q / g.n2, j = q - i * g.n2;
                if (vp_from_index3 && !plan_flags(pl)) {
                    const Tiling t = plan_tiling(pl);
                    const uint32_t row = t.swap ? j : i, col = t.swap ? i : j;
                    const uint32_t ch = col / t.cw, gr = row / t.R;
                    at = ((unsigned long long)(ch * t.ngr + gr) << 6) |
                         ((row - gr * t.R) * t.cw + (col - ch * t.cw));
                }
                double d, lo, hi;
                pair_values<double>(rec1, g.o1 + i, rec2, g.o2 + j, d, lo, hi);
                a = dkey(d);
                b = dkey(lo);
                c = dkey(hi);
            }
            if (vp_from_index3) pwb[k] = at;
        }
        seg_min_fold<NEED>(w, a, b, c, mkey);
        n += (uint64_t)__popcll(__ballot(act));
    }
    if (lane_id() == 0 && n) {
        stat_add(stats, 2, (unsigned long long)n);
        if (vp_from_index3) stat_add(stats, 4, (unsigned long long)n);  // the probes
    }
}

// ------------------------------------------------------------ FP32 pass

// U per quantity: the FP32 upper bound tightened by the exact values found
// so far (rounded up to FP32); -inf for quantities the round does not fold
// (their vpU) and for lo once it is known to be +0, its smallest value
__device__ __forceinline__ void bounds_U(const float *vpU, const unsigned long long *mkey,
                                         size_t w, float U[3]) {
#pragma unroll
    for (int f = 0; f < 3; f++)
        U[f] = fminf(vpU[3 * w + f], ulp_up((float)dunkey(mkey[3 * w + f])));
    if (mkey[3 * w + 1] == dkey(0.0)) U[1] = -INFINITY;
}

// ------------------------------------------------------- candidate test
//
// k_mx_mark keeps a pair when a lower bound dl of its FP64 TriDist value d
// does not exclude it from holding the minimum of a quantity the round
// folds, i.e. dl <= T for T = max over the needed q of T_q, with
//   T_d  = U_d,
//   T_lo = U_lo + ph1 + ph2   (lo = max(d - ph1 - ph2, 0) <= U_lo implies
//                              d <= U_lo + ph1 + ph2, U_lo >= 0),
//   T_hi = U_hi - h1 - h2     (hi = d + h1 + h2 <= U_hi),
// each rounded upwards: the column triangle's part is a per-lane constant of
// the column chunk (MarkCol), so a pair pays one fused multiply-add per
// quantity. Rounding: with H = 1 + 8u and G = 1 - 16u (u = 2^-24),
//   fl(c H + fl(b H + fl(a H)))       >= (a + b + c)(1 + 4u),
//   fl(-c G + fl(-b G + fl(a H)))     >= a - b - c + 3u a
// for a, b, c >= 0 -- above the exact thresholds by more than the FP64
// rounding of the quantities themselves (a few 2^-53 relative). U = -inf
// (a quantity not folded, or lo known to be +0) gives T_q = -inf. The lower
// bounds of mark_pair are not clamped at 0 (a smaller bound only keeps more
// pairs).
constexpr float kTH = 1.f + 8.f * kU;
constexpr float kTG = 1.f - 16.f * kU;

// per-lane constants of a column triangle L
struct MarkCol {
    float clo, chi;  // column parts of T_lo, T_hi
    float rm;        // r_L + 2M, the guard's bound for the row's plane offset
    float x0, y0, z0, x1, y1, z1, x2, y2, z2;  // L's vertices less its centre
};

__device__ __forceinline__ MarkCol mark_col(const TriS &L, const TriV &V, float ulo_h,
                                            float uhi_h, float M) {
    MarkCol c;
    c.clo = fmaf(L.ph, kTH, ulo_h);
    c.chi = fmaf(L.h, -kTG, uhi_h);
    c.rm = L.r + 2.f * M;
    c.x0 = V.x0 - L.cx; c.y0 = V.y0 - L.cy; c.z0 = V.z0 - L.cz;
    c.x1 = V.x1 - L.cx; c.y1 = V.y1 - L.cy; c.z1 = V.z1 - L.cz;
    c.x2 = V.x2 - L.cx; c.y2 = V.y2 - L.cy; c.z2 = V.z2 - L.cz;
    return c;
}

// T for a pair: the row triangle's terms added to the column constants
template <uint32_t NEED>
__device__ __forceinline__ float mark_T(const TriS &S, const MarkCol &c, float ud) {
    float T = -INFINITY;
    if (NEED & 1u) T = ud;
    if (NEED & 2u) T = fmaxf(T, fmaf(S.ph, kTH, c.clo));
    if (NEED & 4u) T = fmaxf(T, fmaf(S.h, -kTG, c.chi));
    return T;
}

// Candidate test of one pair (row triangle S, column triangle L) against T:
//  1. the disc bound (described above; the sines are not clamped at 1 and
//     1 - c^2 + 16u is formed by one fused multiply-add, which only lowers
//     the bound or rounds once where two roundings were);
//  2. for pairs it keeps, the gap between the projections of the actual
//     vertices on the centre axis (the vertex-level bound of pair_interval()).
//     The disc guard, which holds for all but a few in 10^5 of the pairs
//     reaching this step, already shows the triangles disjoint for TriDist,
//     so the vertex-level plane tests are not repeated; a pair without the
//     guard stays a candidate. The projections are taken on the unnormalised
//     centre difference d and scaled by 1/dc once, the row's as d . (v - cS)
//     + |d|^2 from vertices staged less their centre; the rounding of the
//     vertex differences, the dot products, |d|^2 and the reciprocal is a few
//     u (dc + r_S + r_L + cm_S + cm_L), inside the margin.
// M is the voxel pair's margin (mark_margin, at least every pair's m); the
// row's r_S + 2M comes staged in place of its cm (k_mx_mark), the column's
// r_L + 2M in c.rm, so the guard is two compares.
// Operation count per pair through step 1 (add/sub/mul/sqrt/rcp/min/max/
// compare = 1, fused multiply-add = 2, abs free): centre difference 3, dc 6,
// the two plane offsets 10, guard 3, sines and disc gap 16, comparison 1 --
// 39; T adds 5 in rounds that fold lo and hi (44).
//
// Called by every lane of the wave (cells outside the grid run on in-bounds
// data; the caller masks their bits); returns the mask of the lanes whose
// pair is kept. The conditions are lane masks in scalar registers
// throughout, and the vertex-level bound is a wave-uniform branch, taken when
// some lane needs it, so no lane mask is saved and restored per pair.
template <typename RV>
__device__ __forceinline__ uint64_t mark_pair(const TriS &S, const TriS &L, const MarkCol &c,
                                              float M, float T, RV &&row_v) {
#pragma clang fp contract(fast)
    const float dx = S.cx - L.cx, dy = S.cy - L.cy, dz = S.cz - L.cz;
    const float dd = dx * dx + dy * dy + dz * dz;
    const float dc = __builtin_amdgcn_sqrtf(dd);
    const float eL = L.nx * dx + L.ny * dy + L.nz * dz;
    const float eS = S.nx * dx + S.ny * dy + S.nz * dz;
    // the lanes with the guard (S.cm: r_S + 2M as staged)
    const uint64_t g = lanes_gt(fabsf(eL), S.cm) | lanes_gt(fabsf(eS), c.rm);
    // the disc bound for every lane (no branch: nearly every iteration has a
    // lane with the guard); without the guard the bound is 0 (a select on
    // the guard's mask). (dc = 0 gives no guard; its NaN cosines are dropped
    // by fmaxf.)
    const float inv = __builtin_amdgcn_rcpf(dc);
    const float cl = eL * inv, cs = eS * inv;
    const float sl = __builtin_amdgcn_sqrtf(fmaxf(fmaf(-cl, cl, 1.f + 16.f * kU), 16.f * kU));
    const float ss = __builtin_amdgcn_sqrtf(fmaxf(fmaf(-cs, cs, 1.f + 16.f * kU), 16.f * kU));
    const float dg = fmaf(-L.r, sl, fmaf(-S.r, ss, dc - M));
    float dl;
    asm("v_cndmask_b32_e64 %0, 0, %1, %2" : "=v"(dl) : "v"(dg), "s"(g));
    const uint64_t keep = lanes_le(dl, T);
    uint64_t take = keep;
    if (keep & g) {
        // the row's vertices less its centre (as staged): S's vertices lie
        // beyond L's along d = cS - cL, at d . (v - cS) + |d|^2 from L's centre
        const TriV v = row_v();
        const float s0 = dx * v.x0 + dy * v.y0 + dz * v.z0;
        const float s1 = dx * v.x1 + dy * v.y1 + dz * v.z1;
        const float s2 = dx * v.x2 + dy * v.y2 + dz * v.z2;
        const float l0 = dx * c.x0 + dy * c.y0 + dz * c.z0;
        const float l1 = dx * c.x1 + dy * c.y1 + dz * c.z1;
        const float l2 = dx * c.x2 + dy * c.y2 + dz * c.z2;
        const float gap =
            ((fminf(s0, fminf(s1, s2)) + dd) - fmaxf(l0, fmaxf(l1, l2))) * inv - M;
        take = keep & (~g | lanes_le(gap, T));
    }
    return take;
}

// Per object pair, the smallest U_d of its tiled voxel pairs (FP32 upper
// bound, tightened by the probes' exact values) as an order-preserving key
// (obj_u starts at the key of +inf). EMPTY and SMALL voxel pairs give no
// bound here (their pairs are all evaluated in FP64 anyway).
__global__ void __launch_bounds__(kBlock) k_mx_obj_u(
    uint64_t n_vp, const VpPlan *__restrict__ plan, const float *__restrict__ vpU,
    const unsigned long long *__restrict__ mkey, const uint32_t *__restrict__ vp_pair,
    uint32_t *__restrict__ obj_u) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_vp || plan_flags(plan[w])) return;
    const float u = fminf(vpU[3 * w], ulp_up((float)dunkey(mkey[3 * w])));
    atomicMin(&obj_u[vp_pair[w]], fkey(u));
}

// A tiled voxel pair's figures for the mark pass: its thresholds and where
// its probes lie in its candidate masks (k_mx_prep).
struct MarkPrep {
    float ud;            // U_d
    float ulo_h, uhi_h;  // U_lo, U_hi times (1 + 8u)
    uint32_t pbits;      // bit of probe f in its word at bits 8f..8f+5
    uint32_t pw[3];      // word of probe f (~0: none)
    uint32_t pad;
};
static_assert(sizeof(MarkPrep) == 32, "MarkPrep layout");

// The mark pass's thresholds per tiled voxel pair, one thread each (worked
// out by each voxel pair's wave they took a chain of dependent loads and
// about a hundred wave instructions before its walk could start).
//
// U per quantity: the FP32 upper bound tightened by the exact values found
// so far (rounded up to FP32); -inf for lo once it is known to be +0, its
// smallest value; lo without an FP32 bound this round (lo_free) takes only
// the probes' exact values; then the caps of k_mx_mark (bit 0: the voxel
// pair's range, bit 1: the object pair's smallest U_d). The probes' words
// and bits come from k_mx_eval (pwb).
__global__ void __launch_bounds__(kBlock) k_mx_prep(
    uint64_t n_vp, const VpPlan *__restrict__ plan, const float *__restrict__ vpU,
    const unsigned long long *__restrict__ mkey, const unsigned long long *__restrict__ pwb,
    const double *__restrict__ vp_min, const double *__restrict__ vp_max,
    const uint32_t *__restrict__ vp_pair, const uint32_t *__restrict__ obj_u, int caps,
    int lo_free, MarkPrep *__restrict__ prep) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_vp || plan_flags(plan[w])) return;
    const size_t w3 = 3 * w;
    const unsigned long long q1 = mkey[w3 + 1];
    float U0 = fminf(vpU[w3], ulp_up((float)dunkey(mkey[w3])));
    float U1 = q1 == dkey(0.0) ? -INFINITY
                               : (lo_free ? ulp_up((float)dunkey(q1))
                                          : fminf(vpU[w3 + 1], ulp_up((float)dunkey(q1))));
    float U2 = fminf(vpU[w3 + 2], ulp_up((float)dunkey(mkey[w3 + 2])));
    if (caps & 2) U0 = fminf(U0, funkey(obj_u[vp_pair[w]]));
    if (caps & 1) {
        U1 = fminf(U1, ulp_up((float)vp_min[w]));
        U2 = fminf(U2, ulp_up((float)vp_max[w]));
    }
    MarkPrep m;
    m.ud = U0;
    m.ulo_h = U1 * kTH;
    m.uhi_h = U2 * kTH;
    m.pbits = 0;
    m.pad = 0;
#pragma unroll
    for (int f = 0; f < 3; f++) {
        const unsigned long long at = pwb[w3 + f];
        m.pw[f] = at == ~0ull ? 0xffffffffu : (uint32_t)(at >> 6);
        m.pbits |= (uint32_t)(at & 63u) << (8 * f);
    }
    prep[w] = m;
}

// Candidate bit masks: word k of voxel pair w (at mask[woff[w] + k]) holds
// the ballot of iteration k of the tiled walk; count[w] = number of bits.
// A pair is a candidate when mark_pair keeps it; the probes (already
// evaluated) never are: their bits are cleared from the ballot words (word
// and bit worked out once per voxel pair, a scalar test per iteration
// instead of three vector compares per pair). SMALL voxel pairs:
// every pair, no mask. NEED: the quantities of the round (U is -inf for the
// others). The thresholds come from k_mx_prep.
//
// Two caps tighten U where a round's fold does not need every voxel pair's
// exact minima (both keep every pair that can change a folded value):
//  * a round that folds lo and hi takes min(range min, min lo) and
//    min(range max, min hi) (SpatialJoin fold at Hausdorff level 2): a pair
//    whose lo exceeds the voxel pair's current range min cannot lower it,
//    nor can one whose hi exceeds the range max, so U_lo and U_hi are capped
//    by the range (vp_min / vp_max, rounded up to FP32; the minima found over
//    the pairs that pass are then exact wherever they are below the range,
//    and min(range, found) equals min(range, exact) in every case);
//  * the last round folds only each object pair's minimum d over its voxel
//    pairs ([d, d] per voxel pair, the object range their minimum), so U_d is
//    capped by the smallest U_d of the object pair's voxel pairs (obj_u, keys
//    of k_mx_obj_u): a pair above it cannot hold the object's minimum. A
//    voxel pair's own minimum may then stay above its exact value, which
//    changes neither the object range nor which voxel pairs the eviction
//    keeps (a voxel pair whose exact minimum equals the object's has a pair
//    at that value, below the cap, so it is found exactly).
template <uint32_t NEED, bool CULL>
__global__ void __launch_bounds__(kBlock) k_mx_mark(
    uint64_t n_vp, const VpPlan *__restrict__ plan, const float4 *__restrict__ rec1,
    const float4 *__restrict__ bnd1, const float4 *__restrict__ rec2,
    const float4 *__restrict__ bnd2, const MarkPrep *__restrict__ prep,
    const uint64_t *__restrict__ woff,
    const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const float4 *__restrict__ sph1, const float4 *__restrict__ sph2, int slot,
    unsigned long long *__restrict__ mask, uint32_t *__restrict__ count,
    uint32_t *__restrict__ tested) {
    const uint32_t w = wave_item();
    if (w >= n_vp) return;
    const int lane = lane_id();
    // Every figure of the voxel pair is requested before any is used: one
    // wait for all of them (the empty asm takes them all before the first
    // branch; left alone, the compiler moves each load next to its first
    // use, behind the EMPTY / SMALL branch, and the loads wait one after
    // another). The figures of k_mx_prep are not written for EMPTY and
    // SMALL voxel pairs, which do not use them.
    const VpPlan pl = plan[w];
    const MarkPrep pr = prep[w];
    const uint32_t vw1 = vp_v1[w], vw2 = vp_v2[w];
    const uint64_t wo = woff[w];
    asm volatile("" ::"s"(pl.shape), "s"(pl.o1), "s"(pl.o2), "s"(pl.n1), "s"(pl.n2), "s"(pl.M),
                 "s"(pr.ud), "s"(pr.pw[0]), "s"(wo), "s"(vw1), "s"(vw2));
    const uint32_t fl = plan_flags(pl);
    if (fl & MX_EMPTY) {
        if (lane == 0) count[w] = tested[w] = 0;
        return;
    }
    const VpLod g = plan_lod(pl);
    if (fl & MX_SMALL) {
        if (lane == 0) {
            count[w] = g.n1 * g.n2;
            tested[w] = 0;
        }
        return;
    }
    const float ud = pr.ud, ulo_h = pr.ulo_h, uhi_h = pr.uhi_h;
    const Tiling t = plan_tiling(pl);
    unsigned long long *__restrict__ out = mask + wo;
    // the probes' words and bits in the tiling
    uint32_t pword[3];
    unsigned long long pbit[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        pword[f] = pr.pw[f];
        pbit[f] = pr.pw[f] == 0xffffffffu ? 0ull : 1ull << ((pr.pbits >> (8 * f)) & 63u);
    }
    // the words of the current staged block, lane k holding word k, stored
    // together after the block; each lane counts the bits of the words it
    // stores
    uint32_t wlo = 0, whi = 0;
    uint32_t cl = 0;
    uint32_t sk = 0;  // pairs of the iterations passed over (not tested one by one)
    MarkCol mc;
    __shared__ RowStage stage[kWavesPerBlock];
    // The row test against the column voxel's sphere (k_vox_sph; every
    // column triangle of the walk lies in it): a row whose plane has the
    // whole sphere on one side (by more than twice the margin, the disc
    // guard with the sphere for the column triangle) and whose disc is
    // farther from the sphere than the largest T of its pairs cannot hold a
    // minimum with any column triangle -- each pair's FP64 TriDist value is
    // then at least the triangles' distance (the face test shows them
    // disjoint), which is at least the disc-to-sphere gap D - r_S sin - R.
    // The row's T takes the column terms' extremes over the voxel (largest
    // ph, smallest h); the fused multiply-adds are monotone, so it is at
    // least every pair's T.
    // The sphere comes by scalar loads and stays in scalar registers.
    typedef __attribute__((address_space(4))) const f4v cst_f4;
    const cst_f4 *const sp =
        (const cst_f4 *)((t.swap ? sph1 : sph2) + kSph * (5 * (size_t)(t.swap ? vw1 : vw2) + slot));
    const f4v sa = sp[0], sb = sp[1];
    const float M = pl.M;  // the voxel pair's margin (k_mx_plan)
    auto row_test = [&](const float4 *pr) -> bool {
#pragma clang fp contract(fast)
        const float s_mc = fmaf(32.f * kU, sb.x, 1e-6f);
        const float s_tlo = fmaf(sb.y, kTH, ulo_h), s_thi = fmaf(sb.z, -kTG, uhi_h);
        const float dx = pr[0].x - sa.x, dy = pr[0].y - sa.y, dz = pr[0].z - sa.z;
        const float D = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
        const float m = fmaf(32.f * kU, D + pr[1].w, s_mc);
        const float e = pr[1].x * dx + pr[1].y * dy + pr[1].z * dz;
        const float cs = e * __builtin_amdgcn_rcpf(D);
        const float ss = __builtin_amdgcn_sqrtf(fmaxf(fmaf(-cs, cs, 1.f + 16.f * kU), 16.f * kU));
        const float lb = fmaf(-pr[0].w, ss, D - m) - sa.w;
        float T = -INFINITY;
        if (NEED & 1u) T = ud;
        if (NEED & 2u) T = fmaxf(T, fmaf(pr[2].x, kTH, s_tlo));
        if (NEED & 4u) T = fmaxf(T, fmaf(pr[2].y, -kTG, s_thi));
        return fabsf(e) > sa.w + 2.f * m && lb > T;
    };
    for_tiles<true, CULL>(
        g, t, lane, rec1, bnd1, rec2, bnd2, &stage[threadIdx.x / kWave],
        [&](const TriS &L, const TriV &LV, bool) { mc = mark_col(L, LV, ulo_h, uhi_h, M); },
        [&](const TriS &A, const TriS &B, const PairAt &p, auto &&va, auto &&vc, auto swap,
            bool skip) {
            constexpr bool col_is_a = decltype(swap)::value;
            const TriS &S = col_is_a ? B : A;
            const TriS &L = col_is_a ? A : B;
            unsigned long long m = 0;
            if (skip) sk += (uint32_t)__popcll(p.last ? p.cvl : p.cvm);
            else if constexpr (col_is_a)
                m = mark_pair(S, L, mc, M, mark_T<NEED>(S, mc, ud), vc);
            else
                m = mark_pair(S, L, mc, M, mark_T<NEED>(S, mc, ud), va);
            // word k of the block into lane k (v_writelane: the ballot's
            // scalar registers written into one lane, no compare or select)
            const uint32_t mlo = (uint32_t)m, mhi = (uint32_t)(m >> 32);
            const uint32_t kk = p.k;
            uint32_t a = wlo, b = whi;
            asm volatile("v_writelane_b32 %0, %1, m0" : "+v"(a) : "s"(mlo), "{m0}"(kk));
            asm volatile("v_writelane_b32 %0, %1, m0" : "+v"(b) : "s"(mhi), "{m0}"(kk));
            wlo = a;
            whi = b;
        },
        row_test,
        [&](uint32_t w0, uint32_t nb, uint64_t vmask) {
            // lane l holds word w0 + l, less the bits of cells outside the
            // grid; the probes' bits are cleared (they are evaluated already)
            unsigned long long x =
                (((unsigned long long)(uint32_t)whi << 32) | (unsigned long long)(uint32_t)wlo) &
                vmask;
#pragma unroll
            for (int f = 0; f < 3; f++)
                if (pword[f] == w0 + (uint32_t)lane) x &= ~pbit[f];
            if ((uint32_t)lane < nb) {
                out[w0 + lane] = x;
                cl += (uint32_t)__popcll(x);
            }
        },
        // the row's r + 2M staged in place of its cm (the guard's bound)
        [&](float4 *r, bool) { r[1].w = r[0].w + 2.f * M; });
    const uint32_t c = wave_sum_u32(cl);
    if (lane == 0) {
        count[w] = c;
        tested[w] = g.n1 * g.n2 - sk;
    }
}

// Bit masks (and SMALL voxel pairs) to the flat work list at offs[w]. One
// wave per kWave consecutive voxel pairs. Their mask words are contiguous in
// memory (woff is the exclusive scan of the word counts in voxel-pair order),
// so the wave reads them as one stream, 64 words a pass, and turns the set
// bits of each pass into items with every lane busy: lane i of an item pass
// takes set bit number base + i of the pass -- its word by binary search over
// the pass's prefix counts, its bit by halving the word -- and the word's
// voxel pair by binary search over the wave's word offsets (a voxel pair
// averages about 8 items; one pass per voxel pair left most lanes idle). A
// voxel pair's items land at offs[w] + their rank among its items: the set
// bits of the stream before them less those before its first word. SMALL
// voxel pairs (every pair an item, no words) go through passes of their own
// over the wave's concatenated pair counts. The items of a voxel pair are
// contiguous, which is all the FP64 stage's segmented fold needs.
__global__ void __launch_bounds__(kBlock) k_mx_expand(
    uint64_t n_vp, const VpPlan *__restrict__ plan, const uint64_t *__restrict__ woff,
    const unsigned long long *__restrict__ mask, const uint32_t *__restrict__ count,
    const uint64_t *__restrict__ offs, uint32_t *__restrict__ item_vp,
    uint32_t *__restrict__ item_q) {
    const uint64_t w0 = (uint64_t)wave_item() * kWave;
    if (w0 >= n_vp) return;
    const int lane = lane_id();
    const uint64_t wl = w0 + (uint64_t)lane;
    const uint32_t last = (uint32_t)min((uint64_t)(kWave - 1), n_vp - 1 - w0);
    // the lane's voxel pair: its items (SMALL), word offset and tiling
    uint32_t small = 0, nw = 0, meta = 0, n2 = 0, ngr = 1;
    uint64_t pos = 0, wo = 0;
    if ((uint32_t)lane <= last) {
        const VpPlan pl = plan[wl];
        const uint32_t c = count[wl];
        pos = offs[wl];
        wo = woff[wl];
        n2 = pl.n2;
        if (plan_flags(pl) & MX_SMALL) {
            small = c;
        } else if (!(plan_flags(pl) & MX_EMPTY)) {
            const Tiling t = plan_tiling(pl);
            nw = pl.nw;
            ngr = t.ngr;
            meta = t.cw | (t.R << 8) | (t.swap ? 1u << 16 : 0u);
        }
    }
    struct VpTab {
        uint64_t pos;    // items' offset
        uint32_t wo;     // first word, from the wave's first
        uint32_t first;  // set bits of the stream before its first word
        uint32_t meta;   // cw | R << 8 | swap << 16
        uint32_t n2, ngr;
        uint32_t sincl;  // inclusive prefix of the SMALL items
    };
    __shared__ VpTab s_vp[kWavesPerBlock][kWave];
    __shared__ unsigned long long s_word[kWavesPerBlock][kWave];
    __shared__ uint32_t s_incl[kWavesPerBlock][kWave], s_row0[kWavesPerBlock][kWave],
        s_col0[kWavesPerBlock][kWave], s_slot[kWavesPerBlock][kWave];
    const int wv = threadIdx.x / kWave;
    VpTab *const vt = s_vp[wv];
    unsigned long long *const sw = s_word[wv];
    uint32_t *const si = s_incl[wv];
    uint32_t *const sr = s_row0[wv];
    uint32_t *const sc = s_col0[wv];
    uint32_t *const ss = s_slot[wv];
    const uint64_t W0 = (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)wo) |
                        ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(wo >> 32)) << 32);
    const uint64_t wend = wo + nw;
    const uint32_t nwt =
        (uint32_t)(((uint64_t)__builtin_amdgcn_readlane((uint32_t)wend, last) |
                    ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(wend >> 32), last) << 32)) -
                   W0);
    const uint32_t sincl = wave_incl_sum_u32(small);
    const uint32_t stotal = (uint32_t)__builtin_amdgcn_readlane((int)sincl, kWave - 1);
    // lanes past the last voxel pair get an offset no word reaches, so that
    // the owner search (the last voxel pair starting at or before a word)
    // never picks them
    vt[lane] = VpTab{pos, (uint32_t)lane <= last ? (uint32_t)(wo - W0) : 0xffffffffu, 0u, meta,
                     n2, ngr, sincl};
    wave_sync();
    // SMALL voxel pairs: item q of voxel pair s is its pair q
    for (uint32_t b0 = 0; b0 < stotal; b0 += kWave) {
        const uint32_t idx = b0 + (uint32_t)lane;
        if (idx < stotal) {
            uint32_t s = 0;  // the voxel pair: the number of prefixes <= idx
#pragma unroll
            for (uint32_t st = kWave / 2; st >= 1; st >>= 1)
                if (vt[s + st - 1].sincl <= idx) s += st;
            const uint32_t q = idx - (s ? vt[s - 1].sincl : 0u);
            item_vp[vt[s].pos + q] = (uint32_t)(w0 + s);
            item_q[vt[s].pos + q] = q;
        }
    }
    // the word stream
    uint32_t run = 0;  // set bits of the stream before the pass
    for (uint32_t base = 0; base < nwt; base += kWave) {
        const uint32_t k = base + (uint32_t)lane;
        unsigned long long m = 0;
        uint32_t s = 0, kk = 1, r0 = 0, c0 = 0;
        if (k < nwt) {
            m = mask[W0 + k];
#pragma unroll
            for (uint32_t st = kWave / 2; st >= 1; st >>= 1)
                if (s + st < (uint32_t)kWave && vt[s + st].wo <= k) s += st;
            kk = k - vt[s].wo;
            const uint32_t g = vt[s].ngr, mt = vt[s].meta;
            // kk / g through the reciprocal, then one correction each way (the
            // quotient, a column chunk, is small enough for the estimate to be
            // within one)
            uint32_t ch = (uint32_t)((float)kk * __builtin_amdgcn_rcpf((float)g));
            if (ch * g > kk) ch--;
            else if ((ch + 1u) * g <= kk) ch++;
            r0 = (kk - ch * g) * ((mt >> 8) & 255u);
            c0 = ch * (mt & 255u);
        }
        const uint32_t cnt = (uint32_t)__popcll(m);
        const uint32_t incl = wave_incl_sum_u32(cnt);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
        if (k < nwt && kk == 0) vt[s].first = run + incl - cnt;
        sw[lane] = m;
        si[lane] = incl;
        sr[lane] = r0;
        sc[lane] = c0;
        ss[lane] = s;
        wave_sync();
        for (uint32_t b0 = 0; b0 < total; b0 += kWave) {
            const uint32_t idx = b0 + (uint32_t)lane;
            if (idx < total) {
                // word: the number of prefixes <= idx (< 64 as si[63] > idx)
                uint32_t W = 0;
#pragma unroll
                for (uint32_t st = kWave / 2; st >= 1; st >>= 1)
                    if (si[W + st - 1] <= idx) W += st;
                unsigned long long x = sw[W];
                uint32_t rank = idx - (si[W] - (uint32_t)__popcll(x)), b = 0;
#pragma unroll
                for (uint32_t h = 32; h >= 1; h >>= 1) {
                    const uint32_t cb = (uint32_t)__popcll(x & ((1ull << h) - 1ull));
                    if (rank >= cb) {
                        rank -= cb;
                        x >>= h;
                        b += h;
                    }
                }
                const uint32_t sl = ss[W];
                const VpTab e = vt[sl];
                const uint32_t cw = e.meta & 255u;
                // b / cw through the reciprocal: b < 64 and cw <= 64 keep
                // (b + 1/2) / cw at least 1/128 from an integer
                const uint32_t rin =
                    (uint32_t)(((float)b + 0.5f) * __builtin_amdgcn_rcpf((float)cw));
                const uint32_t row = sr[W] + rin, col = sc[W] + (b - rin * cw);
                const uint64_t p = e.pos + (uint64_t)(run + idx - e.first);
                item_vp[p] = (uint32_t)(w0 + sl);
                item_q[p] = (e.meta >> 16) & 1u ? col * e.n2 + row : row * e.n2 + col;
            }
        }
        wave_sync();
        run += total;
    }
}


// ------------------------------------------------------ persistent drain

// queue of (voxel pair, triangle pair) items per wave
constexpr uint32_t kDrainQueue = 2 * kWave;

// Persistent FP64 stage: every wave takes every G-th voxel pair (G waves in
// the grid) until none are left, turns each one's
// candidate bit masks (or, for SMALL voxel pairs, every pair) into items in a
// wave-local LDS queue, and evaluates the queue 64 items at a time -- full
// lanes however few candidates a voxel pair has, and no global work list.
// Minima fold as in k_mx_eval (the queue keeps voxel-pair order).
__global__ void __launch_bounds__(kBlock) k_mx_drain(
    uint64_t n_vp, const VpPlan *__restrict__ plan, const float4 *__restrict__ rec1,
    const float4 *__restrict__ rec2, const uint64_t *__restrict__ woff,
    const unsigned long long *__restrict__ mask, unsigned long long *__restrict__ mkey,
    unsigned long long *__restrict__ stats) {
    __shared__ uint32_t qw[kWavesPerBlock][kDrainQueue], qq[kWavesPerBlock][kDrainQueue];
    const int lane = lane_id();
    const int wv = threadIdx.x / kWave;
    uint32_t *mw = qw[wv], *mq = qq[wv];
    uint32_t qn = 0;      // items in the queue (wave-uniform)
    uint64_t evals = 0;
    const unsigned long long inf_key = dkey(INFINITY);

    // append the lanes' items whose `take` is set (at most 64)
    auto push = [&](bool take, uint32_t w, uint32_t q) {
        const uint64_t m = __ballot(take);
        if (take) {
            const uint32_t at = qn + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            mw[at] = w;
            mq[at] = q;
        }
        qn += (uint32_t)__popcll(m);
    };
    // the voxel pair being queued: SMALL pairs go as they are; otherwise its
    // mask words are taken 64 at a time (one per lane) and their set bits
    // popped one per lane and round
    __shared__ uint16_t tabs[kWavesPerBlock][kWave];
    uint16_t *tab = tabs[wv];
    uint32_t cur = kNoProbe, n2 = 1, total = 0, k = 0, nw = 0, wb = 0;
    uint32_t row0 = 0, col0 = 0;   // this lane's word: first row and column of its tile
    unsigned long long m = 0;      // this lane's word: bits not yet queued
    bool small = false, done = false;
    Tiling t{};
    const unsigned long long *in = nullptr;
    // voxel pairs interleaved over the waves of the grid (a shared atomic
    // counter would serialise on one address)
    uint32_t wnext = __builtin_amdgcn_readfirstlane((uint32_t)wave_global());
    const uint32_t wstride = gridDim.x * kWavesPerBlock;
    auto load_group = [&]() {
        const uint32_t kw = wb + lane;
        m = kw < nw ? in[kw] : 0ull;
        const uint32_t ch = kw / t.ngr, gr = kw - ch * t.ngr;
        row0 = gr * t.R;
        col0 = ch * t.cw;
    };
    for (;;) {
        // fill the queue to 64 items or until the work is gone
        while (qn < (uint32_t)kWave && !done) {
            if (cur == kNoProbe) {
                const uint32_t w = wnext;
                wnext += wstride;
                if (w >= n_vp) {
                    done = true;
                    break;
                }
                const VpPlan pl = plan[w];
                const uint32_t fl = plan_flags(pl);
                if (fl & MX_EMPTY) continue;
                cur = w;
                n2 = pl.n2;
                small = (fl & MX_SMALL) != 0;
                if (small) {
                    total = pl.n1 * pl.n2;
                    k = 0;
                } else {
                    t = plan_tiling(pl);
                    nw = pl.nw;
                    in = mask + woff[w];
                    wb = 0;
                    wave_sync();
                    const uint32_t r = (uint32_t)lane / t.cw;
                    tab[lane] = (uint16_t)((r << 8) | ((uint32_t)lane - r * t.cw));
                    wave_sync();
                    load_group();
                }
            }
            if (small) {
                push(k + lane < total, cur, k + lane);
                k += kWave;
                if (k >= total) cur = kNoProbe;
            } else if (!__any(m != 0ull)) {
                wb += kWave;
                if (wb >= nw) cur = kNoProbe;
                else load_group();
            } else {
                const bool has = m != 0ull;
                uint32_t q = 0;
                if (has) {
                    const uint32_t b = (uint32_t)__ffsll((long long)m) - 1u;
                    m &= m - 1ull;
                    const uint32_t e = tab[b];
                    const uint32_t row = row0 + (e >> 8), col = col0 + (e & 255u);
                    q = t.swap ? col * n2 + row : row * n2 + col;
                }
                push(has, cur, q);
            }
        }
        if (qn == 0) break;
        // evaluate the first (at most 64) queued items, then shift the rest down
        const uint32_t cnt = min(qn, (uint32_t)kWave);
        wave_sync();
        uint32_t w = kNoProbe;
        unsigned long long a = inf_key, b = inf_key, c = inf_key;
        if ((uint32_t)lane < cnt) {
            w = mw[lane];
            const uint32_t q = mq[lane];
            const VpPlan &pl = plan[w];
            const VpLod g{pl.o1, pl.o2, 0u, pl.n2};
            const uint32_t i = q / g.n2, j = q - i * g.n2;
            double d, lo, hi;
            pair_values<double>(rec1, g.o1 + i, rec2, g.o2 + j, d, lo, hi);
            a = dkey(d);
            b = dkey(lo);
            c = dkey(hi);
        }
        seg_min_fold(w, a, b, c, mkey);
        evals += cnt;
        const uint32_t rest = qn - cnt;
        uint32_t mw2 = 0, mq2 = 0;
        if ((uint32_t)lane < rest) {
            mw2 = mw[cnt + lane];
            mq2 = mq[cnt + lane];
        }
        wave_sync();
        if ((uint32_t)lane < rest) {
            mw[lane] = mw2;
            mq[lane] = mq2;
        }
        qn = rest;
    }
    if (lane == 0 && evals) stat_add(stats, 2, (unsigned long long)evals);
}

// per LOD slot: the mask words every voxel pair of the list would need --
// a bound for any later round, whose list is a subset (one atomic per block
// and slot)
__global__ void __launch_bounds__(kBlock) k_words_bound(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint64_t *__restrict__ off1, const uint32_t *__restrict__ cnt1,
    const uint64_t *__restrict__ off2, const uint32_t *__restrict__ cnt2, uint32_t max_rows,
    unsigned long long *__restrict__ words5) {
    __shared__ unsigned long long part[5][kWavesPerBlock];
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int s = 0; s < 5; s++) {
        uint64_t n = 0;
        if (w < n_vp) {
            const VpLod g = vp_lod(vp_v1[w], vp_v2[w], off1, cnt1, off2, cnt2, s);
            if (g.n1 && g.n2) n = mx_words(mx_tiling(g, max_rows));
        }
        n = wave_sum_u64(n);
        if (lane_id() == 0) part[s][threadIdx.x / kWave] = n;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        unsigned long long a = 0;
        for (int k = 0; k < kWavesPerBlock; k++) a += part[threadIdx.x][k];
        if (a) atomicAdd(&words5[threadIdx.x], a);
    }
}

__global__ void k_zero_u32(uint32_t *__restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Minima of two keys over the lanes of one segment (lanes sorted by seg;
// 0xffffffff marks an empty lane), folded with one atomic per segment.
template <typename K>
__device__ __forceinline__ void seg_min2(uint32_t seg, K a, K b, K none, K *__restrict__ ka,
                                         K *__restrict__ kb) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t os = (uint32_t)__shfl_down((int)seg, o, kWave);
        const K oa = __shfl_down(a, o, kWave);
        const K ob = __shfl_down(b, o, kWave);
        if (lane + o < kWave && os == seg) {
            a = min(a, oa);
            b = min(b, ob);
        }
    }
    const uint32_t ps = (uint32_t)__shfl_up((int)seg, 1, kWave);
    if ((lane == 0 || ps != seg) && seg != 0xffffffffu) {
        if (a != none) atomicMin(&ka[seg], a);
        if (b != none) atomicMin(&kb[seg], b);
    }
}

// fold the exact minima into the voxel pair's range (same rules as k_refine);
// voxel pairs of one object pair are contiguous in the list, so the object
// accumulators take one atomic per object pair and wave
template <typename R>
__global__ void __launch_bounds__(kBlock) k_mx_fold(
    uint64_t n_vp, const uint32_t *__restrict__ vp_v1, const uint32_t *__restrict__ vp_v2,
    const uint32_t *__restrict__ vp_pair, R *__restrict__ vp_min, R *__restrict__ vp_max,
    const uint32_t *__restrict__ cnt1, const uint32_t *__restrict__ cnt2,
    const uint32_t *__restrict__ vobj1, const uint32_t *__restrict__ vobj2,
    const float *__restrict__ ohd1, const float *__restrict__ ohd2, int slot, int is_last,
    int level, const uint32_t *__restrict__ flags, const unsigned long long *__restrict__ mkey,
    typename Key<R>::type *__restrict__ acc_min, typename Key<R>::type *__restrict__ acc_max,
    unsigned long long *__restrict__ stats, const uint32_t *__restrict__ tested) {
    using K = typename Key<R>::type;
    const K none = Key<R>::enc((R)__builtin_inff());
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t seg = 0xffffffffu;
    K kmin = none, kmax = none;
    uint32_t tl = 0;  // the mark pass's pairs tested one by one
    if (w < n_vp) {
        tl = tested[w];
        const uint32_t v1 = vp_v1[w], v2 = vp_v2[w];
        seg = vp_pair[w];
        const R omin = vp_min[w], omax = vp_max[w];
        if (flags[w] & MX_EMPTY) {
            // an empty voxel keeps its range (the intersection join: no values)
            if (!(level & kFoldIntersect)) kmin = Key<R>::enc(omin);
        } else {
            const R d = (R)dunkey(mkey[3 * w]), lo = (R)dunkey(mkey[3 * w + 1]),
                    hi = (R)dunkey(mkey[3 * w + 2]);
            R nmin = omin, nmax = omax;
            ObjSlack<R> x;
            if ((level & 3) == 1) x = obj_slack<R>(ohd1, ohd2, vobj1[v1], vobj2[v2], slot);
            fold_range<R>(is_last, level, d, lo, hi, x, nmin, nmax);
            vp_min[w] = nmin;
            vp_max[w] = nmax;
            kmin = Key<R>::enc(nmin);
            kmax = Key<R>::enc(nmax);
        }
    }
    seg_min2<K>(seg, kmin, kmax, none, acc_min, acc_max);
    // the tested pairs, block-reduced into counter 3 (after the fold, so that
    // the block's barrier does not hold back its loads)
    __shared__ unsigned long long part[kWavesPerBlock];
    const uint64_t tv = wave_sum_u64((uint64_t)tl);
    if (lane_id() == 0) part[threadIdx.x / kWave] = tv;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0;
        for (int k = 0; k < kWavesPerBlock; k++) a += part[k];
        if (a) stat_add(stats, 3, a);
    }
}

}  // namespace tdp
