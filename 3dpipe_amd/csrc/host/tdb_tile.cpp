// Tile model, RAW .dt reader/writer and the seeded synthetic generator.
#include "tdb_tile.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace tdb {

void Tile::finalize_boxes() {
    obj_box.assign(6 * n_obj, 0.f);
    for (size_t o = 0; o < n_obj; o++) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t v = obj_vox_off[o]; v < obj_vox_off[o + 1]; v++)
            for (int d = 0; d < 3; d++) {
                lo[d] = std::min(lo[d], vox_box[9 * v + d]);
                hi[d] = std::max(hi[d], vox_box[9 * v + 3 + d]);
            }
        for (int d = 0; d < 3; d++) {
            obj_box[6 * o + d] = lo[d];
            obj_box[6 * o + 3 + d] = hi[d];
        }
    }
}

// ------------------------------------------------------------------ RAW I/O
//
// Layout (src/storage/tile_IO.cpp:79-172 writer, :13-49 + himesh_wrapper.cpp:17-85 reader):
//   u8 type = 1
//   per object: u64 data_size; data (per LOD slot, per voxel: float tri[9n], float hd[2n]);
//               u64 voxel_count; 5 x {f32 hausdorff, f32 proxy};
//               voxel_count x {f32 low[3], high[3], core[3]; 5 x {u64 offset, u64 count}};
//               f32 transform[3]

bool read_raw(const std::string &path, Tile &t, std::string &err) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) { err = "cannot open " + path; return false; }
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<char> buf(sz > 0 ? (size_t)sz : 0);
    size_t got = sz > 0 ? std::fread(buf.data(), 1, buf.size(), f) : 0;
    std::fclose(f);
    if (sz < 1 || got != buf.size()) { err = "short read: " + path; return false; }
    if (buf[0] != 1) { err = "not a RAW tile (type byte " + std::to_string((int)buf[0]) + "); COMPRESSED tiles need the PPMC decoder, which this library does not include"; return false; }
    t = Tile();
    t.obj_vox_off.push_back(0);
    size_t pos = 1;
    // every size read from the file is checked against the bytes left
    // (subtraction form: a huge count cannot wrap the position back)
    auto rd = [&](void *dst, size_t n) {
        if (n > buf.size() - pos) return false;
        std::memcpy(dst, buf.data() + pos, n);
        pos += n;
        return true;
    };
    while (pos < buf.size()) {
        uint64_t dsize = 0, nvox = 0;
        if (!rd(&dsize, 8) || dsize > buf.size() - pos) { err = "truncated object data"; return false; }
        const char *data = buf.data() + pos;
        pos += dsize;
        float hd[10];
        if (!rd(&nvox, 8) || !rd(hd, 40)) { err = "truncated object meta"; return false; }
        // each voxel's meta is 36 + 80 bytes; voxel ids are 32-bit
        if (nvox > (buf.size() - pos) / 116) { err = "truncated voxel meta"; return false; }
        if (nvox > (uint64_t)UINT32_MAX - t.n_vox()) { err = "too many voxels"; return false; }
        t.obj_hd.insert(t.obj_hd.end(), hd, hd + 10);
        size_t v0 = t.n_vox();
        std::vector<std::array<uint64_t, 10>> lodinfo(nvox);
        for (uint64_t v = 0; v < nvox; v++) {
            float box[9];
            if (!rd(box, 36) || !rd(lodinfo[v].data(), 80)) { err = "truncated voxel meta"; return false; }
            t.vox_box.insert(t.vox_box.end(), box, box + 9);
            t.vox_obj.push_back((uint32_t)t.n_obj);
        }
        float tr[3];
        if (!rd(tr, 12)) { err = "truncated transform"; return false; }
        t.obj_transform.insert(t.obj_transform.end(), tr, tr + 3);
        // triangle records, ordered LOD slot -> voxel like the data block
        t.vox_tri_off.resize(5 * (v0 + nvox));
        t.vox_tri_cnt.resize(5 * (v0 + nvox));
        for (int s = 0; s < kLods; s++)
            for (uint64_t v = 0; v < nvox; v++) {
                uint64_t off = lodinfo[v][2 * s], cnt = lodinfo[v][2 * s + 1];
                if (off > dsize || cnt > (dsize - off) / 44) { err = "voxel block outside object data"; return false; }
                if (cnt > UINT32_MAX) { err = "voxel triangle count beyond 32 bits"; return false; }
                t.vox_tri_off[5 * (v0 + v) + s] = t.n_tri();
                t.vox_tri_cnt[5 * (v0 + v) + s] = (uint32_t)cnt;
                // the block sits at any byte offset of the file: copied, not
                // read through float pointers
                const char *tp = data + off, *hp = tp + 36 * cnt;
                for (uint64_t i = 0; i < cnt; i++) {
                    float tri[9], hd2[2];
                    std::memcpy(tri, tp + 36 * i, 36);
                    std::memcpy(hd2, hp + 8 * i, 8);
                    const float r[12] = {tri[0], tri[1], tri[2], hd2[0], tri[3], tri[4],
                                         tri[5], hd2[1], tri[6], tri[7], tri[8], 0.f};
                    t.tri.insert(t.tri.end(), r, r + 12);
                }
            }
        t.n_obj++;
        t.obj_vox_off.push_back((uint32_t)t.n_vox());
    }
    t.finalize_boxes();
    return true;
}

bool write_raw(const std::string &path, const Tile &t, std::string &err) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) { err = "cannot create " + path; return false; }
    char type = 1;
    bool ok = std::fwrite(&type, 1, 1, f) == 1;
    std::vector<char> data;
    for (size_t o = 0; o < t.n_obj && ok; o++) {
        uint32_t va = t.obj_vox_off[o], vb = t.obj_vox_off[o + 1];
        std::vector<uint64_t> off(5 * (vb - va)), cnt(5 * (vb - va));
        data.clear();
        for (int s = 0; s < kLods; s++)
            for (uint32_t v = va; v < vb; v++) {
                uint64_t n = t.vox_tri_cnt[5 * v + s], first = t.vox_tri_off[5 * v + s];
                off[5 * (v - va) + s] = data.size();
                cnt[5 * (v - va) + s] = n;
                size_t base = data.size();
                data.resize(base + 44 * n);
                float *tp = (float *)(data.data() + base);
                float *hp = tp + 9 * n;
                for (uint64_t i = 0; i < n; i++) {
                    const float *r = &t.tri[12 * (first + i)];
                    const float tri9[9] = {r[0], r[1], r[2], r[4], r[5], r[6], r[8], r[9], r[10]};
                    std::memcpy(tp + 9 * i, tri9, 36);
                    hp[2 * i] = r[3];
                    hp[2 * i + 1] = r[7];
                }
            }
        uint64_t dsize = data.size(), nvox = vb - va;
        ok = ok && std::fwrite(&dsize, 8, 1, f) == 1;
        ok = ok && (dsize == 0 || std::fwrite(data.data(), 1, dsize, f) == dsize);
        ok = ok && std::fwrite(&nvox, 8, 1, f) == 1;
        ok = ok && std::fwrite(&t.obj_hd[10 * o], 4, 10, f) == 10;
        for (uint32_t v = va; v < vb && ok; v++) {
            ok = std::fwrite(&t.vox_box[9 * v], 4, 9, f) == 9;
            for (int s = 0; s < kLods && ok; s++) {
                ok = std::fwrite(&off[5 * (v - va) + s], 8, 1, f) == 1 &&
                     std::fwrite(&cnt[5 * (v - va) + s], 8, 1, f) == 1;
            }
        }
        ok = ok && std::fwrite(&t.obj_transform[3 * o], 4, 3, f) == 3;
    }
    if (std::fclose(f) != 0) ok = false;
    if (!ok) err = "write failed: " + path;
    return ok;
}

// ------------------------------------------------------------- generator

namespace {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// splitmix64 stream keyed by (seed, stream id)
struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t stream) : s(mix64(seed * 0x9E3779B97F4A7C15ULL + mix64(stream + 0x632BE59BD9B4E019ULL))) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ULL;
        return mix64(s);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * uni(); }
    int below(int n) { return (int)(uni() * n) % n; }
    // standard normal by Box-Muller
    double normal() {
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) { return a * (1.0 / norm(a)); }

V3 random_unit(Rng &r) {
    V3 v{r.normal(), r.normal(), r.normal()};
    double n = norm(v);
    if (n < 1e-12) return {1, 0, 0};
    return v * (1.0 / n);
}
V3 any_perp(V3 t) {
    V3 a = std::fabs(t.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
    return unit(cross(t, a));
}
// rotate v about unit axis k by angle a (Rodrigues)
V3 rotate(V3 v, V3 k, double a) {
    double c = std::cos(a), s = std::sin(a);
    return v * c + cross(k, v) * s + k * (dot(k, v) * (1 - c));
}

// closest distance from point p to triangle (a, b, c), by region classification
double point_tri_dist(V3 p, V3 a, V3 b, V3 c) {
    V3 ab = b - a, ac = c - a, ap = p - a;
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return norm(p - a);
    V3 bp = p - b;
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return norm(p - b);
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        double v = d1 / (d1 - d3);
        return norm(p - (a + ab * v));
    }
    V3 cp = p - c;
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return norm(p - c);
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        double w = d2 / (d2 - d6);
        return norm(p - (a + ac * w));
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return norm(p - (b + (c - b) * w));
    }
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom, w = vc * denom;
    return norm(p - (a + ab * v + ac * w));
}

// A mesh with its LOD ladder: shared vertices, faces per LOD slot. LOD
// vertices are always a subset of the LOD-100 vertices.
struct LodMesh {
    std::vector<V3> v;
    std::vector<std::array<int, 3>> f[kLods];
    std::vector<uint8_t> is_orig[kLods];  // face identical to a LOD-100 face
};

struct Voxelized {
    int nvox = 0;
    std::vector<int> core_face;              // per voxel: LOD-100 face whose barycentre is the core
    std::vector<int> face_vox[kLods];        // per slot, per face: voxel
    std::vector<double> hd[kLods];           // per slot, per face: proxy, hausdorff
    std::vector<int> order[kLods];           // per slot: faces by voxel, then face index
    std::vector<uint32_t> vcnt[kLods];       // per slot, per voxel: faces
    // decode shortcuts (set_order): per slot the largest proxy / hausdorff
    // bound (-1 when the slot has no faces); per voxel the distinct vertices
    // of its LOD-100 faces (vox_vert_off[v] .. [v + 1]); the vertices LOD-100
    // faces use, ascending
    double hd_max[kLods][2] = {};
    std::vector<uint32_t> vox_vert_off, vox_vert, used_vert;
};

// the record order of an object: slot -> voxel -> face; and the decode
// shortcuts
void set_order(const LodMesh &m, Voxelized &vz) {
    for (int s = 0; s < kLods; s++) {
        const size_t nf = vz.face_vox[s].size();
        vz.vcnt[s].assign(vz.nvox, 0);
        for (size_t i = 0; i < nf; i++) vz.vcnt[s][vz.face_vox[s][i]]++;
        std::vector<uint32_t> at(vz.nvox + 1, 0);
        for (int v = 0; v < vz.nvox; v++) at[v + 1] = at[v] + vz.vcnt[s][v];
        vz.order[s].assign(nf, 0);
        for (size_t i = 0; i < nf; i++) vz.order[s][at[vz.face_vox[s][i]]++] = (int)i;
        vz.hd_max[s][0] = vz.hd_max[s][1] = -1;
        if (s < 4)
            for (size_t i = 0; i < nf; i++) {
                vz.hd_max[s][0] = std::max(vz.hd_max[s][0], vz.hd[s][2 * i]);
                vz.hd_max[s][1] = std::max(vz.hd_max[s][1], vz.hd[s][2 * i + 1]);
            }
    }
    std::vector<std::vector<uint32_t>> vv(vz.nvox);
    std::vector<uint8_t> used(m.v.size(), 0);
    for (size_t i = 0; i < m.f[4].size(); i++)
        for (int k = 0; k < 3; k++) {
            vv[vz.face_vox[4][i]].push_back((uint32_t)m.f[4][i][k]);
            used[m.f[4][i][k]] = 1;
        }
    vz.vox_vert_off.assign(1, 0);
    vz.vox_vert.clear();
    for (auto &l : vv) {
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
        vz.vox_vert.insert(vz.vox_vert.end(), l.begin(), l.end());
        vz.vox_vert_off.push_back((uint32_t)vz.vox_vert.size());
    }
    vz.used_vert.clear();
    for (size_t i = 0; i < used.size(); i++)
        if (used[i]) vz.used_vert.push_back((uint32_t)i);
}

V3 bary(const LodMesh &m, const std::array<int, 3> &f) {
    return (m.v[f[0]] + m.v[f[1]] + m.v[f[2]]) * (1.0 / 3.0);
}

// Voxel partition following the reference's rule: facets/voxel_tris + 1 voxels
// (a single voxel when that is <= 3; src/spatial/himesh_skeleton.cpp:72-185),
// cores snapped to facet barycentres, faces assigned to the nearest core.
// Per-triangle bounds (proxy, hausdorff) are conservative for the lower and
// upper distance bounds of the refinement:
//   proxy: every point of the voxel's LOD-100 surface lies within `proxy` of
//          some LOD triangle of the same voxel (exact max over the vertices of
//          each assigned original face, distance to a triangle being convex);
//   hausdorff: every point of a LOD triangle lies within `hausdorff` of the
//          LOD-100 surface (0 for unchanged faces, else longest edge / sqrt(3),
//          the vertices being LOD-100 vertices).
Voxelized voxelize(const LodMesh &m, int voxel_tris) {
    Voxelized vz;
    const auto &F = m.f[4];
    const int nf = (int)F.size();
    std::vector<V3> bc(nf);
    for (int i = 0; i < nf; i++) bc[i] = bary(m, F[i]);
    int nv = nf / voxel_tris + 1;
    if (nv <= 3) nv = 1;
    std::vector<int> cores;
    if (nv == 1) {
        V3 mean{0, 0, 0};
        for (auto &b : bc) mean = mean + b;
        mean = mean * (1.0 / nf);
        int best = 0;
        double bd = INFINITY;
        for (int i = 0; i < nf; i++) {
            double d = dot(bc[i] - mean, bc[i] - mean);
            if (d < bd) { bd = d; best = i; }
        }
        cores.push_back(best);
    } else {
        // farthest point seeding from face 0, then Lloyd iterations
        std::vector<double> dmin(nf, INFINITY);
        cores.push_back(0);
        for (int k = 1; k < nv; k++) {
            int last = cores.back(), far = 0;
            double fd = -1;
            for (int i = 0; i < nf; i++) {
                double d = dot(bc[i] - bc[last], bc[i] - bc[last]);
                if (d < dmin[i]) dmin[i] = d;
                if (dmin[i] > fd) { fd = dmin[i]; far = i; }
            }
            cores.push_back(far);
        }
        std::vector<V3> ctr(nv);
        for (int k = 0; k < nv; k++) ctr[k] = bc[cores[k]];
        std::vector<int> asg(nf);
        for (int it = 0; it < 5; it++) {
            for (int i = 0; i < nf; i++) {
                int best = 0;
                double bd = INFINITY;
                for (int k = 0; k < nv; k++) {
                    double d = dot(bc[i] - ctr[k], bc[i] - ctr[k]);
                    if (d < bd) { bd = d; best = k; }
                }
                asg[i] = best;
            }
            std::vector<V3> sum(nv, V3{0, 0, 0});
            std::vector<int> cnt(nv, 0);
            for (int i = 0; i < nf; i++) { sum[asg[i]] = sum[asg[i]] + bc[i]; cnt[asg[i]]++; }
            for (int k = 0; k < nv; k++)
                if (cnt[k]) ctr[k] = sum[k] * (1.0 / cnt[k]);
        }
        // snap to the member barycentre nearest the cluster mean
        std::vector<double> bd(nv, INFINITY);
        for (int i = 0; i < nf; i++) {
            double d = dot(bc[i] - ctr[asg[i]], bc[i] - ctr[asg[i]]);
            if (d < bd[asg[i]]) { bd[asg[i]] = d; cores[asg[i]] = i; }
        }
        std::sort(cores.begin(), cores.end());
        cores.erase(std::unique(cores.begin(), cores.end()), cores.end());
    }
    auto nearest_core = [&](V3 p) {
        int best = 0;
        double bdist = INFINITY;
        for (int k = 0; k < (int)cores.size(); k++) {
            V3 c = bc[cores[k]];
            double d = dot(p - c, p - c);
            if (d < bdist) { bdist = d; best = k; }
        }
        return best;
    };
    // final LOD-100 assignment; drop cores that end up without faces
    std::vector<int> a100(nf);
    std::vector<int> count(cores.size(), 0);
    for (int i = 0; i < nf; i++) { a100[i] = nearest_core(bc[i]); count[a100[i]]++; }
    std::vector<int> keep;
    for (int k = 0; k < (int)cores.size(); k++)
        if (count[k] > 0) keep.push_back(cores[k]);
    cores = keep;
    vz.nvox = (int)cores.size();
    vz.core_face = cores;
    for (int s = 0; s < kLods; s++) {
        vz.face_vox[s].resize(m.f[s].size());
        for (size_t i = 0; i < m.f[s].size(); i++) vz.face_vox[s][i] = nearest_core(bary(m, m.f[s][i]));
        vz.hd[s].assign(2 * m.f[s].size(), 0.0);
    }
    // bounds for LOD slots below 100
    std::vector<std::vector<int>> orig(vz.nvox);
    for (int i = 0; i < nf; i++) orig[vz.face_vox[4][i]].push_back(i);
    for (int s = 0; s < 4; s++) {
        std::vector<std::vector<int>> lod(vz.nvox);
        for (size_t i = 0; i < m.f[s].size(); i++) lod[vz.face_vox[s][i]].push_back((int)i);
        for (int v = 0; v < vz.nvox; v++) {
            if (lod[v].empty()) continue;
            std::vector<V3> lbc(lod[v].size());
            for (size_t j = 0; j < lod[v].size(); j++) lbc[j] = bary(m, m.f[s][lod[v][j]]);
            for (int T : orig[v]) {
                // the LOD triangle covering T best among the few nearest by
                // barycentre: any choice is valid, this one keeps it tight
                V3 tb = bc[T];
                constexpr int kNear = 6;
                int near[kNear] = {0};
                double nd[kNear];
                int nn = 0;
                for (size_t j = 0; j < lod[v].size(); j++) {
                    double d = dot(tb - lbc[j], tb - lbc[j]);
                    if (nn < kNear) { near[nn] = (int)j; nd[nn++] = d; }
                    else if (d < nd[kNear - 1]) { near[kNear - 1] = (int)j; nd[kNear - 1] = d; }
                    else continue;
                    for (int a = nn - 1; a > 0 && nd[a] < nd[a - 1]; a--) {
                        std::swap(nd[a], nd[a - 1]);
                        std::swap(near[a], near[a - 1]);
                    }
                }
                int bj = near[0];
                double best = INFINITY;
                for (int c = 0; c < nn; c++) {
                    const auto &t = m.f[s][lod[v][near[c]]];
                    double r = 0;
                    for (int k = 0; k < 3; k++)
                        r = std::max(r, point_tri_dist(m.v[F[T][k]], m.v[t[0]], m.v[t[1]], m.v[t[2]]));
                    if (r < best) { best = r; bj = near[c]; }
                }
                double &ph = vz.hd[s][2 * lod[v][bj]];
                ph = std::max(ph, best);
            }
        }
        // hausdorff of a LOD triangle: distance from a barycentric grid of
        // samples (m subdivisions per edge) to the voxel's LOD-100 faces (a
        // subset of the surface, so an upper bound), plus the distance from any
        // point of the triangle to the nearest sample (longest edge / (m sqrt 3));
        // never above longest edge / sqrt 3 (its vertices lie on the surface)
        for (size_t i = 0; i < m.f[s].size(); i++) {
            if (m.is_orig[s][i]) continue;
            const auto &t = m.f[s][i];
            const V3 a = m.v[t[0]], b = m.v[t[1]], c = m.v[t[2]];
            const double e = std::max({norm(b - a), norm(c - b), norm(a - c)});
            double h = e / std::sqrt(3.0);
            const int sub = std::min(6, (int)std::ceil(e / 3.0));
            if (sub >= 2) {
                const auto &ov = orig[vz.face_vox[s][i]];
                double worst = 0;
                for (int p = 0; p <= sub && worst < h; p++)
                    for (int q = 0; p + q <= sub; q++) {
                        const double u = (double)p / sub, w = (double)q / sub;
                        const V3 x = a + (b - a) * u + (c - a) * w;
                        double dmin = INFINITY;
                        for (int T : ov)
                            dmin = std::min(dmin, point_tri_dist(x, m.v[F[T][0]], m.v[F[T][1]],
                                                                 m.v[F[T][2]]));
                        worst = std::max(worst, dmin);
                    }
                h = std::min(h, worst + e / (sub * std::sqrt(3.0)));
            }
            vz.hd[s][2 * i + 1] = h;
        }
    }
    set_order(m, vz);
    return vz;
}

// Where one object's arrays go: either straight into the tile or into
// per-object scratch vectors.
struct ObjSink {
    float *vox_box;      // 9 per voxel
    uint32_t *cnt;       // 5 per voxel (triangles per LOD slot)
    float *rec;          // 12 per triangle, slot -> voxel order
    float *hd;           // 10
    float *transform;    // 3
};

inline float ulpf(float x) {
    float a = std::fabs(x);
    return std::nextafter(a, INFINITY) - a;
}

size_t total_faces(const LodMesh &m) {
    size_t n = 0;
    for (int s = 0; s < kLods; s++) n += m.f[s].size();
    return n;
}

// Voxel boxes from the LOD-100 faces' vertices, cores from the core face's
// vertices, and the max-distance proxy point: the LOD-100 vertex nearest the
// box centre (the first such in vertex order).
void emit_boxes(const LodMesh &m, const std::vector<float> &vf, const Voxelized &vz,
                float *vox_box, float *transform) {
    const int nv = vz.nvox;
    auto P = [&](int i, int d) { return vf[3 * i + d]; };
    for (int v = 0; v < nv; v++) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t q = vz.vox_vert_off[v]; q < vz.vox_vert_off[v + 1]; q++)
            for (int d = 0; d < 3; d++) {
                const float c = P((int)vz.vox_vert[q], d);
                lo[d] = std::min(lo[d], c);
                hi[d] = std::max(hi[d], c);
            }
        for (int d = 0; d < 3; d++) {
            vox_box[9 * v + d] = lo[d];
            vox_box[9 * v + 3 + d] = hi[d];
        }
    }
    for (int v = 0; v < nv; v++) {
        const auto &cf = m.f[4][vz.core_face[v]];
        for (int d = 0; d < 3; d++) {
            double s = ((double)P(cf[0], d) + (double)P(cf[1], d) + (double)P(cf[2], d)) / 3.0;
            vox_box[9 * v + 6 + d] = (float)s;
        }
    }
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i : vz.used_vert)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], P((int)i, d));
            hi[d] = std::max(hi[d], P((int)i, d));
        }
    float c[3] = {0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2])};
    float best = INFINITY;
    int bi = 0;
    for (uint32_t i : vz.used_vert) {
        float dx = P((int)i, 0) - c[0], dy = P((int)i, 1) - c[1], dz = P((int)i, 2) - c[2];
        float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < best) { best = d2; bi = (int)i; }
    }
    for (int d = 0; d < 3; d++) transform[d] = P(bi, d);
}

// The margin of the records' Hausdorff bounds: covers rounding the vertices
// to float (bounds in `vz` are multiplied by `scale`).
double record_margin(const std::vector<float> &vf, double scale) {
    float maxc = 0;
    for (float c : vf) maxc = std::max(maxc, std::fabs(c));
    return 8.0 * ulpf(maxc) + 1e-6 * scale;
}

// the (proxy hausdorff, hausdorff) pair of face i at slot s (0 at LOD 100)
inline void face_hd(const Voxelized &vz, int s, int i, double scale, double margin, float &ph,
                    float &h) {
    ph = 0;
    h = 0;
    if (s < 4) {
        ph = (float)(vz.hd[s][2 * i] * scale + margin);
        h = (float)(vz.hd[s][2 * i + 1] * scale + margin);
    }
}

// Per-voxel counts and the object Hausdorff pairs: the largest of the
// records' (hausdorff, proxy) values, 0 at LOD 100 and for a slot without
// faces. A record's value is the float of (bound * scale + margin), a
// non-decreasing map of the bound, so the largest record value is the map
// of the largest bound.
void emit_counts_hd(const Voxelized &vz, double scale, double margin, uint32_t *cnt, float *hd) {
    for (int s = 0; s < kLods; s++) {
        hd[2 * s] = 0;
        hd[2 * s + 1] = 0;
        for (int v = 0; v < vz.nvox; v++) cnt[5 * v + s] = vz.vcnt[s][v];
        if (s < 4 && !vz.order[s].empty()) {
            hd[2 * s] = std::max(0.f, (float)(vz.hd_max[s][1] * scale + margin));      // hausdorff
            hd[2 * s + 1] = std::max(0.f, (float)(vz.hd_max[s][0] * scale + margin));  // proxy
        }
    }
}

// the triangle records of slot s (voxel -> face order)
// the records of faces first .. first + n - 1 of slot s's record order
void emit_slot_range(const LodMesh &m, const float *vf, const Voxelized &vz, double scale,
                     double margin, int s, size_t first, size_t n, float *rec) {
    auto P = [&](int i, int d) { return vf[3 * i + d]; };
    size_t w = 0;
    for (size_t q = first; q < first + n; q++) {
        const int i = vz.order[s][q];
        const auto &f = m.f[s][i];
        float ph, h;
        face_hd(vz, s, i, scale, margin, ph, h);
        float r[12] = {P(f[0], 0), P(f[0], 1), P(f[0], 2), ph, P(f[1], 0), P(f[1], 1),
                       P(f[1], 2), h,          P(f[2], 0), P(f[2], 1), P(f[2], 2), 0.f};
        std::memcpy(rec + 12 * w, r, sizeof r);
        w++;
    }
}

void emit_slot(const LodMesh &m, const float *vf, const Voxelized &vz, double scale,
               double margin, int s, float *rec) {
    emit_slot_range(m, vf, vz, scale, margin, s, 0, vz.order[s].size(), rec);
}

// Triangle records (slot -> voxel -> face), per-voxel counts and the object
// Hausdorff pairs; bounds in `vz` are multiplied by `scale` and widened by a
// margin that covers rounding the vertices to float.
void emit_records(const LodMesh &m, const std::vector<float> &vf, const Voxelized &vz,
                  double scale, uint32_t *cnt, float *rec, float *hd) {
    const double margin = record_margin(vf, scale);
    emit_counts_hd(vz, scale, margin, cnt, hd);
    size_t w = 0;
    for (int s = 0; s < kLods; s++) {
        emit_slot(m, vf.data(), vz, scale, margin, s, rec + 12 * w);
        w += vz.order[s].size();
    }
}

// Build the object's arrays from float vertices and a voxel partition.
void emit_object(const LodMesh &m, const std::vector<float> &vf, const Voxelized &vz,
                 double scale, const ObjSink &out) {
    emit_boxes(m, vf, vz, out.vox_box, out.transform);
    emit_records(m, vf, vz, scale, out.cnt, out.rec, out.hd);
}

// Per-object layout inside a tile: voxel and triangle counts are known
// before any geometry is produced, so objects are written in place.
struct ObjShape {
    uint32_t nvox;
    uint64_t ntri;
};

void layout(Tile &t, const std::vector<ObjShape> &shape, std::vector<uint64_t> &tri_off) {
    t = Tile();
    t.n_obj = shape.size();
    t.obj_vox_off.assign(shape.size() + 1, 0);
    tri_off.assign(shape.size() + 1, 0);
    for (size_t o = 0; o < shape.size(); o++) {
        t.obj_vox_off[o + 1] = t.obj_vox_off[o] + shape[o].nvox;
        tri_off[o + 1] = tri_off[o] + shape[o].ntri;
    }
    size_t nvox = t.obj_vox_off.back();
    t.vox_box.resize(9 * nvox);
    t.vox_obj.resize(nvox);
    t.vox_tri_off.resize(5 * nvox);
    t.vox_tri_cnt.resize(5 * nvox);
    t.tri.resize(12 * tri_off.back());
    t.obj_hd.resize(10 * shape.size());
    t.obj_transform.resize(3 * shape.size());
}

ObjSink sink_for(Tile &t, const std::vector<uint64_t> &tri_off, size_t o) {
    uint32_t v0 = t.obj_vox_off[o];
    return ObjSink{&t.vox_box[9 * v0], &t.vox_tri_cnt[5 * v0], &t.tri[12 * tri_off[o]],
                   &t.obj_hd[10 * o], &t.obj_transform[3 * o]};
}

// vox_tri_cnt holds per-voxel counts as [voxel][slot]; turn them into the
// record index of each (voxel, slot) block, objects laid out slot -> voxel
void finish_object(Tile &t, const std::vector<uint64_t> &tri_off, size_t o) {
    uint32_t v0 = t.obj_vox_off[o], nv = t.obj_vox_off[o + 1] - v0;
    uint64_t cur = tri_off[o];
    for (int s = 0; s < kLods; s++)
        for (uint32_t v = 0; v < nv; v++) {
            t.vox_obj[v0 + v] = (uint32_t)o;
            t.vox_tri_off[5 * (v0 + v) + s] = cur;
            cur += t.vox_tri_cnt[5 * (v0 + v) + s];
        }
}

// ---- icosphere with a nested LOD ladder
//
// Subdivision level 3 of an icosahedron (642 vertices, 1280 faces). Lower LODs
// replace groups of 4 sibling faces by their parent (which reduces the count
// by 3) following one seeded permutation, so LODs are nested:
// 1280 / 1025 / 767 / 512 / 257 faces for 100 / 80 / 60 / 40 / 20 %.
struct IcoTopo {
    int levels;
    std::vector<V3> dir;                              // unit vertex directions
    std::vector<std::vector<std::array<int, 3>>> f;   // faces per subdivision level
};

IcoTopo make_ico(int subdiv) {
    IcoTopo T;
    T.levels = subdiv;
    const double p = (1.0 + std::sqrt(5.0)) / 2.0;
    double vs[12][3] = {{-1, p, 0}, {1, p, 0}, {-1, -p, 0}, {1, -p, 0}, {0, -1, p}, {0, 1, p},
                        {0, -1, -p}, {0, 1, -p}, {p, 0, -1}, {p, 0, 1}, {-p, 0, -1}, {-p, 0, 1}};
    for (auto &v : vs) T.dir.push_back(unit(V3{v[0], v[1], v[2]}));
    int fs[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                     {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                     {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                     {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    T.f.emplace_back();
    for (auto &x : fs) T.f[0].push_back({x[0], x[1], x[2]});
    std::map<std::pair<int, int>, int> mid;
    auto midpoint = [&](int a, int b) {
        auto key = std::make_pair(std::min(a, b), std::max(a, b));
        auto it = mid.find(key);
        if (it != mid.end()) return it->second;
        T.dir.push_back(unit(T.dir[a] + T.dir[b]));
        int id = (int)T.dir.size() - 1;
        mid[key] = id;
        return id;
    };
    for (int l = 0; l < subdiv; l++) {
        std::vector<std::array<int, 3>> nf;
        for (auto &t : T.f[l]) {
            int ab = midpoint(t[0], t[1]), bc = midpoint(t[1], t[2]), ca = midpoint(t[2], t[0]);
            nf.push_back({t[0], ab, ca});
            nf.push_back({t[1], bc, ab});
            nf.push_back({t[2], ca, bc});
            nf.push_back({ab, bc, ca});
        }
        T.f.push_back(nf);
    }
    return T;
}

void shuffle(std::vector<int> &a, Rng &r) {
    for (int i = (int)a.size() - 1; i > 0; i--) std::swap(a[i], a[r.below(i + 1)]);
}

LodMesh make_nucleus_proto(const IcoTopo &T, double noise, Rng &r) {
    LodMesh m;
    m.v.resize(T.dir.size());
    for (size_t i = 0; i < T.dir.size(); i++) m.v[i] = T.dir[i] * (1.0 + r.uni(-noise, noise));
    const int L = T.levels;  // finest level index
    const int nfine = (int)T.f[L].size();
    const double frac[kLods] = {0.2, 0.4, 0.6, 0.8, 1.0};
    // permutations of coarsening groups at each level above the finest
    std::vector<std::vector<int>> rank(L);
    for (int l = 0; l < L; l++) {
        std::vector<int> perm(T.f[l].size());
        for (size_t i = 0; i < perm.size(); i++) perm[i] = (int)i;
        shuffle(perm, r);
        rank[l].assign(perm.size(), 0);
        for (size_t i = 0; i < perm.size(); i++) rank[l][perm[i]] = (int)i;
    }
    for (int s = 0; s < kLods; s++) {
        int target = (int)std::lround(frac[s] * nfine);
        // number of coarsened groups per level, finest groups first
        std::vector<int> k(L, 0);
        int count = nfine;
        for (int l = L - 1; l >= 0 && count > target; l--) {
            int groups = (int)T.f[l].size();
            int need = (int)std::lround((count - target) / 3.0);
            k[l] = std::min(groups, need);
            count -= 3 * k[l];
            if (k[l] < groups) break;
        }
        // a face at level l is emitted if its group is not coarsened and every
        // ancestor group is coarsened... walk down from level 0
        std::function<void(int, int)> emit = [&](int l, int face) {
            bool coarse = (l < L) && (rank[l][face] < k[l]);
            if (l == L || coarse) {
                m.f[s].push_back(T.f[l][face]);
                m.is_orig[s].push_back(l == L);
                return;
            }
            for (int c = 0; c < 4; c++) emit(l + 1, 4 * face + c);
        };
        for (size_t face = 0; face < T.f[0].size(); face++) emit(0, (int)face);
    }
    return m;
}

// ---- tube vessel swept along a random polyline
struct Vessel {
    std::vector<V3> pts;   // polyline vertices (segments + 1)
    LodMesh mesh;
};

// with_mesh = false: the centreline only (the mesh takes no random draws)
Vessel make_vessel(const GenParams &p, Rng &r, bool with_mesh = true) {
    Vessel V;
    const int segs = p.vessel_segments, S = p.vessel_sides;
    const int rps = p.rings_min + (p.rings_max > p.rings_min ? r.below(p.rings_max - p.rings_min + 1) : 0);
    const double radius = r.uni(p.vessel_r_min, p.vessel_r_max);
    V3 start{r.uni(0, p.cube), r.uni(0, p.cube), r.uni(0, p.cube)};
    std::vector<V3> dirs;
    dirs.push_back(random_unit(r));
    for (int s = 1; s < segs; s++) {
        V3 axis = random_unit(r);
        axis = axis - dirs.back() * dot(axis, dirs.back());
        if (norm(axis) < 1e-9) axis = any_perp(dirs.back());
        axis = unit(axis);
        dirs.push_back(unit(rotate(dirs.back(), axis, r.uni(0, p.turn_max_deg) * M_PI / 180.0)));
    }
    V.pts.push_back(start);
    for (int s = 0; s < segs; s++) V.pts.push_back(V.pts.back() + dirs[s] * p.segment_length);
    if (!with_mesh) return V;
    const int R = segs * rps + 1;
    std::vector<V3> ctr(R), tan(R);
    for (int j = 0; j < R; j++) {
        int s = std::min(j / rps, segs - 1);
        double u = (j - s * rps) * (p.segment_length / rps);
        ctr[j] = V.pts[s] + dirs[s] * u;
        tan[j] = dirs[s];
        if (j % rps == 0 && j > 0 && j < R - 1) tan[j] = unit(dirs[s - 1] + dirs[s]);
    }
    V3 N = any_perp(tan[0]);
    auto &m = V.mesh;
    for (int j = 0; j < R; j++) {
        N = unit(N - tan[j] * dot(N, tan[j]));
        V3 B = cross(tan[j], N);
        for (int k = 0; k < S; k++) {
            double a = 2.0 * M_PI * k / S;
            m.v.push_back(ctr[j] + (N * std::cos(a) + B * std::sin(a)) * radius);
        }
    }
    const double frac[kLods] = {0.2, 0.4, 0.6, 0.8, 1.0};
    for (int s = 0; s < kLods; s++) {
        int ks = frac[s] > 0.5 ? 1 : 2;
        if (S % ks) ks = 1;
        int gaps = (int)std::lround(frac[s] * (R - 1) * ks);
        gaps = std::max(1, std::min(R - 1, gaps));
        if (s == 4) { ks = 1; gaps = R - 1; }
        std::vector<int> ring(gaps + 1);
        for (int i = 0; i <= gaps; i++) ring[i] = (int)std::lround((double)i * (R - 1) / gaps);
        const int Sl = S / ks;
        for (int i = 0; i < gaps; i++) {
            int j0 = ring[i], j1 = ring[i + 1];
            bool orig = (ks == 1) && (j1 == j0 + 1);
            for (int q = 0; q < Sl; q++) {
                int k0 = q * ks, k1 = ((q + 1) % Sl) * ks;
                int a = j0 * S + k0, b = j0 * S + k1, c = j1 * S + k0, d = j1 * S + k1;
                m.f[s].push_back({a, b, c});
                m.f[s].push_back({b, d, c});
                m.is_orig[s].push_back(orig);
                m.is_orig[s].push_back(orig);
            }
        }
    }
    return V;
}

// uniformly random rotation matrix (from a random unit quaternion)
void random_rotation(Rng &r, double R[9]) {
    double q[4] = {r.normal(), r.normal(), r.normal(), r.normal()};
    double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (double &x : q) x /= n;
    double w = q[0], x = q[1], y = q[2], z = q[3];
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - z * w);     R[2] = 2 * (x * z + y * w);
    R[3] = 2 * (x * y + z * w);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - x * w);
    R[6] = 2 * (x * z - y * w);     R[7] = 2 * (y * z + x * w);     R[8] = 1 - 2 * (x * x + y * y);
}

// nucleus instances: placement, shape, size and rotation, drawn per nucleus
struct NucleusInst {
    V3 center;
    int proto;
    double radius;
    double R[9];
};

struct NucleusBank {
    std::vector<LodMesh> protos;
    std::vector<Voxelized> pvz;
    std::vector<NucleusInst> inst;
};

size_t nucleus_count(const GenParams &p) {
    size_t nn = p.n_vessels > 0 ? (size_t)p.n_vessels * p.nuclei_per_vessel
                                : (size_t)p.nuclei_per_vessel;
    if (p.max_nuclei > 0 && (size_t)p.max_nuclei < nn) nn = (size_t)p.max_nuclei;
    return nn;
}

// prototypes and instances (vs: the vessels' centrelines, when nuclei follow vessels)
void build_nuclei(const GenParams &p, const std::vector<Vessel> &vs, NucleusBank &b) {
    IcoTopo topo = make_ico(p.subdiv);
    const int np = std::max(1, p.prototypes);
    b.protos.assign(np, LodMesh());
    b.pvz.assign(np, Voxelized());
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < np; k++) {
        Rng r(p.seed, 2000000000ULL + (uint64_t)k);
        b.protos[k] = make_nucleus_proto(topo, p.noise, r);
        b.pvz[k] = voxelize(b.protos[k], p.voxel_tris);
    }
    const size_t nn = nucleus_count(p);
    b.inst.assign(nn, NucleusInst());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)nn; i++) {
        Rng r(p.nuclei_seed ? p.nuclei_seed : p.seed, 3000000000ULL + (uint64_t)i);
        NucleusInst &in = b.inst[i];
        if (p.n_vessels > 0) {
            const Vessel &V = vs[i / p.nuclei_per_vessel];
            int segs = (int)V.pts.size() - 1;
            double a = r.uni(0, segs);
            int s = std::min((int)a, segs - 1);
            V3 t = unit(V.pts[s + 1] - V.pts[s]);
            V3 c = V.pts[s] + (V.pts[s + 1] - V.pts[s]) * (a - s);
            V3 w = random_unit(r);
            w = w - t * dot(w, t);
            if (norm(w) < 1e-9) w = any_perp(t);
            in.center = c + unit(w) * r.uni(p.place_min, p.place_max);
        } else {
            in.center = V3{r.uni(0, p.cube), r.uni(0, p.cube), r.uni(0, p.cube)};
        }
        in.proto = r.below(np);
        in.radius = r.uni(p.nucleus_r_min, p.nucleus_r_max);
        random_rotation(r, in.R);
    }
}

// float vertices of nucleus i
void nucleus_vertices(const NucleusBank &b, size_t i, std::vector<float> &vf) {
    const NucleusInst &in = b.inst[i];
    const LodMesh &m = b.protos[in.proto];
    vf.resize(3 * m.v.size());
    for (size_t q = 0; q < m.v.size(); q++) {
        V3 x = m.v[q] * in.radius;
        const double *R = in.R;
        vf[3 * q] = (float)(R[0] * x.x + R[1] * x.y + R[2] * x.z + in.center.x);
        vf[3 * q + 1] = (float)(R[3] * x.x + R[4] * x.y + R[5] * x.z + in.center.y);
        vf[3 * q + 2] = (float)(R[6] * x.x + R[7] * x.y + R[8] * x.z + in.center.z);
    }
}

}  // namespace

void generate(const GenParams &p, Tile &nuclei, Tile &vessels) {
#ifdef _OPENMP
    int saved = omp_get_max_threads();
    if (p.threads > 0) omp_set_num_threads(p.threads);
#endif
    // vessels: meshes and partitions first (counts), then written in place
    const size_t nvs = (size_t)std::max(0, p.n_vessels);
    std::vector<Vessel> vs(nvs);
    std::vector<Voxelized> vvz(nvs);
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < (long)nvs; i++) {
        Rng r(p.seed, 1000000000ULL + (uint64_t)i);
        vs[i] = make_vessel(p, r, p.emit_vessels != 0);
        if (p.emit_vessels) vvz[i] = voxelize(vs[i].mesh, p.voxel_tris);
    }
    const size_t nvs_out = p.emit_vessels ? nvs : 0;
    std::vector<ObjShape> vshape(nvs_out);
    for (size_t i = 0; i < nvs_out; i++) vshape[i] = {(uint32_t)vvz[i].nvox, total_faces(vs[i].mesh)};
    std::vector<uint64_t> vtri;
    layout(vessels, vshape, vtri);
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < (long)nvs_out; i++) {
        const LodMesh &m = vs[i].mesh;
        std::vector<float> vf(3 * m.v.size());
        for (size_t k = 0; k < m.v.size(); k++) {
            vf[3 * k] = (float)m.v[k].x;
            vf[3 * k + 1] = (float)m.v[k].y;
            vf[3 * k + 2] = (float)m.v[k].z;
        }
        emit_object(m, vf, vvz[i], 1.0, sink_for(vessels, vtri, i));
        finish_object(vessels, vtri, i);
    }
    vessels.finalize_boxes();
    // nuclei: instances of the prototype bank, written in place
    NucleusBank bank;
    build_nuclei(p, vs, bank);
    const size_t nn = bank.inst.size();
    std::vector<ObjShape> nshape(nn);
    for (size_t i = 0; i < nn; i++) {
        const int k = bank.inst[i].proto;
        nshape[i] = {(uint32_t)bank.pvz[k].nvox, total_faces(bank.protos[k])};
    }
    std::vector<uint64_t> ntri;
    layout(nuclei, nshape, ntri);
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)nn; i++) {
        const int k = bank.inst[i].proto;
        std::vector<float> vf;
        nucleus_vertices(bank, i, vf);
        emit_object(bank.protos[k], vf, bank.pvz[k], bank.inst[i].radius,
                    sink_for(nuclei, ntri, i));
        finish_object(nuclei, ntri, i);
    }
    nuclei.finalize_boxes();
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
}

// ------------------------------------------------------------ object sources

namespace {

// the generator's nuclei, decoded one object at a time
class NucleusSource final : public ObjectSource {
  public:
    explicit NucleusSource(const GenParams &p) {
        std::vector<Vessel> vs;
        if (p.n_vessels > 0) {
            vs.resize(p.n_vessels);
#pragma omp parallel for schedule(dynamic, 1)
            for (long i = 0; i < (long)p.n_vessels; i++) {
                Rng r(p.seed, 1000000000ULL + (uint64_t)i);
                vs[i] = make_vessel(p, r, false);
            }
        }
        build_nuclei(p, vs, b_);
    }
    size_t n_obj() const override { return b_.inst.size(); }
    uint32_t n_vox(size_t o) const override { return (uint32_t)b_.pvz[b_.inst[o].proto].nvox; }
    uint64_t n_tri(size_t o) const override { return total_faces(b_.protos[b_.inst[o].proto]); }
    void meta(size_t o, float box[6], float tr[3]) const override {
        const int k = b_.inst[o].proto;
        std::vector<float> vf, vb(9 * (size_t)b_.pvz[k].nvox);
        nucleus_vertices(b_, o, vf);
        emit_boxes(b_.protos[k], vf, b_.pvz[k], vb.data(), tr);
        for (int d = 0; d < 3; d++) {
            box[d] = INFINITY;
            box[3 + d] = -INFINITY;
        }
        for (int v = 0; v < b_.pvz[k].nvox; v++)
            for (int d = 0; d < 3; d++) {
                box[d] = std::min(box[d], vb[9 * v + d]);
                box[3 + d] = std::max(box[3 + d], vb[9 * v + 3 + d]);
            }
    }
    void decode(size_t o, const ObjArrays &out) const override {
        const int k = b_.inst[o].proto;
        std::vector<float> vf;
        float tr[3];
        nucleus_vertices(b_, o, vf);
        emit_boxes(b_.protos[k], vf, b_.pvz[k], out.vox_box, tr);
        emit_records(b_.protos[k], vf, b_.pvz[k], b_.inst[o].radius, out.cnt, out.rec, out.hd);
    }
    // state: the float vertices, then the records' margin (as a double's two halves)
    size_t state_floats(size_t o) const override {
        return 3 * b_.protos[b_.inst[o].proto].v.size() + 2;
    }
    void decode_meta(size_t o, const ObjArrays &out, float *state) const override {
        const int k = b_.inst[o].proto;
        std::vector<float> vf;
        float tr[3];
        nucleus_vertices(b_, o, vf);
        emit_boxes(b_.protos[k], vf, b_.pvz[k], out.vox_box, tr);
        const double scale = b_.inst[o].radius, margin = record_margin(vf, scale);
        emit_counts_hd(b_.pvz[k], scale, margin, out.cnt, out.hd);
        if (state) {
            std::memcpy(state, vf.data(), 4 * vf.size());
            std::memcpy(state + vf.size(), &margin, sizeof margin);
        }
    }
    void decode_slot(size_t o, int slot, float *rec, const float *state) const override {
        const int k = b_.inst[o].proto;
        const double scale = b_.inst[o].radius;
        if (state) {
            double margin;
            std::memcpy(&margin, state + 3 * b_.protos[k].v.size(), sizeof margin);
            emit_slot(b_.protos[k], state, b_.pvz[k], scale, margin, slot, rec);
            return;
        }
        std::vector<float> vf;
        nucleus_vertices(b_, o, vf);
        emit_slot(b_.protos[k], vf.data(), b_.pvz[k], scale, record_margin(vf, scale), slot, rec);
    }
    void decode_slot_voxel(size_t o, int slot, uint32_t v, float *rec,
                           const float *state) const override {
        if (!state) {
            ObjectSource::decode_slot_voxel(o, slot, v, rec, state);
            return;
        }
        const int k = b_.inst[o].proto;
        const Voxelized &vz = b_.pvz[k];
        size_t first = 0;
        for (uint32_t u = 0; u < v; u++) first += vz.vcnt[slot][u];
        double margin;
        std::memcpy(&margin, state + 3 * b_.protos[k].v.size(), sizeof margin);
        emit_slot_range(b_.protos[k], state, vz, b_.inst[o].radius, margin, slot, first,
                        vz.vcnt[slot][v], rec);
    }

  private:
    NucleusBank b_;
};

// an in-memory tile's objects (records copied out in slot -> voxel order)
class TileSource final : public ObjectSource {
  public:
    explicit TileSource(const Tile &t) : t_(t) {}
    size_t n_obj() const override { return t_.n_obj; }
    uint32_t n_vox(size_t o) const override { return t_.obj_vox_off[o + 1] - t_.obj_vox_off[o]; }
    uint64_t n_tri(size_t o) const override {
        uint64_t n = 0;
        for (uint32_t v = t_.obj_vox_off[o]; v < t_.obj_vox_off[o + 1]; v++)
            for (int s = 0; s < kLods; s++) n += t_.vox_tri_cnt[5 * v + s];
        return n;
    }
    void meta(size_t o, float box[6], float tr[3]) const override {
        std::memcpy(box, &t_.obj_box[6 * o], 6 * sizeof(float));
        std::memcpy(tr, &t_.obj_transform[3 * o], 3 * sizeof(float));
    }
    void decode(size_t o, const ObjArrays &out) const override {
        const uint32_t v0 = t_.obj_vox_off[o], nv = t_.obj_vox_off[o + 1] - v0;
        std::memcpy(out.vox_box, &t_.vox_box[9 * (size_t)v0], 36 * (size_t)nv);
        std::memcpy(out.cnt, &t_.vox_tri_cnt[5 * (size_t)v0], 20 * (size_t)nv);
        std::memcpy(out.hd, &t_.obj_hd[10 * o], 40);
        size_t w = 0;
        for (int s = 0; s < kLods; s++)
            for (uint32_t v = 0; v < nv; v++) {
                const size_t n = t_.vox_tri_cnt[5 * (size_t)(v0 + v) + s];
                std::memcpy(out.rec + 12 * w, &t_.tri[12 * t_.vox_tri_off[5 * (size_t)(v0 + v) + s]],
                            48 * n);
                w += n;
            }
    }
    void decode_meta(size_t o, const ObjArrays &out, float *) const override {
        const uint32_t v0 = t_.obj_vox_off[o], nv = t_.obj_vox_off[o + 1] - v0;
        std::memcpy(out.vox_box, &t_.vox_box[9 * (size_t)v0], 36 * (size_t)nv);
        std::memcpy(out.cnt, &t_.vox_tri_cnt[5 * (size_t)v0], 20 * (size_t)nv);
        std::memcpy(out.hd, &t_.obj_hd[10 * o], 40);
    }
    void decode_slot(size_t o, int slot, float *rec, const float *) const override {
        const uint32_t v0 = t_.obj_vox_off[o], nv = t_.obj_vox_off[o + 1] - v0;
        size_t w = 0;
        for (uint32_t v = 0; v < nv; v++) {
            const size_t q = 5 * (size_t)(v0 + v) + slot, n = t_.vox_tri_cnt[q];
            std::memcpy(rec + 12 * w, &t_.tri[12 * t_.vox_tri_off[q]], 48 * n);
            w += n;
        }
    }

  private:
    const Tile &t_;
};

}  // namespace

// generic forms: a whole decode into scratch
void ObjectSource::decode_meta(size_t o, const ObjArrays &out, float *) const {
    std::vector<float> rec(12 * n_tri(o));
    decode(o, ObjArrays{out.vox_box, out.cnt, out.hd, rec.data()});
}

void ObjectSource::decode_slot_voxel(size_t o, int slot, uint32_t v, float *rec,
                                     const float *state) const {
    const uint32_t nv = n_vox(o);
    std::vector<float> box(9 * (size_t)nv), hd(10), all(12 * n_tri(o));
    std::vector<uint32_t> cnt(5 * (size_t)nv);
    decode_meta(o, ObjArrays{box.data(), cnt.data(), hd.data(), nullptr});
    size_t at = 0;
    for (uint32_t u = 0; u < v; u++) at += cnt[5 * u + slot];
    decode_slot(o, slot, all.data(), state);
    std::memcpy(rec, all.data() + 12 * at, 48 * (size_t)cnt[5 * v + slot]);
}

void ObjectSource::decode_slot(size_t o, int slot, float *rec, const float *) const {
    const uint32_t nv = n_vox(o);
    std::vector<float> box(9 * (size_t)nv), hd(10), all(12 * n_tri(o));
    std::vector<uint32_t> cnt(5 * (size_t)nv);
    decode(o, ObjArrays{box.data(), cnt.data(), hd.data(), all.data()});
    size_t at = 0, n = 0;
    for (int s = 0; s < kLods; s++)
        for (uint32_t v = 0; v < nv; v++) {
            if (s < slot) at += cnt[5 * v + s];
            if (s == slot) n += cnt[5 * v + s];
        }
    std::memcpy(rec, all.data() + 12 * at, 48 * n);
}

std::unique_ptr<ObjectSource> nucleus_source(const GenParams &p) {
    return std::unique_ptr<ObjectSource>(new NucleusSource(p));
}

std::unique_ptr<ObjectSource> tile_source(const Tile &t) {
    return std::unique_ptr<ObjectSource>(new TileSource(t));
}

}  // namespace tdb
