// Thin C entry points over the reference's own CGAL-free sources, compiled
// where they lie under the reference tree by oracle/Makefile into
// oracle/_ref/libtdref.so. Test infrastructure only: it exists so the C
// restatement in tdp_oracle.c can be checked against the reference itself
// (TriDist/MeshDist: src/geometry/computing_cpu.cpp:403-481; MeshInt:
// :482-523; TriInt: src/geometry/TriInt.cpp:400; aab::distance:
// src/spatial/himesh_box.cpp:149-168; octree: src/index/octree.cpp:33-365).
#include <cstdint>
#include <vector>

#include "geometry.h"
#include "index.h"

using namespace tdbase;

namespace {
struct BoxSet {
    std::vector<weighted_aab *> boxes;
    OctreeNode *tree = nullptr;
};
}  // namespace

extern "C" {

void ref_tridist(const float *S, const float *T, size_t n, float *out) {
    for (size_t i = 0; i < n; i++) out[i] = TriDist(S + 9 * i, T + 9 * i);
}

void ref_meshdist(const float *tri, const float *hd, const uint64_t *os, size_t P, float *out) {
    for (size_t p = 0; p < P; p++) {
        const uint64_t *q = os + 4 * p;
        result_container r = MeshDist(tri + 9 * q[0], tri + 9 * q[2], q[1], q[3],
                                      hd ? hd + 2 * q[0] : nullptr, hd ? hd + 2 * q[2] : nullptr);
        out[3 * p] = r.distance;
        out[3 * p + 1] = r.min_dist;
        out[3 * p + 2] = r.max_dist;
    }
}

void ref_triint(const float *S, const float *T, size_t n, uint8_t *out) {
    for (size_t i = 0; i < n; i++) out[i] = TriInt(S + 9 * i, T + 9 * i) ? 1 : 0;
}

// MeshInt per voxel pair (os as for ref_meshdist): intersected, and the
// first intersecting pair it found
void ref_meshint(const float *tri, const float *hd, const uint64_t *os, size_t P, uint8_t *out) {
    for (size_t p = 0; p < P; p++) {
        const uint64_t *q = os + 4 * p;
        result_container r = MeshInt(tri + 9 * q[0], tri + 9 * q[2], q[1], q[3],
                                     hd ? hd + 2 * q[0] : nullptr, hd ? hd + 2 * q[2] : nullptr);
        out[p] = r.intersected ? 1 : 0;
    }
}

void ref_aab_distance(const float *a, const float *b, float *out2) {
    aab A(a[0], a[1], a[2], a[3], a[4], a[5]);
    aab B(b[0], b[1], b[2], b[3], b[4], b[5]);
    range r = A.distance(B);
    out2[0] = r.mindist;
    out2[1] = r.maxdist;
}

void *ref_boxset_build(size_t n, const float *low, const float *high, const float *transform,
                       int leaf) {
    BoxSet *s = new BoxSet();
    for (size_t i = 0; i < n; i++) {
        weighted_aab *w = new weighted_aab();
        w->set_box(low[3 * i], low[3 * i + 1], low[3 * i + 2], high[3 * i], high[3 * i + 1],
                   high[3 * i + 2]);
        w->id = (int)i;
        for (int d = 0; d < 3; d++) w->transform[d] = transform[3 * i + d];
        s->boxes.push_back(w);
    }
    s->tree = build_octree(s->boxes, leaf);
    return s;
}

void ref_boxset_free(void *h) {
    BoxSet *s = (BoxSet *)h;
    delete s->tree;
    for (auto *w : s->boxes) delete w;
    delete s;
}

// self >= 0 queries with the stored box of object `self` (excluded by pointer,
// as in a self-join); otherwise with a fresh box.
static size_t emit(std::vector<std::pair<int, range>> &c, int *ids, float *mn, float *mx,
                   size_t cap) {
    for (size_t i = 0; i < c.size() && i < cap; i++) {
        ids[i] = c[i].first;
        mn[i] = c[i].second.mindist;
        mx[i] = c[i].second.maxdist;
    }
    return c.size();
}

size_t ref_query_knn(void *h, const float *ql, const float *qh, const float *qt, int self, int k,
                     int *ids, float *mn, float *mx, size_t cap) {
    BoxSet *s = (BoxSet *)h;
    weighted_aab q;
    q.set_box(ql[0], ql[1], ql[2], qh[0], qh[1], qh[2]);
    weighted_aab *qp = self >= 0 ? s->boxes[self] : &q;
    float t[3] = {qt[0], qt[1], qt[2]};
    std::vector<std::pair<int, range>> c;
    float minmax = DBL_MAX;
    s->tree->query_knn_fast(qp, t, c, minmax, k);
    return emit(c, ids, mn, mx, cap);
}

// query_intersect ids in the order the octree returns them (duplicates kept)
size_t ref_query_intersect(void *h, const float *ql, const float *qh, int self, int *ids,
                           size_t cap) {
    BoxSet *s = (BoxSet *)h;
    weighted_aab q;
    q.set_box(ql[0], ql[1], ql[2], qh[0], qh[1], qh[2]);
    weighted_aab *qp = self >= 0 ? s->boxes[self] : &q;
    std::vector<int> c;
    s->tree->query_intersect(qp, c);
    for (size_t i = 0; i < c.size() && i < cap; i++) ids[i] = c[i];
    return c.size();
}

size_t ref_query_within(void *h, const float *ql, const float *qh, const float *qt, int self,
                        float tau, int *ids, float *mn, float *mx, size_t cap) {
    BoxSet *s = (BoxSet *)h;
    weighted_aab q;
    q.set_box(ql[0], ql[1], ql[2], qh[0], qh[1], qh[2]);
    weighted_aab *qp = self >= 0 ? s->boxes[self] : &q;
    float t[3] = {qt[0], qt[1], qt[2]};
    std::vector<std::pair<int, range>> c;
    s->tree->query_within(qp, t, c, tau);
    return emit(c, ids, mn, mx, cap);
}
}
