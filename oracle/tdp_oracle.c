/*
 * CPU oracle for the spatial-join refinement path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
 * product (3dpipe_amd) never links, loads or calls it.
 *
 * It restates, in plain C, what the reference computes on the `tdbase join
 * -q nn|within` path:
 *   - triangle distance / per-voxel-pair bounds   (oracle_tridist.inc)
 *   - aab::distance                               src/spatial/himesh_box.cpp:149-168
 *   - the octree: addObject                       src/index/octree.cpp:33-102
 *                 add_candidate                   src/index/octree.cpp:120-154
 *                 query_knn_fast                  src/index/octree.cpp:215-291
 *                 query_within                    src/index/octree.cpp:293-330
 *                 query_intersect                 src/index/octree.cpp:332-349
 *                 build_octree (leaf 10)          src/storage/tile.cpp:76-85, tile_IO.cpp:47
 *   - the RAW tile reader                         src/storage/tile_IO.cpp:13-49,
 *                                                 src/spatial/himesh_wrapper.cpp:17-85,137-161
 *   - the kNN / within / intersect join           (oracle_join.inc)
 *
 * Two arithmetic forms: REAL=float reproduces the reference's FP32 operation
 * order (build with -ffp-contract=off, IEEE division and sqrt), REAL=double is
 * the same algorithm in FP64 over the same float inputs.
 *
 * Pinning: TriDist/MeshDist (float) and the octree candidate lists (order
 * included) are checked against the reference's own compiled sources
 * (oracle/_ref, built by oracle/Makefile) and against golden fixtures in
 * tests/golden; the FP64 TriDist against closed forms and dense sampling; the
 * join drivers (which need CGAL in the reference and cannot be compiled) by
 * invariants and by brute force at the final LOD on small inputs.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CAT2_(a, b) a##b
#define CAT_(a, b) CAT2_(a, b)

/* ------------------------------------------------------------------ tile */

typedef struct {
    float low[3], high[3], core[3];
    uint64_t off[5], cnt[5]; /* per LOD slot: byte offset into data, #triangles */
} orc_voxel;

typedef struct {
    int nvox;
    orc_voxel *vox;
    const char *data; /* start of the object's data block */
    float hd[5], phd[5];
    float box_low[3], box_high[3];
    float transform[3];
} orc_object;

struct octnode;
typedef struct {
    size_t n;
    orc_object *obj;
    char *buf;
    size_t bufsize;
    float space_low[3], space_high[3];
    struct octnode *tree;
} orc_tile;

static int lod_slot(int lod) {
    if (lod == 20) return 0;
    if (lod == 40) return 1;
    if (lod == 60) return 2;
    if (lod == 80) return 3;
    if (lod == 100) return 4;
    return -1; /* any other value decodes to empty voxels */
}

static void voxel_lod(const orc_object *o, int v, int slot, const float **tri, const float **hd,
                      size_t *n) {
    if (slot < 0) { *n = 0; return; }
    const orc_voxel *vx = &o->vox[v];
    *n = vx->cnt[slot];
    *tri = (const float *)(o->data + vx->off[slot]);
    *hd = (const float *)(o->data + vx->off[slot] + 9 * vx->cnt[slot] * sizeof(float));
}
static float obj_hausdorff(const orc_object *o, int slot) { return slot < 0 ? 0.f : o->hd[slot]; }
static float obj_proxy(const orc_object *o, int slot) { return slot < 0 ? 0.f : o->phd[slot]; }

/* ---------------------------------------------------------------- results */

typedef struct {
    uint64_t tri_pairs, voxel_pairs, object_pairs;
    int rounds, lod_count;
} orc_stats;

typedef struct {
    int target, obj2;
    double mn, mx;
    int round; /* 0 = filter, r = r-th LOD round */
} res_item;

typedef struct {
    size_t n_targets;
    res_item *items;
    size_t n, cap;
    orc_stats stats;
} orc_result;

static orc_result *res_new(size_t n_targets) {
    orc_result *r = (orc_result *)calloc(1, sizeof(orc_result));
    r->n_targets = n_targets;
    return r;
}
static void res_push(orc_result *R, int target, int obj2, double mn, double mx) {
    if (R->n == R->cap) {
        R->cap = R->cap ? 2 * R->cap : 1024;
        R->items = (res_item *)realloc(R->items, R->cap * sizeof(res_item));
    }
    res_item it = {target, obj2, mn, mx, R->stats.rounds};
    R->items[R->n++] = it;
}

/* ---------------------------------------------------------------- aab */

static void box_reset(float low[3], float high[3]) {
    for (int i = 0; i < 3; i++) {
        low[i] = INFINITY;
        high[i] = -INFINITY;
    }
}
static void box_update(float low[3], float high[3], const float *pl, const float *ph) {
    for (int i = 0; i < 3; i++) {
        if (low[i] > pl[i]) low[i] = pl[i];
        if (high[i] < ph[i]) high[i] = ph[i];
    }
}
/* strict overlap test of aab::intersect (himesh_box.cpp:111-115) */
static int box_intersect(const float *al, const float *ah, const float *bl, const float *bh) {
    return !(bl[0] >= ah[0] || bh[0] <= al[0] || bl[1] >= ah[1] || bh[1] <= al[1] ||
             bl[2] >= ah[2] || bh[2] <= al[2]);
}
/* aab::distance: a = this, b = argument */
static void aab_distance(const float *al, const float *ah, const float *bl, const float *bh,
                         float *mn, float *mx) {
    float rmin = 0, rmax = 0;
    for (int i = 0; i < 3; i++) {
        float t1 = al[i] - bh[i];
        float t2 = ah[i] - bl[i];
        rmax += (t1 + t2) * (t1 + t2) / 4;
        if (t2 < 0) rmin += t2 * t2;
        else if (t1 > 0) rmin += t1 * t1;
    }
    *mn = sqrtf(rmin);
    *mx = sqrtf(rmax);
}

/* ---------------------------------------------------------------- octree */

typedef struct octnode {
    float low[3], high[3];
    uint32_t size;
    long tile_size;
    int level, is_leaf, can_split;
    struct octnode *child[8];
    int *objs;
    size_t nobj, cap;
} octnode;

static octnode *oct_new(const float *low, const float *high, int level, long ts) {
    octnode *n = (octnode *)calloc(1, sizeof(octnode));
    memcpy(n->low, low, sizeof n->low);
    memcpy(n->high, high, sizeof n->high);
    n->level = level;
    n->tile_size = ts;
    n->is_leaf = 1;
    n->can_split = 1;
    return n;
}
static void oct_free(octnode *n) {
    if (!n) return;
    if (!n->is_leaf)
        for (int i = 0; i < 8; i++) oct_free(n->child[i]);
    free(n->objs);
    free(n);
}
static void oct_push(octnode *n, int id) {
    if (n->nobj == n->cap) {
        n->cap = n->cap ? 2 * n->cap : 8;
        n->objs = (int *)realloc(n->objs, n->cap * sizeof(int));
    }
    n->objs[n->nobj++] = id;
}

static void oct_add(octnode *n, const orc_tile *t, int id) {
    const orc_object *o = &t->obj[id];
    n->size += 1;
    if (n->size == 1) {
        oct_push(n, id);
        return;
    }
    if (n->is_leaf) {
        oct_push(n, id);
        /* The undo decision (a split placing objects 2 (size - 1) times or
           more) depends only on the child boxes, so it is counted before any
           child is built; the reference builds the children recursively and
           then deletes them, which does not end (or takes exponential time)
           where more objects than a leaf holds share a region. The same tree
           results. The depth cap (more halvings than a float box can take)
           changes no tree. */
        if ((long)n->size > n->tile_size && n->can_split && n->level < 320) {
            float mid[3], clo[8][3], chi[8][3];
            for (int i = 0; i < 3; i++) mid[i] = (n->low[i] + n->high[i]) / 2;
            for (int c = 0; c < 8; c++)
                for (int i = 0; i < 3; i++) {
                    int upper = (c >> i) & 1;
                    clo[c][i] = upper ? mid[i] : n->low[i];
                    chi[c][i] = upper ? n->high[i] : mid[i];
                }
            uint32_t total = 0;
            for (size_t k = 0; k < n->nobj; k++) {
                const orc_object *v = &t->obj[n->objs[k]];
                for (int c = 0; c < 8; c++)
                    total += box_intersect(clo[c], chi[c], v->box_low, v->box_high) ? 1u : 0u;
            }
            if (total >= 2 * (n->size - 1)) {
                n->can_split = 0;
            } else {
                for (int c = 0; c < 8; c++)
                    n->child[c] = oct_new(clo[c], chi[c], n->level + 1, n->tile_size);
                for (size_t k = 0; k < n->nobj; k++) {
                    const orc_object *v = &t->obj[n->objs[k]];
                    for (int c = 0; c < 8; c++)
                        if (box_intersect(n->child[c]->low, n->child[c]->high, v->box_low,
                                          v->box_high))
                            oct_add(n->child[c], t, n->objs[k]);
                }
                n->nobj = 0;
                n->is_leaf = 0;
            }
        } else if ((double)n->size > 1.5 * (double)n->tile_size) {
            n->can_split = 1;
        }
    } else {
        for (int c = 0; c < 8; c++)
            if (box_intersect(n->child[c]->low, n->child[c]->high, o->box_low, o->box_high))
                oct_add(n->child[c], t, id);
    }
}

typedef struct {
    int *id;
    float *mn, *mx;
    size_t n, cap;
} cand_list;

static void cand_free(cand_list *c) {
    free(c->id);
    free(c->mn);
    free(c->mx);
    memset(c, 0, sizeof *c);
}
static void cand_insert(cand_list *c, size_t pos, int id, float mn, float mx) {
    if (c->n == c->cap) {
        c->cap = c->cap ? 2 * c->cap : 16;
        c->id = (int *)realloc(c->id, c->cap * sizeof(int));
        c->mn = (float *)realloc(c->mn, c->cap * sizeof(float));
        c->mx = (float *)realloc(c->mx, c->cap * sizeof(float));
    }
    memmove(c->id + pos + 1, c->id + pos, (c->n - pos) * sizeof(int));
    memmove(c->mn + pos + 1, c->mn + pos, (c->n - pos) * sizeof(float));
    memmove(c->mx + pos + 1, c->mx + pos, (c->n - pos) * sizeof(float));
    c->id[pos] = id;
    c->mn[pos] = mn;
    c->mx[pos] = mx;
    c->n++;
}
static void cand_erase(cand_list *c, size_t pos) {
    memmove(c->id + pos, c->id + pos + 1, (c->n - pos - 1) * sizeof(int));
    memmove(c->mn + pos, c->mn + pos + 1, (c->n - pos - 1) * sizeof(float));
    memmove(c->mx + pos, c->mx + pos + 1, (c->n - pos - 1) * sizeof(float));
    c->n--;
}

/* add_candidate (octree.cpp:120-154): sorted insertion by max, ties after
 * existing entries; dedupe by id; prune the tail against the k-th max */
static int add_candidate(cand_list *c, int id, float mn, float mx, int k) {
    if ((int)c->n < k || mn < c->mx[k - 1]) {
        size_t loc = 0;
        for (; loc < c->n; loc++) {
            if (c->id[loc] == id) return 0;
            if (c->mx[loc] > mx) {
                cand_insert(c, loc, id, mn, mx);
                break;
            }
        }
        if (loc == c->n) cand_insert(c, c->n, id, mn, mx);
        int process = (int)loc > k - 1 ? (int)loc : k - 1;
        int size = (int)c->n;
        for (int i = process + 1; i < size;) {
            if (c->mn[i] >= c->mx[k - 1]) {
                cand_erase(c, i);
                size--;
                continue;
            }
            i++;
        }
        return (int)loc < k;
    }
    return 0;
}

/* The best-first queue of query_knn_fast is a std::priority_queue ordered by
 * std::greater on the node's min distance. Candidate order depends on how it
 * breaks ties, so this is the binary-heap push / pop (sift-up, then
 * hole-to-leaf descent followed by sift-up) of the GNU libstdc++ heap
 * algorithms, with comp(a,b) = a.mindist > b.mindist. */
typedef struct {
    octnode *node;
    float d;
} heap_ent;
typedef struct {
    heap_ent *a;
    size_t n, cap;
} nheap;

static void heap_sift_up(heap_ent *a, size_t hole, size_t top, heap_ent v) {
    size_t parent = (hole - 1) / 2;
    while (hole > top && a[parent].d > v.d) {
        a[hole] = a[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[hole] = v;
}
static void heap_push(nheap *h, heap_ent v) {
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 64;
        h->a = (heap_ent *)realloc(h->a, h->cap * sizeof(heap_ent));
    }
    h->a[h->n++] = v;
    heap_sift_up(h->a, h->n - 1, 0, v);
}
static heap_ent heap_pop(nheap *h) {
    heap_ent top = h->a[0];
    if (h->n > 1) {
        size_t len = h->n - 1;
        heap_ent v = h->a[len];
        h->a[len] = h->a[0];
        size_t hole = 0, second = 0;
        while (second < (len - 1) / 2) {
            second = 2 * (second + 1);
            if (h->a[second].d > h->a[second - 1].d) second--;
            h->a[hole] = h->a[second];
            hole = second;
        }
        if ((len & 1) == 0 && second == (len - 2) / 2) {
            second = 2 * (second + 1);
            h->a[hole] = h->a[second - 1];
            hole = second - 1;
        }
        heap_sift_up(h->a, hole, 0, v);
    }
    h->n--;
    return top;
}

static float transform_dist(const float *a, const float *b) {
    return sqrtf((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                 (a[2] - b[2]) * (a[2] - b[2]));
}

static void octree_knn(octnode *root, const orc_tile *t, const float *ql, const float *qh,
                       const float *qt, int self, int k, cand_list *out) {
    nheap h = {0};
    float mn, mx;
    aab_distance(root->low, root->high, ql, qh, &mn, &mx);
    heap_ent e0 = {root, mn};
    heap_push(&h, e0);
    float upper = INFINITY;
    while (h.n > 0) {
        heap_ent e = heap_pop(&h);
        if (e.d > upper) break;
        octnode *n = e.node;
        if (n->is_leaf) {
            for (size_t i = 0; i < n->nobj; i++) {
                int id = n->objs[i];
                if (id == self) continue;
                const orc_object *o = &t->obj[id];
                aab_distance(o->box_low, o->box_high, ql, qh, &mn, &mx);
                mx = transform_dist(qt, o->transform);
                if (add_candidate(out, id, mn, mx, k) && (int)out->n >= k) upper = out->mx[k - 1];
            }
        } else {
            for (int c = 0; c < 8; c++) {
                aab_distance(n->child[c]->low, n->child[c]->high, ql, qh, &mn, &mx);
                if (mn <= upper) {
                    heap_ent ce = {n->child[c], mn};
                    heap_push(&h, ce);
                }
            }
        }
    }
    free(h.a);
}

static void octree_within(octnode *n, const orc_tile *t, const float *ql, const float *qh,
                          const float *qt, int self, float tau, cand_list *out) {
    float mn, mx;
    aab_distance(n->low, n->high, ql, qh, &mn, &mx);
    if (!(mn <= tau)) return;
    if (n->is_leaf) {
        for (size_t i = 0; i < n->nobj; i++) {
            int id = n->objs[i];
            if (id == self) continue;
            const orc_object *o = &t->obj[id];
            aab_distance(o->box_low, o->box_high, ql, qh, &mn, &mx);
            mx = transform_dist(qt, o->transform);
            if (mn <= tau) {
                int exist = 0;
                for (size_t c = 0; c < out->n; c++)
                    if (out->id[c] == id) { exist = 1; break; }
                if (!exist) cand_insert(out, out->n, id, mn, mx);
            }
        }
    } else {
        for (int c = 0; c < 8; c++) octree_within(n->child[c], t, ql, qh, qt, self, tau, out);
    }
}

/* query_intersect: objects whose boxes overlap the query box (strictly),
 * collected depth-first; mbb_intersect (IntersectionJoin.cpp:20-33) then
 * sorts the ids and skips duplicates */
static void octree_intersect_rec(octnode *n, const orc_tile *t, const float *ql, const float *qh,
                                 int self, int **ids, size_t *cnt, size_t *cap) {
    if (!box_intersect(n->low, n->high, ql, qh)) return;
    if (n->is_leaf) {
        for (size_t i = 0; i < n->nobj; i++) {
            int id = n->objs[i];
            if (id == self) continue;
            const orc_object *o = &t->obj[id];
            if (!box_intersect(o->box_low, o->box_high, ql, qh)) continue;
            if (*cnt == *cap) {
                *cap = *cap ? 2 * *cap : 16;
                *ids = (int *)realloc(*ids, *cap * sizeof(int));
            }
            (*ids)[(*cnt)++] = id;
        }
    } else {
        for (int c = 0; c < 8; c++) octree_intersect_rec(n->child[c], t, ql, qh, self, ids, cnt, cap);
    }
}

static int cmp_int(const void *a, const void *b) {
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* sorted, duplicates removed; returns the count (the caller frees *ids) */
static size_t octree_intersect(octnode *root, const orc_tile *t, const float *ql, const float *qh,
                               int self, int **ids) {
    size_t cnt = 0, cap = 0;
    *ids = NULL;
    octree_intersect_rec(root, t, ql, qh, self, ids, &cnt, &cap);
    if (cnt > 1) qsort(*ids, cnt, sizeof(int), cmp_int);
    size_t w = 0;
    for (size_t i = 0; i < cnt; i++)
        if (w == 0 || (*ids)[w - 1] != (*ids)[i]) (*ids)[w++] = (*ids)[i];
    return w;
}

static void tile_build_octree(orc_tile *t, int leaf) {
    t->tree = oct_new(t->space_low, t->space_high, 0, leaf);
    for (size_t i = 0; i < t->n; i++) oct_add(t->tree, t, (int)i);
}

/* ------------------------------------------------------- RAW tile reader */

static int rd(const char *buf, size_t size, size_t *pos, void *dst, size_t n) {
    if (*pos + n > size) return -1;
    memcpy(dst, buf + *pos, n);
    *pos += n;
    return 0;
}

orc_tile *orc_tile_load(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 1) { fclose(f); return NULL; }
    orc_tile *t = (orc_tile *)calloc(1, sizeof(orc_tile));
    t->bufsize = (size_t)sz;
    t->buf = (char *)malloc(t->bufsize);
    if (fread(t->buf, 1, t->bufsize, f) != t->bufsize || t->buf[0] != 1) {
        fclose(f);
        free(t->buf);
        free(t);
        return NULL;
    }
    fclose(f);
    size_t pos = 1, cap = 0;
    box_reset(t->space_low, t->space_high);
    while (pos < t->bufsize) {
        if (t->n == cap) {
            cap = cap ? 2 * cap : 1024;
            t->obj = (orc_object *)realloc(t->obj, cap * sizeof(orc_object));
        }
        orc_object *o = &t->obj[t->n];
        memset(o, 0, sizeof *o);
        uint64_t dsize = 0, nvox = 0;
        if (rd(t->buf, t->bufsize, &pos, &dsize, 8)) goto bad;
        o->data = t->buf + pos;
        if (pos + dsize > t->bufsize) goto bad;
        pos += dsize;
        if (rd(t->buf, t->bufsize, &pos, &nvox, 8)) goto bad;
        for (int l = 0; l < 5; l++) {
            if (rd(t->buf, t->bufsize, &pos, &o->hd[l], 4)) goto bad;
            if (rd(t->buf, t->bufsize, &pos, &o->phd[l], 4)) goto bad;
        }
        o->nvox = (int)nvox;
        o->vox = (orc_voxel *)calloc(nvox ? nvox : 1, sizeof(orc_voxel));
        box_reset(o->box_low, o->box_high);
        for (uint64_t v = 0; v < nvox; v++) {
            orc_voxel *vx = &o->vox[v];
            if (rd(t->buf, t->bufsize, &pos, vx->low, 12)) goto bad;
            if (rd(t->buf, t->bufsize, &pos, vx->high, 12)) goto bad;
            if (rd(t->buf, t->bufsize, &pos, vx->core, 12)) goto bad;
            for (int l = 0; l < 5; l++) {
                if (rd(t->buf, t->bufsize, &pos, &vx->off[l], 8)) goto bad;
                if (rd(t->buf, t->bufsize, &pos, &vx->cnt[l], 8)) goto bad;
                if (vx->off[l] + 44 * vx->cnt[l] > dsize) goto bad;
            }
            box_update(o->box_low, o->box_high, vx->low, vx->high);
        }
        if (rd(t->buf, t->bufsize, &pos, o->transform, 12)) goto bad;
        box_update(t->space_low, t->space_high, o->box_low, o->box_high);
        t->n++;
    }
    tile_build_octree(t, 10);
    return t;
bad:
    for (size_t i = 0; i <= t->n && i < cap; i++) free(t->obj[i].vox);
    free(t->obj);
    free(t->buf);
    free(t);
    return NULL;
}

void orc_tile_free(orc_tile *t) {
    if (!t) return;
    oct_free(t->tree);
    for (size_t i = 0; i < t->n; i++) free(t->obj[i].vox);
    free(t->obj);
    free(t->buf);
    free(t);
}
size_t orc_tile_count(const orc_tile *t) { return t ? t->n : 0; }

/* ----------------------------------------------- generic (float / double) */

#define REAL float
#define SFX _f
#define SQRT sqrtf
#include "oracle_tridist.inc"
#include "oracle_join.inc"
#undef REAL
#undef SFX
#undef SQRT

#define REAL double
#define SFX _d
#define SQRT sqrt
#include "oracle_tridist.inc"
#include "oracle_join.inc"
#undef REAL
#undef SFX
#undef SQRT

/* ------------------------------------------------------------- C API */

void orc_tridist_f(const float *S, const float *T, size_t n, float *out) {
    for (size_t i = 0; i < n; i++) {
        float s[9], t[9];
        load_tri_f(s, S + 9 * i);
        load_tri_f(t, T + 9 * i);
        out[i] = tridist_f(s, t);
    }
}
void orc_tridist_d(const float *S, const float *T, size_t n, double *out) {
    for (size_t i = 0; i < n; i++) {
        double s[9], t[9];
        load_tri_d(s, S + 9 * i);
        load_tri_d(t, T + 9 * i);
        out[i] = tridist_d(s, t);
    }
}
/* triangles given in double directly (for closed-form tests off the float grid) */
void orc_tridist_dd(const double *S, const double *T, size_t n, double *out) {
    for (size_t i = 0; i < n; i++) out[i] = tridist_d(S + 9 * i, T + 9 * i);
}

/* MeshDist per voxel pair over an AoS triangle array (9 floats) with a
 * parallel per-triangle (proxy, hausdorff) array; os = 4 u64 per pair:
 * offset1, n1, offset2, n2 (in triangles) */
void orc_meshdist_f(const float *tri, const float *hd, const uint64_t *os, size_t P, float *out) {
    for (size_t p = 0; p < P; p++) {
        const uint64_t *q = os + 4 * p;
        meshdist_f(tri + 9 * q[0], hd ? hd + 2 * q[0] : NULL, q[1], tri + 9 * q[2],
                   hd ? hd + 2 * q[2] : NULL, q[3], out + 3 * p);
    }
}
void orc_meshdist_d(const float *tri, const float *hd, const uint64_t *os, size_t P,
                    double *out) {
    for (size_t p = 0; p < P; p++) {
        const uint64_t *q = os + 4 * p;
        meshdist_d(tri + 9 * q[0], hd ? hd + 2 * q[0] : NULL, q[1], tri + 9 * q[2],
                   hd ? hd + 2 * q[2] : NULL, q[3], out + 3 * p);
    }
}

void orc_voxel_bound_f(const float *v1, const float *v2, float *out2) {
    voxel_bound_f(v1, v1 + 3, v1 + 6, v2, v2 + 3, v2 + 6, out2, out2 + 1);
}
void orc_voxel_bound_d(const float *v1, const float *v2, double *out2) {
    voxel_bound_d(v1, v1 + 3, v1 + 6, v2, v2 + 3, v2 + 6, out2, out2 + 1);
}
void orc_aab_distance(const float *a, const float *b, float *out2) {
    aab_distance(a, a + 3, b, b + 3, out2, out2 + 1);
}

/* kNN status of one target's m candidate ranges (2 doubles each, min and
 * max) with cand_left still wanted: 1 = CONFIRMED, 2 = REMOVE, 0 = KEEP,
 * in FP32 (the ranges rounded to float) or FP64 */
void orc_knn_status(const double *ranges, int m, int cand_left, int fp64, int *status) {
    if (fp64) {
        range_d *r = (range_d *)malloc((m ? m : 1) * sizeof(range_d));
        for (int i = 0; i < m; i++) { r[i].mn = ranges[2 * i]; r[i].mx = ranges[2 * i + 1]; }
        knn_status_d(r, m, cand_left, status);
        free(r);
    } else {
        range_f *r = (range_f *)malloc((m ? m : 1) * sizeof(range_f));
        for (int i = 0; i < m; i++) {
            r[i].mn = (float)ranges[2 * i];
            r[i].mx = (float)ranges[2 * i + 1];
        }
        knn_status_f(r, m, cand_left, status);
        free(r);
    }
}

/* Octree over bare boxes (no tile data) for candidate-order tests: n boxes
 * given as low[3n], high[3n], transform[3n]. */
typedef struct {
    orc_tile t;
} orc_boxset;

orc_tile *orc_boxset_build(size_t n, const float *low, const float *high, const float *transform,
                           int leaf) {
    orc_tile *t = (orc_tile *)calloc(1, sizeof(orc_tile));
    t->n = n;
    t->obj = (orc_object *)calloc(n ? n : 1, sizeof(orc_object));
    box_reset(t->space_low, t->space_high);
    for (size_t i = 0; i < n; i++) {
        memcpy(t->obj[i].box_low, low + 3 * i, 12);
        memcpy(t->obj[i].box_high, high + 3 * i, 12);
        memcpy(t->obj[i].transform, transform + 3 * i, 12);
        box_update(t->space_low, t->space_high, t->obj[i].box_low, t->obj[i].box_high);
    }
    tile_build_octree(t, leaf);
    return t;
}

/* query the octree of tile t; self = index excluded (-1 for none).
 * Returns the candidate count (written up to cap). */
size_t orc_query_knn(const orc_tile *t, const float *ql, const float *qh, const float *qt,
                     int self, int k, int *ids, float *mn, float *mx, size_t cap) {
    cand_list c = {0};
    octree_knn(t->tree, t, ql, qh, qt, self, k, &c);
    size_t n = c.n;
    for (size_t i = 0; i < n && i < cap; i++) {
        ids[i] = c.id[i];
        mn[i] = c.mn[i];
        mx[i] = c.mx[i];
    }
    cand_free(&c);
    return n;
}
size_t orc_query_within(const orc_tile *t, const float *ql, const float *qh, const float *qt,
                        int self, float tau, int *ids, float *mn, float *mx, size_t cap) {
    cand_list c = {0};
    octree_within(t->tree, t, ql, qh, qt, self, tau, &c);
    size_t n = c.n;
    for (size_t i = 0; i < n && i < cap; i++) {
        ids[i] = c.id[i];
        mn[i] = c.mn[i];
        mx[i] = c.mx[i];
    }
    cand_free(&c);
    return n;
}
size_t orc_query_intersect(const orc_tile *t, const float *ql, const float *qh, int self,
                           int *ids, size_t cap) {
    int *c = NULL;
    size_t n = octree_intersect(t->tree, t, ql, qh, self, &c);
    for (size_t i = 0; i < n && i < cap; i++) ids[i] = c[i];
    free(c);
    return n;
}
/* box / transform of object i of a loaded tile (6 + 3 floats) */
void orc_tile_object_box(const orc_tile *t, size_t i, float *out9) {
    memcpy(out9, t->obj[i].box_low, 12);
    memcpy(out9 + 3, t->obj[i].box_high, 12);
    memcpy(out9 + 6, t->obj[i].transform, 12);
}

/* LOD list post-processing of parse_args (query_context.h:183-188): sort, then
 * std::unique without erase, which leaves the tail as it was. */
int orc_lod_list(const int *in, int n, int *out) {
    for (int i = 0; i < n; i++) out[i] = in[i];
    for (int i = 1; i < n; i++) {
        int v = out[i], j = i - 1;
        while (j >= 0 && out[j] > v) { out[j + 1] = out[j]; j--; }
        out[j + 1] = v;
    }
    if (n > 0) {
        int w = 0;
        for (int r = 1; r < n; r++)
            if (out[r] != out[w]) out[++w] = out[r];
        /* elements past w keep their previous values */
    }
    return n;
}

/* query: 1 = kNN, 0 = within, 2 = intersect;
 * precision: 0 = FP32 reference arithmetic, 1 = FP64 */
orc_result *orc_join(const orc_tile *t1, const orc_tile *t2, int query, int k, float tau,
                     const int *lods_in, int nlods, int level, int precision,
                     uint64_t max_objects1) {
    const int is_knn = query == 1;
    if (!t1 || !t2 || (is_knn && k < 1) || query < 0 || query > 2) return NULL;
    int *lods = (int *)malloc((nlods ? nlods : 1) * sizeof(int));
    orc_lod_list(lods_in, nlods, lods);
    orc_result *r;
    if (query == 2)
        r = precision == 0 ? join_intersect_f(t1, t2, lods, nlods, level, (size_t)max_objects1)
                           : join_intersect_d(t1, t2, lods, nlods, level, (size_t)max_objects1);
    else
        r = precision == 0
                ? join_f(t1, t2, is_knn, k, tau, lods, nlods, level, (size_t)max_objects1)
                : join_d(t1, t2, is_knn, k, tau, lods, nlods, level, (size_t)max_objects1);
    free(lods);
    return r;
}

size_t orc_result_count(const orc_result *r) { return r ? r->n : 0; }
size_t orc_result_targets(const orc_result *r) { return r ? r->n_targets : 0; }
/* flat report list in report order: target, obj2, min, max */
void orc_result_get(const orc_result *r, int32_t *target, int32_t *obj2, double *mn,
                    double *mx) {
    for (size_t i = 0; i < r->n; i++) {
        target[i] = r->items[i].target;
        obj2[i] = r->items[i].obj2;
        mn[i] = r->items[i].mn;
        mx[i] = r->items[i].mx;
    }
}
/* the round of every report (0 = filter, r = r-th LOD round) */
void orc_result_rounds(const orc_result *r, uint8_t *round) {
    for (size_t i = 0; i < r->n; i++) round[i] = (uint8_t)r->items[i].round;
}
/* tri_pairs, voxel_pairs, object_pairs, rounds */
/* threads the join's refinement folds run on (OpenMP) */
int orc_max_threads(void) { return omp_get_max_threads(); }

void orc_result_stats(const orc_result *r, uint64_t *out4) {
    out4[0] = r->stats.tri_pairs;
    out4[1] = r->stats.voxel_pairs;
    out4[2] = r->stats.object_pairs;
    out4[3] = (uint64_t)r->stats.rounds;
}
void orc_result_free(orc_result *r) {
    if (!r) return;
    free(r->items);
    free(r);
}
