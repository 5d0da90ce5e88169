// libtdbase.so: the host join engine behind include/tdbase.h. It generates
// candidates on the host (octree, OpenMP over targets) and runs the filter and
// every LOD round on the device through include/tdp.h.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "tdb_internal.hpp"

// Device copies are keyed by tile; a tile that is freed drops its copies in
// every engine that holds one, and a closing engine unregisters itself, so a
// new tile allocated at a recycled address never meets a stale device copy.
static std::mutex g_registry;

static void set_err(char *err, size_t len, const std::string &msg) {
    if (err && len) std::snprintf(err, len, "%s", msg.c_str());
}


// the records of the next round's voxels at `lod` that the tiles lack
static int feed_round(tdp_ctx *ctx, tdp_join *j, int lod, LazyFeed &f, std::string &err) {
    const int slot = tdb::lod_slot(lod);
    for (int side = 1; side <= 2; side++) {
        const uint32_t *vox = nullptr;
        size_t n = 0;
        uint64_t nrec = 0;
        int rc = tdp_join_lazy_needs(j, lod, side, &vox, &n, &nrec);
        if (rc != TDP_OK) {
            err = tdp_last_error(ctx);
            return rc;
        }
        if (n) {
            const int w = side - 1;
            const uint32_t *cnt = f.vox_tri_cnt[w];
            if (!cnt || slot < 0) {
                err = "lazy feed: no record counts for side " + std::to_string(side);
                return TDP_ERR_STATE;
            }
            // packed offsets from the host's own counts
            f.off.resize(n);
            uint64_t at = 0;
            for (size_t k = 0; k < n; k++) {
                f.off[k] = at;
                at += cnt[5 * (size_t)vox[k] + slot];
            }
            if (at != nrec) {
                err = "lazy feed: the host's record counts differ from the tile's";
                return TDP_ERR_STATE;
            }
            if (f.cap[w] < nrec) {
                const size_t cap = std::max<size_t>(nrec, f.cap[w] + f.cap[w] / 2);
                if (f.buf[w]) tdp_host_free(f.ctx, f.buf[w]);
                f.buf[w] = nullptr;
                f.cap[w] = 0;
                rc = tdp_host_alloc(f.ctx, 48 * cap, (void **)&f.buf[w]);
                if (rc != TDP_OK) {
                    err = tdp_last_error(f.ctx);
                    return rc;
                }
                f.cap[w] = cap;
            }
            const auto t0 = Clock::now();
            if (!f.pack(side, slot, vox, n, f.off.data(), f.buf[w], err)) return TDP_ERR_STATE;
            f.pack_ms += ms_since(t0);
            f.put_records += nrec;
        }
        const auto t1 = Clock::now();
        rc = tdp_join_lazy_put(j, lod, side, n ? f.buf[side - 1] : nullptr);
        f.put_ms += ms_since(t1);
        if (rc != TDP_OK) {
            err = tdp_last_error(ctx);
            return rc;
        }
    }
    return TDP_OK;
}

int tdb_run_join(tdp_ctx *ctx, tdp_tile *d1, tdp_tile *d2, const tdb_join_opts *o,
                 const std::vector<uint32_t> &prefix, const std::vector<uint32_t> &ids,
                 const int32_t *k_t, const float *tau_t, tdb_result *res, std::string &err,
                 LazyFeed *feed, const uint8_t *within_t) {
    const size_t N = prefix.size() - 1, P = ids.size();
    res->n_targets = N;
    res->stats[4] = P;

    tdp_join_params prm;
    prm.query = o->query == TDB_QUERY_NN       ? TDP_QUERY_KNN
                : o->query == TDB_QUERY_WITHIN ? TDP_QUERY_WITHIN
                                               : TDP_QUERY_INTERSECT;
    prm.k = o->k;
    prm.within_dist = o->within_dist;
    prm.precision = o->precision;
    prm.hausdorff_level = o->hausdorff_level;
    tdp_join *j = nullptr;
    auto t0 = Clock::now();
    int rc = within_t ? tdp_join_create_kinds(ctx, d1, d2, &prm, N, prefix.data(), ids.data(),
                                              within_t, k_t, tau_t, &j)
                      : tdp_join_create_multi(ctx, d1, d2, &prm, N, prefix.data(), ids.data(),
                                              k_t, tau_t, &j);
    if (rc != TDP_OK) {
        err = tdp_last_error(ctx);
        return rc;
    }
    res->timing[7] = ms_since(t0);
    t0 = Clock::now();
    uint64_t n_vp = 0;
    rc = tdp_join_filter(j, &n_vp);
    res->timing[2] = ms_since(t0);
    t0 = Clock::now();
    const std::vector<int> lods = tdb_lod_list(o);
    const int highest = lods.empty() ? 0 : lods.back();
    std::string ferr;
    for (size_t li = 0; rc == TDP_OK && li < lods.size(); li++) {
        if (n_vp == 0) break;
        if (feed) {
            rc = feed_round(ctx, j, lods[li], *feed, ferr);
            if (rc != TDP_OK) break;
        }
        rc = tdp_join_refine(j, lods[li], lods[li] == highest, &n_vp);
        res->stats[5]++;
    }
    res->timing[3] = ms_since(t0);
    if (rc != TDP_OK && !ferr.empty()) {
        err = ferr;
        tdp_join_destroy(j);
        return rc;
    }
    t0 = Clock::now();
    // the per-pair results come down into pinned memory (kept by the context
    // between joins): a DMA at full rate, no page faults
    void *pin = nullptr;
    if (rc == TDP_OK) rc = tdp_host_alloc(ctx, 17 * (P ? P : 1), &pin);
    std::unique_ptr<void, std::function<void(void *)>> pin_hold(
        pin, [ctx](void *p) { tdp_host_free(ctx, p); });
    double *const mn = (double *)pin, *const mx = mn + P;
    uint8_t *const rnd = (uint8_t *)(mx + P);
    if (rc == TDP_OK) rc = tdp_join_results(j, rnd, mn, mx);
    uint64_t st[7] = {0, 0, 0, 0, 0, 0, 0};
    if (rc == TDP_OK) rc = tdp_join_stats_n(j, st, 7);
    if (rc == TDP_OK) tdp_join_kernel_ms(j, &res->timing[5], &res->timing[4]);
    if (rc == TDP_OK) tdp_join_stage_ms(j, res->stage_ms, res->stage_n, TDP_STAGE_COUNT);
    if (rc != TDP_OK) {
        err = tdp_last_error(ctx);
        tdp_join_destroy(j);
        return rc;
    }
    tdp_join_destroy(j);
    res->timing[8] = ms_since(t0);
    t0 = Clock::now();
    for (int k = 0; k < 4; k++) res->stats[k] = st[k];
    res->stats[6] = st[4];
    res->stats[7] = st[5];
    res->stats[8] = st[6];
    // per target: reported pairs ordered by (round, list position); targets
    // in parallel, each written at its offset
    std::vector<uint64_t> at(N + 1, 0);
#pragma omp parallel for schedule(static)
    for (long t = 0; t < (long)N; t++) {
        uint64_t c = 0;
        for (uint32_t p = prefix[t]; p < prefix[t + 1]; p++) c += rnd[p] != 255;
        at[t + 1] = c;
    }
    for (size_t t = 0; t < N; t++) at[t + 1] += at[t];
    const size_t n_rep = at[N];
    res->target.resize(n_rep);
    res->obj2.resize(n_rep);
    res->rmin.resize(n_rep);
    res->rmax.resize(n_rep);
    res->round.resize(n_rep);
#pragma omp parallel for schedule(dynamic, 1024)
    for (long t = 0; t < (long)N; t++) {
        uint64_t w = at[t];
        if (w == at[t + 1]) continue;
        // rounds are few: one pass per round present, list order inside each
        uint8_t r = 255, next = 0;
        for (uint32_t p = prefix[t]; p < prefix[t + 1]; p++) r = std::min(r, rnd[p]);
        while (r != 255) {
            next = 255;
            for (uint32_t p = prefix[t]; p < prefix[t + 1]; p++) {
                if (rnd[p] == r) {
                    res->target[w] = (int32_t)t;
                    res->obj2[w] = (int32_t)ids[p];
                    res->rmin[w] = mn[p];
                    res->rmax[w] = mx[p];
                    res->round[w] = r;
                    w++;
                } else if (rnd[p] > r && rnd[p] < next) {
                    next = rnd[p];
                }
            }
            r = next;
        }
    }
    res->timing[9] = ms_since(t0);
    return TDP_OK;
}

tdp_tile_desc tdb_desc_of(const tdb::Tile &t) {
    tdp_tile_desc d;
    d.n_obj = t.n_obj;
    d.n_vox = t.n_vox();
    d.n_tri = t.n_tri();
    d.obj_vox_off = t.obj_vox_off.data();
    d.obj_hd = t.obj_hd.data();
    d.vox_box = t.vox_box.data();
    d.vox_tri_off = t.vox_tri_off.data();
    d.vox_tri_cnt = t.vox_tri_cnt.data();
    d.tri = t.tri.data();
    d.obj_box = t.obj_box.data();
    d.obj_transform = t.obj_transform.data();
    return d;
}

// Candidates with the device where it applies (kNN and within, tile 2 of any
// size, both tiles resident): the
// device lists, and the octree for the targets the device leaves to it
// (built only when there are such targets). Elsewhere the octree for all
// targets.
int tdb_gen_candidates_dev(tdp_ctx *ctx, const tdp_tile *d1, const tdp_tile *d2,
                           const tdb_tile *t1, const tdb_tile *t2, bool self, size_t N,
                           const tdb_join_opts *o, const int32_t *k_t, const float *tau_t,
                           std::vector<uint32_t> &prefix, std::vector<uint32_t> &ids,
                           uint64_t *n_fallback, std::string &err, bool device_within) {
    if (n_fallback) *n_fallback = 0;
    const bool knn = o->query == TDB_QUERY_NN,
               within = o->query == TDB_QUERY_WITHIN && device_within;
    if (!ctx || !d1 || !d2 || !(knn || within) || N == 0) {
        tdb_gen_candidates(t1, t2, self, N, o, k_t, tau_t, prefix, ids);
        return TDP_OK;
    }
    std::vector<int32_t> kv;
    std::vector<float> tv;
    std::vector<float> lbox;
    std::vector<uint32_t> loff, lobj;
    if (knn) {
        kv.resize(N);
        for (size_t t = 0; t < N; t++) kv[t] = k_t ? k_t[t] : o->k;
    } else {
        tv.resize(N);
        for (size_t t = 0; t < N; t++) tv[t] = tau_t ? tau_t[t] : o->within_dist;
        t2->octree().leaves(lbox, loff, lobj);
    }
    std::vector<uint32_t> dpre(N + 1);
    std::vector<uint8_t> fb(N);
    // room for about 16 per target; asked again with the exact count when short
    size_t room = 16 * N + 1024;
    std::unique_ptr<uint32_t[]> dids(new uint32_t[room]);
    uint64_t n_ids = 0;
    auto run = [&]() {
        if (knn)
            return tdp_knn_candidates(ctx, d1, N, d2, self ? 1 : 0, kv.data(), dpre.data(),
                                      dids.get(), room, &n_ids, fb.data());
        return tdp_within_candidates(ctx, d1, N, d2, self ? 1 : 0, tv.data(), loff.size() - 1,
                                     lbox.data(), loff.data(), lobj.data(), dpre.data(),
                                     dids.get(), room, &n_ids, fb.data());
    };
    int rc = run();
    if (rc == TDP_ERR_ARG && n_ids > room) {
        room = n_ids;
        dids.reset(new uint32_t[room]);
        rc = run();
    }
    if (rc != TDP_OK) {
        err = tdp_last_error(ctx);
        return rc;
    }
    std::vector<long> redo;
    for (size_t t = 0; t < N; t++)
        if (fb[t]) redo.push_back((long)t);
    if (n_fallback) *n_fallback = redo.size();
    std::vector<std::vector<uint32_t>> host(redo.size());
    if (redo.empty()) {
        prefix.swap(dpre);
        ids.assign(dids.get(), dids.get() + n_ids);
        return TDP_OK;
    }
    const tdb::Octree &tree = t2->octree();
#ifdef _OPENMP
    const int nt = o->threads > 0 ? o->threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
#endif
    for (long r = 0; r < (long)redo.size(); r++) {
        const long i = redo[r];
        std::vector<std::pair<int, tdb::Range>> c;
        if (knn)
            tree.knn(&t1->t.obj_box[6 * i], &t1->t.obj_transform[3 * i], self ? (int)i : -1,
                     kv[i], c);
        else
            tree.within(&t1->t.obj_box[6 * i], &t1->t.obj_transform[3 * i], self ? (int)i : -1,
                        tv[i], c);
        for (auto &x : c) host[r].push_back((uint32_t)x.first);
    }
    prefix.assign(N + 1, 0);
    size_t r = 0;
    for (size_t t = 0; t < N; t++) {
        const size_t n = fb[t] ? host[r++].size() : dpre[t + 1] - dpre[t];
        prefix[t + 1] = prefix[t] + (uint32_t)n;
    }
    ids.resize(prefix[N]);
    r = 0;
    for (size_t t = 0; t < N; t++) {
        if (fb[t]) std::copy(host[r].begin(), host[r].end(), ids.begin() + prefix[t]), r++;
        else std::copy(dids.get() + dpre[t], dids.get() + dpre[t + 1], ids.begin() + prefix[t]);
    }
    return TDP_OK;
}

extern "C" {

// ------------------------------------------------------------------ tiles

int tdb_tile_load(const char *path, tdb_tile **out, char *err, size_t err_len) {
    if (!path || !out) return -1;
    *out = nullptr;
    std::unique_ptr<tdb_tile> t(new tdb_tile());
    std::string msg;
    if (!tdb::read_raw(path, t->t, msg)) {
        set_err(err, err_len, msg);
        return -1;
    }
    *out = t.release();
    return 0;
}

int tdb_tile_save(const tdb_tile *t, const char *path, char *err, size_t err_len) {
    if (!t || !path) return -1;
    std::string msg;
    if (!tdb::write_raw(path, t->t, msg)) {
        set_err(err, err_len, msg);
        return -1;
    }
    return 0;
}

void tdb_tile_free(tdb_tile *t) {
    if (!t) return;
    {
        std::lock_guard<std::mutex> g(g_registry);
        for (tdb_engine *e : t->holders) {
            auto it = e->resident.find(t);
            if (it != e->resident.end()) {
                tdp_tile_free(it->second);
                e->resident.erase(it);
            }
        }
    }
    delete t;
}

int tdb_tile_clone(const tdb_tile *t, tdb_tile **out) {
    if (!t || !out) return -1;
    tdb_tile *c = new tdb_tile();
    c->t = t->t;
    *out = c;
    return 0;
}

void tdb_tile_size(const tdb_tile *t, uint64_t *out3) {
    out3[0] = t->t.n_obj;
    out3[1] = t->t.n_vox();
    out3[2] = t->t.n_tri();
}

void tdb_tile_views(const tdb_tile *t, const uint32_t **obj_vox_off, const float **obj_box,
                    const float **obj_transform, const float **obj_hd, const float **vox_box,
                    const uint64_t **vox_tri_off, const uint32_t **vox_tri_cnt,
                    const float **tri) {
    if (obj_vox_off) *obj_vox_off = t->t.obj_vox_off.data();
    if (obj_box) *obj_box = t->t.obj_box.data();
    if (obj_transform) *obj_transform = t->t.obj_transform.data();
    if (obj_hd) *obj_hd = t->t.obj_hd.data();
    if (vox_box) *vox_box = t->t.vox_box.data();
    if (vox_tri_off) *vox_tri_off = t->t.vox_tri_off.data();
    if (vox_tri_cnt) *vox_tri_cnt = t->t.vox_tri_cnt.data();
    if (tri) *tri = t->t.tri.data();
}

void tdb_gen_defaults(tdb_gen_params *p) {
    tdb::GenParams g;
    p->seed = g.seed;
    p->n_vessels = g.n_vessels;
    p->nuclei_per_vessel = g.nuclei_per_vessel;
    p->cube = g.cube;
    p->vessel_r_min = g.vessel_r_min;
    p->vessel_r_max = g.vessel_r_max;
    p->vessel_segments = g.vessel_segments;
    p->segment_length = g.segment_length;
    p->turn_max_deg = g.turn_max_deg;
    p->vessel_sides = g.vessel_sides;
    p->rings_min = g.rings_min;
    p->rings_max = g.rings_max;
    p->nucleus_r_min = g.nucleus_r_min;
    p->nucleus_r_max = g.nucleus_r_max;
    p->noise = g.noise;
    p->subdiv = g.subdiv;
    p->place_min = g.place_min;
    p->place_max = g.place_max;
    p->prototypes = g.prototypes;
    p->voxel_tris = g.voxel_tris;
    p->threads = g.threads;
    p->max_nuclei = 0;
    p->nuclei_seed = 0;
    p->emit_vessels = 1;
}

int tdb_generate(const tdb_gen_params *p, tdb_tile **nuclei, tdb_tile **vessels) {
    if (!p || !nuclei || !vessels) return -1;
    if (p->n_vessels < 0 || p->nuclei_per_vessel < 0 || p->vessel_segments < 1 ||
        p->vessel_sides < 3 || p->rings_min < 1 || p->rings_max < p->rings_min ||
        p->subdiv < 0 || p->subdiv > 6 || p->voxel_tris < 1 || p->prototypes < 1)
        return -1;
    tdb::GenParams g;
    g.seed = p->seed;
    g.n_vessels = p->n_vessels;
    g.nuclei_per_vessel = p->nuclei_per_vessel;
    g.cube = p->cube;
    g.vessel_r_min = p->vessel_r_min;
    g.vessel_r_max = p->vessel_r_max;
    g.vessel_segments = p->vessel_segments;
    g.segment_length = p->segment_length;
    g.turn_max_deg = p->turn_max_deg;
    g.vessel_sides = p->vessel_sides;
    g.rings_min = p->rings_min;
    g.rings_max = p->rings_max;
    g.nucleus_r_min = p->nucleus_r_min;
    g.nucleus_r_max = p->nucleus_r_max;
    g.noise = p->noise;
    g.subdiv = p->subdiv;
    g.place_min = p->place_min;
    g.place_max = p->place_max;
    g.prototypes = p->prototypes;
    g.voxel_tris = p->voxel_tris;
    g.threads = p->threads;
    g.max_nuclei = p->max_nuclei;
    g.nuclei_seed = p->nuclei_seed;
    g.emit_vessels = p->emit_vessels;
    std::unique_ptr<tdb_tile> n(new tdb_tile()), v(new tdb_tile());
    tdb::generate(g, n->t, v->t);
    *nuclei = n.release();
    *vessels = v.release();
    return 0;
}

// ------------------------------------------------------------------ engine

int tdb_engine_open(int device, tdb_engine **out) {
    if (!out) return -1;
    tdb_engine *e = new tdb_engine();
    e->device = device;
    int rc = tdp_open(device, &e->ctx);
    if (rc != TDP_OK) {
        e->err = e->ctx ? tdp_last_error(e->ctx) : "tdp_open failed";
        if (e->ctx) tdp_close(e->ctx);
        e->ctx = nullptr;
        *out = e;  // so the message can be read; the engine cannot join
        return rc;
    }
    *out = e;
    return 0;
}

void tdb_engine_close(tdb_engine *e) {
    if (!e) return;
    {
        std::lock_guard<std::mutex> g(g_registry);
        for (auto &kv : e->resident) {
            auto &h = kv.first->holders;
            h.erase(std::remove(h.begin(), h.end(), e), h.end());
            tdp_tile_free(kv.second);
        }
        e->resident.clear();
    }
    if (e->ctx) tdp_close(e->ctx);
    delete e;
}

const char *tdb_engine_error(const tdb_engine *e) { return e ? e->err.c_str() : "null engine"; }

int tdb_engine_device(tdb_engine *e, char *name, size_t len, int *cu) {
    if (!e || !e->ctx) return -1;
    return tdp_device_info(e->ctx, name, len, cu);
}

void tdb_engine_evict(tdb_engine *e, const tdb_tile *t) {
    if (!e || !t) return;
    std::lock_guard<std::mutex> g(g_registry);
    auto it = e->resident.find(t);
    if (it == e->resident.end()) return;
    tdp_tile_free(it->second);
    e->resident.erase(it);
    auto &h = t->holders;
    h.erase(std::remove(h.begin(), h.end(), e), h.end());
}

int tdb_engine_get_tuning(tdb_engine *e, tdp_tuning *out) {
    if (!e || !e->ctx) return TDP_ERR_ARG;
    return tdp_get_tuning(e->ctx, out);
}

int tdb_engine_set_tuning(tdb_engine *e, const tdp_tuning *t) {
    if (!e || !e->ctx) return TDP_ERR_ARG;
    const int rc = tdp_set_tuning(e->ctx, t);
    if (rc != TDP_OK) e->err = tdp_last_error(e->ctx);
    return rc;
}

static int resident(tdb_engine *e, const tdb_tile *t, tdp_tile **out) {
    std::lock_guard<std::mutex> g(g_registry);
    auto it = e->resident.find(t);
    if (it != e->resident.end()) {
        *out = it->second;
        return 0;
    }
    const tdp_tile_desc d = tdb_desc_of(t->t);
    int rc = tdp_tile_upload(e->ctx, &d, out);
    if (rc != TDP_OK) {
        e->err = tdp_last_error(e->ctx);
        return rc;
    }
    e->resident[t] = *out;
    t->holders.push_back(e);
    return 0;
}

void tdb_join_defaults(tdb_join_opts *o) {
    o->query = TDB_QUERY_NN;
    o->k = 1;
    o->within_dist = 1000.0f;
    o->lods = nullptr;
    o->n_lods = 0;
    o->hausdorff_level = 2;
    o->max_objects1 = 0;
    o->precision = TDP_PREC_MIXED;
    o->threads = 0;
}

}  // extern "C"

// LOD list post-processing of parse_args (query_context.h:183-188): sort,
// then std::unique without erase, which keeps the old tail.
std::vector<int> tdb_lod_list(const tdb_join_opts *o) {
    std::vector<int> l;
    if (o->lods && o->n_lods > 0) l.assign(o->lods, o->lods + o->n_lods);
    else
        for (int v = 20; v <= 100; v += 20) l.push_back(v);
    std::sort(l.begin(), l.end());
    std::unique(l.begin(), l.end());
    return l;
}

static size_t n_targets(const tdb_tile *t1, const tdb_join_opts *o) {
    size_t n = t1->t.n_obj;
    if (o->max_objects1 > 0 && o->max_objects1 < n) n = (size_t)o->max_objects1;
    return n;
}

// candidate lists of every target (KNNJoin.cpp:552-573, WithinJoin.cpp:79-97)
void tdb_gen_candidates(const tdb_tile *t1, const tdb_tile *t2, bool self, size_t N,
                        const tdb_join_opts *o, const int32_t *k_t, const float *tau_t,
                        std::vector<uint32_t> &prefix, std::vector<uint32_t> &ids) {
    const tdb::Octree &tree = t2->octree();
    std::vector<std::vector<uint32_t>> per(N);
#ifdef _OPENMP
    const int nt = o->threads > 0 ? o->threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
#endif
    for (long i = 0; i < (long)N; i++) {
        std::vector<std::pair<int, tdb::Range>> c;
        const float *box = &t1->t.obj_box[6 * i];
        const float *tr = &t1->t.obj_transform[3 * i];
        if (o->query == TDB_QUERY_INTERSECT) {
            std::vector<int> ids;
            tree.intersect(box, self ? (int)i : -1, ids);
            per[i].assign(ids.begin(), ids.end());
            continue;
        }
        if (o->query == TDB_QUERY_NN)
            tree.knn(box, tr, self ? (int)i : -1, k_t ? k_t[i] : o->k, c);
        else
            tree.within(box, tr, self ? (int)i : -1, tau_t ? tau_t[i] : o->within_dist, c);
        per[i].reserve(c.size());
        for (auto &x : c) per[i].push_back((uint32_t)x.first);
    }
    prefix.assign(N + 1, 0);
    for (size_t i = 0; i < N; i++) prefix[i + 1] = prefix[i] + (uint32_t)per[i].size();
    ids.resize(prefix[N]);
    for (size_t i = 0; i < N; i++) std::copy(per[i].begin(), per[i].end(), ids.begin() + prefix[i]);
}

extern "C" {

int64_t tdb_candidates(const tdb_tile *t1, const tdb_tile *t2, const tdb_join_opts *o,
                       uint32_t *prefix, uint32_t *ids, size_t cap) {
    if (!t1 || !o) return -1;
    const bool self = (t2 == nullptr || t2 == t1);
    std::vector<uint32_t> pf, id;
    tdb_gen_candidates(t1, self ? t1 : t2, self, n_targets(t1, o), o, nullptr, nullptr, pf, id);
    if (prefix) std::copy(pf.begin(), pf.end(), prefix);
    if (ids) std::copy(id.begin(), id.begin() + std::min(cap, id.size()), ids);
    return (int64_t)id.size();
}

int64_t tdb_engine_candidates(tdb_engine *e, const tdb_tile *t1, const tdb_tile *t2,
                              const tdb_join_opts *o, uint32_t *prefix, uint32_t *ids, size_t cap,
                              uint64_t *n_fallback) {
    if (!e || !t1 || !o || !e->ctx) return -1;
    const bool self = (t2 == nullptr || t2 == t1);
    const tdb_tile *T2 = self ? t1 : t2;
    tdp_tile *d1 = nullptr, *d2 = nullptr;
    if (resident(e, t1, &d1) != 0 || resident(e, T2, &d2) != 0) return -1;
    std::vector<uint32_t> pf, id;
    if (tdb_gen_candidates_dev(e->ctx, d1, d2, t1, T2, self, n_targets(t1, o), o, nullptr, nullptr,
                               pf, id, n_fallback, e->err) != TDP_OK)
        return -1;
    if (prefix) std::copy(pf.begin(), pf.end(), prefix);
    if (ids) std::copy(id.begin(), id.begin() + std::min(cap, id.size()), ids);
    return (int64_t)id.size();
}

int tdb_engine_autotune(tdb_engine *e, const tdb_tile *t1, const tdb_tile *t2,
                        const tdb_join_opts *o, size_t sample, tdp_tuning *best,
                        tdp_tuning *tried, double *ms, int cap) {
    if (!e || !e->ctx || !t1 || !o) return -1;
    tdp_tuning base;
    if (tdp_get_tuning(e->ctx, &base) != TDP_OK) return -1;
    // the engine's tuning first, then a fixed grid -- FP64 stage: work list
    // (0), persistent (1), persistent in a captured round (2)
    std::vector<tdp_tuning> cand{base};
    auto same = [](const tdp_tuning &a, const tdp_tuning &b) {
        return a.small_pairs == b.small_pairs && a.max_rows == b.max_rows && a.drain == b.drain &&
               a.drain_blocks == b.drain_blocks && a.eval_blocks == b.eval_blocks &&
               a.graph == b.graph && a.row_cull == b.row_cull;
    };
    for (int form : {0, 1, 2})
        for (int rows : {64, 4})
            for (int small : {8, 64}) {
                tdp_tuning t = base;
                t.drain = form >= 1;
                t.graph = form == 2;
                t.max_rows = rows;
                t.small_pairs = small;
                t.drain_blocks = form ? 8 * 256 : 0;
                t.eval_blocks = 4096;
                if (!same(t, base)) cand.push_back(t);
            }
    tdb_join_opts so = *o;
    if (sample && (so.max_objects1 == 0 || so.max_objects1 > sample)) so.max_objects1 = sample;
    double best_ms = 1e300, base_ms = 1e300;
    tdp_tuning pick = base;
    int n = 0;
    for (const tdp_tuning &t : cand) {
        if (tdp_set_tuning(e->ctx, &t) != TDP_OK) continue;
        tdb_result *r = nullptr;
        double dev = 1e300;
        // twice: the first run also warms the buffer cache
        for (int rep = 0; rep < 2; rep++) {
            if (tdb_join(e, t1, t2, &so, &r) != 0) {
                tdp_set_tuning(e->ctx, &base);
                return -1;
            }
            // wall time of filter and refinement: host synchronisation counts
            dev = std::min(dev, r->timing[2] + r->timing[3]);
            tdb_result_free(r);
        }
        if (tried && n < cap) tried[n] = t;
        if (ms && n < cap) ms[n] = dev;
        if (n == 0) base_ms = dev;
        n++;
        if (dev < best_ms) {
            best_ms = dev;
            pick = t;
        }
    }
    // the engine's tuning stays unless another is clearly faster (timing
    // noise on the sample is a few percent)
    if (base_ms <= best_ms * 1.03) pick = base;
    tdp_set_tuning(e->ctx, &pick);
    if (best) *best = pick;
    return n;
}

int tdb_join(tdb_engine *e, const tdb_tile *t1, const tdb_tile *t2, const tdb_join_opts *o,
             tdb_result **out) {
    if (!e || !t1 || !o || !out) return -1;
    *out = nullptr;
    if (!e->ctx) {
        e->err = "engine has no device: " + e->err;
        return TDP_ERR_HIP;
    }
    if (o->query != TDB_QUERY_NN && o->query != TDB_QUERY_WITHIN &&
        o->query != TDB_QUERY_INTERSECT) {
        e->err = "query must be nn, within or intersect";
        return TDP_ERR_ARG;
    }
    if (o->query == TDB_QUERY_NN && o->k < 1) {
        e->err = "--knn must be at least 1";
        return TDP_ERR_ARG;
    }
    const bool self = (t2 == nullptr || t2 == t1);
    const tdb_tile *T2 = self ? t1 : t2;
    std::unique_ptr<tdb_result> res(new tdb_result());
    const auto t_start = Clock::now();

    auto t0 = Clock::now();
    tdp_tile *d1 = nullptr, *d2 = nullptr;
    int rc = resident(e, t1, &d1);
    if (rc == 0) rc = resident(e, T2, &d2);
    if (rc != 0) return rc;
    res->timing[6] = ms_since(t0);

    t0 = Clock::now();
    std::vector<uint32_t> prefix, ids;
    rc = tdb_gen_candidates_dev(e->ctx, d1, d2, t1, T2, self, n_targets(t1, o), o, nullptr, nullptr,
                                prefix, ids, nullptr, e->err);
    if (rc != TDP_OK) return rc;
    res->timing[1] = ms_since(t0);
    rc = tdb_run_join(e->ctx, d1, d2, o, prefix, ids, nullptr, nullptr, res.get(), e->err);
    if (rc != TDP_OK) return rc;
    res->timing[0] = ms_since(t_start);
    *out = res.release();
    return 0;
}

uint64_t tdb_result_count(const tdb_result *r) { return r ? r->target.size() : 0; }
uint64_t tdb_result_targets(const tdb_result *r) { return r ? r->n_targets : 0; }

void tdb_result_get(const tdb_result *r, int32_t *target, int32_t *obj2, double *rmin,
                    double *rmax, uint8_t *round) {
    const long n = (long)r->target.size();
    // in slices over the threads (the caller's arrays are often fresh)
    const long sl = 1 << 16;
#pragma omp parallel for schedule(static) if (n > 4 * sl)
    for (long a = 0; a < n; a += sl) {
        const long b = std::min(n, a + sl);
        if (target) std::copy(r->target.begin() + a, r->target.begin() + b, target + a);
        if (obj2) std::copy(r->obj2.begin() + a, r->obj2.begin() + b, obj2 + a);
        if (rmin) std::copy(r->rmin.begin() + a, r->rmin.begin() + b, rmin + a);
        if (rmax) std::copy(r->rmax.begin() + a, r->rmax.begin() + b, rmax + a);
        if (round) std::copy(r->round.begin() + a, r->round.begin() + b, round + a);
    }
}

void tdb_result_timing(const tdb_result *r, double *out7) {
    for (int k = 0; k < 7; k++) out7[k] = r->timing[k];
}
int tdb_result_timing_n(const tdb_result *r, double *out, int n) {
    if (!r || !out || n < 0) return 0;
    n = std::min(n, 10);
    for (int k = 0; k < n; k++) out[k] = r->timing[k];
    return n;
}
void tdb_result_stats(const tdb_result *r, uint64_t *out7) {
    for (int k = 0; k < 7; k++) out7[k] = r->stats[k];
}
int tdb_result_stats_n(const tdb_result *r, uint64_t *out, int n) {
    if (!r || !out || n < 0) return 0;
    n = std::min(n, 9);
    for (int k = 0; k < n; k++) out[k] = r->stats[k];
    return n;
}
void tdb_result_stages(const tdb_result *r, double *ms, uint64_t *launches, int n) {
    for (int k = 0; k < n; k++) {
        if (ms) ms[k] = k < TDP_STAGE_COUNT ? r->stage_ms[k] : 0.0;
        if (launches) launches[k] = k < TDP_STAGE_COUNT ? r->stage_n[k] : 0;
    }
}
void tdb_result_free(tdb_result *r) { delete r; }

}  // extern "C"
