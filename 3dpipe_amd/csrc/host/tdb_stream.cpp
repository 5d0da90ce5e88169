// Streaming join: tile sets larger than the device budget, joined chunk by
// chunk with host decoding, pinned copies and device joins overlapped.
//
// The reference streams voxel-pair chunks through pinned slots with one
// producer thread packing triangles and the join thread computing
// (KNNJoin.cpp:1524-1811 with the producer/consumer at :1586-1617,
// packing_data_stream SpatialJoin.cpp:169-248, and the filter's _reduceIO
// pipeline computing_gpu.cu:4417-4862). Here the unit streamed is a chunk of
// targets: the targets, taken in Morton order of their box centres so that
// neighbours share chunks, together with every object their candidate lists
// name, form one device tile; the chunk's join (filter + every LOD round)
// runs on that tile alone, so nothing a chunk needs lives outside it and a
// chunk's results are those of the same targets in one big join.
//
// Workers -- each a device context (stream) with its own reserved tiles,
// pinned staging and join buffers -- take chunks in turn; a worker decodes
// its chunk on host threads (ObjectSource, the LOD decode of the reference's
// HiMesh_Wrapper::decode_to, himesh_wrapper.cpp:137-161), fills its tiles
// and joins them, so one worker's host work runs beside another's kernels.
// By default a chunk's tiles go over without their triangle records
// (tdp_tile_fill_meta_async): before each LOD round the device lists the
// voxels the round's live voxel pairs read, and the worker decodes just those
// objects' records at that LOD (ObjectSource::decode_slot) and hands the
// listed voxels' records over (tdp_join_lazy_put) -- the reference's per-LOD
// packing of live candidates' voxels (KNNJoin.cpp:1584-1617,
// SpatialJoin.cpp:169-248). Voxel pairs pruned at LOD 20 never ship their
// higher LODs. The device budget covers every worker's tiles and join
// buffers; a chunk's join that would pass its worker's cap is split into
// target ranges on the same tile.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "tdb_internal.hpp"

struct tdb_source {
    std::unique_ptr<tdb::ObjectSource> src;
};

namespace {

// device bytes of one object in a reserved tile: records + bounds, voxel
// arrays (with the per-round lists of a tile filled without records),
// object arrays
inline uint64_t obj_bytes(const tdb::ObjectSource &s, size_t o) {
    return 96 * s.n_tri(o) + 132 * (uint64_t)s.n_vox(o) + 48;
}

// a device tile holding the index's object boxes and proxy points only (no
// voxels, no records): what the device candidate lists read
int upload_index(tdp_ctx *ctx, const tdb_tile &ix, tdp_tile **out) {
    const size_t n = ix.t.n_obj;
    std::vector<uint32_t> vox_off(n + 1, 0u);
    std::vector<float> hd(10 * n, 0.f);
    tdp_tile_desc d{};
    d.n_obj = n;
    d.obj_vox_off = vox_off.data();
    d.obj_hd = hd.data();
    d.obj_box = ix.t.obj_box.data();
    d.obj_transform = ix.t.obj_transform.data();
    return tdp_tile_upload(ctx, &d, out);
}

struct Chunk {
    std::vector<uint32_t> targets;      // targets (ids in source 1), local 0..n-1
    std::vector<uint32_t> objs[2];      // objects of tile 1 (self: targets, then the rest) / tile 2
    std::vector<uint32_t> prefix, ids;  // candidate lists in local ids of tile 2 (self: tile 1)
    uint64_t n_vox[2] = {0, 0}, n_tri[2] = {0, 0};
};

// pinned arrays of one tile
struct Staging {
    uint32_t *obj_vox_off = nullptr;
    float *obj_hd = nullptr, *vox_box = nullptr, *tri = nullptr;
    uint64_t *vox_tri_off = nullptr;
    uint32_t *vox_tri_cnt = nullptr;
    tdp_tile_desc desc{};
};

uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    auto spread = [](uint64_t v) {
        v &= 0x1fffff;
        v = (v | v << 32) & 0x1f00000000ffffULL;
        v = (v | v << 16) & 0x1f0000ff0000ffULL;
        v = (v | v << 8) & 0x100f00f00f00f00fULL;
        v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
        v = (v | v << 2) & 0x1249249249249249ULL;
        return v;
    };
    return spread(x) | spread(y) << 1 | spread(z) << 2;
}

// targets in Morton order of their box centres
std::vector<uint32_t> morton_order(const std::vector<float> &box, size_t n) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            const float c = 0.5f * (box[6 * i + d] + box[6 * i + 3 + d]);
            lo[d] = std::min(lo[d], c);
            hi[d] = std::max(hi[d], c);
        }
    std::vector<std::pair<uint64_t, uint32_t>> key(n);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) {
        uint32_t q[3];
        for (int d = 0; d < 3; d++) {
            const float c = 0.5f * (box[6 * i + d] + box[6 * i + 3 + d]);
            const double span = hi[d] > lo[d] ? (double)hi[d] - lo[d] : 1.0;
            q[d] = (uint32_t)std::min(2097151.0, std::max(0.0, (c - lo[d]) / span * 2097151.0));
        }
        key[i] = {morton3(q[0], q[1], q[2]), (uint32_t)i};
    }
    std::sort(key.begin(), key.end());
    std::vector<uint32_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = key[i].second;
    return order;
}

// Greedy chunks over the targets in order: a chunk takes targets while its
// tiles (targets + every object their lists name) fit slot_bytes.
bool plan_chunks(const tdb::ObjectSource &s1, const tdb::ObjectSource &s2, bool self,
                 const std::vector<uint32_t> &order, const std::vector<uint32_t> &prefix,
                 const std::vector<uint32_t> &ids, uint64_t slot_bytes, uint64_t max_targets,
                 std::vector<Chunk> &chunks, std::string &err) {
    const size_t n2 = s2.n_obj();
    std::vector<int32_t> seen1(self ? n2 : s1.n_obj(), -1), seen2(self ? 0 : n2, -1);
    int32_t cur = 0;
    Chunk c;
    uint64_t bytes = 0;
    auto close = [&]() {
        if (c.targets.empty()) return;
        chunks.push_back(std::move(c));
        c = Chunk();
        bytes = 0;
        cur++;
    };
    for (size_t k = 0; k < order.size(); k++) {
        const uint32_t t = order[k];
        // bytes this target adds
        uint64_t add = 0;
        std::vector<int32_t> &seenT = seen1;
        std::vector<int32_t> &seenC = self ? seen1 : seen2;
        if (seenT[t] != cur) add += obj_bytes(s1, t);
        for (uint32_t p = prefix[k]; p < prefix[k + 1]; p++) {
            const uint32_t o = ids[p];
            if (seenC[o] != cur && !(self && o == t)) add += obj_bytes(s2, o);
        }
        if (!c.targets.empty() && (bytes + add > slot_bytes || c.targets.size() >= max_targets)) {
            close();
            k--;  // this target starts the next chunk
            continue;
        }
        if (add > slot_bytes) {
            err = "stream join: one target with its candidates needs " + std::to_string(add) +
                  " device bytes, more than a tile slot (" + std::to_string(slot_bytes) + ")";
            return false;
        }
        c.targets.push_back(t);
        if (seenT[t] != cur) {
            seenT[t] = cur;
            c.objs[0].push_back(t);  // order fixed when the chunk is finished
        }
        for (uint32_t p = prefix[k]; p < prefix[k + 1]; p++) {
            const uint32_t o = ids[p];
            if (seenC[o] != cur) {
                seenC[o] = cur;
                c.objs[self ? 0 : 1].push_back(o);
            }
        }
        bytes += add;
    }
    close();
    return true;
}

// Local numbering of a planned chunk: tile 1 holds the targets first (self:
// then the other objects in first-seen order); the lists are renumbered.
void finish_chunk(Chunk &c, bool self, const std::vector<uint32_t> &kpos,
                  const std::vector<uint32_t> &prefix, const std::vector<uint32_t> &ids,
                  const tdb::ObjectSource &s1, const tdb::ObjectSource &s2,
                  std::vector<int32_t> &loc) {
    if (self) {
        std::vector<uint32_t> all;
        all.swap(c.objs[0]);
        c.objs[0] = c.targets;
        for (uint32_t t : c.targets) loc[t] = 1;  // mark targets
        for (uint32_t o : all)
            if (loc[o] != 1) c.objs[0].push_back(o);
        for (uint32_t t : c.targets) loc[t] = -1;
    } else {
        c.objs[0] = c.targets;
    }
    const std::vector<uint32_t> &T2 = self ? c.objs[0] : c.objs[1];
    for (size_t i = 0; i < T2.size(); i++) loc[T2[i]] = (int32_t)i;
    c.prefix.assign(c.targets.size() + 1, 0);
    for (size_t i = 0; i < c.targets.size(); i++) {
        const uint32_t k = kpos[c.targets[i]];
        for (uint32_t p = prefix[k]; p < prefix[k + 1]; p++) c.ids.push_back((uint32_t)loc[ids[p]]);
        c.prefix[i + 1] = (uint32_t)c.ids.size();
    }
    for (uint32_t o : T2) loc[o] = -1;
    for (int w = 0; w < (self ? 1 : 2); w++) {
        const tdb::ObjectSource &s = w == 0 ? s1 : s2;
        for (uint32_t o : c.objs[w]) {
            c.n_vox[w] += s.n_vox(o);
            c.n_tri[w] += s.n_tri(o);
        }
    }
}

int alloc_staging(tdp_ctx *ctx, size_t n_obj, size_t n_vox, size_t n_tri, Staging &st) {
    int rc = tdp_host_alloc(ctx, 4 * (n_obj + 1), (void **)&st.obj_vox_off);
    if (rc == TDP_OK) rc = tdp_host_alloc(ctx, 40 * n_obj, (void **)&st.obj_hd);
    if (rc == TDP_OK) rc = tdp_host_alloc(ctx, 36 * n_vox, (void **)&st.vox_box);
    if (rc == TDP_OK) rc = tdp_host_alloc(ctx, 40 * n_vox, (void **)&st.vox_tri_off);
    if (rc == TDP_OK) rc = tdp_host_alloc(ctx, 20 * n_vox, (void **)&st.vox_tri_cnt);
    if (rc == TDP_OK) rc = tdp_host_alloc(ctx, 48 * n_tri, (void **)&st.tri);
    return rc;
}

void free_staging(tdp_ctx *ctx, Staging &st) {
    for (void *p : {(void *)st.obj_vox_off, (void *)st.obj_hd, (void *)st.vox_box,
                    (void *)st.vox_tri_off, (void *)st.vox_tri_cnt, (void *)st.tri})
        if (p) tdp_host_free(ctx, p);
    st = Staging();
}

// the objects' arrays into pinned staging, laid out as one tile (meta_only:
// without the records, which then come round by round)
// meta_only: without the records, which then come round by round, and the
// objects' decode state for those later decodes in state (object i's at
// state_off[i])
void decode_tile(const tdb::ObjectSource &s, const std::vector<uint32_t> &objs, Staging &st,
                 int threads, bool meta_only, std::vector<float> *state = nullptr,
                 std::vector<uint64_t> *state_off = nullptr) {
    const size_t n = objs.size();
    std::vector<uint64_t> toff(n + 1, 0);
    st.obj_vox_off[0] = 0;
    for (size_t i = 0; i < n; i++) {
        st.obj_vox_off[i + 1] = st.obj_vox_off[i] + s.n_vox(objs[i]);
        toff[i + 1] = toff[i] + s.n_tri(objs[i]);
    }
    if (meta_only && state) {
        state_off->assign(n + 1, 0);
        for (size_t i = 0; i < n; i++) (*state_off)[i + 1] = (*state_off)[i] + s.state_floats(objs[i]);
        if (state->size() < state_off->back()) state->resize(state_off->back());
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
    for (long i = 0; i < (long)n; i++) {
        const uint32_t v0 = st.obj_vox_off[i], nv = st.obj_vox_off[i + 1] - v0;
        tdb::ObjArrays out{st.vox_box + 9 * (size_t)v0, st.vox_tri_cnt + 5 * (size_t)v0,
                           st.obj_hd + 10 * (size_t)i, meta_only ? nullptr : st.tri + 12 * toff[i]};
        if (meta_only)
            s.decode_meta(objs[i], out, state ? state->data() + (*state_off)[i] : nullptr);
        else
            s.decode(objs[i], out);
        uint64_t cur = toff[i];
        for (int sl = 0; sl < tdb::kLods; sl++)
            for (uint32_t v = 0; v < nv; v++) {
                st.vox_tri_off[5 * (size_t)(v0 + v) + sl] = cur;
                cur += st.vox_tri_cnt[5 * (size_t)(v0 + v) + sl];
            }
    }
    tdp_tile_desc &d = st.desc;
    d = tdp_tile_desc();
    d.n_obj = n;
    d.n_vox = st.obj_vox_off[n];
    d.n_tri = toff[n];
    d.obj_vox_off = st.obj_vox_off;
    d.obj_hd = st.obj_hd;
    d.vox_box = st.vox_box;
    d.vox_tri_off = st.vox_tri_off;
    d.vox_tri_cnt = st.vox_tri_cnt;
    d.tri = meta_only ? nullptr : st.tri;
}

// one chunk decoded on the host, per tile: pinned arrays, and (lazy) the
// chunk-local object of each voxel and the objects' decode state
struct Decoded {
    Staging stg[2];
    std::vector<uint32_t> vobj[2];
    std::vector<float> state[2];
    std::vector<uint64_t> state_off[2];
};

// one pipeline worker: a context (stream) with its tile slots, two decode
// sets (the next chunk is decoded while the current one joins) and record
// feed, and what it did
struct Worker {
    tdp_ctx *ctx = nullptr;
    tdp_tile *tile[2] = {nullptr, nullptr};
    Decoded set[2];
    LazyFeed feed;
    double decode_ms = 0, copy_ms = 0, join_ms = 0;
    uint64_t h2d = 0, n_decoded = 0, splits = 0, meta_bytes = 0, all_records = 0;
};

// the first failure of any worker (the others stop at their next chunk)
struct Progress {
    std::mutex mu;
    std::atomic<bool> failed{false};
    std::string err;
    void fail(const std::string &e) {
        std::lock_guard<std::mutex> g(mu);
        if (!failed) err = e;
        failed = true;
    }
};

}  // namespace

extern "C" {

int tdb_source_generate(const tdb_gen_params *p, tdb_source **out) {
    if (!p || !out) return -1;
    *out = nullptr;
    if (p->n_vessels < 0 || p->nuclei_per_vessel < 0 || p->subdiv < 0 || p->subdiv > 6 ||
        p->voxel_tris < 1 || p->prototypes < 1)
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
    tdb_source *s = new tdb_source();
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    if (p->threads > 0) omp_set_num_threads(p->threads);
#endif
    s->src = tdb::nucleus_source(g);
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    *out = s;
    return 0;
}

int tdb_source_tile(const tdb_tile *t, tdb_source **out) {
    if (!t || !out) return -1;
    tdb_source *s = new tdb_source();
    s->src = tdb::tile_source(t->t);
    *out = s;
    return 0;
}

uint64_t tdb_source_objects(const tdb_source *s) { return s ? s->src->n_obj() : 0; }

void tdb_source_free(tdb_source *s) { delete s; }

int tdb_source_decode(const tdb_source *s, const uint64_t *objects, size_t n, tdb_tile **out) {
    if (!s || !out || (n && !objects)) return -1;
    *out = nullptr;
    const tdb::ObjectSource &src = *s->src;
    for (size_t i = 0; i < n; i++)
        if (objects[i] >= src.n_obj()) return -1;
    std::unique_ptr<tdb_tile> t(new tdb_tile());
    tdb::Tile &T = t->t;
    T.n_obj = n;
    T.obj_vox_off.assign(n + 1, 0);
    std::vector<uint64_t> toff(n + 1, 0);
    for (size_t i = 0; i < n; i++) {
        T.obj_vox_off[i + 1] = T.obj_vox_off[i] + src.n_vox(objects[i]);
        toff[i + 1] = toff[i] + src.n_tri(objects[i]);
    }
    const size_t nv = T.obj_vox_off[n];
    T.vox_box.resize(9 * nv);
    T.vox_obj.resize(nv);
    T.vox_tri_off.resize(5 * nv);
    T.vox_tri_cnt.resize(5 * nv);
    T.tri.resize(12 * toff[n]);
    T.obj_hd.resize(10 * n);
    T.obj_transform.resize(3 * n);
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < (long)n; i++) {
        const uint32_t v0 = T.obj_vox_off[i], cnt = T.obj_vox_off[i + 1] - v0;
        src.decode(objects[i], tdb::ObjArrays{&T.vox_box[9 * (size_t)v0], &T.vox_tri_cnt[5 * (size_t)v0],
                                              &T.obj_hd[10 * (size_t)i], &T.tri[12 * toff[i]]});
        float box[6];
        src.meta(objects[i], box, &T.obj_transform[3 * i]);
        uint64_t cur = toff[i];
        for (int sl = 0; sl < tdb::kLods; sl++)
            for (uint32_t v = 0; v < cnt; v++) {
                T.vox_obj[v0 + v] = (uint32_t)i;
                T.vox_tri_off[5 * (size_t)(v0 + v) + sl] = cur;
                cur += T.vox_tri_cnt[5 * (size_t)(v0 + v) + sl];
            }
    }
    T.finalize_boxes();
    *out = t.release();
    return 0;
}

void tdb_stream_defaults(tdb_stream_opts *o) {
    o->hbm_budget = 0;
    o->tile_share = 0.5;
    o->max_chunk_targets = 0;
    o->decode_threads = 0;
    o->overlap = 1;
    o->lazy = 1;
    o->workers = 0;
}

void tdb_result_stream(const tdb_result *r, double *out12) {
    for (int k = 0; k < 12; k++) out12[k] = r ? r->stream[k] : 0.0;
}

int tdb_result_stream_n(const tdb_result *r, double *out, int n) {
    if (!out || n < 0) return 0;
    n = std::min(n, 19);
    for (int k = 0; k < n; k++) out[k] = r ? r->stream[k] : 0.0;
    return n;
}

int tdb_stream_join(tdb_engine *e, const tdb_source *src1, const tdb_source *src2,
                    const tdb_join_opts *o, const tdb_stream_opts *so, tdb_result **out) {
    if (!e || !src1 || !o || !so || !out) return -1;
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
    if (so->tile_share <= 0 || so->tile_share >= 1) {
        e->err = "tile_share must lie in (0, 1)";
        return TDP_ERR_ARG;
    }
    const bool self = src2 == nullptr || src2 == src1;
    const tdb::ObjectSource &s1 = *src1->src, &s2 = self ? s1 : *src2->src;
    const auto t_start = Clock::now();
    std::unique_ptr<tdb_result> res(new tdb_result());
    size_t N = s1.n_obj();
    if (o->max_objects1 > 0 && o->max_objects1 < N) N = (size_t)o->max_objects1;
    res->n_targets = N;
    const int nt = so->decode_threads > 0 ? so->decode_threads
#ifdef _OPENMP
                                          : omp_get_max_threads();
#else
                                          : 1;
#endif

    // index: object boxes and proxy points of both sides (a light decode)
    auto t0 = Clock::now();
    tdb_tile idx1, idx2;
    auto index = [&](const tdb::ObjectSource &s, size_t n, tdb_tile &ix) {
        ix.t.n_obj = n;
        ix.t.obj_box.resize(6 * n);
        ix.t.obj_transform.resize(3 * n);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
#endif
        for (long i = 0; i < (long)n; i++) s.meta(i, &ix.t.obj_box[6 * i], &ix.t.obj_transform[3 * i]);
    };
    index(s1, self ? s1.n_obj() : N, idx1);
    if (!self) index(s2, s2.n_obj(), idx2);
    const tdb_tile &I2 = self ? idx1 : idx2;
    res->stream[8] = ms_since(t0);

    // candidate lists of all targets, in target order: kNN and within lists
    // on the device from the two sides' boxes and proxy points (the octree
    // for the targets it leaves, and the within walk's leaves), intersection
    // lists from the octree
    t0 = Clock::now();
    std::vector<uint32_t> order = morton_order(idx1.t.obj_box, N);
    std::vector<uint32_t> prefix, ids;
    {
        std::vector<uint32_t> p0, i0;
        res->stream[16] = -1;
        if (o->query == TDB_QUERY_NN || o->query == TDB_QUERY_WITHIN) {
            tdp_tile *d1 = nullptr, *d2 = nullptr;
            int crc = upload_index(e->ctx, idx1, &d1);
            if (crc == TDP_OK) crc = self ? TDP_OK : upload_index(e->ctx, I2, &d2);
            uint64_t n_fb = 0;
            if (crc == TDP_OK)
                crc = tdb_gen_candidates_dev(e->ctx, d1, self ? d1 : d2, &idx1, &I2, self, N, o,
                                             nullptr, nullptr, p0, i0, &n_fb, e->err, true);
            if (d2) tdp_tile_free(d2);
            if (d1) tdp_tile_free(d1);
            if (crc != TDP_OK) {
                if (e->err.empty()) e->err = tdp_last_error(e->ctx);
                return crc;
            }
            res->stream[16] = (double)n_fb;
        } else {
            tdb_gen_candidates(&idx1, &I2, self, N, o, nullptr, nullptr, p0, i0);
        }
        // lists in chunk-plan order (position k holds target order[k])
        prefix.assign(N + 1, 0);
        for (size_t k = 0; k < N; k++) prefix[k + 1] = prefix[k] + (p0[order[k] + 1] - p0[order[k]]);
        ids.resize(prefix[N]);
        for (size_t k = 0; k < N; k++)
            std::copy(i0.begin() + p0[order[k]], i0.begin() + p0[order[k] + 1], ids.begin() + prefix[k]);
    }
    res->timing[1] = ms_since(t0);
    res->stream[9] = res->timing[1];
    res->stats[4] = ids.size();

    // budget: two tile slots + the compute context's join buffers
    uint64_t budget = so->hbm_budget;
    if (!budget) {
        uint64_t fr = 0, tot = 0;
        if (tdp_device_memory(e->ctx, &fr, &tot) != TDP_OK) {
            e->err = tdp_last_error(e->ctx);
            return TDP_ERR_HIP;
        }
        budget = fr / 10 * 9;
    }
    const bool lazy = so->lazy != 0;
    // fewer workers (larger tile slots) when a target with its candidates
    // does not fit the slot of as many as were asked for
    t0 = Clock::now();
    int n_workers = so->overlap ? std::max(1, so->workers > 0 ? so->workers : 4) : 1;
    const uint64_t max_targets = so->max_chunk_targets ? so->max_chunk_targets : UINT64_MAX;
    std::vector<Chunk> chunks;
    for (;; n_workers--) {
        const uint64_t slot_bytes = (uint64_t)(budget * so->tile_share / n_workers);
        chunks.clear();
        e->err.clear();
        if (plan_chunks(s1, s2, self, order, prefix, ids, slot_bytes, max_targets, chunks, e->err))
            break;
        if (n_workers == 1) return TDP_ERR_NOMEM;
    }
    {
        std::vector<uint32_t> kpos(s1.n_obj(), 0);
        for (size_t k = 0; k < N; k++) kpos[order[k]] = (uint32_t)k;
        // chunks side by side, each thread with its own renumbering scratch
        // (finish_chunk leaves it as it found it)
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
        {
            std::vector<int32_t> loc(std::max(s1.n_obj(), s2.n_obj()), -1);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 2)
#endif
            for (long i = 0; i < (long)chunks.size(); i++)
                finish_chunk(chunks[i], self, kpos, prefix, ids, s1, s2, loc);
        }
    }
    std::vector<uint32_t>().swap(prefix);
    std::vector<uint32_t>().swap(ids);
    res->stream[17] = ms_since(t0);
    t0 = Clock::now();
    const int n_tiles = self ? 1 : 2;
    size_t cap_obj[2] = {1, 1}, cap_vox[2] = {1, 1}, cap_tri[2] = {1, 1};
    for (const Chunk &c : chunks)
        for (int w = 0; w < n_tiles; w++) {
            cap_obj[w] = std::max(cap_obj[w], c.objs[w].size());
            cap_vox[w] = std::max<size_t>(cap_vox[w], c.n_vox[w]);
            cap_tri[w] = std::max<size_t>(cap_tri[w], c.n_tri[w]);
        }

    // workers: each its own context (stream), tiles and staging, the
    // engine's tuning; they take chunks in turn
    std::vector<Worker> W(n_workers);
    auto cleanup = [&]() {
        for (Worker &w : W) {
            if (w.ctx) w.feed.release();
            for (int k = 0; k < 2; k++) {
                if (w.tile[k]) tdp_tile_free(w.tile[k]);
                for (Decoded &d : w.set)
                    if (w.ctx) free_staging(w.ctx, d.stg[k]);
            }
            if (w.ctx) tdp_close(w.ctx);
        }
    };
    tdp_tuning tune;
    int rc = tdp_get_tuning(e->ctx, &tune);
    uint64_t tiles_bytes = 0;
    for (int k = 0; k < n_workers && rc == TDP_OK; k++) {
        Worker &w = W[k];
        rc = tdp_open(e->device, &w.ctx);
        if (rc == TDP_OK) rc = tdp_set_tuning(w.ctx, &tune);
        for (int t = 0; t < n_tiles && rc == TDP_OK; t++) {
            rc = tdp_tile_reserve(w.ctx, cap_obj[t], cap_vox[t], cap_tri[t], &w.tile[t]);
            for (int b = 0; b < 2 && rc == TDP_OK; b++)
                rc = alloc_staging(w.ctx, cap_obj[t], cap_vox[t], lazy ? 0 : cap_tri[t],
                                   w.set[b].stg[t]);
        }
        uint64_t used = 0;
        if (w.ctx) tdp_memory_usage(w.ctx, &used, nullptr, 0);
        tiles_bytes += used;
        w.feed.ctx = w.ctx;
    }
    if (rc == TDP_OK && tiles_bytes >= budget) {
        e->err = "stream join: the tile slots alone exceed the device budget";
        rc = TDP_ERR_NOMEM;
    }
    if (rc != TDP_OK) {
        if (e->err.empty()) {
            tdp_ctx *c = nullptr;
            for (Worker &w : W)
                if (w.ctx) c = w.ctx;
            e->err = std::string("stream join: worker set-up") +
                     (c ? std::string(": ") + tdp_last_error(c) : std::string());
        }
        cleanup();
        return rc;
    }
    for (Worker &w : W) tdp_set_memory_limit(w.ctx, tiles_bytes / n_workers + (budget - tiles_bytes) / n_workers);
    res->stream[18] = ms_since(t0);

    const size_t nc = chunks.size();
    Progress pg;
    std::vector<tdb_result> parts(nc);
    const int nt_w = std::max(1, nt / n_workers);

    // chunk i's arrays into decode set D: without records (lazy) or whole;
    // returns the host ms it took
    auto decode = [&](size_t i, Decoded &D) -> double {
        const auto a = Clock::now();
        const Chunk &c = chunks[i];
        for (int t = 0; t < n_tiles; t++) {
            decode_tile(t == 0 ? s1 : s2, c.objs[t], D.stg[t], nt_w, lazy, &D.state[t],
                        &D.state_off[t]);
            if (lazy) {
                const Staging &st = D.stg[t];
                D.vobj[t].resize(st.desc.n_vox);
                for (size_t o = 0; o < st.desc.n_obj; o++)
                    for (uint32_t v = st.obj_vox_off[o]; v < st.obj_vox_off[o + 1]; v++)
                        D.vobj[t][v] = (uint32_t)o;
            }
        }
        return ms_since(a);
    };
    auto copy = [&](size_t i, Worker &w, const Decoded &D) -> bool {
        const auto a = Clock::now();
        for (int t = 0; t < n_tiles; t++) {
            const tdp_tile_desc &d = D.stg[t].desc;
            if ((lazy ? tdp_tile_fill_meta_async(w.tile[t], &d) : tdp_tile_fill_async(w.tile[t], &d)) !=
                TDP_OK)
                return false;
            const uint64_t meta = 100 * d.n_vox + 44 * d.n_obj + 4;
            w.meta_bytes += meta;
            w.h2d += meta + (lazy ? 0 : 48 * d.n_tri);
            w.all_records += d.n_tri;
        }
        if (tdp_sync(w.ctx) != TDP_OK) return false;
        w.copy_ms += ms_since(a);
        return true;
    };
    // lazy: the records of a round's voxels, decoded per object at that slot
    auto set_feed = [&](size_t i, Worker &w, const Decoded &D) {
        w.feed.vox_tri_cnt[0] = D.stg[0].vox_tri_cnt;
        w.feed.vox_tri_cnt[1] = self ? nullptr : D.stg[1].vox_tri_cnt;
        w.feed.pack = [&, i](int side, int slot, const uint32_t *vox, size_t n, const uint64_t *off,
                             float *out, std::string &err) -> bool {
            const int t = side - 1;
            if (t >= n_tiles) {
                err = "lazy feed: side 2 of a self-join";
                return false;
            }
            const tdb::ObjectSource &s = t == 0 ? s1 : s2;
            const std::vector<uint32_t> &objs = chunks[i].objs[t];
            const Staging &st = D.stg[t];
            const std::vector<uint32_t> &vo = D.vobj[t];
            // runs of one object in the (ascending) list
            std::vector<size_t> run{0};
            for (size_t k = 1; k < n; k++)
                if (vo[vox[k]] != vo[vox[k - 1]]) run.push_back(k);
            run.push_back(n);
#ifdef _OPENMP
#pragma omp parallel num_threads(nt_w)
#endif
            {
                std::vector<float> tmp;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
                for (long r = 0; r < (long)run.size() - 1; r++) {
                    const size_t a = run[r], b = run[r + 1];
                    const uint32_t o = vo[vox[a]];
                    const uint32_t v0 = st.obj_vox_off[o], nv = st.obj_vox_off[o + 1] - v0;
                    const float *sto = D.state_off[t][o + 1] > D.state_off[t][o]
                                           ? D.state[t].data() + D.state_off[t][o]
                                           : nullptr;
                    if (b - a == nv) {  // every voxel: the slot's block as decoded
                        s.decode_slot(objs[o], slot, out + 12 * off[a], sto);
                        continue;
                    }
                    if (sto) {  // some voxels: each decoded alone from the object's state
                        for (size_t k = a; k < b; k++)
                            s.decode_slot_voxel(objs[o], slot, vox[k] - v0, out + 12 * off[k], sto);
                        continue;
                    }
                    const uint64_t base = st.vox_tri_off[5 * (size_t)v0 + slot];
                    const size_t last = 5 * (size_t)(v0 + nv - 1) + slot;
                    tmp.resize(12 * (st.vox_tri_off[last] + st.vox_tri_cnt[last] - base));
                    s.decode_slot(objs[o], slot, tmp.data(), sto);
                    for (size_t k = a; k < b; k++) {
                        const size_t q = 5 * (size_t)vox[k] + slot;
                        std::memcpy(out + 12 * off[k], tmp.data() + 12 * (st.vox_tri_off[q] - base),
                                    48 * (size_t)st.vox_tri_cnt[q]);
                    }
                }
            }
            return true;
        };
    };
    auto join = [&](size_t i, Worker &w, const Decoded &D) -> bool {
        const auto a = Clock::now();
        const Chunk &c = chunks[i];
        tdp_tile *d1 = w.tile[0], *d2 = self ? d1 : w.tile[1];
        if (lazy) set_feed(i, w, D);
        const double feed0 = w.feed.pack_ms + w.feed.put_ms;
        // target ranges [lo, hi); halved while a join passes the memory cap
        std::vector<std::pair<size_t, size_t>> todo{{0, c.targets.size()}};
        tdb_result &R = parts[i];
        while (!todo.empty()) {
            const auto [lo, hi] = todo.back();
            todo.pop_back();
            std::vector<uint32_t> pf(hi + 1, 0);
            for (size_t t = lo; t < hi; t++) pf[t + 1] = c.prefix[t + 1] - c.prefix[lo];
            std::vector<uint32_t> id(c.ids.begin() + c.prefix[lo], c.ids.begin() + c.prefix[hi]);
            tdb_result r;
            std::string err;
            int jr = tdb_run_join(w.ctx, d1, d2, o, pf, id, nullptr, nullptr, &r, err,
                                  lazy ? &w.feed : nullptr);
            if (jr == TDP_ERR_NOMEM && hi - lo > 1) {
                w.splits++;
                const size_t mid = lo + (hi - lo) / 2;
                todo.push_back({mid, hi});
                todo.push_back({lo, mid});
                continue;
            }
            if (jr != TDP_OK) {
                pg.fail("stream join, chunk " + std::to_string(i) + ": " + err);
                return false;
            }
            // back to source ids
            const std::vector<uint32_t> &T2 = self ? c.objs[0] : c.objs[1];
            for (size_t q = 0; q < r.target.size(); q++) {
                R.target.push_back((int32_t)c.targets[r.target[q]]);
                R.obj2.push_back((int32_t)T2[r.obj2[q]]);
            }
            R.rmin.insert(R.rmin.end(), r.rmin.begin(), r.rmin.end());
            R.rmax.insert(R.rmax.end(), r.rmax.begin(), r.rmax.end());
            R.round.insert(R.round.end(), r.round.begin(), r.round.end());
            for (int k = 2; k < 10; k++) if (k != 6) R.timing[k] += r.timing[k];
            for (int k = 0; k < 9; k++)
                if (k != 4) R.stats[k] = k == 5 ? std::max(R.stats[k], r.stats[k]) : R.stats[k] + r.stats[k];
            for (int k = 0; k < TDP_STAGE_COUNT; k++) {
                R.stage_ms[k] += r.stage_ms[k];
                R.stage_n[k] += r.stage_n[k];
            }
        }
        w.join_ms += ms_since(a) - (w.feed.pack_ms + w.feed.put_ms - feed0);
        return true;
    };

    const auto t_pipe = Clock::now();
    std::atomic<size_t> next{0};
    // a worker decodes the next chunk it takes (on a helper thread) while it
    // copies and joins the current one
    auto run_worker = [&](Worker &w) {
        size_t cur = next.fetch_add(1);
        if (cur >= nc) return;
        w.decode_ms += decode(cur, w.set[0]);
        int b = 0;
        for (;;) {
            if (pg.failed) return;
            const size_t nxt = so->overlap ? next.fetch_add(1) : nc;
            double ahead_ms = 0;
            std::thread helper;
            if (nxt < nc)
                helper = std::thread([&, nxt, b] { ahead_ms = decode(nxt, w.set[1 - b]); });
            bool ok = copy(cur, w, w.set[b]);
            if (!ok) pg.fail(std::string("stream join: copy: ") + tdp_last_error(w.ctx));
            else ok = join(cur, w, w.set[b]);
            if (helper.joinable()) helper.join();
            w.decode_ms += ahead_ms;
            w.n_decoded += chunks[cur].objs[0].size() + (self ? 0 : chunks[cur].objs[1].size());
            if (!ok) return;
            if (nxt >= nc) {
                // without overlap: the next chunk in turn
                if (so->overlap) return;
                cur = next.fetch_add(1);
                if (cur >= nc) return;
                w.decode_ms += decode(cur, w.set[b]);
                continue;
            }
            cur = nxt;
            b = 1 - b;
        }
    };
    if (n_workers == 1) {
        run_worker(W[0]);
    } else {
        std::vector<std::thread> th;
        for (Worker &w : W) th.emplace_back([&run_worker, &w] { run_worker(w); });
        for (auto &t : th) t.join();
    }
    const double wall = ms_since(t_pipe);
    double decode_ms = 0, copy_ms = 0, join_ms = 0;
    uint64_t h2d = 0, n_decoded = 0, splits = 0, meta_bytes = 0, all_records = 0, put_records = 0,
             peak = 0;
    for (Worker &w : W) {
        decode_ms += w.decode_ms + w.feed.pack_ms;
        copy_ms += w.copy_ms + w.feed.put_ms;
        join_ms += w.join_ms;
        h2d += w.h2d + 48 * w.feed.put_records;
        n_decoded += w.n_decoded;
        splits += w.splits;
        meta_bytes += w.meta_bytes;
        all_records += w.all_records;
        put_records += lazy ? w.feed.put_records : w.all_records;
        uint64_t p = 0;
        tdp_memory_usage(w.ctx, nullptr, &p, 0);
        peak += p;
    }
    cleanup();
    if (pg.failed) {
        e->err = pg.err;
        return TDP_ERR_HIP;
    }

    // results in target order, each target's pairs in chunk order and in
    // report order within a chunk: a counting sort by target (stable)
    size_t total = 0;
    int32_t t_hi = -1;
    for (const tdb_result &r : parts) {
        total += r.target.size();
        for (int32_t t : r.target) t_hi = std::max(t_hi, t);
    }
    std::vector<uint64_t> at((size_t)(t_hi + 1) + 1, 0);
    for (const tdb_result &r : parts)
        for (int32_t t : r.target) at[(size_t)t + 1]++;
    for (size_t t = 1; t < at.size(); t++) at[t] += at[t - 1];
    res->target.resize(total);
    res->obj2.resize(total);
    res->rmin.resize(total);
    res->rmax.resize(total);
    res->round.resize(total);
    for (const tdb_result &r : parts)
        for (size_t q = 0; q < r.target.size(); q++) {
            const uint64_t w = at[r.target[q]]++;
            res->target[w] = r.target[q];
            res->obj2[w] = r.obj2[q];
            res->rmin[w] = r.rmin[q];
            res->rmax[w] = r.rmax[q];
            res->round[w] = r.round[q];
        }
    for (const tdb_result &r : parts) {
        for (int k = 2; k < 10; k++) if (k != 6) res->timing[k] += r.timing[k];
        for (int k = 0; k < 9; k++)
            if (k != 4) res->stats[k] = k == 5 ? std::max(res->stats[k], r.stats[k]) : res->stats[k] + r.stats[k];
        for (int k = 0; k < TDP_STAGE_COUNT; k++) {
            res->stage_ms[k] += r.stage_ms[k];
            res->stage_n[k] += r.stage_n[k];
        }
    }
    res->timing[6] = copy_ms;
    const double serial = decode_ms + copy_ms + join_ms;
    const double slowest = std::max({decode_ms, copy_ms, join_ms});
    res->stream[0] = (double)nc;
    res->stream[1] = (double)h2d;
    res->stream[2] = decode_ms;
    res->stream[3] = copy_ms;
    res->stream[4] = join_ms;
    res->stream[5] = wall;
    res->stream[6] = serial > slowest ? std::min(1.0, std::max(0.0, (serial - wall) / (serial - slowest))) : 1.0;
    res->stream[7] = (double)peak;
    res->stream[10] = (double)n_decoded;
    res->stream[11] = (double)splits;
    res->stream[12] = 48.0 * (double)all_records;
    res->stream[13] = 48.0 * (double)put_records;
    res->stream[14] = (double)meta_bytes;
    res->stream[15] = (double)n_workers;
    res->timing[0] = ms_since(t_start);
    *out = res.release();
    return 0;
}

}  // extern "C"
