// Serving loop behind the tdb_server_* entry points of include/tdbase.h: join
// queries (fresh target tiles) against one resident tile, with continuous
// batching.
//
// One preparer and E executors, each a device context (stream):
//   preparer  takes the queries waiting at that moment -- as many as fit the
//             device-memory budget and share the LOD list and Hausdorff level
//             (kNN and within queries together, each target keeping its
//             kind's rules; intersection queries among themselves) -- merges
//             their tiles, builds the candidates (per-target kind, k and
//             within distance) and uploads the merged tile on its own stream;
//   executor  runs a prepared batch as ONE join on its stream and splits the
//             result lists back per query.
// A query that arrives while batches run is admitted as soon as a tile slot
// and the memory budget allow: the preparer builds its batch beside the
// running joins and a free executor starts it at once, so the joins of
// different batches overlap on the device (E + 1 tile slots: one per
// executor and one being prepared). Each target's join depends only on that
// target and the resident tile, so a query's results are the same alone and
// in any mix.
//
// By default a batch's tile goes up without its triangle records: before
// each LOD round the executor's join lists the query voxels that round reads
// and copies just their records at that LOD from the query tiles into pinned
// memory for the device (tdp_join_lazy_needs / tdp_join_lazy_put), the
// reference's per-LOD packing of live candidates' voxels
// (KNNJoin.cpp:1584-1617). Otherwise every record goes up with the tile,
// through pinned chunks copied by host threads while the previous chunk's
// DMA runs.
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cstring>
#include <algorithm>
#include <condition_variable>
#include <array>
#include <deque>
#include <thread>
#include <unordered_map>

#include "tdb_internal.hpp"

namespace {

struct Query {
    int64_t ticket = 0;
    tdb_tile *tile = nullptr;  // owned by the server from submit on
    tdb_join_opts o{};
    std::vector<int> lods;
    Clock::time_point arrival;
    uint64_t dev_bytes = 0;
};

struct Batch {
    std::vector<Query> qs;
    tdb_tile merged;
    tdp_tile *dev = nullptr;
    int slot = -1;  // the device tile slot dev lives in
    std::vector<uint32_t> prefix, ids;
    std::vector<int32_t> k_t;
    std::vector<float> tau_t;
    std::vector<uint8_t> within_t;  // kNN and within queries together: 1 for within targets
    std::vector<size_t> first;  // first target of each query in the merged tile
    std::vector<size_t> vfirst, tfirst;  // its first voxel and first record there
    uint64_t dev_bytes = 0;
    std::string err;
};

struct Done {
    tdb_result *res = nullptr;  // owned by the caller after tdb_server_wait
    double latency_ms = 0;
    int rc = 0;
    std::string err;
};

// device bytes a query needs while it is in flight: its tile (48-byte records
// plus 48-byte bound records, voxel and object arrays) and the join state,
// taken as half the tile again
uint64_t estimate_bytes(const tdb::Tile &t) {
    const uint64_t tile = (uint64_t)t.n_tri() * 96 + (uint64_t)t.n_vox() * 100 +
                          (uint64_t)t.n_obj * 64;
    return tile + tile / 2 + ((uint64_t)16 << 20);
}

// queries that may share a join: the same LOD list and Hausdorff level, and
// the same query type -- or kNN and within together when mix_kinds
bool same_kind(const Query &a, const Query &b, bool mix_kinds) {
    const bool kinds = a.o.query == b.o.query ||
                       (mix_kinds && a.o.query != TDB_QUERY_INTERSECT &&
                        b.o.query != TDB_QUERY_INTERSECT);
    return kinds && a.lods == b.lods &&
           a.o.hausdorff_level == b.o.hausdorff_level;
}

// the batch's tile without its triangle records (those go to the device
// straight from each query's buffer)
void merge_tiles(const std::vector<Query> &qs, tdb::Tile &m, std::vector<size_t> &first,
                 std::vector<size_t> &vfirst, std::vector<size_t> &tfirst) {
    size_t no = 0, nv = 0, nt = 0;
    for (auto &q : qs) {
        no += q.tile->t.n_obj;
        nv += q.tile->t.n_vox();
        nt += q.tile->t.n_tri();
    }
    m = tdb::Tile();
    m.n_obj = no;
    m.obj_vox_off.reserve(no + 1);
    m.obj_vox_off.push_back(0);
    m.obj_box.reserve(6 * no);
    m.obj_transform.reserve(3 * no);
    m.obj_hd.reserve(10 * no);
    m.vox_box.reserve(9 * nv);
    m.vox_obj.reserve(nv);
    m.vox_tri_off.reserve(5 * nv);
    m.vox_tri_cnt.reserve(5 * nv);
    size_t oo = 0, vo = 0, to = 0;
    for (auto &q : qs) {
        const tdb::Tile &t = q.tile->t;
        first.push_back(oo);
        vfirst.push_back(vo);
        tfirst.push_back(to);
        for (size_t i = 1; i <= t.n_obj; i++) m.obj_vox_off.push_back((uint32_t)(vo + t.obj_vox_off[i]));
        m.obj_box.insert(m.obj_box.end(), t.obj_box.begin(), t.obj_box.end());
        m.obj_transform.insert(m.obj_transform.end(), t.obj_transform.begin(), t.obj_transform.end());
        m.obj_hd.insert(m.obj_hd.end(), t.obj_hd.begin(), t.obj_hd.end());
        m.vox_box.insert(m.vox_box.end(), t.vox_box.begin(), t.vox_box.end());
        for (uint32_t v : t.vox_obj) m.vox_obj.push_back((uint32_t)(oo + v));
        for (uint64_t r : t.vox_tri_off) m.vox_tri_off.push_back(to + r);
        m.vox_tri_cnt.insert(m.vox_tri_cnt.end(), t.vox_tri_cnt.begin(), t.vox_tri_cnt.end());
        oo += t.n_obj;
        vo += t.n_vox();
        to += t.n_tri();
    }
}

}  // namespace

struct tdb_server {
    tdb_server_opts opts{};
    int n_exec = 1;
    tdp_ctx *prep_ctx = nullptr;
    std::vector<tdp_ctx *> run_ctx;  // one per executor
    std::vector<LazyFeed> feed;      // one per executor
    const tdb_tile *resident = nullptr;
    tdp_tile *res_dev = nullptr;
    uint64_t budget = 0, res_bytes = 0;

    std::mutex mu;
    std::condition_variable cv;
    std::deque<Query> pending;
    std::deque<std::unique_ptr<Batch>> ready;
    std::unordered_map<int64_t, Done> done;
    int64_t next_ticket = 0;
    uint64_t in_use = 0;        // bytes of prepared and running batches
    int running = 0;            // batches being joined
    bool stop = false, prep_exit = false;
    std::string err;
    uint64_t n_batches = 0, n_queries = 0, max_batch_seen = 0, peak_bytes = 0;
    int max_running = 0;
    // device tile slots (guarded by mu: busy) and the preparer's pinned staging
    static constexpr size_t kChunkTri = (size_t)1 << 20;  // records per staging chunk
    std::vector<tdp_tile *> slot;
    std::vector<std::array<size_t, 3>> slot_cap;
    std::vector<char> slot_busy;
    char *meta_pin = nullptr;
    size_t meta_cap = 0;
    float *stage[2] = {nullptr, nullptr};
    int upload(Batch &b);
    // host time of the stages (ms, summed over batches): merge, upload,
    // candidates (preparer); join, record packing and puts (executors);
    // bytes to the device
    double t_merge = 0, t_upload = 0, t_cand = 0, t_join = 0, t_pack = 0, t_put = 0;
    uint64_t h2d_meta = 0, h2d_rec = 0, rec_all = 0, mixed_batches = 0;
    std::thread prep;
    std::vector<std::thread> exec;

    void preparer();
    void executor(int e);
};

// the merged tile into a free slot through pinned staging (preparer thread)
int tdb_server::upload(Batch &b) {
    const tdb::Tile &m = b.merged.t;
    const size_t no = m.n_obj, nv = m.n_vox();
    size_t nt = 0;
    for (auto &q : b.qs) nt += q.tile->t.n_tri();
    int s = -1;
    {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] {
            for (char x : slot_busy)
                if (!x) return true;
            return false;
        });
        while (slot_busy[s + 1]) s++;
        s++;
        slot_busy[s] = 1;
    }
    b.slot = s;
    int rc = TDP_OK;
    size_t *cap = slot_cap[s].data();
    if (!slot[s] || cap[0] < no || cap[1] < nv || cap[2] < nt) {
        if (slot[s]) tdp_tile_free(slot[s]);
        slot[s] = nullptr;
        cap[0] = std::max(cap[0], std::max<size_t>(no, 1));
        cap[1] = std::max(cap[1], std::max<size_t>(nv, 1));
        cap[2] = std::max(cap[2], std::max<size_t>(nt, 1));
        rc = tdp_tile_reserve(prep_ctx, cap[0], cap[1], cap[2], &slot[s]);
        if (rc != TDP_OK) {
            cap[0] = cap[1] = cap[2] = 0;
            return rc;
        }
    }
    if (!stage[0] && !opts.lazy) {
        for (int k = 0; k < 2 && rc == TDP_OK; k++)
            rc = tdp_host_alloc(prep_ctx, 48 * kChunkTri, (void **)&stage[k]);
        if (rc != TDP_OK) return rc;
    }
    // object and voxel arrays into one pinned block
    const size_t sz[8] = {4 * (no + 1), 40 * no, 36 * nv, 40 * nv, 20 * nv, 24 * no, 12 * no, 0};
    size_t need = 0, at[8];
    for (int k = 0; k < 7; k++) {
        at[k] = need;
        need += (sz[k] + 63) / 64 * 64;
    }
    if (meta_cap < need) {
        if (meta_pin) tdp_host_free(prep_ctx, meta_pin);
        meta_pin = nullptr;
        meta_cap = 0;
        rc = tdp_host_alloc(prep_ctx, need, (void **)&meta_pin);
        if (rc != TDP_OK) return rc;
        meta_cap = need;
    }
    const void *src[7] = {m.obj_vox_off.data(), m.obj_hd.data(),      m.vox_box.data(),
                          m.vox_tri_off.data(), m.vox_tri_cnt.data(), m.obj_box.data(),
                          m.obj_transform.data()};
    for (int k = 0; k < 7; k++)
        if (sz[k]) std::memcpy(meta_pin + at[k], src[k], sz[k]);
    tdp_tile_desc d{};
    d.n_obj = no;
    d.n_vox = nv;
    d.n_tri = nt;
    d.obj_vox_off = (const uint32_t *)(meta_pin + at[0]);
    d.obj_hd = (const float *)(meta_pin + at[1]);
    d.vox_box = (const float *)(meta_pin + at[2]);
    d.vox_tri_off = (const uint64_t *)(meta_pin + at[3]);
    d.vox_tri_cnt = (const uint32_t *)(meta_pin + at[4]);
    d.obj_box = (const float *)(meta_pin + at[5]);
    d.obj_transform = (const float *)(meta_pin + at[6]);
    h2d_meta += need;
    rec_all += 48 * (uint64_t)nt;
    if (opts.lazy) {
        // the records come round by round, from the query tiles
        rc = tdp_tile_fill_meta_async(slot[s], &d);
        if (rc == TDP_OK) rc = tdp_sync(prep_ctx);
        if (rc == TDP_OK) b.dev = slot[s];
        return rc;
    }
    h2d_rec += 48 * (uint64_t)nt;
    rc = tdp_tile_fill_async(slot[s], &d);
    // triangle records: host threads copy chunk k while chunk k-1 is on the wire
    uint64_t mk[2] = {0, 0};
    size_t k = 0, first = 0;
#ifdef _OPENMP
    const int nth = opts.threads > 0 ? opts.threads : std::min(16, omp_get_max_threads());
#endif
    for (size_t qi = 0; rc == TDP_OK && qi < b.qs.size(); qi++) {
        const tdb::Tile &t = b.qs[qi].tile->t;
        for (size_t off = 0; rc == TDP_OK && off < t.n_tri(); off += kChunkTri, k++) {
            const size_t n = std::min(kChunkTri, t.n_tri() - off);
            float *buf = stage[k % 2];
            if (k >= 2) rc = tdp_wait_mark(prep_ctx, mk[k % 2]);
            if (rc != TDP_OK) break;
            const float *from = t.tri.data() + 12 * off;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nth) schedule(static)
#endif
            for (long p = 0; p < 16; p++) {
                const size_t a = n * p / 16, e = n * (p + 1) / 16;
                if (e > a) std::memcpy(buf + 12 * a, from + 12 * a, 48 * (e - a));
            }
            rc = tdp_tile_fill_tri_async(slot[s], first, n, buf);
            if (rc == TDP_OK) rc = tdp_mark(prep_ctx, &mk[k % 2]);
            first += n;
        }
    }
    if (rc == TDP_OK) rc = tdp_sync(prep_ctx);
    if (rc == TDP_OK) b.dev = slot[s];
    return rc;
}

void tdb_server::preparer() {
    for (;;) {
        std::unique_ptr<Batch> b(new Batch());
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (stop && pending.empty()) {
                    prep_exit = true;
                    cv.notify_all();
                    return;
                }
                // batches in flight (running or ready) at most one more than
                // the executors; without overlap, one at a time
                const bool room = opts.overlap ? (int)ready.size() + running < n_exec + 1
                                               : (ready.empty() && in_use == 0);
                if (!pending.empty() && room) {
                    const Query &q = pending.front();
                    if (in_use == 0 || res_bytes + in_use + q.dev_bytes <= budget) break;
                }
                cv.wait(lk);
            }
            // the queries waiting now, in arrival order, as far as they fit
            const Query head = pending.front();
            uint64_t bytes = 0;
            for (auto it = pending.begin(); it != pending.end();) {
                if ((int)b->qs.size() >= std::max(1, opts.max_batch)) break;
                if (!same_kind(*it, head, opts.mix_kinds != 0)) {
                    ++it;
                    continue;
                }
                if (!b->qs.empty() && res_bytes + in_use + bytes + it->dev_bytes > budget) break;
                bytes += it->dev_bytes;
                b->qs.push_back(*it);
                it = pending.erase(it);
            }
            b->dev_bytes = bytes;
            in_use += bytes;
            peak_bytes = std::max(peak_bytes, res_bytes + in_use);
        }
        // merge, candidates, upload: outside the lock, beside the running join
        auto t0 = Clock::now();
        merge_tiles(b->qs, b->merged.t, b->first, b->vfirst, b->tfirst);
        const size_t N = b->merged.t.n_obj;
        b->k_t.resize(N);
        b->tau_t.resize(N);
        bool mixed = false;
        for (auto &q : b->qs) mixed = mixed || q.o.query != b->qs[0].o.query;
        if (mixed) b->within_t.assign(N, 0);
        for (size_t q = 0; q < b->qs.size(); q++) {
            const size_t e = q + 1 < b->qs.size() ? b->first[q + 1] : N;
            const bool within = b->qs[q].o.query == TDB_QUERY_WITHIN;
            for (size_t t = b->first[q]; t < e; t++) {
                // (a within target's k only feeds the kNN candidate pass)
                b->k_t[t] = within && mixed ? 1 : b->qs[q].o.k;
                b->tau_t[t] = b->qs[q].o.within_dist;
                if (mixed) b->within_t[t] = within ? 1 : 0;
            }
        }
        tdb_join_opts o = b->qs[0].o;
        o.threads = opts.threads;
        const double ms_merge = ms_since(t0);
        t0 = Clock::now();
        double ms_upload = 0;
        if (upload(*b) != TDP_OK) {
            b->err = tdp_last_error(prep_ctx);
        } else {
            ms_upload = ms_since(t0);
            t0 = Clock::now();
            // within lists from the host octree here: the preparer runs while
            // the executor's joins occupy the device, and a device pass would
            // queue behind them
            auto lists = [&](const tdb_join_opts &oq, std::vector<uint32_t> &pf,
                             std::vector<uint32_t> &id) {
                if (tdb_gen_candidates_dev(prep_ctx, b->dev, res_dev, &b->merged, resident, false,
                                           N, &oq, b->k_t.data(), b->tau_t.data(), pf, id,
                                           nullptr, b->err, false) != TDP_OK &&
                    b->err.empty())
                    b->err = "candidate generation failed";
            };
            if (!mixed) {
                lists(o, b->prefix, b->ids);
            } else {
                // each target's list of its own kind
                std::vector<uint32_t> pk, ik, pw, iw;
                tdb_join_opts ok = o, ow = o;
                ok.query = TDB_QUERY_NN;
                ow.query = TDB_QUERY_WITHIN;
                lists(ok, pk, ik);
                if (b->err.empty()) lists(ow, pw, iw);
                if (b->err.empty()) {
                    b->prefix.assign(N + 1, 0);
                    b->ids.clear();
                    for (size_t t = 0; t < N; t++) {
                        const auto &p = b->within_t[t] ? pw : pk;
                        const auto &id = b->within_t[t] ? iw : ik;
                        b->ids.insert(b->ids.end(), id.begin() + p[t], id.begin() + p[t + 1]);
                        b->prefix[t + 1] = (uint32_t)b->ids.size();
                    }
                }
            }
        }
        const double ms_cand = ms_since(t0);
        {
            std::lock_guard<std::mutex> lk(mu);
            t_merge += ms_merge;
            t_upload += ms_upload;
            t_cand += ms_cand;
            mixed_batches += mixed ? 1 : 0;
            ready.push_back(std::move(b));
        }
        cv.notify_all();
    }
}

void tdb_server::executor(int e) {
    tdp_ctx *ctx = run_ctx[e];
    LazyFeed &f = feed[e];
    for (;;) {
        std::unique_ptr<Batch> b;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return !ready.empty() || (prep_exit && ready.empty()); });
            if (ready.empty()) return;
            b = std::move(ready.front());
            ready.pop_front();
            running++;
            max_running = std::max(max_running, running);
        }
        cv.notify_all();  // the preparer may form the next batch now
        tdb_result all;
        std::string msg = b->err;
        int rc = b->err.empty() ? TDP_OK : TDP_ERR_HIP;
        if (rc == TDP_OK) {
            tdb_join_opts o = b->qs[0].o;
            o.lods = b->qs[0].lods.data();
            o.n_lods = (int)b->qs[0].lods.size();
            o.precision = opts.precision;
            o.max_objects1 = 0;
            // query voxels' records at a round's slot, from the query tiles
            const tdb::Tile &m = b->merged.t;
            const Batch &B = *b;
            f.vox_tri_cnt[0] = m.vox_tri_cnt.data();
            f.vox_tri_cnt[1] = nullptr;
            f.pack = [&](int side, int slot, const uint32_t *vox, size_t n, const uint64_t *off,
                         float *out, std::string &err) -> bool {
                if (side != 1) {
                    err = "serving feed: the resident tile holds its records";
                    return false;
                }
#ifdef _OPENMP
                const int nth = opts.threads > 0 ? opts.threads : std::min(16, omp_get_max_threads());
#pragma omp parallel for num_threads(nth) schedule(dynamic, 256)
#endif
                for (long k = 0; k < (long)n; k++) {
                    const size_t v = vox[k];
                    const size_t q = std::upper_bound(B.vfirst.begin(), B.vfirst.end(), v) -
                                     B.vfirst.begin() - 1;
                    const size_t at = 5 * v + slot;
                    const tdb::Tile &t = B.qs[q].tile->t;
                    std::memcpy(out + 12 * off[k], t.tri.data() + 12 * (m.vox_tri_off[at] - B.tfirst[q]),
                                48 * (size_t)m.vox_tri_cnt[at]);
                }
                return true;
            };
            const double pk0 = f.pack_ms, pt0 = f.put_ms;
            const uint64_t pr0 = f.put_records;
            const auto t0 = Clock::now();
            rc = tdb_run_join(ctx, b->dev, res_dev, &o, b->prefix, b->ids, b->k_t.data(),
                              b->tau_t.data(), &all, msg, opts.lazy ? &f : nullptr,
                              b->within_t.empty() ? nullptr : b->within_t.data());
            const double ms = ms_since(t0);
            std::lock_guard<std::mutex> lk(mu);
            t_join += ms - (f.pack_ms - pk0) - (f.put_ms - pt0);
            t_pack += f.pack_ms - pk0;
            t_put += f.put_ms - pt0;
            h2d_rec += 48 * (f.put_records - pr0);
        }
        b->dev = nullptr;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (b->slot >= 0) slot_busy[b->slot] = 0;
            running--;
        }
        cv.notify_all();
        // per query results and latencies
        const auto now = Clock::now();
        std::vector<std::pair<int64_t, Done>> out;
        for (size_t q = 0; q < b->qs.size(); q++) {
            Done d;
            d.rc = rc;
            d.err = msg;
            d.latency_ms = std::chrono::duration<double, std::milli>(now - b->qs[q].arrival).count();
            if (rc == TDP_OK) {
                const size_t lo = b->first[q], n = b->qs[q].tile->t.n_obj, hi = lo + n;
                tdb_result *r = new tdb_result();
                r->n_targets = n;
                // the joined lists are grouped by target in ascending order
                const auto fi = std::lower_bound(all.target.begin(), all.target.end(), (int32_t)lo);
                const auto la = std::lower_bound(all.target.begin(), all.target.end(), (int32_t)hi);
                for (size_t k = fi - all.target.begin(); k < (size_t)(la - all.target.begin()); k++) {
                    r->target.push_back(all.target[k] - (int32_t)lo);
                    r->obj2.push_back(all.obj2[k]);
                    r->rmin.push_back(all.rmin[k]);
                    r->rmax.push_back(all.rmax[k]);
                    r->round.push_back(all.round[k]);
                }
                std::copy(all.timing, all.timing + 10, r->timing);
                r->timing[0] = d.latency_ms;
                std::copy(all.stage_ms, all.stage_ms + TDP_STAGE_COUNT, r->stage_ms);
                std::copy(all.stage_n, all.stage_n + TDP_STAGE_COUNT, r->stage_n);
                std::copy(all.stats, all.stats + 7, r->stats);
                d.res = r;
            }
            out.emplace_back(b->qs[q].ticket, d);
            tdb_tile_free(b->qs[q].tile);
            b->qs[q].tile = nullptr;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto &kv : out) done[kv.first] = kv.second;
            in_use -= b->dev_bytes;
            n_batches++;
            n_queries += b->qs.size();
            max_batch_seen = std::max<uint64_t>(max_batch_seen, b->qs.size());
        }
        cv.notify_all();
    }
}

extern "C" {

void tdb_server_defaults(tdb_server_opts *o) {
    o->device = 0;
    o->hbm_budget = 0;
    o->max_batch = 64;
    o->overlap = 1;
    o->precision = TDP_PREC_MIXED;
    o->threads = 0;
    o->lazy = 1;
    o->executors = 0;
    o->mix_kinds = 1;
}

int tdb_server_open(const tdb_server_opts *o, const tdb_tile *resident, tdb_server **out) {
    if (!o || !resident || !out) return TDP_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<tdb_server> s(new tdb_server());
    s->opts = *o;
    s->resident = resident;
    s->n_exec = o->overlap ? (o->executors > 0 ? o->executors : 2) : 1;
    s->run_ctx.assign(s->n_exec, nullptr);
    s->feed.resize(s->n_exec);
    int rc = TDP_OK;
    for (int e = 0; e < s->n_exec && rc == TDP_OK; e++) rc = tdp_open(o->device, &s->run_ctx[e]);
    if (rc == TDP_OK) rc = tdp_open(o->device, &s->prep_ctx);
    uint64_t free_b = 0, total_b = 0;
    if (rc == TDP_OK) rc = tdp_device_memory(s->prep_ctx, &free_b, &total_b);
    const tdp_tile_desc d = tdb_desc_of(resident->t);
    if (rc == TDP_OK) rc = tdp_tile_upload(s->prep_ctx, &d, &s->res_dev);
    if (rc != TDP_OK) {
        tdp_ctx *c = s->prep_ctx;
        for (tdp_ctx *x : s->run_ctx)
            if (!c && x) c = x;
        s->err = c ? tdp_last_error(c) : "tdp_open failed";
        if (s->res_dev) tdp_tile_free(s->res_dev);
        if (s->prep_ctx) tdp_close(s->prep_ctx);
        for (tdp_ctx *x : s->run_ctx)
            if (x) tdp_close(x);
        s->run_ctx.clear();
        s->prep_ctx = nullptr;
        *out = s.release();  // so that the message can be read
        return rc;
    }
    for (int e = 0; e < s->n_exec; e++) s->feed[e].ctx = s->run_ctx[e];
    s->slot.assign(s->n_exec + 1, nullptr);
    s->slot_cap.assign(s->n_exec + 1, {0, 0, 0});
    s->slot_busy.assign(s->n_exec + 1, 0);
    s->res_bytes = tdp_tile_bytes(s->res_dev);
    s->budget = o->hbm_budget ? o->hbm_budget : free_b / 10 * 9 + s->res_bytes;
    tdb_server *p = s.get();
    s->prep = std::thread([p] { p->preparer(); });
    for (int e = 0; e < s->n_exec; e++) s->exec.emplace_back([p, e] { p->executor(e); });
    *out = s.release();
    return TDP_OK;
}

int64_t tdb_server_submit(tdb_server *s, tdb_tile *query, const tdb_join_opts *o) {
    if (!s || !query || !o || s->run_ctx.empty()) return TDP_ERR_ARG;
    if ((o->query != TDB_QUERY_NN && o->query != TDB_QUERY_WITHIN &&
         o->query != TDB_QUERY_INTERSECT) ||
        (o->query == TDB_QUERY_NN && o->k < 1)) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "tdb_server_submit: bad query options";
        return TDP_ERR_ARG;
    }
    Query q;
    q.tile = query;
    q.o = *o;
    q.lods = tdb_lod_list(o);
    q.o.lods = nullptr;
    q.o.n_lods = 0;
    q.arrival = Clock::now();
    q.dev_bytes = estimate_bytes(query->t);
    int64_t t;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            s->err = "tdb_server_submit: server is closing";
            return TDP_ERR_STATE;
        }
        if (s->res_bytes + q.dev_bytes > s->budget) {
            s->err = "tdb_server_submit: the query alone exceeds the device-memory budget";
            return TDP_ERR_NOMEM;
        }
        t = q.ticket = s->next_ticket++;
        s->pending.push_back(q);
    }
    s->cv.notify_all();
    return t;
}

int tdb_server_wait(tdb_server *s, int64_t ticket, tdb_result **out, double *latency_ms) {
    if (!s || !out) return TDP_ERR_ARG;
    *out = nullptr;
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket < 0 || ticket >= s->next_ticket) return TDP_ERR_ARG;
    s->cv.wait(lk, [&] { return s->done.count(ticket) > 0; });
    Done d = s->done[ticket];
    s->done.erase(ticket);
    if (latency_ms) *latency_ms = d.latency_ms;
    if (d.rc != TDP_OK) {
        s->err = d.err;
        return d.rc;
    }
    *out = d.res;
    return TDP_OK;
}

const char *tdb_server_error(const tdb_server *s) { return s ? s->err.c_str() : "null server"; }

void tdb_server_stats(tdb_server *s, uint64_t *out4) {
    std::lock_guard<std::mutex> lk(s->mu);
    out4[0] = s->n_batches;
    out4[1] = s->n_queries;
    out4[2] = s->max_batch_seen;
    out4[3] = s->peak_bytes;
}

void tdb_server_stage_ms(tdb_server *s, double *out4) {
    std::lock_guard<std::mutex> lk(s->mu);
    out4[0] = s->t_merge;
    out4[1] = s->t_upload;
    out4[2] = s->t_cand;
    out4[3] = s->t_join;
}

int tdb_server_figures(tdb_server *s, double *out, int n) {
    if (!s || !out || n < 0) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    const double v[13] = {s->t_merge, s->t_upload, s->t_cand, s->t_join, s->t_pack, s->t_put,
                          (double)s->h2d_meta, (double)s->h2d_rec, (double)s->rec_all,
                          (double)s->n_exec, (double)s->max_running, (double)s->n_batches,
                          (double)s->mixed_batches};
    n = std::min(n, 13);
    for (int k = 0; k < n; k++) out[k] = v[k];
    return n;
}

void tdb_server_close(tdb_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->prep.joinable()) s->prep.join();
    for (auto &t : s->exec)
        if (t.joinable()) t.join();
    for (auto &kv : s->done) delete kv.second.res;
    for (auto &q : s->pending) tdb_tile_free(q.tile);
    for (tdp_tile *t : s->slot)
        if (t) tdp_tile_free(t);
    for (int k = 0; k < 2; k++)
        if (s->stage[k]) tdp_host_free(s->prep_ctx, s->stage[k]);
    if (s->meta_pin) tdp_host_free(s->prep_ctx, s->meta_pin);
    for (LazyFeed &f : s->feed)
        if (f.ctx) f.release();
    if (s->res_dev) tdp_tile_free(s->res_dev);
    if (s->prep_ctx) tdp_close(s->prep_ctx);
    for (tdp_ctx *x : s->run_ctx)
        if (x) tdp_close(x);
    delete s;
}

}  // extern "C"
