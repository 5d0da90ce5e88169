// Serving loop behind the tdb_server_* entry points of include/tdbase.h: join
// queries (fresh target tiles) against one resident tile, with continuous
// batching.
//
// Two threads and two device contexts (streams):
//   preparer  takes the queries waiting at that moment -- as many as fit the
//             device-memory budget and share the query type, LOD list and
//             Hausdorff level -- merges their tiles, builds the candidates
//             (per-target k / within distance) and uploads the merged tile on
//             its own stream;
//   executor  runs each prepared batch as ONE join on the other stream and
//             splits the result lists back per query.
// While the executor runs batch n, the preparer builds batch n+1 from the
// queries that arrived meanwhile (host merge, octree, upload overlap the
// kernels), and batch n+1 starts as soon as batch n ends. Each target's join
// depends only on that target and the resident tile, so a query's results
// are the same alone and in any mix.
//
// Uploads go into two device tiles reserved once (grown when a batch needs
// more) that the preparer and the executor hand back and forth, through
// pinned staging: the object and voxel arrays in one pinned block, the
// triangle records in chunks copied by host threads into two pinned buffers
// that alternate while the previous chunk's DMA runs.
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cstring>
#include <algorithm>
#include <condition_variable>
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
    std::vector<size_t> first;  // first target of each query in the merged tile
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

bool same_kind(const Query &a, const Query &b) {
    return a.o.query == b.o.query && a.lods == b.lods &&
           a.o.hausdorff_level == b.o.hausdorff_level;
}

// the batch's tile without its triangle records (those go to the device
// straight from each query's buffer)
void merge_tiles(const std::vector<Query> &qs, tdb::Tile &m, std::vector<size_t> &first) {
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
    tdp_ctx *run_ctx = nullptr, *prep_ctx = nullptr;
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
    bool stop = false, prep_exit = false;
    std::string err;
    uint64_t n_batches = 0, n_queries = 0, max_batch_seen = 0, peak_bytes = 0;
    // device tile slots (guarded by mu: busy) and the preparer's pinned staging
    static constexpr size_t kChunkTri = (size_t)1 << 20;  // records per staging chunk
    tdp_tile *slot[2] = {nullptr, nullptr};
    size_t slot_cap[2][3] = {{0, 0, 0}, {0, 0, 0}};
    bool slot_busy[2] = {false, false};
    char *meta_pin = nullptr;
    size_t meta_cap = 0;
    float *stage[2] = {nullptr, nullptr};
    int upload(Batch &b);
    // host time of the stages (ms, summed over batches): merge, upload,
    // candidates (preparer); join (executor)
    double t_merge = 0, t_upload = 0, t_cand = 0, t_join = 0;
    std::thread prep, exec;

    void preparer();
    void executor();
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
        cv.wait(lk, [&] { return !slot_busy[0] || !slot_busy[1]; });
        s = slot_busy[0] ? 1 : 0;
        slot_busy[s] = true;
    }
    b.slot = s;
    int rc = TDP_OK;
    size_t *cap = slot_cap[s];
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
    if (!stage[0]) {
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
                // look ahead at most one batch; without overlap, only when idle
                const bool room = opts.overlap ? ready.empty() : (ready.empty() && in_use == 0);
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
                if (!same_kind(*it, head)) {
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
        merge_tiles(b->qs, b->merged.t, b->first);
        const size_t N = b->merged.t.n_obj;
        b->k_t.resize(N);
        b->tau_t.resize(N);
        for (size_t q = 0; q < b->qs.size(); q++) {
            const size_t e = q + 1 < b->qs.size() ? b->first[q + 1] : N;
            for (size_t t = b->first[q]; t < e; t++) {
                b->k_t[t] = b->qs[q].o.k;
                b->tau_t[t] = b->qs[q].o.within_dist;
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
            if (tdb_gen_candidates_dev(prep_ctx, b->dev, res_dev, &b->merged, resident, false, N,
                                       &o, b->k_t.data(), b->tau_t.data(), b->prefix, b->ids,
                                       nullptr, b->err, false) != TDP_OK &&
                b->err.empty())
                b->err = "candidate generation failed";
        }
        const double ms_cand = ms_since(t0);
        {
            std::lock_guard<std::mutex> lk(mu);
            t_merge += ms_merge;
            t_upload += ms_upload;
            t_cand += ms_cand;
            ready.push_back(std::move(b));
        }
        cv.notify_all();
    }
}

void tdb_server::executor() {
    for (;;) {
        std::unique_ptr<Batch> b;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return !ready.empty() || (prep_exit && ready.empty()); });
            if (ready.empty()) return;
            b = std::move(ready.front());
            ready.pop_front();
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
            const auto t0 = Clock::now();
            rc = tdb_run_join(run_ctx, b->dev, res_dev, &o, b->prefix, b->ids, b->k_t.data(),
                              b->tau_t.data(), &all, msg);
            const double ms = ms_since(t0);
            std::lock_guard<std::mutex> lk(mu);
            t_join += ms;
        }
        b->dev = nullptr;
        if (b->slot >= 0) {
            std::lock_guard<std::mutex> lk(mu);
            slot_busy[b->slot] = false;
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
                const auto f = std::lower_bound(all.target.begin(), all.target.end(), (int32_t)lo);
                const auto l = std::lower_bound(all.target.begin(), all.target.end(), (int32_t)hi);
                for (size_t k = f - all.target.begin(); k < (size_t)(l - all.target.begin()); k++) {
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
}

int tdb_server_open(const tdb_server_opts *o, const tdb_tile *resident, tdb_server **out) {
    if (!o || !resident || !out) return TDP_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<tdb_server> s(new tdb_server());
    s->opts = *o;
    s->resident = resident;
    int rc = tdp_open(o->device, &s->run_ctx);
    if (rc == TDP_OK) rc = tdp_open(o->device, &s->prep_ctx);
    uint64_t free_b = 0, total_b = 0;
    if (rc == TDP_OK) rc = tdp_device_memory(s->run_ctx, &free_b, &total_b);
    const tdp_tile_desc d = tdb_desc_of(resident->t);
    if (rc == TDP_OK) rc = tdp_tile_upload(s->run_ctx, &d, &s->res_dev);
    if (rc != TDP_OK) {
        s->err = s->run_ctx ? tdp_last_error(s->run_ctx) : "tdp_open failed";
        if (s->res_dev) tdp_tile_free(s->res_dev);
        if (s->prep_ctx) tdp_close(s->prep_ctx);
        if (s->run_ctx) tdp_close(s->run_ctx);
        s->run_ctx = s->prep_ctx = nullptr;
        *out = s.release();  // so that the message can be read
        return rc;
    }
    s->res_bytes = tdp_tile_bytes(s->res_dev);
    s->budget = o->hbm_budget ? o->hbm_budget : free_b / 10 * 9 + s->res_bytes;
    tdb_server *p = s.get();
    s->prep = std::thread([p] { p->preparer(); });
    s->exec = std::thread([p] { p->executor(); });
    *out = s.release();
    return TDP_OK;
}

int64_t tdb_server_submit(tdb_server *s, tdb_tile *query, const tdb_join_opts *o) {
    if (!s || !query || !o || !s->run_ctx) return TDP_ERR_ARG;
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

void tdb_server_close(tdb_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->prep.joinable()) s->prep.join();
    if (s->exec.joinable()) s->exec.join();
    for (auto &kv : s->done) delete kv.second.res;
    for (auto &q : s->pending) tdb_tile_free(q.tile);
    for (int k = 0; k < 2; k++) {
        if (s->slot[k]) tdp_tile_free(s->slot[k]);
        if (s->stage[k]) tdp_host_free(s->prep_ctx, s->stage[k]);
    }
    if (s->meta_pin) tdp_host_free(s->prep_ctx, s->meta_pin);
    if (s->res_dev) tdp_tile_free(s->res_dev);
    if (s->prep_ctx) tdp_close(s->prep_ctx);
    if (s->run_ctx) tdp_close(s->run_ctx);
    delete s;
}

}  // extern "C"
