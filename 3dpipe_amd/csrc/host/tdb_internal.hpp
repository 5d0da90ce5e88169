// Internals shared by the host join engine (tdb_capi.cpp) and the serving
// loop (tdb_serve.cpp): the opaque types of include/tdbase.h and the join
// driver both run.
#pragma once

#include <chrono>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/tdbase.h"
#include "../../../include/tdp.h"
#include "tdb_octree.hpp"
#include "tdb_tile.hpp"

struct tdb_tile {
    tdb::Tile t;
    mutable std::once_flag tree_once;
    mutable std::unique_ptr<tdb::Octree> tree;
    // engines holding a device copy of this tile (guarded by g_registry)
    mutable std::vector<tdb_engine *> holders;
    const tdb::Octree &octree() const {
        std::call_once(tree_once, [this] {
            tree.reset(new tdb::Octree(t.obj_box.data(), t.obj_transform.data(), t.n_obj, 10));
        });
        return *tree;
    }
};

struct tdb_engine {
    int device = 0;
    tdp_ctx *ctx = nullptr;
    std::string err;
    std::unordered_map<const tdb_tile *, tdp_tile *> resident;
};

// Leaves new elements uninitialised on resize (they are all written
// afterwards, in parallel, which also spreads the page faults over threads).
template <typename T>
struct NoInit : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInit<U>;
    };
    NoInit() = default;
    template <typename U>
    NoInit(const NoInit<U> &) {}
    template <typename U>
    void construct(U *) noexcept {}
    template <typename U, typename... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
template <typename T>
using ResultVec = std::vector<T, NoInit<T>>;

struct tdb_result {
    uint64_t n_targets = 0;
    ResultVec<int32_t> target, obj2;
    ResultVec<double> rmin, rmax;
    ResultVec<uint8_t> round;
    double timing[10] = {};
    double stage_ms[TDP_STAGE_COUNT] = {};
    uint64_t stage_n[TDP_STAGE_COUNT] = {};
    uint64_t stats[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double stream[19] = {};  // streaming figures (tdb_result_stream_n)
};

using Clock = std::chrono::steady_clock;
static double ms_since(Clock::time_point a) {
    return std::chrono::duration<double, std::milli>(Clock::now() - a).count();
}


// the device description of a host tile (arrays stay owned by the tile)
tdp_tile_desc tdb_desc_of(const tdb::Tile &t);

// the --lods list as the reference's parser leaves it (sorted, unique without erase)
std::vector<int> tdb_lod_list(const tdb_join_opts *o);

// candidate lists of targets [0, n_targets) of t1 against t2 (octree); k_t and
// tau_t, when given, hold each target's k / within distance
void tdb_gen_candidates(const tdb_tile *t1, const tdb_tile *t2, bool self, size_t n_targets,
                        const tdb_join_opts *o, const int32_t *k_t, const float *tau_t,
                        std::vector<uint32_t> &prefix, std::vector<uint32_t> &ids);

// the same, on the device where it applies (kNN, and within when
// device_within, tile 2 of at most 1024 objects; the octree for targets whose
// lists the device leaves to it)
int tdb_gen_candidates_dev(tdp_ctx *ctx, const tdp_tile *d1, const tdp_tile *d2,
                           const tdb_tile *t1, const tdb_tile *t2, bool self, size_t N,
                           const tdb_join_opts *o, const int32_t *k_t, const float *tau_t,
                           std::vector<uint32_t> &prefix, std::vector<uint32_t> &ids,
                           uint64_t *n_fallback, std::string &err, bool device_within = true);

// Records for tiles filled without them (tdp_tile_fill_meta_async): before
// each LOD round the join lists, per side, the voxels that round reads and
// the tile does not hold; pack(side, slot, voxels, n, packed_off, out) writes
// their records at that slot into out (pinned, 12 floats per record), voxel
// k's first at packed_off[k], and returns false on failure (err set). The
// pinned buffers are the feed's, grown as needed (free them with
// release()).
struct LazyFeed {
    std::function<bool(int side, int slot, const uint32_t *vox, size_t n, const uint64_t *off,
                       float *out, std::string &err)>
        pack;
    // per side: the tile's record count of each voxel at each slot (5 per voxel)
    const uint32_t *vox_tri_cnt[2] = {nullptr, nullptr};
    tdp_ctx *ctx = nullptr;  // owner of the pinned buffers
    float *buf[2] = {nullptr, nullptr};
    size_t cap[2] = {0, 0};  // records
    std::vector<uint64_t> off;
    double pack_ms = 0, put_ms = 0;
    uint64_t put_records = 0;
    void release() {
        for (int k = 0; k < 2; k++) {
            if (buf[k]) tdp_host_free(ctx, buf[k]);
            buf[k] = nullptr;
            cap[k] = 0;
        }
    }
};

// Filter + LOD rounds on the device for the given candidates, results into
// res (targets numbered 0..N-1, reported pairs ordered by (round, list
// position)); timing [2..5] and stats filled. feed: the records of tiles
// filled without them, round by round (nullptr: the tiles hold theirs).
// Returns a TDP_* status.
// within_t: kNN and within targets in one join (1 = within; o->query is
// then either), or nullptr for o->query's kind throughout.
int tdb_run_join(tdp_ctx *ctx, tdp_tile *d1, tdp_tile *d2, const tdb_join_opts *o,
                 const std::vector<uint32_t> &prefix, const std::vector<uint32_t> &ids,
                 const int32_t *k_t, const float *tau_t, tdb_result *res, std::string &err,
                 LazyFeed *feed = nullptr, const uint8_t *within_t = nullptr);
