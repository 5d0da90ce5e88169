/*
 * tdbase.h - host layer of the MI355X spatial-join path (libtdbase.so).
 *
 * C++ host code (tiles, RAW .dt I/O, the seeded generator, octree candidate
 * generation, the LOD loop) that drives the device through include/tdp.h. It
 * is the operator behind the `tdbase join` command of the reference
 * (src/tools/tdbase.cpp:482-549 -> SpatialJoin::join, SpatialJoin.cpp:463-484
 * -> nearest_neighbor_chunking / within_chunking). C ABI, integer status
 * codes (0 ok, < 0 error, message in tdb_engine_error / the err buffer).
 */
#ifndef TDBASE_H
#define TDBASE_H

#include <stddef.h>
#include <stdint.h>

#include "tdp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tdb_tile tdb_tile;
typedef struct tdb_engine tdb_engine;
typedef struct tdb_result tdb_result;

/* ---- tiles ---- */

/* Load a RAW .dt tile (type byte 1). Replaces Tile::load
 * (src/storage/tile_IO.cpp:13-49) with the RAW branch of the HiMesh_Wrapper
 * constructor (src/spatial/himesh_wrapper.cpp:17-85). COMPRESSED tiles are
 * rejected with an error. */
int tdb_tile_load(const char *path, tdb_tile **out, char *err, size_t err_len);
/* Write a RAW .dt tile, byte-compatible with Tile::dump_raw_stream
 * (src/storage/tile_IO.cpp:79-172). */
int tdb_tile_save(const tdb_tile *t, const char *path, char *err, size_t err_len);
void tdb_tile_free(tdb_tile *t);
/* an independent copy of a tile */
int tdb_tile_clone(const tdb_tile *t, tdb_tile **out);
/* n_obj, n_vox, n_tri */
void tdb_tile_size(const tdb_tile *t, uint64_t *out3);
/* read-only views of the tile arrays (valid while the tile lives):
 * obj_vox_off[n_obj+1], obj_box[6n_obj], obj_transform[3n_obj], obj_hd[10n_obj],
 * vox_box[9n_vox], vox_tri_off[5n_vox], vox_tri_cnt[5n_vox], tri[12n_tri] */
void tdb_tile_views(const tdb_tile *t, const uint32_t **obj_vox_off, const float **obj_box,
                    const float **obj_transform, const float **obj_hd, const float **vox_box,
                    const uint64_t **vox_tri_off, const uint32_t **vox_tri_cnt,
                    const float **tri);

/* Seeded synthetic data (no counterpart in the reference's join path; the
 * reference's simulator, src/tools/simulator.cpp, replicates prototype OFF
 * meshes instead): icosphere nuclei with radial noise, tube vessels along
 * random polylines, a 5-level LOD ladder, voxels of ~voxel_tris triangles and
 * conservative per-triangle Hausdorff bounds. */
typedef struct {
    uint64_t seed;
    int n_vessels;            /* 0: nuclei only, uniform in the cube */
    int nuclei_per_vessel;    /* with n_vessels == 0: total nuclei */
    double cube;
    double vessel_r_min, vessel_r_max;
    int vessel_segments;
    double segment_length;
    double turn_max_deg;
    int vessel_sides;
    int rings_min, rings_max; /* rings per segment, uniform */
    double nucleus_r_min, nucleus_r_max;
    double noise;             /* radial noise amplitude (fraction) */
    int subdiv;               /* icosphere subdivision level */
    double place_min, place_max; /* nucleus distance from its vessel's centreline */
    int prototypes;           /* nucleus shape bank */
    int voxel_tris;
    int threads;
    int64_t max_nuclei;       /* > 0: only the first max_nuclei nuclei (same objects) */
    uint64_t nuclei_seed;     /* != 0: nucleus placements from this seed (vessels and the
                                 nucleus shape bank still follow seed) */
    int emit_vessels;         /* 0: the vessel tile is left empty */
} tdb_gen_params;

void tdb_gen_defaults(tdb_gen_params *p);
int tdb_generate(const tdb_gen_params *p, tdb_tile **nuclei, tdb_tile **vessels);

/* ---- engine and join ---- */

/* Engine on one GPU: owns the device context and the device-resident copies
 * of the tiles it has joined (uploaded once, reused across joins). */
int tdb_engine_open(int device, tdb_engine **out);
void tdb_engine_close(tdb_engine *e);
const char *tdb_engine_error(const tdb_engine *e);
/* device name (for reports) */
int tdb_engine_device(tdb_engine *e, char *name, size_t len, int *cu_count);
/* drop the device copy of a tile (it is uploaded again on next use) */
void tdb_engine_evict(tdb_engine *e, const tdb_tile *t);
/* the engine's kernel tuning (tdp_tuning of tdp.h; performance only) */
int tdb_engine_get_tuning(tdb_engine *e, tdp_tuning *out);
int tdb_engine_set_tuning(tdb_engine *e, const tdp_tuning *t);


#define TDB_QUERY_NN 0
#define TDB_QUERY_WITHIN 1
#define TDB_QUERY_INTERSECT 2  /* SpatialJoin::intersect (IntersectionJoin.cpp:68-161) */

/* the `tdbase join` option set (query_context.h:141-192) */
typedef struct {
    int query;               /* TDB_QUERY_NN, TDB_QUERY_WITHIN or TDB_QUERY_INTERSECT */
    int k;                   /* --knn */
    float within_dist;       /* --within_dist */
    const int *lods;         /* --lods; NULL or n_lods == 0 -> 20 40 60 80 100 */
    int n_lods;
    int hausdorff_level;     /* --hausdorff_level, default 2 */
    uint64_t max_objects1;   /* --max_objects1 (0 = all) */
    int precision;           /* TDP_PREC_* of tdp.h: 0 ref32, 1 fp64, 2 mixed */
    int threads;             /* host threads for candidate generation (0 = all) */
} tdb_join_opts;

void tdb_join_defaults(tdb_join_opts *o);

/* Join tile1 (targets) against tile2 (NULL = self-join, objects never match
 * themselves). Replaces SpatialJoin::join for one tile pair
 * (SpatialJoin.cpp:435-484). */
int tdb_join(tdb_engine *e, const tdb_tile *t1, const tdb_tile *t2, const tdb_join_opts *o,
             tdb_result **out);

/* Autotuner: run the join of the first sample_targets targets (all if 0)
 * under each tuning of a fixed candidate set -- FP64 stage as a work list,
 * persistent, or persistent in a graph-captured round; rows per wave
 * iteration; the small-pair threshold -- timing filter and refinement on the
 * host clock (so host synchronisation counts), and keep the fastest on the
 * engine -- the engine's own tuning, tried first, stays unless another is
 * more than 3% faster (results never depend on the tuning). tried/ms
 * (optional, room for cap entries): each candidate in the order tried and
 * its time. Returns the number of candidates tried, or < 0. */
int tdb_engine_autotune(tdb_engine *e, const tdb_tile *t1, const tdb_tile *t2,
                        const tdb_join_opts *o, size_t sample_targets, tdp_tuning *best,
                        tdp_tuning *tried, double *ms, int cap);

/* candidate lists alone (octree query_knn_fast / query_within per target):
 * prefix[n_targets+1] and ids; returns the number of ids (written up to cap) */
int64_t tdb_candidates(const tdb_tile *t1, const tdb_tile *t2, const tdb_join_opts *o,
                       uint32_t *prefix, uint32_t *ids, size_t cap);
/* the same lists as tdb_join builds them: on the engine's device for kNN
 * (tdp_knn_candidates) and within (tdp_within_candidates), tile 2 of any
 * size, the octree for the targets the
 * device leaves to it (*n_fallback of them) and for everything else; returns
 * the number of ids or -1 */
int64_t tdb_engine_candidates(tdb_engine *e, const tdb_tile *t1, const tdb_tile *t2,
                              const tdb_join_opts *o, uint32_t *prefix, uint32_t *ids, size_t cap,
                              uint64_t *n_fallback);

/* results: every reported (target, object) pair, grouped by target, in
 * report order within a target; with the object distance range at report
 * time and the round (0 = filter, r = r-th LOD round) */
uint64_t tdb_result_count(const tdb_result *r);
uint64_t tdb_result_targets(const tdb_result *r);
void tdb_result_get(const tdb_result *r, int32_t *target, int32_t *obj2, double *rmin,
                    double *rmax, uint8_t *round);
/* timing (ms): [0] join total, [1] candidate generation, [2] filter stage,
 * [3] refinement, [4] filter kernels (device), [5] refine kernels (device),
 * [6] tile upload */
void tdb_result_timing(const tdb_result *r, double *out7);
/* the first n (<= 10) timing figures: the seven above, then [7] join set-up
 * (state and candidate lists to the device), [8] result download, [9] result
 * assembly on the host. Returns the number written. */
int tdb_result_timing_n(const tdb_result *r, double *out, int n);
/* [0] triangle pairs covered, [1] voxel pairs evaluated, [2] FP64 TriDist
 * evaluations (mixed form), [3] voxel pairs after the filter, [4] candidate
 * object pairs, [5] rounds, [6] triangle pairs excluded by the FP32 bounds
 * (mixed form) */
void tdb_result_stats(const tdb_result *r, uint64_t *out7);
/* the same and, at [7], the triangle pairs the mixed form's candidate pass
 * tested one by one, at [8] the FP64 evaluations of the probes (part of [2];
 * tdp_join_stats_n); the first min(n, 9). No counterpart in the reference
 * (statistics). */
int tdb_result_stats_n(const tdb_result *r, uint64_t *out, int n);
/* device time (ms) and launches per refinement stage, TDP_STAGE_* order of
 * tdp.h (n up to TDP_STAGE_COUNT) */
void tdb_result_stages(const tdb_result *r, double *ms, uint64_t *launches, int n);
void tdb_result_free(tdb_result *r);

/* ---- streaming join: object sets larger than the device budget ----
 *
 * A source produces any object's arrays on demand (the LOD decode of the
 * reference, HiMesh_Wrapper::decode_to, himesh_wrapper.cpp:137-161, done per
 * chunk instead of per tile). tdb_stream_join joins chunks of targets --
 * taken in Morton order of their box centres, each with every object its
 * candidate lists name -- within a device-memory budget, decoding chunk i+2
 * on host threads into pinned buffers while chunk i+1 is copied to the device
 * on one stream and chunk i is joined on another. Replaces the reference's
 * chunked pipeline (nearest_neighbor_chunking / within_chunking,
 * KNNJoin.cpp:1524-1811, WithinJoin.cpp:644-813, with the packing producer
 * KNNJoin.cpp:1586-1617 and SpatialJoin.cpp:169-248). The result equals
 * tdb_join's on the same objects.
 *
 * Workers (each with its own device stream, tile slots and pinned staging)
 * take chunks in turn, so one worker's host work runs beside another's
 * kernels. With `lazy` a chunk's tiles go to the device without their
 * triangle records: before each LOD round the device lists the voxels that
 * the round's live voxel pairs read, and only those voxels' records at that
 * LOD are decoded (ObjectSource::decode_slot) and shipped -- the reference's
 * per-LOD packing of live candidates' voxels (KNNJoin.cpp:1584-1617). */
typedef struct tdb_source tdb_source;
/* the generator's nuclei (the vessel tile is not produced), object for object
 * equal to tdb_generate's nuclei tile */
int tdb_source_generate(const tdb_gen_params *p, tdb_source **out);
/* the objects of a tile (which must outlive the source) */
int tdb_source_tile(const tdb_tile *t, tdb_source **out);
uint64_t tdb_source_objects(const tdb_source *s);
/* objects[0..n) decoded into a new tile, in that order */
int tdb_source_decode(const tdb_source *s, const uint64_t *objects, size_t n, tdb_tile **out);
void tdb_source_free(tdb_source *s);

typedef struct {
    uint64_t hbm_budget;        /* device bytes the join may hold: both tile slots and the
                                   join buffers (0 = 90% of the free memory) */
    double tile_share;          /* share of the budget for the workers' tile slots, in (0, 1) */
    uint64_t max_chunk_targets; /* most targets per chunk (0 = as many as fit a slot) */
    int decode_threads;         /* host threads decoding, shared by the workers (0 = all) */
    int overlap;                /* 1: `workers` workers side by side; 0: one, stages in turn */
    int lazy;                   /* 1: records shipped per LOD round for the voxels it reads
                                   (default); 0: every record of every chunk object */
    int workers;                /* workers when overlapping (0 = 4); fewer when a target with
                                   its candidates would not fit their tile slots */
} tdb_stream_opts;

void tdb_stream_defaults(tdb_stream_opts *o);
/* Join source 1's objects (targets: the first max_objects1) against source
 * 2's (NULL = self-join). */
int tdb_stream_join(tdb_engine *e, const tdb_source *s1, const tdb_source *s2,
                    const tdb_join_opts *o, const tdb_stream_opts *so, tdb_result **out);
/* figures of a streamed result: [0] chunks, [1] bytes copied to the device,
 * [2] decode ms (host, summed over chunks), [3] copy ms (submit to done),
 * [4] join ms, [5] pipeline wall ms, [6] overlap fraction ((sum of the three
 * stages - wall) / (sum - slowest stage): 1 = the wall equals the slowest
 * stage, 0 = no overlap), [7] peak device bytes held, [8] index ms (object
 * boxes), [9] candidate ms, [10] objects decoded, [11] chunk joins split to
 * fit the budget */
void tdb_result_stream(const tdb_result *r, double *out12);
/* the same figures and more, n of them (at most 19; returns how many): [0..11]
 * as above with [2] decode ms including the per-round record decodes and [3]
 * copy ms including the per-round record puts, then [12] record bytes of the
 * chunks' objects (what shipping every record would copy), [13] record bytes
 * shipped, [14] voxel and object array bytes shipped, [15] workers, [16]
 * targets whose kNN or within candidate lists came from the host octree (the
 * device builds the others; -1 when the octree built them all), [17] chunk plan ms,
 * [18] worker set-up ms (contexts, tile slots, pinned staging) */
int tdb_result_stream_n(const tdb_result *r, double *out, int n);

/* ---- serving: queries against one resident tile, continuously batched ----
 *
 * No counterpart in the reference (which runs one join per process). A
 * server keeps `resident` (tile 2 of every query) on the device and joins
 * each submitted query tile (tile 1) against it. Queries waiting when a tile
 * slot frees up are merged into one join -- those sharing the LOD list and
 * Hausdorff level (kNN and within queries together, each target with its
 * kind's rules, tdp_join_create_kinds; intersection queries among
 * themselves), as many as fit the device-memory budget; k and the within
 * distance may differ per query. Batches are prepared (tile merge,
 * octree candidates, upload) on their own stream while others run, and up to
 * `executors` batches are joined at once, each on its own stream, so a query
 * that arrives while others are on the device starts beside them. With
 * `lazy` a batch's triangle records go to the device round by round, only
 * for the query voxels each LOD round reads. A query's result is the same
 * alone and in any batch. */
typedef struct tdb_server tdb_server;
typedef struct {
    int device;
    uint64_t hbm_budget;  /* device bytes the server may hold, resident tile included
                             (0 = the resident tile + 90% of the free memory) */
    int max_batch;        /* most queries in one join (1 = one at a time) */
    int overlap;          /* 1: prepare the next batch while one runs */
    int precision;        /* TDP_PREC_* of tdp.h */
    int threads;          /* host threads for candidates and record copies (0 = all) */
    int lazy;             /* 1: query records per LOD round for the voxels it reads
                             (default); 0: every record with the tile */
    int executors;        /* batches joined at once when overlapping (0 = 2) */
    int mix_kinds;        /* 1: kNN and within queries share joins (default); 0: a batch
                             holds one query type */
} tdb_server_opts;

void tdb_server_defaults(tdb_server_opts *o);
/* resident must outlive the server */
int tdb_server_open(const tdb_server_opts *o, const tdb_tile *resident, tdb_server **out);
/* Queue a query: the server takes ownership of `query` (it frees it). Uses
 * o->query, k, within_dist, lods, hausdorff_level. Returns a ticket >= 0, or
 * a negative status. */
int64_t tdb_server_submit(tdb_server *s, tdb_tile *query, const tdb_join_opts *o);
/* Block until the ticket's query is done; *out as tdb_join's (free with
 * tdb_result_free), *latency_ms = submit to completion. */
int tdb_server_wait(tdb_server *s, int64_t ticket, tdb_result **out, double *latency_ms);
const char *tdb_server_error(const tdb_server *s);
/* [0] batches run, [1] queries done, [2] most queries in a batch,
 * [3] peak device bytes held (resident tile included) */
void tdb_server_stats(tdb_server *s, uint64_t *out4);
/* host time (ms, summed over batches) of [0] tile merge, [1] upload, [2]
 * candidates (the preparer) and [3] joins (the executor) */
void tdb_server_stage_ms(tdb_server *s, double *out4);
/* more figures, n of them (at most 13; returns how many): [0..3] as
 * tdb_server_stage_ms with [3] excluding [4] the record copies into pinned
 * memory and [5] the record puts (executors, ms), [6] bytes of voxel and
 * object arrays and [7] of triangle records sent to the device, [8] record
 * bytes of the queries' tiles (what shipping every record would send),
 * [9] executors, [10] most batches joined at once, [11] batches run,
 * [12] batches holding both kNN and within queries */
int tdb_server_figures(tdb_server *s, double *out, int n);
/* finishes the queued queries, then stops */
void tdb_server_close(tdb_server *s);

#ifdef __cplusplus
}
#endif
#endif /* TDBASE_H */
