/*
 * tdp.h - device layer of the MI355X spatial-join refinement path (libtdp.so).
 *
 * A C ABI over hand-written gfx950 HIP kernels: plain pointers and sizes, no
 * C++ or torch types, integer status codes (0 = ok, < 0 = error; the message
 * is in tdp_last_error). Nothing here calls exit(). One host thread drives a
 * context at a time; the library owns all device memory it allocates.
 *
 * It replaces the GPU entry points the reference's join engine calls through
 * src/include/geometry.h (the free functions at geometry.h:209-531 and the
 * dispatcher geometry_computer, geometry.h:533-572), implemented in
 * src/geometry/computing_gpu.cu. Each entry point names the reference
 * interface it stands in for.
 */
#ifndef TDP_H
#define TDP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TDP_OK 0
#define TDP_ERR_ARG -1     /* invalid argument */
#define TDP_ERR_HIP -2     /* HIP runtime error (no device, launch failure, ...) */
#define TDP_ERR_NOMEM -3   /* device or pinned allocation failed */
#define TDP_ERR_STATE -4   /* call out of order */

/* arithmetic form of the distance path */
#define TDP_PREC_REF32 0  /* the reference's FP32 arithmetic, bit for bit */
#define TDP_PREC_FP64 1   /* the same algorithm in FP64 ("plain") */
#define TDP_PREC_MIXED 2  /* FP32 bulk + FP64 re-check; results identical to FP64 */

#define TDP_QUERY_KNN 0
#define TDP_QUERY_WITHIN 1
#define TDP_QUERY_INTERSECT 2  /* object pairs whose meshes touch (distance 0) */

typedef struct tdp_ctx tdp_ctx;
typedef struct tdp_tile tdp_tile;
typedef struct tdp_join tdp_join;

/* Context on one HIP device.
 * Replaces: initialize()/get_gpus()/init_gpu (src/geometry/cuda_common.cu:15-73,
 * src/include/mygpu.h:22-29) and initialize_refine_stream
 * (src/geometry/computing_gpu.cu:37-43). */
int tdp_open(int device, tdp_ctx **out);
int tdp_close(tdp_ctx *ctx);
/* message of the last failed call on ctx (never NULL; "" if none) */
const char *tdp_last_error(const tdp_ctx *ctx);
/* device name and compute-unit count (for reports) */
int tdp_device_info(tdp_ctx *ctx, char *name, size_t name_len, int *cu_count);
/* free and total device memory in bytes (for memory budgets) */
int tdp_device_memory(tdp_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);

/* Device-resident tile. Triangles are 48-byte records
 * {v0.xyz, proxy_hausdorff, v1.xyz, hausdorff, v2.xyz, 0}; voxel boxes are
 * 9 floats {low.xyz, high.xyz, core.xyz}. Per (voxel, LOD slot 0..4 for LOD
 * 20..100): the first record index and the record count.
 * Replaces: allocateVoxelsForAll (computing_gpu.cu:1708-1748) for the voxel
 * boxes, and the per-chunk packing + H2D of triangles and Hausdorff values
 * (SpatialJoin.cpp:169-248, computing_gpu.cu:6540-6658): the whole tile is
 * uploaded once and stays resident for every LOD. */
typedef struct {
    size_t n_obj, n_vox, n_tri;
    const uint32_t *obj_vox_off; /* n_obj + 1 */
    const float *obj_hd;         /* 10 per object: (hausdorff, proxy) per LOD slot */
    const float *vox_box;        /* 9 per voxel */
    const uint64_t *vox_tri_off; /* 5 per voxel */
    const uint32_t *vox_tri_cnt; /* 5 per voxel */
    const float *tri;            /* 12 per triangle */
    const float *obj_box;        /* 6 per object {low, high}, or NULL */
    const float *obj_transform;  /* 3 per object (the max-distance proxy point), or NULL */
} tdp_tile_desc;

int tdp_tile_upload(tdp_ctx *ctx, const tdp_tile_desc *desc, tdp_tile **out);
/* The same with the triangle records in n_parts host buffers, concatenated
 * in order (desc->tri is ignored): tiles assembled from several sources, such
 * as a batch of query tiles, go up without a host-side copy. No counterpart
 * in the reference. */
int tdp_tile_upload_parts(tdp_ctx *ctx, const tdp_tile_desc *desc, const float *const *tri_parts,
                          const size_t *part_n_tri, size_t n_parts, tdp_tile **out);
int tdp_tile_free(tdp_tile *tile);
size_t tdp_tile_bytes(const tdp_tile *tile);

/* ---- streaming: a device-memory cap, pinned staging, tiles refilled in place ----
 *
 * Replaces the reference's chunked streaming (the producer/consumer of
 * KNNJoin.cpp:1586-1617 packing into pinned slots, SpatialJoin.cpp:169-248,
 * and the per-chunk cudaMalloc + H2D + kernel + D2H of
 * computing_gpu.cu:6540-6658 on refine_streams): a tile reserved once per
 * pipeline slot is refilled for each chunk from pinned host buffers by
 * copies queued on the filling context's stream, while joins run on another
 * context. */
/* Cap on the device bytes ctx holds (tiles it uploaded or reserved, join
 * buffers, caches); an allocation beyond it first releases the cache, then
 * fails with TDP_ERR_NOMEM. 0 = no cap (the default). */
int tdp_set_memory_limit(tdp_ctx *ctx, uint64_t bytes);
/* bytes held now and the most held (since open, or the last reset) */
int tdp_memory_usage(tdp_ctx *ctx, uint64_t *used, uint64_t *peak, int reset_peak);
/* pinned host memory for the asynchronous fills and fast downloads; freed
 * blocks are kept by the context (up to 1 GiB) and handed out again by later
 * calls, so a per-call buffer costs no pinning after the first call. Free
 * every block before tdp_close. */
int tdp_host_alloc(tdp_ctx *ctx, size_t bytes, void **out);
int tdp_host_free(tdp_ctx *ctx, void *p);
/* wait until the work queued on ctx's stream (fills) has finished */
int tdp_sync(tdp_ctx *ctx);
/* a point in ctx's stream (after the work queued so far): *ticket names it;
 * tdp_wait_mark blocks until the stream has passed it (for reusing a pinned
 * buffer whose copy was queued before the mark) */
int tdp_mark(tdp_ctx *ctx, uint64_t *ticket);
int tdp_wait_mark(tdp_ctx *ctx, uint64_t ticket);
/* an empty device tile with room for n_obj objects, n_vox voxels and n_tri
 * triangle records */
int tdp_tile_reserve(tdp_ctx *ctx, size_t n_obj, size_t n_vox, size_t n_tri, tdp_tile **out);
/* Refill a reserved tile (sizes within its room). The arrays of desc must be
 * pinned (tdp_host_alloc) and stay untouched until tdp_sync: the copies, the
 * voxel owners and the triangle bounds are queued on the stream of the
 * context that reserved the tile and the call returns at once. desc->tri may
 * be NULL: the records then come in pieces through tdp_tile_fill_tri_async.
 * obj_box / obj_transform are copied when given (for device candidates). */
int tdp_tile_fill_async(tdp_tile *tile, const tdp_tile_desc *desc);
/* records [first, first + n) of the tile being filled, from pinned memory */
int tdp_tile_fill_tri_async(tdp_tile *tile, size_t first, size_t n, const float *records);

/* ---- records per LOD round: only the voxels the live voxel pairs still need ----
 *
 * Replaces the reference's per-LOD packing of the triangles of live
 * candidates' voxels into the chunk buffers before each round's kernels
 * (build_chunks_by_pair + the producer thread, KNNJoin.cpp:1584-1617;
 * packing_data_stream, SpatialJoin.cpp:169-248): a tile filled by
 * tdp_tile_fill_meta_async holds its voxel and object arrays but no
 * records; before each tdp_join_refine the caller asks, per side, for the
 * voxels that round reads and the tile does not hold yet, and hands their
 * records over packed in pinned memory: one DMA takes them to device
 * staging, and a kernel writes them into place with their bound records.
 *
 * Refill a reserved tile without its records (desc->tri ignored; arrays
 * pinned, as for tdp_tile_fill_async). */
int tdp_tile_fill_meta_async(tdp_tile *tile, const tdp_tile_desc *desc);
/* Voxels of join side 1 (t1) or 2 (t2) that the next round at `lod` reads
 * and the tile does not hold: ascending voxel ids in *voxels (pinned memory
 * the join owns, valid until the next call for that side), *n_voxels of
 * them holding *n_records records at that LOD. A self-join lists its one
 * tile on side 1 (side 2 lists nothing). A tile with records lists nothing.
 * The next call on the side is tdp_join_lazy_put (the listed voxels count
 * as held once it has run; listed again by a later call if it failed). */
int tdp_join_lazy_needs(tdp_join *j, int lod, int side, const uint32_t **voxels,
                        size_t *n_voxels, uint64_t *n_records);
/* The listed voxels' records at `lod`, voxel after voxel in list order, each
 * voxel's records in its (voxel, LOD) order: 12 floats per record, in
 * pinned memory from tdp_host_alloc that stays untouched until the next
 * tdp_join_refine on j has returned (the copy is queued). tdp_join_refine
 * fails with TDP_ERR_STATE on a side whose records for its LOD were not put. */
int tdp_join_lazy_put(tdp_join *j, int lod, int side, const float *records);
/* records a tile holds: all of them, or for a tile filled without records
 * those put since the fill */
int tdp_tile_put_records(const tdp_tile *tile, uint64_t *records);

/* Triangle-triangle distance for n independent pairs; S and T are 9 floats
 * per triangle (host memory), out is n values (host) in the given precision:
 * REF32 -> float* (bit-identical to the reference's TriDist), FP64 and MIXED
 * -> double*.
 * Replaces: TriDist (computing_cpu.cpp:403-439) / TriDist_kernel
 * (computing_gpu.cu:566-610). */
int tdp_tridist_batch(tdp_ctx *ctx, int precision, const float *S, const float *T, size_t n,
                      void *out);

/* Per voxel pair: minimum triangle distance, minimum lower bound
 * max(d - ph1 - ph2, 0) and minimum upper bound d + h1 + h2 over the n1 x n2
 * triangle pairs. tri: 12-float records (host), offset_size: 4 x u64 per pair
 * {offset1, n1, offset2, n2} in records; out: 3 values per pair (float for
 * REF32, double otherwise).
 * Replaces: MeshDist_batch_gpu_op_shm_flat_stream (computing_gpu.cu:6540-6658)
 * and its kernel TriDist_cuda_op_shm_flat_fused (:1069-1152). */
int tdp_meshdist_batch(tdp_ctx *ctx, int precision, const float *tri, size_t n_tri,
                       const uint64_t *offset_size, size_t n_pairs, void *out);

/* tdp_meshdist_batch on device-resident arrays, allocating no device memory.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc). tri
 * (12-float records), offset_size (4 x u64 per pair) and out (3 values per
 * pair, float for REF32, double otherwise) are device pointers the caller
 * owns; the triangles' bound records and the counters live in a fixed
 * per-GPU array of the library (12 MiB, held until the process ends), so
 * n_tri may be at most 262144 (TDP_ERR_ARG beyond, before any work). Pairs
 * reaching past n_tri records are refused with TDP_ERR_ARG (checked on the
 * device). Results equal tdp_meshdist_batch's bit for bit. The call returns
 * when out is written. Calls on one GPU share its array: within a process
 * they are serialised by a lock; calls from threads that bypass it, or
 * work queued on other streams that reads the array, would race. */
int tdp_meshdist_batch_noalloc(tdp_ctx *ctx, int precision, const float *tri, size_t n_tri,
                               const uint64_t *offset_size, size_t n_pairs, void *out);

/* kNN join status (evaluate_knn_kernel, computing_gpu.cu:105-151, via
 * evaluation_kernel :6905-6950): per target t with candidates
 * [prefix[t], prefix[t+1]) and confirmed[t] already confirmed:
 * status 1 = CONFIRMED, 2 = REMOVE, 0 = KEEP. Host arrays; min/max are double
 * and compared in double (exact for values that came from float). */
int tdp_knn_status(tdp_ctx *ctx, size_t n_targets, const uint32_t *prefix,
                   const int32_t *confirmed, const double *rmin, const double *rmax, int k,
                   int32_t *status);

/* kNN candidate lists on the device, equal to the octree's query_knn_fast
 * lists (octree.cpp:215-291 with add_candidate :120-154) including their
 * order: per target t of t1 (t < n_targets), the objects o of t2 (o != t when
 * self) with box distance < M or among the k_t smallest transform distances,
 * M being the k_t-th smallest, ordered by transform distance. Where that
 * order or the k-th place is decided by equal distances -- which the octree
 * breaks by its visiting order -- or a list is longer than 256, fallback[t]
 * = 1 and the target's list is left empty for the host octree. Both tiles
 * need obj_box and obj_transform; t2 may hold any number of objects below
 * 2^32 - 1 (up to 1024, one wave per target holds them all; beyond, they are
 * binned into a grid of cells and a target reads the cells near it).
 * Host arrays: k_target[n_targets],
 * prefix[n_targets + 1], ids[ids_cap], fallback[n_targets]; *n_ids = the
 * number of ids (if it exceeds ids_cap, nothing is written to ids and the
 * call returns TDP_ERR_ARG: call again with room for *n_ids). */
int tdp_knn_candidates(tdp_ctx *ctx, const tdp_tile *t1, size_t n_targets, const tdp_tile *t2,
                       int self, const int32_t *k_target, uint32_t *prefix, uint32_t *ids,
                       size_t ids_cap, uint64_t *n_ids, uint8_t *fallback);

/* Within candidate lists on the device, equal to the octree's query_within
 * lists (octree.cpp:293-330) including their order: per target t of t1
 * (t < n_targets), the objects o of t2 (o != t when self) whose box distance
 * to the target's box is at most tau_target[t], in the order of the octree's
 * depth-first walk, each once. The caller passes the octree's leaves in
 * depth-first order: leaf_box[6 n_leaves] ({low.xyz, high.xyz}), leaf_off
 * [n_leaves + 1] and leaf_obj[leaf_off[n_leaves]] (each leaf's objects in
 * its order). Lists longer than 256 set fallback[t] = 1 and are left empty
 * for the host octree. Both tiles need obj_box; t2 may hold any number of
 * objects below 2^32 - 1 (beyond 16384 it needs obj_transform as well: its
 * objects are binned into a grid of cells by proxy point, and each listed
 * object's place comes from its first leaf on the walk whose box passes). Host arrays; ids, ids_cap and n_ids as
 * for tdp_knn_candidates. */
int tdp_within_candidates(tdp_ctx *ctx, const tdp_tile *t1, size_t n_targets, const tdp_tile *t2,
                          int self, const float *tau_target, size_t n_leaves,
                          const float *leaf_box, const uint32_t *leaf_off,
                          const uint32_t *leaf_obj, uint32_t *prefix, uint32_t *ids,
                          size_t ids_cap, uint64_t *n_ids, uint8_t *fallback);

/* ---- tuning (performance only: results never depend on it) ---- */
typedef struct {
    int small_pairs;  /* MIXED: voxel pairs of at most this many triangle pairs go to FP64
                         without the FP32 passes */
    int max_rows;     /* MIXED: most rows of a voxel pair's grid per wave iteration of the
                         FP32 passes (1..64; the tile is rows x columns <= 64 lanes) */
    int drain;        /* MIXED: 1 = persistent FP64 stage (candidates queued in LDS per wave,
                         voxel pairs taken from a work counter); 0 = a flat work list */
    int drain_blocks; /* blocks of the persistent stage (0 = 4 per compute unit) */
    int eval_blocks;  /* grid of the work-list FP64 stage (grid-stride) */
    int graph;        /* MIXED: 1 = each LOD round is captured in a hipGraph and launched as one
                         (implies the persistent FP64 stage; one host synchronisation per round) */
    int row_cull;     /* MIXED: 1 = the candidate pass first bounds each row triangle against a
                         sphere around the other voxel and skips the wave iterations whose rows
                         are all ruled out */
} tdp_tuning;
/* the context's tuning (defaults at tdp_open); joins created afterwards use it */
int tdp_get_tuning(tdp_ctx *ctx, tdp_tuning *out);
int tdp_set_tuning(tdp_ctx *ctx, const tdp_tuning *t);

/* ---- device-resident join (filter + LOD refinement) ---- */

/* The intersection join (TDP_QUERY_INTERSECT) follows SpatialJoin::intersect
 * (IntersectionJoin.cpp:68-161) with the per-voxel-pair values of TriInt_cuda
 * (computing_gpu.cu:1441-1512): candidates are the objects whose boxes
 * overlap the target's (sorted by index), voxel pairs those whose boxes
 * overlap; in each LOD round an object pair is reported (range [0, 0]) when
 * a voxel pair's min of d + h1 + h2 is 0, and, at Hausdorff levels 2 / 1,
 * ruled out when every voxel pair's min of max(d - ph1 - ph2, 0) / d minus
 * both objects' proxy Hausdorff is positive. Pairs left after the last LOD
 * are not reported. */
typedef struct {
    int query;            /* TDP_QUERY_KNN, TDP_QUERY_WITHIN or TDP_QUERY_INTERSECT */
    int k;                /* kNN */
    float within_dist;    /* within */
    int precision;        /* TDP_PREC_* */
    int hausdorff_level;  /* 0, 1 or 2 */
} tdp_join_params;

/* Candidate lists from the octree: target t of tile 1 (t < n_targets) has
 * candidates cand_obj[cand_prefix[t] .. cand_prefix[t+1]) of tile 2, in list
 * order. Replaces the device set-up of mbb_knn_op_gpu_one_kernel /
 * mbb_within_op_gpu_one_kernel (KNNJoin.cpp:497-851, WithinJoin.cpp:26-258). */
int tdp_join_create(tdp_ctx *ctx, const tdp_tile *t1, const tdp_tile *t2,
                    const tdp_join_params *params, size_t n_targets, const uint32_t *cand_prefix,
                    const uint32_t *cand_obj, tdp_join **out);

/* The same with a k (kNN) and a within distance per target (host arrays of
 * n_targets; NULL = params->k / params->within_dist for every target), so
 * that queries with different parameters share one join: each target's
 * lists and ranges are what a join of that target alone would give. No
 * counterpart in the reference, where a join has one k / tau (its callers
 * are serving loops that batch concurrent queries). */
int tdp_join_create_multi(tdp_ctx *ctx, const tdp_tile *t1, const tdp_tile *t2,
                          const tdp_join_params *params, size_t n_targets,
                          const uint32_t *cand_prefix, const uint32_t *cand_obj,
                          const int32_t *k_target, const float *tau_target, tdp_join **out);
/* The same with kNN and within targets in one join: within_target[t] = 1
 * marks a within target (its within distance from tau_target or
 * params->within_dist), 0 a kNN target (its k from k_target or params->k);
 * params->query is TDP_QUERY_KNN or TDP_QUERY_WITHIN (its rounds fold the
 * same quantities and evict the same voxel pairs for both kinds). Each
 * target gets the kNN or within rules of the reference at the filter and
 * after every round (KNNJoin.cpp:133-202, 497-851; WithinJoin.cpp:178-252,
 * 769-792), so its results are those of a join of its kind alone. No
 * counterpart in the reference (serving loops merging concurrent queries of
 * both kinds). */
int tdp_join_create_kinds(tdp_ctx *ctx, const tdp_tile *t1, const tdp_tile *t2,
                          const tdp_join_params *params, size_t n_targets,
                          const uint32_t *cand_prefix, const uint32_t *cand_obj,
                          const uint8_t *within_target, const int32_t *k_target,
                          const float *tau_target, tdp_join **out);

/* Filter stage on the device: object-level voxel bounds, the filter-time kNN
 * status / within decisions and compaction of surviving voxel pairs.
 * Replaces compute_voxel_pair_distance_for_all_streaming_pipeline_reduceIO
 * (computing_gpu.cu:4417-4862) and ..._within_streaming_pipeline (:5681-6013)
 * together with the host decisions that follow them. *pairs_out = number of
 * surviving voxel pairs. */
int tdp_join_filter(tdp_join *j, uint64_t *pairs_out);

/* One refinement round at `lod` (20..100; other values decode to empty
 * voxels): triangle distances of every live voxel pair, bound folding,
 * eviction, then the kNN status walk or the within decision. is_last selects
 * exact distances ([d,d]). *pairs_out = live voxel pairs after the round.
 * Replaces one iteration of the LOD loop of nearest_neighbor_chunking
 * (KNNJoin.cpp:1564-1800) / within_chunking (WithinJoin.cpp:663-802). */
int tdp_join_refine(tdp_join *j, int lod, int is_last, uint64_t *pairs_out);

/* Per candidate pair (in cand_prefix order): the round that reported it
 * (0 = filter, r = r-th refinement round, 255 = not reported) and its object
 * distance range when reported (or its last range). */
int tdp_join_results(tdp_join *j, uint8_t *report_round, double *rmin, double *rmax);

/* counters: [0] triangle pairs covered (sum of n1*n2 over evaluated voxel
 * pairs), [1] voxel pairs evaluated, [2] FP64 TriDist evaluations (MIXED),
 * [3] voxel pairs after the filter, [4] triangle pairs excluded by the FP32
 * bounds without any TriDist evaluation (MIXED) */
int tdp_join_stats(tdp_join *j, uint64_t *out5);
/* The same figures and, at [5], the triangle pairs the MIXED form's candidate
 * pass tested one by one (the others lie in rows it ruled out whole against
 * the voxel spheres, or in voxel pairs small enough for FP64 alone), at [6]
 * the FP64 TriDist evaluations of the probes (part of [2]); the first
 * min(n, 7). No counterpart in the reference (statistics). */
int tdp_join_stats_n(tdp_join *j, uint64_t *out, int n);
/* device time (ms, from HIP events) spent in the refine kernels so far */
int tdp_join_kernel_ms(tdp_join *j, double *refine_ms, double *filter_ms);

/* Device time per stage of the refinement rounds so far (ms, from HIP events
 * recorded on the join's stream between the stages) and launches per stage.
 * Stages: */
#define TDP_STAGE_RESET 0   /* per-round bookkeeping: accumulator reset, pair counts */
#define TDP_STAGE_BOUNDS 1  /* MIXED: FP32 upper bounds and probe choice */
#define TDP_STAGE_SCAN 2    /* MIXED: prefix sums, including the host read-back */
#define TDP_STAGE_PROBE 3   /* MIXED: FP64 TriDist of the probes, the mark pass's thresholds */
#define TDP_STAGE_MARK 4    /* MIXED: FP32 candidate test and bit masks */
#define TDP_STAGE_EXPAND 5  /* MIXED: bit masks to the FP64 work list */
#define TDP_STAGE_EVAL 6    /* MIXED: FP64 TriDist of the work list */
#define TDP_STAGE_FOLD 7    /* MIXED: range fold */
#define TDP_STAGE_DIRECT 8  /* REF32 / FP64: the one-pass distance kernel */
#define TDP_STAGE_GRAPH 9   /* MIXED with graph tuning: the whole captured round */
#define TDP_STAGE_COUNT 10
int tdp_join_stage_ms(tdp_join *j, double *ms, uint64_t *launches, int n_stages);
int tdp_join_destroy(tdp_join *j);

#ifdef __cplusplus
}
#endif
#endif /* TDP_H */
