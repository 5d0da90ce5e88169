// Host-side tile model for the join: objects partitioned into voxels, each
// voxel holding its triangles at the five LODs 20/40/60/80/100.
//
// The reference keeps a tile as one RAW buffer and re-links voxel pointers per
// LOD (src/storage/tile_IO.cpp:13-49, src/spatial/himesh_wrapper.cpp:17-85,
// 137-161). Here the same content is held as flat arrays laid out for upload:
// one 48-byte record per triangle {v0.xyz, proxy, v1.xyz, hausdorff, v2.xyz, 0},
// grouped object -> LOD -> voxel, and per (voxel, LOD) an index and a count.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace tdb {

constexpr int kLods = 5;                       // LOD slots 20,40,60,80,100
constexpr int kLodValue[kLods] = {20, 40, 60, 80, 100};
inline int lod_slot(int lod) {
    for (int s = 0; s < kLods; s++)
        if (kLodValue[s] == lod) return s;
    return -1;  // any other value decodes to empty voxels
}

struct Tile {
    size_t n_obj = 0;
    std::vector<uint32_t> obj_vox_off;   // n_obj + 1
    std::vector<float> obj_box;          // 6 per object: low xyz, high xyz (union of voxel boxes)
    std::vector<float> obj_transform;    // 3 per object: the max-distance proxy point
    std::vector<float> obj_hd;           // 10 per object: (hausdorff, proxy) per LOD slot
    std::vector<float> vox_box;          // 9 per voxel: low xyz, high xyz, core xyz
    std::vector<uint32_t> vox_obj;       // owning object of each voxel
    std::vector<uint64_t> vox_tri_off;   // 5 per voxel: first triangle record
    std::vector<uint32_t> vox_tri_cnt;   // 5 per voxel: number of triangles
    std::vector<float> tri;              // 12 floats per triangle record

    size_t n_vox() const { return vox_obj.size(); }
    size_t n_tri() const { return tri.size() / 12; }
    void finalize_boxes();               // obj_box from voxel boxes
};

// RAW .dt format (type byte 1), byte-compatible with the reference's
// dump_raw_stream (src/storage/tile_IO.cpp:79-172).
bool read_raw(const std::string &path, Tile &t, std::string &err);
bool write_raw(const std::string &path, const Tile &t, std::string &err);

struct GenParams {
    uint64_t seed = 0;
    int n_vessels = 10;
    int nuclei_per_vessel = 20;     // with n_vessels == 0: total nuclei in the cube
    double cube = 2000.0;
    double vessel_r_min = 15, vessel_r_max = 25;
    int vessel_segments = 6;
    double segment_length = 100;
    double turn_max_deg = 40;
    int vessel_sides = 32;
    int rings_min = 12, rings_max = 12;
    double nucleus_r_min = 8, nucleus_r_max = 12;
    double noise = 0.15;
    int subdiv = 3;
    double place_min = 30, place_max = 300;
    int prototypes = 256;            // nucleus shape bank
    int voxel_tris = 50;             // triangles per voxel at LOD 100
    int threads = 0;                 // 0 = all
    int64_t max_nuclei = 0;          // > 0: only the first max_nuclei nuclei
    uint64_t nuclei_seed = 0;        // != 0: nucleus placements drawn from this seed (the
                                     // vessels and the shape bank still follow `seed`)
    int emit_vessels = 1;            // 0: leave the vessel tile empty (placements only)
};

// Nuclei around vessels (n_vessels > 0) or nuclei uniform in the cube.
void generate(const GenParams &p, Tile &nuclei, Tile &vessels);

// One object's arrays, as a tile holds them: 9 floats per voxel box, 5 counts
// per voxel ([voxel][slot]), 10 Hausdorff values, and the triangle records in
// slot -> voxel order.
struct ObjArrays {
    float *vox_box;
    uint32_t *cnt;
    float *hd;
    float *rec;
};

// Objects produced on demand -- the streaming join's "decode" step: any
// object's arrays, and its box and max-distance proxy point, without the
// rest of the tile in memory. Calls may run concurrently.
class ObjectSource {
  public:
    virtual ~ObjectSource() = default;
    virtual size_t n_obj() const = 0;
    virtual uint32_t n_vox(size_t o) const = 0;
    virtual uint64_t n_tri(size_t o) const = 0;
    virtual void meta(size_t o, float box[6], float transform[3]) const = 0;
    virtual void decode(size_t o, const ObjArrays &out) const = 0;
    // decode without the records (out.rec unused): voxel boxes, counts and
    // Hausdorff values, as decode writes them; with `state` (state_floats(o)
    // floats) it also leaves what later decode_slot calls reuse
    virtual void decode_meta(size_t o, const ObjArrays &out, float *state = nullptr) const;
    // the records of LOD slot `slot` alone, voxel after voxel, as decode
    // writes that slot's part of out.rec (state: from decode_meta, or null)
    virtual void decode_slot(size_t o, int slot, float *rec, const float *state = nullptr) const;
    // the records of voxel v (of the object's) at LOD slot `slot` alone, as
    // decode_slot writes that voxel's part
    virtual void decode_slot_voxel(size_t o, int slot, uint32_t v, float *rec,
                                   const float *state = nullptr) const;
    virtual size_t state_floats(size_t o) const { return 0; }
};

// the generator's nuclei (equal, object for object, to generate()'s nuclei tile)
std::unique_ptr<ObjectSource> nucleus_source(const GenParams &p);
// the objects of an in-memory tile (which must outlive the source)
std::unique_ptr<ObjectSource> tile_source(const Tile &t);

}  // namespace tdb
