// Octree over object boxes for candidate generation (host side).
//
// Same structure and query semantics as the reference's OctreeNode
// (src/index/octree.cpp:10-365, src/include/index.h:28-57): leaf capacity 10,
// split into 8 equal children with an undo when splitting does not help,
// best-first kNN (query_knn_fast) over a std::priority_queue ordered by the
// node's box distance, and a depth-first within query. The candidate order is
// part of the join result, so tie handling follows the reference exactly.
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace tdb {

struct Range {
    float mindist, maxdist;
};

class Octree {
public:
    // boxes: 6 floats per object {low.xyz, high.xyz}; transform: 3 per object
    Octree(const float *boxes, const float *transform, size_t n, long leaf_size = 10);
    ~Octree();
    Octree(const Octree &) = delete;
    Octree &operator=(const Octree &) = delete;

    // k nearest by the transform-point distance, pruned by box distance;
    // `self` (an object index, or -1) is skipped as the reference skips the
    // query box itself in a self-join
    void knn(const float *qbox, const float *qtransform, int self, int k,
             std::vector<std::pair<int, Range>> &out) const;
    // all objects whose box distance is <= tau, in depth-first order
    void within(const float *qbox, const float *qtransform, int self, float tau,
                std::vector<std::pair<int, Range>> &out) const;
    // all objects whose box overlaps qbox with positive volume, ascending
    // (query_intersect, octree.cpp:332-349, then the sort and de-duplication
    // of mbb_intersect, IntersectionJoin.cpp:20-33)
    void intersect(const float *qbox, int self, std::vector<int> &out) const;
    // the leaves in the depth-first order of within(): boxes (6 floats
    // each), offsets into objs (n_leaves + 1) and each leaf's objects in order
    void leaves(std::vector<float> &box, std::vector<uint32_t> &off,
                std::vector<uint32_t> &objs) const;

private:
    struct Node;
    Node *root_ = nullptr;
    const float *boxes_;
    const float *transform_;
    void add(Node *n, int id);
    void build(Node *n, const std::vector<int> &seq, size_t from, int par_depth);
    void within_rec(const Node *n, const float *qbox, const float *qt, int self, float tau,
                    std::vector<std::pair<int, Range>> &out) const;
    void intersect_rec(const Node *n, const float *qbox, int self, std::vector<int> &out) const;
    void leaves_rec(const Node *n, std::vector<float> &box, std::vector<uint32_t> &off,
                    std::vector<uint32_t> &objs) const;
};

// aab::distance (src/spatial/himesh_box.cpp:149-168): a is the box the
// method is called on, b its argument
Range box_distance(const float *a, const float *b);
// aab::intersect (src/spatial/himesh_box.cpp:95-99): overlap with positive volume
bool box_overlap(const float *a, const float *b);

}  // namespace tdb
