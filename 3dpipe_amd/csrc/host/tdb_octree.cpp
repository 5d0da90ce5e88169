#include "tdb_octree.hpp"

#include <algorithm>

#include <cmath>
#include <functional>
#include <limits>
#include <queue>

namespace tdb {

// Deepest level a node splits at: more halvings than a float box can take
// (its exponent range and mantissa), so the cap changes no tree; it bounds the
// recursion where more than leaf_size objects share one point.
constexpr int kMaxLevel = 320;

struct Octree::Node {
    float box[6];
    uint32_t size = 0;
    long leaf_size;
    int level;
    bool leaf = true, can_split = true;
    Node *child[8] = {nullptr};
    std::vector<int> objs;
    Node(const float *b, int lvl, long ls) : leaf_size(ls), level(lvl) {
        for (int i = 0; i < 6; i++) box[i] = b[i];
    }
    ~Node() {
        if (!leaf)
            for (Node *c : child) delete c;
    }
};

Range box_distance(const float *a, const float *b) {
    float mn = 0, mx = 0;
    for (int i = 0; i < 3; i++) {
        const float t1 = a[i] - b[3 + i];
        const float t2 = a[3 + i] - b[i];
        mx += (t1 + t2) * (t1 + t2) / 4;
        if (t2 < 0) mn += t2 * t2;
        else if (t1 > 0) mn += t1 * t1;
    }
    return Range{std::sqrt(mn), std::sqrt(mx)};
}

// strict overlap (aab::intersect, himesh_box.cpp:111-115)
static bool overlaps(const float *node, const float *obj) {
    return !(obj[0] >= node[3] || obj[3] <= node[0] || obj[1] >= node[4] || obj[4] <= node[1] ||
             obj[2] >= node[5] || obj[5] <= node[2]);
}

Octree::Octree(const float *boxes, const float *transform, size_t n, long leaf_size)
    : boxes_(boxes), transform_(transform) {
    float space[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            if (space[d] > boxes[6 * i + d]) space[d] = boxes[6 * i + d];
            if (space[3 + d] < boxes[6 * i + 3 + d]) space[3 + d] = boxes[6 * i + 3 + d];
        }
    root_ = new Node(space, 0, leaf_size);
    std::vector<int> seq(n);
    for (size_t i = 0; i < n; i++) seq[i] = (int)i;
    // subtrees side by side below the first two levels (64 of them) where the
    // tile is large; small tiles in turn
    const int par = n >= 50000 ? 2 : 0;
#ifdef _OPENMP
#pragma omp parallel if (par > 0)
#pragma omp single
#endif
    build(root_, seq, 0, par);
}

// Adds seq[from..] to n in order, as add() one by one would. A node's subtree
// depends only on the objects it receives, in their order: once n has split,
// every later object goes to the children its box overlaps, so each child's
// part of the rest is filtered out of seq in order and the children are built
// independently (side by side for par_depth more levels).
void Octree::build(Node *n, const std::vector<int> &seq, size_t from, int par_depth) {
    if (par_depth <= 0) {
        for (size_t j = from; j < seq.size(); j++) add(n, seq[j]);
        return;
    }
    size_t i = from;
    while (i < seq.size() && n->leaf) add(n, seq[i++]);
    if (i == seq.size()) return;
    n->size += (uint32_t)(seq.size() - i);
    for (int c = 0; c < 8; c++) {
        auto part = [this, n, c, i, &seq, par_depth]() {
            Node *ch = n->child[c];
            std::vector<int> sub;
            for (size_t j = i; j < seq.size(); j++)
                if (overlaps(ch->box, boxes_ + 6 * seq[j])) sub.push_back(seq[j]);
            build(ch, sub, 0, par_depth - 1);
        };
#ifdef _OPENMP
#pragma omp task default(none) firstprivate(part)
#endif
        part();
    }
#ifdef _OPENMP
#pragma omp taskwait
#endif
}

Octree::~Octree() { delete root_; }

void Octree::add(Node *n, int id) {
    n->size += 1;
    if (n->size == 1) {  // first object of a new node
        n->objs.push_back(id);
        return;
    }
    if (n->leaf) {
        n->objs.push_back(id);
        if ((long)n->size > n->leaf_size && n->can_split && n->level < kMaxLevel) {
            float mid[3], cb[8][6];
            for (int i = 0; i < 3; i++) mid[i] = (n->box[i] + n->box[3 + i]) / 2;
            for (int c = 0; c < 8; c++)
                for (int i = 0; i < 3; i++) {
                    const bool upper = (c >> i) & 1;
                    cb[c][i] = upper ? mid[i] : n->box[i];
                    cb[c][3 + i] = upper ? n->box[3 + i] : mid[i];
                }
            // a split that duplicates most objects is undone until the node
            // has grown by half again; the count of placements does not
            // depend on the children's own splits, so it is taken first and
            // an undone split builds no children (the reference builds them
            // and deletes them, recursively: exponential work where many
            // objects share a region)
            uint32_t placed = 0;
            for (int o : n->objs)
                for (int c = 0; c < 8; c++) placed += overlaps(cb[c], boxes_ + 6 * o) ? 1u : 0u;
            if (placed >= 2 * (n->size - 1)) {
                n->can_split = false;
            } else {
                for (int c = 0; c < 8; c++) n->child[c] = new Node(cb[c], n->level + 1, n->leaf_size);
                for (int o : n->objs)
                    for (int c = 0; c < 8; c++)
                        if (overlaps(n->child[c]->box, boxes_ + 6 * o)) add(n->child[c], o);
                n->objs.clear();
                n->leaf = false;
            }
        } else if ((double)n->size > 1.5 * (double)n->leaf_size) {
            n->can_split = true;
        }
    } else {
        for (int c = 0; c < 8; c++)
            if (overlaps(n->child[c]->box, boxes_ + 6 * id)) add(n->child[c], id);
    }
}

static float point_dist(const float *a, const float *b) {
    return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                     (a[2] - b[2]) * (a[2] - b[2]));
}

// sorted insertion by max distance (ties after existing entries), dedupe by
// id, then drop entries that can no longer be among the k nearest
// (add_candidate, octree.cpp:120-154). Returns whether the new entry landed
// among the first k.
static bool add_candidate(std::vector<std::pair<int, Range>> &c, int id, Range r, int k) {
    if ((int)c.size() < k || r.mindist < c[k - 1].second.maxdist) {
        size_t loc = 0;
        for (; loc < c.size(); loc++) {
            if (c[loc].first == id) return false;
            if (c[loc].second.maxdist > r.maxdist) {
                c.insert(c.begin() + loc, {id, r});
                break;
            }
        }
        if (loc == c.size()) c.push_back({id, r});
        const int from = std::max((int)loc, k - 1);
        for (int i = from + 1; i < (int)c.size();) {
            if (c[i].second.mindist >= c[k - 1].second.maxdist) {
                c.erase(c.begin() + i);
                continue;
            }
            i++;
        }
        return (int)loc < k;
    }
    return false;
}

void Octree::knn(const float *qbox, const float *qt, int self, int k,
                 std::vector<std::pair<int, Range>> &out) const {
    struct Entry {
        const Node *node;
        float mindist;
        bool operator>(const Entry &o) const { return mindist > o.mindist; }
    };
    std::priority_queue<Entry, std::vector<Entry>, std::greater<Entry>> heap;
    heap.push({root_, box_distance(root_->box, qbox).mindist});
    float upper = std::numeric_limits<float>::infinity();
    while (!heap.empty()) {
        const Entry e = heap.top();
        heap.pop();
        if (e.mindist > upper) break;
        const Node *n = e.node;
        if (n->leaf) {
            for (int o : n->objs) {
                if (o == self) continue;
                Range r = box_distance(boxes_ + 6 * o, qbox);
                r.maxdist = point_dist(qt, transform_ + 3 * o);
                if (add_candidate(out, o, r, k) && (int)out.size() >= k)
                    upper = out[k - 1].second.maxdist;
            }
        } else {
            for (const Node *c : n->child) {
                const Range r = box_distance(c->box, qbox);
                if (r.mindist <= upper) heap.push({c, r.mindist});
            }
        }
    }
}

void Octree::within_rec(const Node *n, const float *qbox, const float *qt, int self, float tau,
                        std::vector<std::pair<int, Range>> &out) const {
    if (!(box_distance(n->box, qbox).mindist <= tau)) return;
    if (n->leaf) {
        for (int o : n->objs) {
            if (o == self) continue;
            Range r = box_distance(boxes_ + 6 * o, qbox);
            r.maxdist = point_dist(qt, transform_ + 3 * o);
            if (r.mindist <= tau) {
                bool exist = false;
                for (auto &c : out)
                    if (c.first == o) { exist = true; break; }
                if (!exist) out.push_back({o, r});
            }
        }
    } else {
        for (const Node *c : n->child) within_rec(c, qbox, qt, self, tau, out);
    }
}

void Octree::within(const float *qbox, const float *qt, int self, float tau,
                    std::vector<std::pair<int, Range>> &out) const {
    within_rec(root_, qbox, qt, self, tau, out);
}

void Octree::leaves_rec(const Node *n, std::vector<float> &box, std::vector<uint32_t> &off,
                        std::vector<uint32_t> &objs) const {
    if (n->leaf) {
        box.insert(box.end(), n->box, n->box + 6);
        for (int o : n->objs) objs.push_back((uint32_t)o);
        off.push_back((uint32_t)objs.size());
    } else {
        for (const Node *c : n->child) leaves_rec(c, box, off, objs);
    }
}

void Octree::leaves(std::vector<float> &box, std::vector<uint32_t> &off,
                    std::vector<uint32_t> &objs) const {
    box.clear();
    objs.clear();
    off.assign(1, 0u);
    if (root_) leaves_rec(root_, box, off, objs);
}

bool box_overlap(const float *a, const float *b) {
    return !(b[0] >= a[3] || b[3] <= a[0] || b[1] >= a[4] || b[4] <= a[1] || b[2] >= a[5] ||
             b[5] <= a[2]);
}

void Octree::intersect_rec(const Node *n, const float *qbox, int self, std::vector<int> &out) const {
    if (!box_overlap(n->box, qbox)) return;
    if (n->leaf) {
        for (int o : n->objs)
            if (o != self && box_overlap(boxes_ + 6 * o, qbox)) out.push_back(o);
    } else {
        for (const Node *c : n->child) intersect_rec(c, qbox, self, out);
    }
}

void Octree::intersect(const float *qbox, int self, std::vector<int> &out) const {
    out.clear();
    if (root_) intersect_rec(root_, qbox, self, out);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

}  // namespace tdb
