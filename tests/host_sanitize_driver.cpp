// Host code of the join engine under AddressSanitizer and UndefinedBehavior-
// Sanitizer (built by tests/test_cpu.py with g++ -fsanitize=address,undefined
// from the engine's own sources; no device code). It exercises what runs
// without a GPU: the seeded generator (OpenMP), the RAW writer and reader --
// on the intact file, on truncations of a small one at every length of its
// first 4 KiB and sampled beyond, and on files with
// random bytes overwritten, which must all be refused cleanly or read -- the
// object source's decode, and the octree's kNN / within / intersect queries
// and leaf listing. Exits non-zero on a failed check; the sanitizers abort on
// a memory error.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../3dpipe_amd/csrc/host/tdb_octree.hpp"
#include "../3dpipe_amd/csrc/host/tdb_tile.hpp"

static int failures = 0;
#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            failures++;                                             \
        }                                                           \
    } while (0)

static std::vector<char> slurp(const std::string &p) {
    std::vector<char> b;
    FILE *f = std::fopen(p.c_str(), "rb");
    if (!f) return b;
    std::fseek(f, 0, SEEK_END);
    b.resize((size_t)std::ftell(f));
    std::fseek(f, 0, SEEK_SET);
    if (!b.empty() && std::fread(b.data(), 1, b.size(), f) != b.size()) b.clear();
    std::fclose(f);
    return b;
}

static void spit(const std::string &p, const char *d, size_t n) {
    FILE *f = std::fopen(p.c_str(), "wb");
    if (n) std::fwrite(d, 1, n, f);
    std::fclose(f);
}

static bool same(const tdb::Tile &a, const tdb::Tile &b) {
    return a.n_obj == b.n_obj && a.obj_vox_off == b.obj_vox_off && a.obj_box == b.obj_box &&
           a.obj_transform == b.obj_transform && a.obj_hd == b.obj_hd && a.vox_box == b.vox_box &&
           a.vox_tri_off == b.vox_tri_off && a.vox_tri_cnt == b.vox_tri_cnt && a.tri == b.tri;
}

int main(int argc, char **argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    tdb::GenParams g;
    g.seed = 7;
    g.n_vessels = 3;
    g.nuclei_per_vessel = 4;
    g.cube = 600;
    g.subdiv = 2;
    g.vessel_sides = 8;
    g.rings_min = 2;
    g.rings_max = 3;
    g.prototypes = 4;
    g.voxel_tris = 20;
    tdb::Tile n, v;
    tdb::generate(g, n, v);
    CHECK(n.n_obj == 12 && v.n_obj == 3);

    // write, read back, identical
    std::string err;
    const std::string pn = dir + "/n.dt", pv = dir + "/v.dt", pc = dir + "/c.dt";
    CHECK(tdb::write_raw(pn, n, err));
    CHECK(tdb::write_raw(pv, v, err));
    tdb::Tile n2;
    CHECK(tdb::read_raw(pn, n2, err));
    CHECK(same(n, n2));

    // every truncation of the vessel file: refused (or, at an object
    // boundary, read as fewer objects), never a bad access
    const std::vector<char> bytes = slurp(pv);
    CHECK(!bytes.empty());
    size_t refused = 0, read_ok = 0;
    // every length through the first object's meta, then a sample of the rest
    for (size_t len = 0; len < bytes.size(); len += len < 4096 ? 1 : 61) {
        spit(pc, bytes.data(), len);
        tdb::Tile t;
        if (tdb::read_raw(pc, t, err)) read_ok++;
        else refused++;
        if (len == 1) CHECK(t.n_obj == 0);  // the type byte alone: no objects
    }
    CHECK(refused > 4000);
    // random bytes overwritten (sizes, counts, offsets, data)
    std::mt19937_64 rng(3);
    for (int trial = 0; trial < 1500; trial++) {
        std::vector<char> b = bytes;
        const int k = 1 + (int)(rng() % 8);
        for (int i = 0; i < k; i++) b[1 + rng() % (b.size() - 1)] = (char)(rng() & 0xff);
        spit(pc, b.data(), b.size());
        tdb::Tile t;
        if (tdb::read_raw(pc, t, err)) {
            // what was read is self-consistent
            CHECK(t.obj_vox_off.size() == t.n_obj + 1);
            CHECK(t.vox_tri_off.size() == 5 * t.n_vox());
            for (size_t q = 0; q < 5 * t.n_vox(); q++)
                CHECK(t.vox_tri_off[q] + t.vox_tri_cnt[q] <= t.n_tri());
        }
    }

    // the object source decodes the generator's objects; a decode without
    // records plus one decode per LOD slot gives the whole decode exactly
    // (the generator's source, a tile's, and the generic forms)
    auto src = tdb::nucleus_source(g);
    CHECK(src->n_obj() == n.n_obj);
    auto ts = tdb::tile_source(n);
    struct WholeOnly final : tdb::ObjectSource {  // only decode(): the generic forms
        const tdb::ObjectSource &s;
        explicit WholeOnly(const tdb::ObjectSource &x) : s(x) {}
        size_t n_obj() const override { return s.n_obj(); }
        uint32_t n_vox(size_t o) const override { return s.n_vox(o); }
        uint64_t n_tri(size_t o) const override { return s.n_tri(o); }
        void meta(size_t o, float b[6], float t[3]) const override { s.meta(o, b, t); }
        void decode(size_t o, const tdb::ObjArrays &a) const override { s.decode(o, a); }
    } whole(*src);
    const tdb::ObjectSource *sources[3] = {src.get(), ts.get(), &whole};
    for (const tdb::ObjectSource *S : sources)
        for (size_t o = 0; o < S->n_obj(); o++) {
            const uint32_t nv = S->n_vox(o);
            const uint64_t nt = S->n_tri(o);
            std::vector<float> box(9 * nv), hd(10), rec(12 * nt);
            std::vector<uint32_t> cnt(5 * nv);
            tdb::ObjArrays a{box.data(), cnt.data(), hd.data(), rec.data()};
            S->decode(o, a);
            float b6[6], tr[3];
            S->meta(o, b6, tr);
            for (int i = 0; i < 6; i++) CHECK(b6[i] == n.obj_box[6 * o + i]);
            std::vector<float> box2(9 * nv), hd2(10), rec2(12 * nt, -1.f);
            std::vector<uint32_t> cnt2(5 * nv);
            S->decode_meta(o, tdb::ObjArrays{box2.data(), cnt2.data(), hd2.data(), nullptr});
            CHECK(box2 == box && cnt2 == cnt && hd2 == hd);
            size_t at = 0;
            for (int s = 0; s < tdb::kLods; s++) {
                S->decode_slot(o, s, rec2.data() + 12 * at);
                for (uint32_t v = 0; v < nv; v++) at += cnt[5 * v + s];
            }
            CHECK(at == nt && rec2 == rec);
            // the same with the decode state kept from the meta decode
            std::vector<float> state(S->state_floats(o)), rec3(12 * nt, -1.f);
            S->decode_meta(o, tdb::ObjArrays{box2.data(), cnt2.data(), hd2.data(), nullptr},
                           state.empty() ? nullptr : state.data());
            CHECK(box2 == box && cnt2 == cnt && hd2 == hd);
            at = 0;
            for (int s = 0; s < tdb::kLods; s++) {
                S->decode_slot(o, s, rec3.data() + 12 * at, state.empty() ? nullptr : state.data());
                for (uint32_t v = 0; v < nv; v++) at += cnt[5 * v + s];
            }
            CHECK(rec3 == rec);
        }

    // octree queries over the nuclei (self-join) and the vessels
    for (const tdb::Tile *t : {&n, &v}) {
        tdb::Octree tree(t->obj_box.data(), t->obj_transform.data(), t->n_obj, 10);
        std::vector<float> lb;
        std::vector<uint32_t> lo, lobj;
        tree.leaves(lb, lo, lobj);
        CHECK(lo.size() * 6 == lb.size() + 6);
        for (size_t i = 0; i < n.n_obj; i++) {
            std::vector<std::pair<int, tdb::Range>> c;
            tree.knn(&n.obj_box[6 * i], &n.obj_transform[3 * i], t == &n ? (int)i : -1, 3, c);
            for (auto &e : c) CHECK(e.first >= 0 && (size_t)e.first < t->n_obj);
            tree.within(&n.obj_box[6 * i], &n.obj_transform[3 * i], -1, 150.f, c);
            std::vector<int> ids;
            tree.intersect(&n.obj_box[6 * i], -1, ids);
        }
    }
    // more objects than a leaf holds with one box: the split depth is capped
    // (the reference's recursion would not end)
    {
        const int m = 40;
        std::vector<float> box(6 * m), tr(3 * m);
        for (int i = 0; i < m; i++) {
            const float b[6] = {10, 10, 10, 15, 15, 15}, p[3] = {12, 12, 12};
            std::memcpy(&box[6 * i], b, sizeof b);
            std::memcpy(&tr[3 * i], p, sizeof p);
        }
        tdb::Octree tree(box.data(), tr.data(), m, 10);
        std::vector<std::pair<int, tdb::Range>> c;
        tree.knn(&box[0], &tr[0], 0, 5, c);
        CHECK(c.size() >= 5);
        tree.within(&box[0], &tr[0], 0, 1.f, c);
        CHECK(c.size() == (size_t)(m - 1));
    }
    // an octree without objects answers every query with nothing
    tdb::Octree none(nullptr, nullptr, 0, 10);
    std::vector<std::pair<int, tdb::Range>> c;
    none.knn(&n.obj_box[0], &n.obj_transform[0], -1, 3, c);
    CHECK(c.empty());

    std::printf("host sanitize driver: %d failures, %zu truncations refused, %zu read\n",
                failures, refused, read_ok);
    return failures ? 1 : 0;
}
