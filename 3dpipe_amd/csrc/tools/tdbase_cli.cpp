// `tdbase` command: the `join` operator of the reference (src/tools/tdbase.cpp:
// 482-549, options of src/include/query_context.h:141-192) on the MI355X path,
// plus `generate` to write seeded synthetic RAW tiles.
//
// Option handling follows the reference's parser: exact long names, a short
// option takes the rest of its argument or the next one, a value option may be
// repeated (the LOD list is built from repeated -l/--lods), and unknown
// options and stray arguments are collected and ignored.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../../include/tdbase.h"
#include "../../../include/tdp.h"

namespace {

struct Spec {
    char shortname;
    const char *longname;
    bool takes_value;
};

struct Parsed {
    std::map<std::string, std::vector<std::string>> values;  // by long name
    std::vector<std::string> unknown, stray;
    bool has(const char *n) const { return values.count(n) > 0; }
    std::string str(const char *n, const std::string &def) const {
        auto it = values.find(n);
        return it == values.end() || it->second.empty() ? def : it->second.back();
    }
};

Parsed parse(int argc, char **argv, const std::vector<Spec> &specs) {
    Parsed p;
    auto by_long = [&](const std::string &n) -> const Spec * {
        for (auto &s : specs)
            if (n == s.longname) return &s;
        return nullptr;
    };
    auto by_short = [&](char c) -> const Spec * {
        for (auto &s : specs)
            if (s.shortname && s.shortname == c) return &s;
        return nullptr;
    };
    for (int n = 1; n < argc; n++) {
        std::string a = argv[n];
        if (a == "--") {
            for (int m = n + 1; m < argc; m++) p.stray.push_back(argv[m]);
            break;
        }
        if (a.rfind("--", 0) == 0) {
            std::string opt = a.substr(2), arg;
            size_t eq = opt.find('=');
            if (eq != std::string::npos) {
                arg = opt.substr(eq + 1);
                opt.resize(eq);
            }
            const Spec *s = by_long(opt);
            if (s && !s->takes_value && !arg.empty()) s = nullptr;
            if (s && s->takes_value && arg.empty() && n < argc - 1) arg = argv[++n];
            if (s) p.values[s->longname].push_back(s->takes_value ? arg : "1");
            else p.unknown.push_back(a);
        } else if (a.rfind("-", 0) == 0 && a.size() > 1) {
            std::string opt = a.substr(1);
            bool unknown = false;
            for (size_t m = 0; m < opt.size(); m++) {
                const Spec *s = by_short(opt[m]);
                if (!s) { unknown = true; continue; }
                if (s->takes_value) {
                    std::string arg = opt.substr(m + 1);
                    if (arg.empty() && n < argc - 1) arg = argv[++n];
                    p.values[s->longname].push_back(arg);
                    break;
                }
                p.values[s->longname].push_back("1");
            }
            if (unknown) p.unknown.push_back(a);
        } else {
            p.stray.push_back(a);
        }
    }
    return p;
}

int to_int(const std::string &s, const char *what) {
    char *end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (s.empty() || *end) {
        std::fprintf(stderr, "invalid argument for %s: '%s'\n", what, s.c_str());
        std::exit(1);
    }
    return (int)v;
}
double to_double(const std::string &s, const char *what) {
    char *end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (s.empty() || *end) {
        std::fprintf(stderr, "invalid argument for %s: '%s'\n", what, s.c_str());
        std::exit(1);
    }
    return v;
}

int cmd_join(int argc, char **argv) {
    const std::vector<Spec> specs = {
        {'h', "help", false},        {'t', "threads", true},      {0, "cn", true},
        {'g', "gpu", false},         {'v', "verbose", true},      {0, "print_result", false},
        {0, "aabb", false},          {'c', "counter_clock", false},
        {0, "disable_byte_encoding", false},                     {0, "hausdorff_level", true},
        {'l', "lods", true},         {0, "tile1", true},          {0, "tile2", true},
        {0, "max_objects1", true},   {0, "max_objects2", true},   {'q', "query", true},
        {'k', "knn", true},          {'w', "within_dist", true},
        // options beyond the reference tool's
        {0, "precision", true},      {0, "device", true},         {0, "print_distance", false},
    };
    Parsed p = parse(argc, argv, specs);
    if (p.has("help")) {
        std::printf(
            "tdbase join --tile1 PATH [--tile2 PATH] -q nn|within|intersect [-k K] [-w TAU]\n"
            "            [-l LOD]... [--hausdorff_level 0|1|2] [--max_objects1 N]\n"
            "            [--print_result] [--print_distance] [--precision ref32|fp64|mixed]\n"
            "            [--device D] [-t THREADS]\n");
        return 0;
    }
    if (!p.has("tile1") || !p.has("query")) {
        std::fprintf(stderr, "Invalid Option Exception: option \"%s\" is required\n",
                     !p.has("tile1") ? "tile1" : "query");
        return 1;
    }
    const std::string q = p.str("query", "");
    if (q != "nn" && q != "within" && q != "intersect") {
        std::printf("error query type: %s\n", q.c_str());
        return 0;
    }
    tdb_join_opts o;
    tdb_join_defaults(&o);
    o.query = q == "nn" ? TDB_QUERY_NN : q == "within" ? TDB_QUERY_WITHIN : TDB_QUERY_INTERSECT;
    o.k = to_int(p.str("knn", "1"), "--knn");
    o.within_dist = (float)to_double(p.str("within_dist", "1000"), "--within_dist");
    o.hausdorff_level = to_int(p.str("hausdorff_level", "2"), "--hausdorff_level");
    if (o.hausdorff_level < 0 || o.hausdorff_level > 2) {
        std::fprintf(stderr, "--hausdorff_level must be 0, 1 or 2\n");
        return 1;
    }
    if (p.has("max_objects1")) o.max_objects1 = (uint64_t)to_double(p.str("max_objects1", "0"), "--max_objects1");
    if (p.has("threads")) o.threads = to_int(p.str("threads", "0"), "--threads");
    const std::string prec = p.str("precision", "mixed");
    if (prec == "ref32") o.precision = TDP_PREC_REF32;
    else if (prec == "fp64") o.precision = TDP_PREC_FP64;
    else if (prec == "mixed") o.precision = TDP_PREC_MIXED;
    else {
        std::fprintf(stderr, "--precision must be ref32, fp64 or mixed\n");
        return 1;
    }
    std::vector<int> lods;
    if (p.has("lods"))
        for (auto &s : p.values["lods"]) lods.push_back(to_int(s, "--lods"));
    o.lods = lods.empty() ? nullptr : lods.data();
    o.n_lods = (int)lods.size();
    for (auto &u : p.unknown) std::fprintf(stderr, "ignoring unknown option %s\n", u.c_str());

    char err[512] = {0};
    tdb_tile *t1 = nullptr, *t2 = nullptr;
    if (tdb_tile_load(p.str("tile1", "").c_str(), &t1, err, sizeof err) != 0) {
        std::fprintf(stderr, "%s\n", err);
        return 1;
    }
    if (p.has("tile2") && !p.str("tile2", "").empty()) {
        if (tdb_tile_load(p.str("tile2", "").c_str(), &t2, err, sizeof err) != 0) {
            std::fprintf(stderr, "%s\n", err);
            tdb_tile_free(t1);
            return 1;
        }
    }
    tdb_engine *e = nullptr;
    if (tdb_engine_open(to_int(p.str("device", "0"), "--device"), &e) != 0) {
        std::fprintf(stderr, "GPU unavailable: %s\n", tdb_engine_error(e));
        tdb_engine_close(e);
        tdb_tile_free(t1);
        tdb_tile_free(t2);
        return 2;
    }
    tdb_result *r = nullptr;
    int rc = tdb_join(e, t1, t2, &o, &r);
    if (rc != 0) {
        std::fprintf(stderr, "join failed: %s\n", tdb_engine_error(e));
    } else {
        const uint64_t n = tdb_result_count(r);
        std::vector<int32_t> a(n), b(n);
        std::vector<double> mn(n), mx(n);
        std::vector<uint8_t> rnd(n);
        tdb_result_get(r, a.data(), b.data(), mn.data(), mx.data(), rnd.data());
        // --print_result prints what the reference prints: every report made
        // through report_result (himesh_wrapper.cpp:195-202). The within join
        // appends its refinement-round results without it
        // (WithinJoin.cpp:779), so for -q within only the filter's reports
        // (round 0) are printed; --print_distance prints every result with
        // its distance range.
        const bool dist = p.has("print_distance");
        const bool refine_silent = o.query == TDB_QUERY_WITHIN;
        if (p.has("print_result") || dist)
            for (uint64_t i = 0; i < n; i++) {
                if (dist) std::printf("%d %d %.9g %.9g\n", a[i], b[i], mn[i], mx[i]);
                else if (!refine_silent || rnd[i] == 0) std::printf("%d %d\n", a[i], b[i]);
            }
        double tm[7];
        uint64_t st[7];
        tdb_result_timing(r, tm);
        tdb_result_stats(r, st);
        std::fprintf(stderr,
                     "total:\t%.3f ms\ncandidates:\t%.3f ms\nfilter:\t%.3f ms\nrefine:\t%.3f ms\n"
                     "filter kernels:\t%.3f ms\nrefine kernels:\t%.3f ms\n#objects:\t%llu\n"
                     "results:\t%llu\ntriangle pairs:\t%llu\nvoxel pairs:\t%llu\n",
                     tm[0], tm[1], tm[2], tm[3], tm[4], tm[5],
                     (unsigned long long)tdb_result_targets(r), (unsigned long long)n,
                     (unsigned long long)st[0], (unsigned long long)st[1]);
        tdb_result_free(r);
    }
    tdb_engine_close(e);
    tdb_tile_free(t1);
    tdb_tile_free(t2);
    return rc == 0 ? 0 : 1;
}

int cmd_generate(int argc, char **argv) {
    const std::vector<Spec> specs = {
        {0, "out_nuclei", true},  {0, "out_vessels", true}, {0, "seed", true},
        {0, "nv", true},          {0, "nu", true},          {0, "cube", true},
        {0, "rings_min", true},   {0, "rings_max", true},   {0, "nucleus_r_min", true},
        {0, "nucleus_r_max", true}, {0, "prototypes", true}, {0, "threads", true},
    };
    Parsed p = parse(argc, argv, specs);
    if (!p.has("out_nuclei")) {
        std::fprintf(stderr, "tdbase generate --out_nuclei PATH [--out_vessels PATH] [--seed S] "
                             "[--nv VESSELS] [--nu NUCLEI_PER_VESSEL] [--cube SIDE] ...\n");
        return 1;
    }
    tdb_gen_params g;
    tdb_gen_defaults(&g);
    g.seed = (uint64_t)to_double(p.str("seed", "0"), "--seed");
    g.n_vessels = to_int(p.str("nv", "10"), "--nv");
    g.nuclei_per_vessel = to_int(p.str("nu", "20"), "--nu");
    g.cube = to_double(p.str("cube", "2000"), "--cube");
    g.rings_min = to_int(p.str("rings_min", "12"), "--rings_min");
    g.rings_max = to_int(p.str("rings_max", std::to_string(g.rings_min)), "--rings_max");
    g.nucleus_r_min = to_double(p.str("nucleus_r_min", "8"), "--nucleus_r_min");
    g.nucleus_r_max = to_double(p.str("nucleus_r_max", "12"), "--nucleus_r_max");
    g.prototypes = to_int(p.str("prototypes", "256"), "--prototypes");
    g.threads = to_int(p.str("threads", "0"), "--threads");
    tdb_tile *n = nullptr, *v = nullptr;
    if (tdb_generate(&g, &n, &v) != 0) {
        std::fprintf(stderr, "invalid generator parameters\n");
        return 1;
    }
    char err[512] = {0};
    int rc = 0;
    if (tdb_tile_save(n, p.str("out_nuclei", "").c_str(), err, sizeof err) != 0) rc = 1;
    if (rc == 0 && p.has("out_vessels") &&
        tdb_tile_save(v, p.str("out_vessels", "").c_str(), err, sizeof err) != 0)
        rc = 1;
    if (rc) std::fprintf(stderr, "%s\n", err);
    tdb_tile_free(n);
    tdb_tile_free(v);
    return rc;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2 || (std::strcmp(argv[1], "join") && std::strcmp(argv[1], "generate"))) {
        std::printf("usage: tdbase function [args]\nfunction could be:\ngenerate join\n");
        return 0;
    }
    if (!std::strcmp(argv[1], "join")) return cmd_join(argc - 1, argv + 1);
    return cmd_generate(argc - 1, argv + 1);
}
