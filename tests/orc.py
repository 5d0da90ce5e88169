"""ctypes access to the CPU oracle (oracle/lib/liboracle.so) and to the
reference's own compiled sources (oracle/_ref/libtdref.so). Test
infrastructure: used as the checker only."""
import ctypes as c
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libtdref.so")
P = c.c_void_p
SZ = c.c_size_t


def ptr(a):
    return None if a is None else a.ctypes.data_as(P)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE):
            raise RuntimeError("oracle not built: make -C oracle")
        L = c.CDLL(ORACLE)
        L.orc_tile_load.argtypes = [c.c_char_p]
        L.orc_tile_load.restype = P
        L.orc_tile_free.argtypes = [P]
        L.orc_tile_count.argtypes = [P]
        L.orc_tile_count.restype = SZ
        L.orc_tile_object_box.argtypes = [P, SZ, P]
        L.orc_tridist_f.argtypes = [P, P, SZ, P]
        L.orc_tridist_d.argtypes = [P, P, SZ, P]
        L.orc_tridist_dd.argtypes = [P, P, SZ, P]
        L.orc_meshdist_f.argtypes = [P, P, P, SZ, P]
        L.orc_meshdist_d.argtypes = [P, P, P, SZ, P]
        L.orc_voxel_bound_f.argtypes = [P, P, P]
        L.orc_voxel_bound_d.argtypes = [P, P, P]
        L.orc_aab_distance.argtypes = [P, P, P]
        L.orc_knn_status.argtypes = [P, c.c_int, c.c_int, c.c_int, P]
        L.orc_boxset_build.argtypes = [SZ, P, P, P, c.c_int]
        L.orc_boxset_build.restype = P
        for f in (L.orc_query_knn,):
            f.argtypes = [P, P, P, P, c.c_int, c.c_int, P, P, P, SZ]
            f.restype = SZ
        L.orc_query_within.argtypes = [P, P, P, P, c.c_int, c.c_float, P, P, P, SZ]
        L.orc_query_within.restype = SZ
        L.orc_query_intersect.argtypes = [P, P, P, c.c_int, P, SZ]
        L.orc_query_intersect.restype = SZ
        L.orc_lod_list.argtypes = [P, c.c_int, P]
        L.orc_join.argtypes = [P, P, c.c_int, c.c_int, c.c_float, P, c.c_int, c.c_int, c.c_int,
                               c.c_uint64]
        L.orc_join.restype = P
        L.orc_result_count.argtypes = [P]
        L.orc_result_count.restype = SZ
        L.orc_result_get.argtypes = [P, P, P, P, P]
        L.orc_result_stats.argtypes = [P, P]
        L.orc_result_rounds.argtypes = [P, P]
        L.orc_result_free.argtypes = [P]
        L.orc_max_threads.restype = c.c_int
        _lib = L
    return _lib


_ref = None


def ref():
    """The reference's TriDist / MeshDist / aab / octree, compiled from its
    sources (None when oracle/_ref was not built)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF):
            return None
        L = c.CDLL(REF)
        L.ref_tridist.argtypes = [P, P, SZ, P]
        L.ref_meshdist.argtypes = [P, P, P, SZ, P]
        L.ref_aab_distance.argtypes = [P, P, P]
        L.ref_boxset_build.argtypes = [SZ, P, P, P, c.c_int]
        L.ref_boxset_build.restype = P
        L.ref_boxset_free.argtypes = [P]
        L.ref_query_knn.argtypes = [P, P, P, P, c.c_int, c.c_int, P, P, P, SZ]
        L.ref_query_knn.restype = SZ
        L.ref_query_within.argtypes = [P, P, P, P, c.c_int, c.c_float, P, P, P, SZ]
        L.ref_query_within.restype = SZ
        L.ref_query_intersect.argtypes = [P, P, P, c.c_int, P, SZ]
        L.ref_query_intersect.restype = SZ
        L.ref_triint.argtypes = [P, P, SZ, P]
        L.ref_meshint.argtypes = [P, P, P, SZ, P]
        _ref = L
    return _ref


def tridist(S, T, precision="f32"):
    S = np.ascontiguousarray(S, np.float32).reshape(-1, 9)
    T = np.ascontiguousarray(T, np.float32).reshape(-1, 9)
    if precision == "f32":
        out = np.zeros(len(S), np.float32)
        lib().orc_tridist_f(ptr(S), ptr(T), len(S), ptr(out))
    else:
        out = np.zeros(len(S), np.float64)
        lib().orc_tridist_d(ptr(S), ptr(T), len(S), ptr(out))
    return out


def tridist_dd(S, T):
    S = np.ascontiguousarray(S, np.float64).reshape(-1, 9)
    T = np.ascontiguousarray(T, np.float64).reshape(-1, 9)
    out = np.zeros(len(S), np.float64)
    lib().orc_tridist_dd(ptr(S), ptr(T), len(S), ptr(out))
    return out


def meshdist(tri9, hd, offset_size, precision="f32"):
    tri9 = np.ascontiguousarray(tri9, np.float32).reshape(-1, 9)
    hd = None if hd is None else np.ascontiguousarray(hd, np.float32).reshape(-1, 2)
    os_ = np.ascontiguousarray(offset_size, np.uint64).reshape(-1, 4)
    if precision == "f32":
        out = np.zeros((len(os_), 3), np.float32)
        lib().orc_meshdist_f(ptr(tri9), ptr(hd), ptr(os_), len(os_), ptr(out))
    else:
        out = np.zeros((len(os_), 3), np.float64)
        lib().orc_meshdist_d(ptr(tri9), ptr(hd), ptr(os_), len(os_), ptr(out))
    return out


def lod_list(lods):
    a = np.array(lods, np.int32)
    out = np.zeros_like(a)
    lib().orc_lod_list(ptr(a), len(a), ptr(out))
    return out.tolist()


class OracleTile:
    def __init__(self, path):
        self.h = lib().orc_tile_load(os.fsencode(path))
        if not self.h:
            raise RuntimeError(f"oracle could not load {path}")
        self.n = lib().orc_tile_count(self.h)

    def __del__(self):
        try:
            lib().orc_tile_free(self.h)
        except Exception:
            pass


def join(t1, t2, query="nn", k=1, tau=1000.0, lods=(20, 40, 60, 80, 100), level=2,
         precision="f32", max_objects1=0, rounds=False):
    """Returns (target, obj2, min, max) arrays in report order and stats
    (with rounds=True, also the round of every report: 0 = filter)."""
    L = lib()
    lods = np.array(lods, np.int32)
    qcode = {"nn": 1, "within": 0, "intersect": 2}[query]
    r = L.orc_join(t1.h, (t2 or t1).h, qcode, int(k), float(tau), ptr(lods),
                   len(lods), int(level), 0 if precision == "f32" else 1,
                   int(max_objects1 or 2**63 - 1))
    if not r:
        raise RuntimeError("orc_join failed")
    n = L.orc_result_count(r)
    t = np.zeros(n, np.int32); o = np.zeros(n, np.int32)
    mn = np.zeros(n, np.float64); mx = np.zeros(n, np.float64)
    L.orc_result_get(r, ptr(t), ptr(o), ptr(mn), ptr(mx))
    rd = np.zeros(n, np.uint8)
    L.orc_result_rounds(r, ptr(rd))
    st = np.zeros(4, np.uint64)
    L.orc_result_stats(r, ptr(st))
    L.orc_result_free(r)
    stats = dict(zip(["tri_pairs", "voxel_pairs", "object_pairs", "rounds"], st.tolist()))
    if rounds:
        return t, o, mn, mx, stats, rd
    return t, o, mn, mx, stats


def threads():
    """OpenMP threads of the oracle's join refinement."""
    return int(lib().orc_max_threads())


def group(target, obj2, mn=None, mx=None):
    """Group a report list by target keeping report order."""
    out = {}
    for i, (t, o) in enumerate(zip(target.tolist(), obj2.tolist())):
        e = (o,) if mn is None else (o, float(mn[i]), float(mx[i]))
        out.setdefault(t, []).append(e)
    return out
