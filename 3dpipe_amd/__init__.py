"""MI355X-native spatial-join refinement path (kNN / within over progressively
decoded polyhedra) -- Python bindings over the C ABI.

The product is native: ``lib/libtdp.so`` (hand-written gfx950 HIP kernels,
include/tdp.h) and ``lib/libtdbase.so`` (C++ host join engine, include/tdbase.h),
plus the ``bin/tdbase`` command. This module only loads those libraries with
ctypes and wraps them for tests and benchmarks; it never computes a distance
itself and there is no CPU fallback: without the built libraries, or without a
GPU, the calls raise.

Operator surface mirrored from the reference's ``tdbase join``
(src/tools/tdbase.cpp:482-549, src/include/query_context.h:141-192):
``join(tile1, tile2=None, query="nn"|"within", knn=1, within_dist=1000.0,
lods=(20,40,60,80,100), hausdorff_level=2, max_objects1=None)`` with the added
``precision`` ("ref32" | "fp64" | "mixed").
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
BIN_DIR = os.path.join(_HERE, "bin")

PREC = {"ref32": 0, "fp64": 1, "mixed": 2}
QUERY = {"nn": 0, "within": 1, "intersect": 2}

_c = ctypes
_P = _c.c_void_p


class TdpError(RuntimeError):
    """A failed call into the native join path (message from the library)."""


class _GenParams(_c.Structure):
    _fields_ = [
        ("seed", _c.c_uint64), ("n_vessels", _c.c_int), ("nuclei_per_vessel", _c.c_int),
        ("cube", _c.c_double), ("vessel_r_min", _c.c_double), ("vessel_r_max", _c.c_double),
        ("vessel_segments", _c.c_int), ("segment_length", _c.c_double),
        ("turn_max_deg", _c.c_double), ("vessel_sides", _c.c_int), ("rings_min", _c.c_int),
        ("rings_max", _c.c_int), ("nucleus_r_min", _c.c_double), ("nucleus_r_max", _c.c_double),
        ("noise", _c.c_double), ("subdiv", _c.c_int), ("place_min", _c.c_double),
        ("place_max", _c.c_double), ("prototypes", _c.c_int), ("voxel_tris", _c.c_int),
        ("threads", _c.c_int), ("max_nuclei", _c.c_int64),
        ("nuclei_seed", _c.c_uint64), ("emit_vessels", _c.c_int),
    ]


class _Tuning(_c.Structure):
    _fields_ = [("small_pairs", _c.c_int), ("max_rows", _c.c_int), ("drain", _c.c_int),
                ("drain_blocks", _c.c_int), ("eval_blocks", _c.c_int), ("graph", _c.c_int),
                ("row_cull", _c.c_int)]


class _JoinOpts(_c.Structure):
    _fields_ = [
        ("query", _c.c_int), ("k", _c.c_int), ("within_dist", _c.c_float),
        ("lods", _c.POINTER(_c.c_int)), ("n_lods", _c.c_int), ("hausdorff_level", _c.c_int),
        ("max_objects1", _c.c_uint64), ("precision", _c.c_int), ("threads", _c.c_int),
    ]


_libs = {}


def _load():
    if _libs:
        return _libs["tdbase"], _libs["tdp"]
    tdp_path = os.path.join(LIB_DIR, "libtdp.so")
    tdb_path = os.path.join(LIB_DIR, "libtdbase.so")
    for p in (tdp_path, tdb_path):
        if not os.path.exists(p):
            raise TdpError(f"{p} is not built; run __graft_entry__.build() or make -C 3dpipe_amd")
    tdp = _c.CDLL(tdp_path, mode=_c.RTLD_GLOBAL)
    tdb = _c.CDLL(tdb_path)
    i64, u64, sz = _c.c_int64, _c.c_uint64, _c.c_size_t
    # libtdbase
    tdb.tdb_tile_load.argtypes = [_c.c_char_p, _c.POINTER(_P), _c.c_char_p, sz]
    tdb.tdb_tile_save.argtypes = [_P, _c.c_char_p, _c.c_char_p, sz]
    tdb.tdb_tile_free.argtypes = [_P]
    tdb.tdb_tile_size.argtypes = [_P, _P]
    tdb.tdb_tile_views.argtypes = [_P] + [_c.POINTER(_P)] * 8
    tdb.tdb_gen_defaults.argtypes = [_c.POINTER(_GenParams)]
    tdb.tdb_generate.argtypes = [_c.POINTER(_GenParams), _c.POINTER(_P), _c.POINTER(_P)]
    tdb.tdb_engine_open.argtypes = [_c.c_int, _c.POINTER(_P)]
    tdb.tdb_engine_close.argtypes = [_P]
    tdb.tdb_engine_error.argtypes = [_P]
    tdb.tdb_engine_error.restype = _c.c_char_p
    tdb.tdb_engine_device.argtypes = [_P, _c.c_char_p, sz, _c.POINTER(_c.c_int)]
    tdb.tdb_engine_evict.argtypes = [_P, _P]
    tdb.tdb_join_defaults.argtypes = [_c.POINTER(_JoinOpts)]
    tdb.tdb_join.argtypes = [_P, _P, _P, _c.POINTER(_JoinOpts), _c.POINTER(_P)]
    tdb.tdb_candidates.argtypes = [_P, _P, _c.POINTER(_JoinOpts), _P, _P, sz]
    tdb.tdb_candidates.restype = i64
    tdb.tdb_result_count.argtypes = [_P]
    tdb.tdb_result_count.restype = u64
    tdb.tdb_result_targets.argtypes = [_P]
    tdb.tdb_result_targets.restype = u64
    tdb.tdb_result_get.argtypes = [_P, _P, _P, _P, _P, _P]
    tdb.tdb_result_timing.argtypes = [_P, _P]
    tdb.tdb_result_timing_n.argtypes = [_P, _P, _c.c_int]
    tdb.tdb_result_timing_n.restype = _c.c_int
    tdb.tdb_result_stats.argtypes = [_P, _P]
    tdb.tdb_result_stats_n.argtypes = [_P, _P, _c.c_int]
    tdb.tdb_result_stats_n.restype = _c.c_int
    tdb.tdb_result_free.argtypes = [_P]
    tdb.tdb_result_stages.argtypes = [_P, _P, _P, _c.c_int]
    tdb.tdb_tile_clone.argtypes = [_P, _c.POINTER(_P)]
    tdb.tdb_engine_get_tuning.argtypes = [_P, _P]
    tdb.tdb_engine_autotune.argtypes = [_P, _P, _P, _P, _c.c_size_t, _P, _P, _P, _c.c_int]
    tdb.tdb_engine_set_tuning.argtypes = [_P, _P]
    tdb.tdb_engine_candidates.argtypes = [_P, _P, _P, _P, _P, _P, _c.c_size_t,
                                          _c.POINTER(_c.c_uint64)]
    tdb.tdb_engine_candidates.restype = _c.c_int64
    tdb.tdb_server_defaults.argtypes = [_P]
    tdb.tdb_server_open.argtypes = [_P, _P, _c.POINTER(_P)]
    tdb.tdb_server_submit.argtypes = [_P, _P, _P]
    tdb.tdb_server_submit.restype = _c.c_int64
    tdb.tdb_server_wait.argtypes = [_P, _c.c_int64, _c.POINTER(_P), _c.POINTER(_c.c_double)]
    tdb.tdb_server_error.argtypes = [_P]
    tdb.tdb_server_error.restype = _c.c_char_p
    tdb.tdb_server_stats.argtypes = [_P, _P]
    tdb.tdb_server_close.argtypes = [_P]
    tdb.tdb_server_stage_ms.argtypes = [_P, _P]
    tdb.tdb_server_figures.argtypes = [_P, _P, _c.c_int]
    tdb.tdb_source_generate.argtypes = [_c.POINTER(_GenParams), _c.POINTER(_P)]
    tdb.tdb_source_tile.argtypes = [_P, _c.POINTER(_P)]
    tdb.tdb_source_objects.argtypes = [_P]
    tdb.tdb_source_objects.restype = u64
    tdb.tdb_source_decode.argtypes = [_P, _P, sz, _c.POINTER(_P)]
    tdb.tdb_source_free.argtypes = [_P]
    tdb.tdb_stream_defaults.argtypes = [_P]
    tdb.tdb_stream_join.argtypes = [_P, _P, _P, _c.POINTER(_JoinOpts), _P, _c.POINTER(_P)]
    tdb.tdb_result_stream.argtypes = [_P, _P]
    tdb.tdb_result_stream_n.argtypes = [_P, _P, _c.c_int]
    # libtdp (kernel-level parity ops)
    tdp.tdp_open.argtypes = [_c.c_int, _c.POINTER(_P)]
    tdp.tdp_close.argtypes = [_P]
    tdp.tdp_last_error.argtypes = [_P]
    tdp.tdp_last_error.restype = _c.c_char_p
    tdp.tdp_tridist_batch.argtypes = [_P, _c.c_int, _P, _P, sz, _P]
    tdp.tdp_meshdist_batch.argtypes = [_P, _c.c_int, _P, sz, _P, sz, _P]
    tdp.tdp_meshdist_batch_noalloc.argtypes = [_P, _c.c_int, _P, sz, _P, sz, _P]
    tdp.tdp_knn_status.argtypes = [_P, sz, _P, _P, _P, _P, _c.c_int, _P]
    _libs["tdbase"], _libs["tdp"] = tdb, tdp
    return tdb, tdp


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_P)


# ------------------------------------------------------------------ tiles


class Tile:
    """A tile of polyhedra (objects -> voxels -> triangles at LOD 20..100)."""

    def __init__(self, handle):
        self._h = _P(handle)
        tdb, _ = _load()
        sz = np.zeros(3, np.uint64)
        tdb.tdb_tile_size(self._h, _ptr(sz))
        self.n_obj, self.n_vox, self.n_tri = (int(x) for x in sz)

    @classmethod
    def load(cls, path: str) -> "Tile":
        tdb, _ = _load()
        h, err = _P(), _c.create_string_buffer(512)
        if tdb.tdb_tile_load(os.fsencode(path), _c.byref(h), err, 512) != 0:
            raise TdpError(err.value.decode())
        return cls(h.value)

    def clone(self) -> "Tile":
        tdb, _ = _load()
        h = _P()
        if tdb.tdb_tile_clone(self._h, _c.byref(h)) != 0:
            raise TdpError("tile clone failed")
        return Tile(h.value)

    def save(self, path: str) -> None:
        tdb, _ = _load()
        err = _c.create_string_buffer(512)
        if tdb.tdb_tile_save(self._h, os.fsencode(path), err, 512) != 0:
            raise TdpError(err.value.decode())

    def arrays(self) -> dict:
        """Copies of the tile arrays (see include/tdbase.h tdb_tile_views)."""
        tdb, _ = _load()
        ptrs = [_P() for _ in range(8)]
        tdb.tdb_tile_views(self._h, *[_c.byref(p) for p in ptrs])
        spec = [("obj_vox_off", np.uint32, self.n_obj + 1), ("obj_box", np.float32, 6 * self.n_obj),
                ("obj_transform", np.float32, 3 * self.n_obj), ("obj_hd", np.float32, 10 * self.n_obj),
                ("vox_box", np.float32, 9 * self.n_vox), ("vox_tri_off", np.uint64, 5 * self.n_vox),
                ("vox_tri_cnt", np.uint32, 5 * self.n_vox), ("tri", np.float32, 12 * self.n_tri)]
        out = {}
        for (name, dt, n), p in zip(spec, ptrs):
            if n == 0 or not p.value:
                out[name] = np.zeros(0, dt)
                continue
            buf = (_c.c_char * (n * np.dtype(dt).itemsize)).from_address(p.value)
            out[name] = np.frombuffer(buf, dt, n).copy()
        return out

    def __del__(self):
        try:
            if self._h:
                _libs["tdbase"].tdb_tile_free(self._h)
                self._h = None
        except Exception:
            pass


def _gen_params(seed, n_vessels, nuclei_per_vessel, cube, kw) -> "_GenParams":
    tdb, _ = _load()
    p = _GenParams()
    tdb.tdb_gen_defaults(_c.byref(p))
    p.seed, p.n_vessels, p.nuclei_per_vessel, p.cube = seed, n_vessels, nuclei_per_vessel, cube
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown generator parameter {k}")
        setattr(p, k, v)
    return p


def generate(seed: int = 0, n_vessels: int = 10, nuclei_per_vessel: int = 20,
             cube: float = 2000.0, **kw) -> tuple:
    """Seeded synthetic (nuclei, vessels) tiles; keyword names follow
    tdb_gen_params in include/tdbase.h."""
    tdb, _ = _load()
    p = _gen_params(seed, n_vessels, nuclei_per_vessel, cube, kw)
    n, v = _P(), _P()
    if tdb.tdb_generate(_c.byref(p), _c.byref(n), _c.byref(v)) != 0:
        raise TdpError("invalid generator parameters")
    return Tile(n.value), Tile(v.value)


class Source:
    """Objects decoded on demand for the streaming join (tdb_source_* of
    include/tdbase.h): the generator's nuclei, or a tile's objects."""

    def __init__(self, handle, keep=None):
        self._h = _P(handle)
        self._keep = keep  # a tile the source reads from
        self.n_obj = int(_load()[0].tdb_source_objects(self._h))

    @classmethod
    def generate(cls, seed: int = 0, n_vessels: int = 0, nuclei_per_vessel: int = 1000,
                 cube: float = 2000.0, **kw) -> "Source":
        tdb, _ = _load()
        p = _gen_params(seed, n_vessels, nuclei_per_vessel, cube, kw)
        h = _P()
        if tdb.tdb_source_generate(_c.byref(p), _c.byref(h)) != 0:
            raise TdpError("invalid generator parameters")
        return cls(h.value)

    @classmethod
    def of_tile(cls, tile: Tile) -> "Source":
        tdb, _ = _load()
        h = _P()
        if tdb.tdb_source_tile(tile._h, _c.byref(h)) != 0:
            raise TdpError("tdb_source_tile failed")
        return cls(h.value, keep=tile)

    def decode(self, objects) -> Tile:
        """The given objects as a new tile, in that order."""
        tdb, _ = _load()
        ids = np.ascontiguousarray(objects, np.uint64)
        h = _P()
        if tdb.tdb_source_decode(self._h, _ptr(ids), ids.size, _c.byref(h)) != 0:
            raise TdpError("tdb_source_decode: object out of range")
        return Tile(h.value)

    def __del__(self):
        try:
            if self._h:
                _libs["tdbase"].tdb_source_free(self._h)
                self._h = None
        except Exception:
            pass


class _StreamOpts(_c.Structure):
    _fields_ = [("hbm_budget", _c.c_uint64), ("tile_share", _c.c_double),
                ("max_chunk_targets", _c.c_uint64), ("decode_threads", _c.c_int),
                ("overlap", _c.c_int), ("lazy", _c.c_int), ("workers", _c.c_int)]


# streaming figures of a result (tdb_result_stream_n)
STREAM = ["chunks", "h2d_bytes", "decode_ms", "copy_ms", "join_ms", "wall_ms", "overlap",
          "peak_device_bytes", "index_ms", "candidates_ms", "objects_decoded", "splits",
          "record_bytes_all", "record_bytes_shipped", "meta_bytes", "workers",
          "candidate_host_targets", "plan_ms", "workers_setup_ms"]

# ------------------------------------------------------------------ join


# refinement stages, in the order of TDP_STAGE_* (include/tdp.h)
STAGES = ["reset", "bounds", "scan", "probe", "mark", "expand", "eval", "fold", "direct",
          "graph"]


@dataclass
class JoinResult:
    """Reported (target, object) pairs grouped by target in report order."""
    target: np.ndarray
    obj2: np.ndarray
    rmin: np.ndarray
    rmax: np.ndarray
    round: np.ndarray
    n_targets: int
    timing_ms: dict = field(default_factory=dict)
    stats: dict = field(default_factory=dict)
    # device ms and launches per refinement stage (STAGES order)
    stage_ms: dict = field(default_factory=dict)
    stage_launches: dict = field(default_factory=dict)
    # streaming figures (STREAM names; streamed joins only)
    stream: dict = field(default_factory=dict)

    def lists(self) -> dict:
        """{target: [obj2, ...]} in report order."""
        out = {}
        for t, o in zip(self.target.tolist(), self.obj2.tolist()):
            out.setdefault(t, []).append(o)
        return out


def _read_result(r) -> "JoinResult":
    """Copy a tdb_result into a JoinResult and free it."""
    tdb, _ = _load()
    try:
        n = int(tdb.tdb_result_count(r))
        t = np.zeros(n, np.int32); b = np.zeros(n, np.int32)
        mn = np.zeros(n, np.float64); mx = np.zeros(n, np.float64)
        rd = np.zeros(n, np.uint8)
        tdb.tdb_result_get(r, _ptr(t), _ptr(b), _ptr(mn), _ptr(mx), _ptr(rd))
        tm = np.zeros(10, np.float64)
        st = np.zeros(9, np.uint64)
        tdb.tdb_result_timing_n(r, _ptr(tm), 10)
        tdb.tdb_result_stats_n(r, _ptr(st), 9)
        sm = np.zeros(len(STAGES), np.float64)
        sn = np.zeros(len(STAGES), np.uint64)
        tdb.tdb_result_stages(r, _ptr(sm), _ptr(sn), len(STAGES))
        sf = np.zeros(len(STREAM), np.float64)
        tdb.tdb_result_stream_n(r, _ptr(sf), len(STREAM))
        nt = int(tdb.tdb_result_targets(r))
    finally:
        tdb.tdb_result_free(r)
    timing = dict(zip(["total", "candidates", "filter", "refine", "filter_kernels",
                       "refine_kernels", "upload", "setup", "download", "assembly"], tm.tolist()))
    stats = dict(zip(["tri_pairs", "voxel_pairs", "fp64_evals", "filter_voxel_pairs",
                      "object_pairs", "rounds", "fp32_culled", "mark_tested",
                      "fp64_probe_evals"],
                     (int(x) for x in st)))
    return JoinResult(t, b, mn, mx, rd, nt, timing, stats, dict(zip(STAGES, sm.tolist())),
                      dict(zip(STAGES, (int(x) for x in sn))),
                      dict(zip(STREAM, sf.tolist())) if sf[0] > 0 else {})


def _opts(query, knn, within_dist, lods, hausdorff_level, max_objects1, precision, threads):
    tdb, _ = _load()
    o = _JoinOpts()
    tdb.tdb_join_defaults(_c.byref(o))
    if query not in QUERY:
        raise ValueError("query must be 'nn', 'within' or 'intersect'")
    if precision not in PREC:
        raise ValueError("precision must be ref32, fp64 or mixed")
    o.query = QUERY[query]
    o.k = int(knn)
    o.within_dist = float(within_dist)
    lod_arr = None
    if lods is not None and len(lods):
        lod_arr = (_c.c_int * len(lods))(*[int(x) for x in lods])
        o.lods = _c.cast(lod_arr, _c.POINTER(_c.c_int))
        o.n_lods = len(lods)
    o.hausdorff_level = int(hausdorff_level)
    o.max_objects1 = int(max_objects1 or 0)
    o.precision = PREC[precision]
    o.threads = int(threads)
    return o, lod_arr


class Engine:
    """One GPU with device-resident tiles (include/tdbase.h tdb_engine)."""

    def __init__(self, device: int = 0):
        tdb, _ = _load()
        h = _P()
        rc = tdb.tdb_engine_open(int(device), _c.byref(h))
        self._h = h
        if rc != 0:
            msg = tdb.tdb_engine_error(h).decode() if h.value else "tdb_engine_open failed"
            if h.value:
                tdb.tdb_engine_close(h)
            self._h = None
            raise TdpError(f"no usable GPU: {msg}")

    def device(self) -> tuple:
        tdb, _ = _load()
        buf, cu = _c.create_string_buffer(256), _c.c_int(0)
        tdb.tdb_engine_device(self._h, buf, 256, _c.byref(cu))
        return buf.value.decode(), cu.value

    def evict(self, tile: Tile) -> None:
        _load()[0].tdb_engine_evict(self._h, tile._h)

    def tuning(self) -> dict:
        """The kernel tuning (tdp_tuning of include/tdp.h)."""
        tdb, _ = _load()
        t = _Tuning()
        if tdb.tdb_engine_get_tuning(self._h, _c.byref(t)) != 0:
            raise TdpError(tdb.tdb_engine_error(self._h).decode())
        return {name: getattr(t, name) for name, _ in _Tuning._fields_}

    def set_tuning(self, **kw) -> None:
        tdb, _ = _load()
        t = _Tuning()
        tdb.tdb_engine_get_tuning(self._h, _c.byref(t))
        for k, v in kw.items():
            if not hasattr(t, k):
                raise TypeError(f"unknown tuning parameter {k}")
            setattr(t, k, int(v))
        if tdb.tdb_engine_set_tuning(self._h, _c.byref(t)) != 0:
            raise TdpError(tdb.tdb_engine_error(self._h).decode())

    def autotune(self, tile1: Tile, tile2: Optional[Tile] = None, query: str = "nn",
                 knn: int = 1, within_dist: float = 1000.0, sample_targets: int = 0,
                 precision: str = "mixed") -> dict:
        """Time a sample of the join under each candidate tuning and keep the
        fastest (see tdb_engine_autotune). Returns {"best": tuning,
        "tried": [(tuning, ms), ...]}."""
        tdb, _ = _load()
        o, _keep = _opts(query, knn, within_dist, None, 2, None, precision, 0)
        best = _Tuning()
        tried = (_Tuning * 32)()
        ms = np.zeros(32, np.float64)
        n = tdb.tdb_engine_autotune(self._h, tile1._h, tile2._h if tile2 is not None else None,
                                    _c.byref(o), int(sample_targets), _c.byref(best), tried,
                                    _ptr(ms), 32)
        if n < 0:
            raise TdpError(tdb.tdb_engine_error(self._h).decode())
        as_dict = lambda t: {name: getattr(t, name) for name, _ in _Tuning._fields_}
        return {"best": as_dict(best), "tried": [(as_dict(tried[i]), float(ms[i])) for i in range(n)]}

    def candidates(self, tile1: Tile, tile2: Optional[Tile] = None, query: str = "nn",
                   knn: int = 1, within_dist: float = 1000.0,
                   max_objects1: Optional[int] = None):
        """Candidate lists as the join builds them (on the device for kNN
        and within against at most 1024 objects): (prefix, ids, targets left
        to the host octree)."""
        tdb, _ = _load()
        o, _keep = _opts(query, knn, within_dist, None, 2, max_objects1, "mixed", 0)
        n = min(tile1.n_obj, max_objects1) if max_objects1 else tile1.n_obj
        prefix = np.zeros(n + 1, np.uint32)
        fb = _c.c_uint64(0)
        t2 = tile2._h if tile2 is not None else None
        total = tdb.tdb_engine_candidates(self._h, tile1._h, t2, _c.byref(o), _ptr(prefix), None,
                                          0, _c.byref(fb))
        if total < 0:
            raise TdpError(tdb.tdb_engine_error(self._h).decode())
        ids = np.zeros(max(int(total), 1), np.uint32)
        tdb.tdb_engine_candidates(self._h, tile1._h, t2, _c.byref(o), _ptr(prefix), _ptr(ids),
                                  int(total), _c.byref(fb))
        return prefix, ids[: int(total)], int(fb.value)

    def join(self, tile1: Tile, tile2: Optional[Tile] = None, query: str = "nn", knn: int = 1,
             within_dist: float = 1000.0, lods: Optional[Sequence[int]] = None,
             hausdorff_level: int = 2, max_objects1: Optional[int] = None,
             precision: str = "mixed", threads: int = 0) -> JoinResult:
        tdb, _ = _load()
        o, keep = _opts(query, knn, within_dist, lods, hausdorff_level, max_objects1, precision,
                        threads)
        r = _P()
        rc = tdb.tdb_join(self._h, tile1._h, tile2._h if tile2 is not None else None,
                          _c.byref(o), _c.byref(r))
        if rc != 0:
            raise TdpError(tdb.tdb_engine_error(self._h).decode())
        return _read_result(r)

    def stream_join(self, src1: Source, src2: Optional[Source] = None, query: str = "nn",
                    knn: int = 1, within_dist: float = 1000.0,
                    lods: Optional[Sequence[int]] = None, hausdorff_level: int = 2,
                    max_objects1: Optional[int] = None, precision: str = "mixed",
                    threads: int = 0, hbm_budget: int = 0, tile_share: float = 0.5,
                    max_chunk_targets: int = 0, decode_threads: int = 0,
                    overlap: bool = True, lazy: bool = True, workers: int = 0) -> JoinResult:
        """Chunked join within a device budget (tdb_stream_join); the result
        equals join() on the same objects, with the streaming figures in
        .stream. lazy: records shipped per LOD round for the voxels it reads;
        workers: chunk workers side by side (0 = 4; 1 without overlap)."""
        tdb, _ = _load()
        o, keep = _opts(query, knn, within_dist, lods, hausdorff_level, max_objects1, precision,
                        threads)
        so = _StreamOpts()
        tdb.tdb_stream_defaults(_c.byref(so))
        so.hbm_budget, so.tile_share = int(hbm_budget), float(tile_share)
        so.max_chunk_targets, so.decode_threads = int(max_chunk_targets), int(decode_threads)
        so.overlap = 1 if overlap else 0
        so.lazy, so.workers = 1 if lazy else 0, int(workers)
        r = _P()
        rc = tdb.tdb_stream_join(self._h, src1._h, src2._h if src2 is not None else None,
                                 _c.byref(o), _c.byref(so), _c.byref(r))
        if rc != 0:
            raise TdpError(tdb.tdb_engine_error(self._h).decode())
        return _read_result(r)

    def close(self):
        if self._h:
            _load()[0].tdb_engine_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _ServerOpts(_c.Structure):
    _fields_ = [("device", _c.c_int), ("hbm_budget", _c.c_uint64), ("max_batch", _c.c_int),
                ("overlap", _c.c_int), ("precision", _c.c_int), ("threads", _c.c_int),
                ("lazy", _c.c_int), ("executors", _c.c_int), ("mix_kinds", _c.c_int)]


class Server:
    """Queries joined against one device-resident tile with continuous
    batching (tdb_server_* of include/tdbase.h). submit() takes ownership of
    the query tile."""

    def __init__(self, resident: Tile, device: int = 0, hbm_budget: int = 0,
                 max_batch: int = 64, overlap: bool = True, precision: str = "mixed",
                 threads: int = 0, lazy: bool = True, executors: int = 0,
                 mix_kinds: bool = True):
        tdb, _ = _load()
        o = _ServerOpts()
        tdb.tdb_server_defaults(_c.byref(o))
        o.device, o.hbm_budget, o.max_batch = int(device), int(hbm_budget), int(max_batch)
        o.overlap, o.precision, o.threads = int(bool(overlap)), PREC[precision], int(threads)
        o.lazy, o.executors, o.mix_kinds = int(bool(lazy)), int(executors), int(bool(mix_kinds))
        self._resident = resident  # must outlive the server
        h = _P()
        rc = tdb.tdb_server_open(_c.byref(o), resident._h, _c.byref(h))
        self._h = h
        if rc != 0:
            msg = tdb.tdb_server_error(h).decode() if h.value else "tdb_server_open failed"
            if h.value:
                tdb.tdb_server_close(h)
            self._h = None
            raise TdpError(f"no usable GPU: {msg}")

    def submit(self, tile: Tile, query: str = "nn", knn: int = 1, within_dist: float = 1000.0,
               lods: Optional[Sequence[int]] = None, hausdorff_level: int = 2) -> int:
        tdb, _ = _load()
        o, keep = _opts(query, knn, within_dist, lods, hausdorff_level, None, "mixed", 0)
        t = tdb.tdb_server_submit(self._h, tile._h, _c.byref(o))
        if t < 0:
            raise TdpError(tdb.tdb_server_error(self._h).decode())
        tile._h = None  # owned by the server now
        return int(t)

    def wait(self, ticket: int) -> tuple:
        """(JoinResult, latency ms)"""
        tdb, _ = _load()
        r, lat = _P(), _c.c_double(0)
        rc = tdb.tdb_server_wait(self._h, int(ticket), _c.byref(r), _c.byref(lat))
        if rc != 0:
            raise TdpError(tdb.tdb_server_error(self._h).decode())
        return _read_result(r), lat.value

    def stats(self) -> dict:
        tdb = _load()[0]
        st = np.zeros(4, np.uint64)
        tdb.tdb_server_stats(self._h, _ptr(st))
        ms = np.zeros(4, np.float64)
        tdb.tdb_server_stage_ms(self._h, _ptr(ms))
        out = dict(zip(["batches", "queries", "max_batch", "peak_device_bytes"],
                       (int(x) for x in st)))
        fg = np.zeros(13, np.float64)
        tdb.tdb_server_figures(self._h, _ptr(fg), 13)
        out["stage_ms"] = dict(zip(["merge", "upload", "candidates", "join", "record_copy",
                                    "record_put"], fg[:6].tolist()))
        out["h2d_meta_bytes"], out["h2d_record_bytes"] = int(fg[6]), int(fg[7])
        out["record_bytes_of_queries"] = int(fg[8])
        out["executors"], out["most_batches_at_once"] = int(fg[9]), int(fg[10])
        out["mixed_kind_batches"] = int(fg[12])
        return out

    def close(self):
        if self._h:
            _load()[0].tdb_server_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def candidates(tile1: Tile, tile2: Optional[Tile] = None, query: str = "nn", knn: int = 1,
               within_dist: float = 1000.0, max_objects1: Optional[int] = None):
    """Octree candidate lists per target: (prefix, ids)."""
    tdb, _ = _load()
    o, _keep = _opts(query, knn, within_dist, None, 2, max_objects1, "mixed", 0)
    n = min(tile1.n_obj, max_objects1) if max_objects1 else tile1.n_obj
    prefix = np.zeros(n + 1, np.uint32)
    total = tdb.tdb_candidates(tile1._h, tile2._h if tile2 is not None else None, _c.byref(o),
                               _ptr(prefix), None, 0)
    ids = np.zeros(max(int(total), 1), np.uint32)
    tdb.tdb_candidates(tile1._h, tile2._h if tile2 is not None else None, _c.byref(o),
                       _ptr(prefix), _ptr(ids), int(total))
    return prefix, ids[: int(total)]


# ------------------------------------------------------- kernel-level ops


class Device:
    """Kernel-level parity ops of include/tdp.h on one GPU."""

    def __init__(self, device: int = 0):
        _, tdp = _load()
        h = _P()
        rc = tdp.tdp_open(int(device), _c.byref(h))
        self._h = h
        if rc != 0:
            msg = tdp.tdp_last_error(h).decode() if h.value else "tdp_open failed"
            if h.value:
                tdp.tdp_close(h)
            self._h = None
            raise TdpError(f"no usable GPU: {msg}")

    def _err(self):
        return TdpError(_load()[1].tdp_last_error(self._h).decode())

    def tridist(self, S: np.ndarray, T: np.ndarray, precision: str = "ref32") -> np.ndarray:
        _, tdp = _load()
        S = np.ascontiguousarray(S, np.float32).reshape(-1, 9)
        T = np.ascontiguousarray(T, np.float32).reshape(-1, 9)
        assert S.shape == T.shape
        out = np.zeros(len(S), np.float32 if precision == "ref32" else np.float64)
        if tdp.tdp_tridist_batch(self._h, PREC[precision], _ptr(S), _ptr(T), len(S), _ptr(out)):
            raise self._err()
        return out

    def meshdist(self, rec: np.ndarray, offset_size: np.ndarray,
                 precision: str = "ref32") -> np.ndarray:
        """Per voxel pair (distance, min bound, max bound) over 12-float records."""
        _, tdp = _load()
        rec = np.ascontiguousarray(rec, np.float32).reshape(-1, 12)
        os_ = np.ascontiguousarray(offset_size, np.uint64).reshape(-1, 4)
        out = np.zeros((len(os_), 3), np.float32 if precision == "ref32" else np.float64)
        if tdp.tdp_meshdist_batch(self._h, PREC[precision], _ptr(rec), len(rec), _ptr(os_),
                                  len(os_), _ptr(out)):
            raise self._err()
        return out

    def meshdist_noalloc(self, d_rec: int, n_tri: int, d_offset_size: int, n_pairs: int,
                         d_out: int, precision: str = "ref32") -> None:
        """tdp_meshdist_batch_noalloc: the same per voxel pair values, on
        device pointers the caller owns (records, pairs, output); the library
        allocates no device memory."""
        _, tdp = _load()
        if tdp.tdp_meshdist_batch_noalloc(self._h, PREC[precision], _P(d_rec), int(n_tri),
                                          _P(d_offset_size), int(n_pairs), _P(d_out)):
            raise self._err()

    def knn_status(self, prefix, confirmed, rmin, rmax, k: int) -> np.ndarray:
        _, tdp = _load()
        prefix = np.ascontiguousarray(prefix, np.uint32)
        confirmed = np.ascontiguousarray(confirmed, np.int32)
        rmin = np.ascontiguousarray(rmin, np.float64)
        rmax = np.ascontiguousarray(rmax, np.float64)
        out = np.zeros(int(prefix[-1]), np.int32)
        if tdp.tdp_knn_status(self._h, len(prefix) - 1, _ptr(prefix), _ptr(confirmed),
                              _ptr(rmin), _ptr(rmax), int(k), _ptr(out)):
            raise self._err()
        return out

    def close(self):
        if self._h:
            _load()[1].tdp_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def loaded_libraries() -> list:
    """Paths of the native libraries this process has loaded."""
    return [os.path.join(LIB_DIR, n) for n in ("libtdp.so", "libtdbase.so")] if _libs else []
