"""CPU-side tests: the oracle against the reference's own compiled sources and
closed forms, the C-ABI libraries (load, exported symbols), the RAW format,
the generator and the host octree (no GPU calls)."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import orc
from conftest import ROOT, random_triangles


# ---------------------------------------------------------------- oracle pins

def test_oracle_tridist_matches_reference_build():
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(11)
    S, T = random_triangles(rng, 200_000)
    want = np.zeros(len(S), np.float32)
    ref.ref_tridist(orc.ptr(S), orc.ptr(T), len(S), orc.ptr(want))
    got = orc.tridist(S, T, "f32")
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_oracle_meshdist_and_aab_match_reference_build():
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(12)
    tri = rng.uniform(-30, 30, (3000, 9)).astype(np.float32)
    hd = rng.uniform(0, 2, (3000, 2)).astype(np.float32)
    os_ = np.stack([rng.integers(0, 2900, 400), rng.integers(1, 60, 400),
                    rng.integers(0, 2900, 400), rng.integers(1, 60, 400)], 1).astype(np.uint64)
    os_[:, 1] = np.minimum(os_[:, 1], 3000 - os_[:, 0])
    os_[:, 3] = np.minimum(os_[:, 3], 3000 - os_[:, 2])
    want = np.zeros((400, 3), np.float32)
    ref.ref_meshdist(orc.ptr(tri), orc.ptr(hd), orc.ptr(os_), 400, orc.ptr(want))
    got = orc.meshdist(tri, hd, os_, "f32")
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    a = rng.uniform(-5, 5, (1000, 6)).astype(np.float32)
    b = rng.uniform(-5, 5, (1000, 6)).astype(np.float32)
    a[:, 3:] = a[:, :3] + np.abs(a[:, 3:])
    b[:, 3:] = b[:, :3] + np.abs(b[:, 3:])
    for i in range(1000):
        w = np.zeros(2, np.float32)
        g = np.zeros(2, np.float32)
        ref.ref_aab_distance(orc.ptr(a[i]), orc.ptr(b[i]), orc.ptr(w))
        orc.lib().orc_aab_distance(orc.ptr(a[i]), orc.ptr(b[i]), orc.ptr(g))
        assert np.array_equal(w, g)


def _boxsets(rng, n):
    # integer-grid boxes: many equal distances, so queue tie order matters
    low = np.floor(rng.uniform(0, 40, (n, 3))).astype(np.float32)
    high = (low + np.floor(rng.uniform(1, 6, (n, 3)))).astype(np.float32)
    tr = np.floor((low + high) / 2).astype(np.float32)
    return low, high, tr


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_octree_matches_reference_build(seed):
    """candidate lists of query_knn_fast / query_within, order included"""
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(seed)
    n = 1500
    low, high, tr = _boxsets(rng, n)
    ot = orc.lib().orc_boxset_build(n, orc.ptr(low), orc.ptr(high), orc.ptr(tr), 10)
    rt = ref.ref_boxset_build(n, orc.ptr(low), orc.ptr(high), orc.ptr(tr), 10)
    cap = 4096
    for i in range(0, n, 3):
        for k in (1, 3, 10):
            out = []
            for f, h in ((orc.lib().orc_query_knn, ot), (ref.ref_query_knn, rt)):
                ids = np.zeros(cap, np.int32); mn = np.zeros(cap, np.float32); mx = np.zeros(cap, np.float32)
                c = f(h, orc.ptr(low[i]), orc.ptr(high[i]), orc.ptr(tr[i]), i, k, orc.ptr(ids),
                      orc.ptr(mn), orc.ptr(mx), cap)
                out.append((ids[:c].tolist(), mn[:c].tolist(), mx[:c].tolist()))
            assert out[0] == out[1], (i, k)
        out = []
        for f, h in ((orc.lib().orc_query_within, ot), (ref.ref_query_within, rt)):
            ids = np.zeros(cap, np.int32); mn = np.zeros(cap, np.float32); mx = np.zeros(cap, np.float32)
            c = f(h, orc.ptr(low[i]), orc.ptr(high[i]), orc.ptr(tr[i]), i, 4.0, orc.ptr(ids),
                  orc.ptr(mn), orc.ptr(mx), cap)
            out.append((ids[:c].tolist(), mn[:c].tolist(), mx[:c].tolist()))
        assert out[0] == out[1], i
        # query_intersect + mbb_intersect's sort and de-duplication; and the
        # strict-overlap brute force it amounts to
        ids = np.zeros(cap, np.int32)
        c = orc.lib().orc_query_intersect(ot, orc.ptr(low[i]), orc.ptr(high[i]), i, orc.ptr(ids), cap)
        mine = ids[:c].tolist()
        c = ref.ref_query_intersect(rt, orc.ptr(low[i]), orc.ptr(high[i]), i, orc.ptr(ids), cap)
        assert mine == sorted(set(ids[:c].tolist())), i
        brute = [o for o in range(n) if o != i and
                 not np.any((low[o] >= high[i]) | (high[o] <= low[i]))]
        assert mine == brute, i
    ref.ref_boxset_free(rt)


def _lod_voxel(a, v, slot):
    """triangles (9 floats) and (proxy, hausdorff) of voxel v at a LOD slot"""
    off, cnt = int(a["vox_tri_off"][5 * v + slot]), int(a["vox_tri_cnt"][5 * v + slot])
    rec = a["tri"].reshape(-1, 12)[off:off + cnt]
    return (np.ascontiguousarray(rec[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]]),
            np.ascontiguousarray(rec[:, [3, 7]]))


def _overlap(b1, b2):
    return not np.any((b2[:3] >= b1[3:6]) | (b2[3:6] <= b1[:3]))


def _intersect_bruteforce(ta, tb, self, ref, slot=4):
    """object pairs with overlapping boxes holding a voxel pair (boxes
    overlapping) that the reference's MeshInt finds intersecting at the LOD
    slot; and the subset where the reference's TriInt finds a crossing pair"""
    A, B = ta.arrays(), tb.arrays()
    meshint, triint = set(), set()
    for i in range(ta.n_obj):
        for o in range(tb.n_obj):
            if self and o == i:
                continue
            if not _overlap(A["obj_box"][6 * i:6 * i + 6], B["obj_box"][6 * o:6 * o + 6]):
                continue
            for v1 in range(A["obj_vox_off"][i], A["obj_vox_off"][i + 1]):
                for v2 in range(B["obj_vox_off"][o], B["obj_vox_off"][o + 1]):
                    if not _overlap(A["vox_box"][9 * v1:9 * v1 + 6], B["vox_box"][9 * v2:9 * v2 + 6]):
                        continue
                    t1, h1 = _lod_voxel(A, v1, slot)
                    t2, h2 = _lod_voxel(B, v2, slot)
                    if len(t1) == 0 or len(t2) == 0:
                        continue
                    tri = np.ascontiguousarray(np.concatenate([t1, t2]), np.float32)
                    hd = np.ascontiguousarray(np.concatenate([h1, h2]), np.float32)
                    os_ = np.array([0, len(t1), len(t1), len(t2)], np.uint64)
                    hit = np.zeros(1, np.uint8)
                    ref.ref_meshint(orc.ptr(tri), orc.ptr(hd), orc.ptr(os_), 1, orc.ptr(hit))
                    if hit[0]:
                        meshint.add((i, o))
                    S = np.ascontiguousarray(np.repeat(t1, len(t2), axis=0))
                    T = np.ascontiguousarray(np.tile(t2, (len(t1), 1)))
                    ti = np.zeros(len(S), np.uint8)
                    ref.ref_triint(orc.ptr(S), orc.ptr(T), len(S), orc.ptr(ti))
                    if ti.any():
                        triint.add((i, o))
    return meshint, triint


def test_oracle_intersect_join_pinned_by_reference_meshint(tdp, tiny_tiles):
    """-q intersect: at LOD 100 alone, the oracle's result is exactly the
    object pairs the reference's own MeshInt finds intersecting in a voxel
    pair with overlapping boxes, and it holds every pair the reference's
    TriInt (an exact triangle-triangle test) finds crossing. Over all LODs it
    is the same set at Hausdorff level 0 (never ruled out early) and a subset
    at levels 1 and 2 (whose lower bounds may rule out pairs early)."""
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    pn, pv = tiny_tiles
    ta, tb = tdp.Tile.load(pn), tdp.Tile.load(pv)
    o1, o2 = orc.OracleTile(pn), orc.OracleTile(pv)
    meshint, triint = _intersect_bruteforce(ta, tb, False, ref)
    assert meshint and triint <= meshint
    for prec in ("f32", "f64"):
        t, o, mn, mx, _ = orc.join(o1, o2, "intersect", lods=(100,), level=2, precision=prec)
        assert set(zip(t.tolist(), o.tolist())) == meshint, prec
        assert np.all(mn == 0) and np.all(mx == 0)
        for level in (0, 1, 2):
            t, o, *_ = orc.join(o1, o2, "intersect", level=level, precision=prec)
            got = set(zip(t.tolist(), o.tolist()))
            assert len(got) == len(t)
            assert (got == meshint) if level == 0 else (got <= meshint), (prec, level)
    # self-join of overlapping nuclei: symmetric, self excluded
    n, _ = tdp.generate(seed=5, n_vessels=0, nuclei_per_vessel=60, cube=150.0, subdiv=2,
                        prototypes=8, voxel_tris=20)
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        n.save(os.path.join(d, "n.dt"))
        on = orc.OracleTile(os.path.join(d, "n.dt"))
        meshint, _ = _intersect_bruteforce(n, n, True, ref)
        t, o, *_ = orc.join(on, None, "intersect", lods=(100,), level=2, precision="f32")
        got = set(zip(t.tolist(), o.tolist()))
        assert got == meshint and got == {(b, a) for a, b in got} and len(got) > 10


def test_oracle_golden_fixtures():
    """the oracle against vectors written from the reference's own build"""
    path = os.path.join(ROOT, "tests", "golden", "tridist_ref.npz")
    z = np.load(path)
    got = orc.tridist(z["S"], z["T"], "f32")
    assert np.array_equal(got.view(np.uint32), z["d"].view(np.uint32))
    z = np.load(os.path.join(ROOT, "tests", "golden", "meshdist_ref.npz"))
    got = orc.meshdist(z["tri"], z["hd"], z["os"], "f32")
    assert np.array_equal(got.view(np.uint32), z["out"].view(np.uint32))
    z = np.load(os.path.join(ROOT, "tests", "golden", "octree_ref.npz"))
    low, high, tr = z["low"], z["high"], z["tr"]
    n = len(low)
    ot = orc.lib().orc_boxset_build(n, orc.ptr(low), orc.ptr(high), orc.ptr(tr), 10)
    cap = 4096
    for row in range(len(z["q_index"])):
        i, k = int(z["q_index"][row]), int(z["q_k"][row])
        b, e = z["q_prefix"][row], z["q_prefix"][row + 1]
        ids = np.zeros(cap, np.int32); mn = np.zeros(cap, np.float32); mx = np.zeros(cap, np.float32)
        c = orc.lib().orc_query_knn(ot, orc.ptr(low[i]), orc.ptr(high[i]), orc.ptr(tr[i]), i, k,
                                    orc.ptr(ids), orc.ptr(mn), orc.ptr(mx), cap)
        assert ids[:c].tolist() == z["q_ids"][b:e].tolist()


def test_oracle_fp64_closed_forms():
    """FP64 TriDist against configurations whose distance is known exactly"""
    rng = np.random.default_rng(13)
    n = 2000
    S = np.zeros((n, 9)); T = np.zeros((n, 9)); want = np.zeros(n)
    for i in range(n):
        kind = i % 4
        a = rng.uniform(-5, 5, (3, 3))
        if kind == 0:      # parallel planes: T = S shifted along the normal, overlapping in projection
            a[:, 2] = 0
            h = rng.uniform(0.1, 20)
            S[i] = a.ravel(); T[i] = (a + [0, 0, h]).ravel(); want[i] = h
        elif kind == 1:    # vertex above the interior of the other face
            a[:, 2] = 0
            c = a.mean(0)
            h = rng.uniform(0.1, 20)
            S[i] = a.ravel()
            T[i] = np.array([c + [0, 0, h], c + [3, 1, h + 9], c + [-2, 4, h + 7]]).ravel()
            want[i] = h
        elif kind == 2:    # skew edges crossing in projection: x-axis segment vs y-parallel segment at height h
            h = rng.uniform(0.1, 20)
            S[i] = np.array([[-1, 0, 0], [1, 0, 0], [0, -5, -8]]).ravel()
            T[i] = np.array([[0, -1, h], [0, 1, h], [0, 5, h + 8]]).ravel()
            want[i] = h
        else:              # shared vertex
            S[i] = a.ravel()
            b = rng.uniform(-5, 5, (3, 3)); b[0] = a[0]
            T[i] = b.ravel(); want[i] = 0.0
    got = orc.tridist_dd(S, T)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_oracle_fp64_dense_sampling():
    """FP64 TriDist against the minimum over dense samples of both triangles.
    Every value the reference's TriDist returns is a distance between two
    points of the triangles, so it is never below the true distance (up to the
    sampling resolution). It can be above: the reference's SegPoints computes
    u from the unclamped line parameter t (computing_cpu.cpp:86-98), so in
    about a quarter of random pairs the result overestimates the true
    distance; the oracle reproduces that, as the GPU path must."""
    rng = np.random.default_rng(14)
    m = 40
    u, v = np.meshgrid(np.linspace(0, 1, m + 1), np.linspace(0, 1, m + 1))
    keep = (u + v) <= 1
    u, v = u[keep], v[keep]
    over = 0
    for _ in range(120):
        S = rng.uniform(-3, 3, (3, 3)); T = rng.uniform(-3, 3, (3, 3)) + rng.uniform(-6, 6, 3)
        ps = S[0] + np.outer(u, S[1] - S[0]) + np.outer(v, S[2] - S[0])
        pt = T[0] + np.outer(u, T[1] - T[0]) + np.outer(v, T[2] - T[0])
        d2 = ((ps[:, None, :] - pt[None, :, :]) ** 2).sum(-1)
        sampled = math.sqrt(d2.min())
        d = orc.tridist_dd(S.ravel(), T.ravel())[0]
        res = max(np.linalg.norm(S[1] - S[0]), np.linalg.norm(S[2] - S[0]),
                  np.linalg.norm(T[1] - T[0]), np.linalg.norm(T[2] - T[0])) / m
        assert sampled - d <= 2 * res + 1e-9
        over += d > sampled + 1e-9
    assert 0.05 < over / 120 < 0.5


def test_oracle_fp64_agrees_with_fp32_restatement():
    """the FP64 instantiation is the FP32 text (bit-pinned to the reference)
    with REAL = double; on well-separated triangles both agree to FP32
    rounding"""
    rng = np.random.default_rng(16)
    S = rng.uniform(-5, 5, (50000, 9)).astype(np.float32)
    T = (rng.uniform(-5, 5, (50000, 9)) + rng.uniform(20, 60, (50000, 1))).astype(np.float32)
    d32 = orc.tridist(S, T, "f32").astype(np.float64)
    d64 = orc.tridist(S, T, "f64")
    rel = np.abs(d32 - d64) / d64
    assert np.median(rel) < 1e-6 and np.quantile(rel, 0.999) < 1e-4


def test_oracle_meshdist_is_min_of_tridist():
    """MeshDist per voxel pair = min over its triangle pairs of TriDist and of
    the two bound expressions, in both forms"""
    rng = np.random.default_rng(15)
    tri = rng.uniform(-20, 20, (300, 9)).astype(np.float32)
    hd = rng.uniform(0, 3, (300, 2)).astype(np.float32)
    os_ = np.array([[0, 17, 100, 23], [50, 1, 200, 40], [290, 10, 0, 3]], np.uint64)
    for prec, dt in (("f32", np.float32), ("f64", np.float64)):
        out = orc.meshdist(tri, hd, os_, prec)
        for p, (a, na, b, nb) in enumerate(os_.astype(int)):
            S = np.repeat(tri[a:a + na], nb, 0)
            T = np.tile(tri[b:b + nb], (na, 1))
            d = orc.tridist(S, T, prec)
            ph1 = np.repeat(hd[a:a + na, 0], nb).astype(dt)
            ph2 = np.tile(hd[b:b + nb, 0], na).astype(dt)
            h1 = np.repeat(hd[a:a + na, 1], nb).astype(dt)
            h2 = np.tile(hd[b:b + nb, 1], na).astype(dt)
            assert out[p, 0] == d.min()
            assert out[p, 1] == np.maximum((d - ph1) - ph2, 0).min()
            assert out[p, 2] == ((d + h1) + h2).min()


def test_lod_list_quirk():
    """parse_args sorts the LOD list and calls std::unique without erase"""
    assert orc.lod_list([20, 20, 40]) == [20, 40, 40]
    assert orc.lod_list([100, 20, 60, 20]) == [20, 60, 100, 100]
    assert orc.lod_list([40, 20, 20]) == [20, 40, 40]


# ------------------------------------------------------------- C-ABI libs

def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(td[pb]_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("tdp.h", "libtdp.so"), ("tdbase.h", "libtdbase.so")])
def test_library_exports_every_declared_symbol(tdp, header, lib):
    tdp._load()
    path = os.path.join(ROOT, "3dpipe_amd", "lib", lib)
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    names = _declared(header)
    assert names, header
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_device_calls_fail_loudly_without_gpu(tdp):
    """no GPU here: opening the device raises, nothing falls back to the CPU"""
    import ctypes
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is present")
    with pytest.raises(tdp.TdpError):
        tdp.Engine(0)
    with pytest.raises(tdp.TdpError):
        tdp.Device(0)
    _, v = tdp.generate(seed=0, n_vessels=2, nuclei_per_vessel=1)
    with pytest.raises(tdp.TdpError):
        tdp.Server(v)


def test_generator_nuclei_seed_keeps_vessels(tdp):
    """queries for a resident vessel tile: other nucleus placements, the same
    vessels; emit_vessels=0 leaves the vessel tile empty"""
    n0, v0 = tdp.generate(seed=5, n_vessels=3, nuclei_per_vessel=4)
    n1, v1 = tdp.generate(seed=5, n_vessels=3, nuclei_per_vessel=4, nuclei_seed=77)
    _, v2 = tdp.generate(seed=5, n_vessels=3, nuclei_per_vessel=4, nuclei_seed=77,
                         emit_vessels=0)
    a0, a1 = v0.arrays(), v1.arrays()
    assert all(np.array_equal(a0[k], a1[k]) for k in a0)
    assert not np.array_equal(n0.arrays()["obj_box"], n1.arrays()["obj_box"])
    assert v2.n_obj == 0 and n1.n_obj == 12


@pytest.mark.parametrize("kw", [dict(seed=0, n_vessels=10, nuclei_per_vessel=20, cube=2000.0),
                                dict(seed=3, n_vessels=0, nuclei_per_vessel=300, cube=5000.0,
                                     nucleus_r_min=6.0, nucleus_r_max=14.0)])
def test_source_decodes_the_generated_objects(tdp, kw):
    """the streaming join's decode: every object of a generator source, in
    any order, byte for byte the generated tile's object; a tile source gives
    the tile's own objects back"""
    n, _ = tdp.generate(**kw)
    src = tdp.Source.generate(**kw)
    assert src.n_obj == n.n_obj
    a = n.arrays()
    b = src.decode(np.arange(n.n_obj)).arrays()
    assert all(np.array_equal(a[k], b[k]) for k in a)
    perm = np.random.default_rng(1).permutation(n.n_obj)[: n.n_obj // 3]
    d = src.decode(perm).arrays()
    for k, w in (("obj_box", 6), ("obj_transform", 3), ("obj_hd", 10)):
        assert np.array_equal(a[k].reshape(-1, w)[perm], d[k].reshape(-1, w))
    off = a["obj_vox_off"]
    vb = np.concatenate([a["vox_box"].reshape(-1, 9)[off[o]:off[o + 1]] for o in perm])
    assert np.array_equal(vb, d["vox_box"].reshape(-1, 9))
    t = tdp.Source.of_tile(n).decode(perm).arrays()
    assert all(np.array_equal(t[k], d[k]) for k in d)
    with pytest.raises(tdp.TdpError):
        src.decode([n.n_obj])


# ------------------------------------------------------- RAW format, generator

def test_raw_roundtrip_and_oracle_reads_same_bytes(tdp, small_tiles, tmp_path):
    pn, pv = small_tiles
    a = tdp.Tile.load(pn)
    p2 = str(tmp_path / "again.dt")
    a.save(p2)
    assert open(pn, "rb").read() == open(p2, "rb").read()
    b = tdp.Tile.load(p2)
    A, B = a.arrays(), b.arrays()
    for k in A:
        assert np.array_equal(A[k], B[k]), k
    ot = orc.OracleTile(pn)
    assert ot.n == a.n_obj
    box = np.zeros(9, np.float32)
    for i in range(a.n_obj):
        orc.lib().orc_tile_object_box(ot.h, i, orc.ptr(box))
        assert np.array_equal(box[:6], A["obj_box"][6 * i:6 * i + 6])
        assert np.array_equal(box[6:], A["obj_transform"][3 * i:3 * i + 3])


def test_raw_reader_rejects_bad_files(tdp, tmp_path):
    p = tmp_path / "comp.dt"
    p.write_bytes(b"\x00" + b"\x01" * 64)
    with pytest.raises(tdp.TdpError, match="COMPRESSED"):
        tdp.Tile.load(str(p))
    p = tmp_path / "trunc.dt"
    p.write_bytes(b"\x01" + (1000).to_bytes(8, "little") + b"\x00" * 10)
    with pytest.raises(tdp.TdpError):
        tdp.Tile.load(str(p))
    with pytest.raises(tdp.TdpError):
        tdp.Tile.load(str(tmp_path / "missing.dt"))


def _raw_object(dsize, nvox, data=b"", voxels=b""):
    """one RAW object record with the given header counts"""
    return (dsize.to_bytes(8, "little") + data + nvox.to_bytes(8, "little") + b"\x00" * 40 +
            voxels + b"\x00" * 12)


def test_raw_reader_rejects_counts_that_overflow(tdp, tmp_path):
    """sizes read from a corrupt file are checked against the bytes left, in
    a form that a huge count cannot wrap (2^64 - 8 bytes of data, a voxel
    count beyond the file, a voxel block past the object's data)"""
    cases = {
        "huge_dsize": b"\x01" + (2**64 - 8).to_bytes(8, "little") + b"\x00" * 200,
        "huge_nvox": b"\x01" + _raw_object(0, 2**62),
        "nvox_past_end": b"\x01" + _raw_object(0, 3, voxels=b"\x00" * 116),
    }
    # a voxel whose LOD-20 block claims 2^62 triangles from offset 0
    vox = np.zeros(9, np.float32).tobytes() + np.array([0, 2**62] + [0, 0] * 4, np.uint64).tobytes()
    cases["huge_count"] = b"\x01" + _raw_object(44, 1, data=b"\x00" * 44, voxels=vox)
    vox = np.zeros(9, np.float32).tobytes() + np.array([2**64 - 4, 1] + [0, 0] * 4, np.uint64).tobytes()
    cases["huge_offset"] = b"\x01" + _raw_object(44, 1, data=b"\x00" * 44, voxels=vox)
    for name, blob in cases.items():
        p = tmp_path / f"{name}.dt"
        p.write_bytes(blob)
        with pytest.raises(tdp.TdpError, match="truncated|outside|too many|beyond"):
            tdp.Tile.load(str(p))
    # a well-formed one-voxel object with one triangle at LOD 20 still loads
    vox = np.zeros(9, np.float32).tobytes() + np.array([0, 1] + [0, 0] * 4, np.uint64).tobytes()
    p = tmp_path / "ok.dt"
    p.write_bytes(b"\x01" + _raw_object(44, 1, data=b"\x00" * 44, voxels=vox))
    t = tdp.Tile.load(str(p))
    assert (t.n_obj, t.n_vox, t.n_tri) == (1, 1, 1)


def test_upload_rejects_voxels_over_65535_triangles(tdp):
    """the pair walks number a voxel pair's n1 x n2 triangle pairs in 32 bits,
    so a voxel of more than 65535 triangles at one LOD is refused with a named
    error before anything reaches the device (checked here on a context whose
    device open failed: the check runs before any HIP call)"""
    import ctypes as c
    _, lib = tdp._load()
    ctx = c.c_void_p()
    lib.tdp_open(0, c.byref(ctx))
    assert ctx.value, "tdp_open returns its context even when it fails"

    class Desc(c.Structure):
        _fields_ = [("n_obj", c.c_size_t), ("n_vox", c.c_size_t), ("n_tri", c.c_size_t)] + \
                   [(f, c.c_void_p) for f in ("obj_vox_off", "obj_hd", "vox_box", "vox_tri_off",
                                              "vox_tri_cnt", "tri", "obj_box", "obj_transform")]
    n_tri = 70000
    off = np.array([0, 2], np.uint32)
    hd = np.zeros(10, np.float32)
    vb = np.zeros(18, np.float32)
    cnt = np.array([70000, 0, 0, 0, 0, 1, 1, 1, 1, 1], np.uint32)
    first = np.zeros(10, np.uint64)
    tri = np.zeros(12 * n_tri, np.float32)
    d = Desc(1, 2, n_tri, *[a.ctypes.data for a in (off, hd, vb, first, cnt, tri)], None, None)
    lib.tdp_tile_upload.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p]
    out = c.c_void_p()
    assert lib.tdp_tile_upload(ctx, c.byref(d), c.byref(out)) == -1  # TDP_ERR_ARG
    assert b"65535" in lib.tdp_last_error(ctx)
    cnt[0] = 65535
    cnt[5] = 70000  # second voxel, LOD 20: records outside the array
    assert lib.tdp_tile_upload(ctx, c.byref(d), c.byref(out)) == -1
    assert b"65535" in lib.tdp_last_error(ctx)
    lib.tdp_close(ctx)


def test_host_code_under_address_and_undefined_sanitizers(tmp_path):
    """the host engine's CPU code (generator, RAW reader and writer, object
    source, octree) built with -fsanitize=address,undefined from its sources
    and run over intact, truncated and corrupted tiles and a degenerate box
    set (tests/host_sanitize_driver.cpp); any memory error or undefined
    behaviour aborts it"""
    src = os.path.join(ROOT, "3dpipe_amd", "csrc", "host")
    exe = str(tmp_path / "drv")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-fopenmp",
                    "-o", exe, os.path.join(ROOT, "tests", "host_sanitize_driver.cpp"),
                    os.path.join(src, "tdb_tile.cpp"), os.path.join(src, "tdb_octree.cpp")],
                   check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    p = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "0 failures" in p.stdout


def test_generator_shapes(tdp):
    n, v = tdp.generate(seed=0, n_vessels=10, nuclei_per_vessel=20, cube=2000.0)
    assert (n.n_obj, v.n_obj) == (200, 10)
    an, av = n.arrays(), v.arrays()
    cnt = an["vox_tri_cnt"].reshape(-1, 5)
    per_obj = np.add.reduceat(cnt, an["obj_vox_off"][:-1], axis=0)
    assert np.all(per_obj == [257, 512, 767, 1025, 1280])   # nested icosphere LOD ladder
    nv = np.diff(an["obj_vox_off"])
    assert np.all(nv == 26)                                  # 1280/50 + 1 voxels
    cntv = av["vox_tri_cnt"].reshape(-1, 5)
    assert np.all(np.add.reduceat(cntv, av["obj_vox_off"][:-1], axis=0)[:, 4] == 4608)
    # deterministic for a seed
    n2, _ = tdp.generate(seed=0, n_vessels=10, nuclei_per_vessel=20, cube=2000.0)
    assert np.array_equal(n2.arrays()["tri"], an["tri"])
    # cores inside their voxel boxes
    vb = an["vox_box"].reshape(-1, 9)
    assert np.all(vb[:, 6:9] >= vb[:, 0:3]) and np.all(vb[:, 6:9] <= vb[:, 3:6])


def test_generator_bounds_cover_lod100_surface(tdp):
    """every LOD-100 vertex of a voxel lies within the proxy bound of some LOD
    triangle of the same voxel (what the refinement's lower bounds rely on)"""
    n, _ = tdp.generate(seed=5, n_vessels=2, nuclei_per_vessel=3, cube=500.0)
    a = n.arrays()
    rec = a["tri"].reshape(-1, 12).astype(np.float64)
    off = a["vox_tri_off"].reshape(-1, 5)
    cnt = a["vox_tri_cnt"].reshape(-1, 5)

    def point_tri(p, a, b, c):
        # closest point by projection onto the plane, else onto the edges
        ab, ac = b - a, c - a
        nrm = np.cross(ab, ac)
        q = p - np.dot(p - a, nrm) / np.dot(nrm, nrm) * nrm
        w = np.linalg.lstsq(np.stack([ab, ac], 1), q - a, rcond=None)[0]
        if w[0] >= 0 and w[1] >= 0 and w[0] + w[1] <= 1:
            return np.linalg.norm(p - q)

        def seg(p, x, y):
            t = np.clip(np.dot(p - x, y - x) / np.dot(y - x, y - x), 0, 1)
            return np.linalg.norm(p - (x + t * (y - x)))
        return min(seg(p, a, b), seg(p, b, c), seg(p, c, a))

    for v in range(0, n.n_vox, 7):
        full = rec[off[v, 4]:off[v, 4] + cnt[v, 4]]
        for s in range(4):
            lod = rec[off[v, s]:off[v, s] + cnt[v, s]]
            if len(lod) == 0:
                continue
            for t in full[:: max(1, len(full) // 8)]:
                for p in (t[0:3], t[4:7], t[8:11]):
                    gap = min(point_tri(p, l[0:3], l[4:7], l[8:11]) - l[3] for l in lod)
                    assert gap <= 1e-3


def test_host_octree_matches_oracle(tdp, small_tiles):
    """candidate lists of the product's octree equal the oracle's (order
    included) for kNN, within and a self-join"""
    pn, pv = small_tiles
    n, v = tdp.Tile.load(pn), tdp.Tile.load(pv)
    on, ov = orc.OracleTile(pn), orc.OracleTile(pv)
    cap = 4096
    for (t1, t2, o1, o2, self_) in ((n, v, on, ov, False), (n, None, on, on, True)):
        for query, k, tau in (("nn", 3, 0.0), ("nn", 10, 0.0), ("within", 1, 400.0)):
            prefix, ids = tdp.candidates(t1, t2, query=query, knn=k, within_dist=tau)
            box = np.zeros(9, np.float32)
            for i in range(t1.n_obj):
                orc.lib().orc_tile_object_box(o1.h, i, orc.ptr(box))
                ci = np.zeros(cap, np.int32); mn = np.zeros(cap, np.float32); mx = np.zeros(cap, np.float32)
                if query == "nn":
                    c = orc.lib().orc_query_knn(o2.h, orc.ptr(box[:3]), orc.ptr(box[3:6]),
                                                orc.ptr(box[6:]), i if self_ else -1, k,
                                                orc.ptr(ci), orc.ptr(mn), orc.ptr(mx), cap)
                else:
                    c = orc.lib().orc_query_within(o2.h, orc.ptr(box[:3]), orc.ptr(box[3:6]),
                                                   orc.ptr(box[6:]), i if self_ else -1, tau,
                                                   orc.ptr(ci), orc.ptr(mn), orc.ptr(mx), cap)
                assert ids[prefix[i]:prefix[i + 1]].tolist() == ci[:c].tolist(), (query, i)


def test_host_octree_built_by_subtrees_matches_oracle(tdp, tmp_path):
    """a tile of 60,000 objects (octree subtrees built side by side): kNN and
    within lists, order included, equal the oracle octree's, which inserts
    the objects one by one"""
    n, _ = tdp.generate(seed=11, n_vessels=0, nuclei_per_vessel=60000, cube=12000.0, subdiv=0,
                        nucleus_r_min=6.0, nucleus_r_max=14.0)
    assert n.n_obj >= 50000
    pth = str(tmp_path / "n60k.dt")
    n.save(pth)
    on = orc.OracleTile(pth)
    cap = 4096
    box = np.zeros(9, np.float32)
    ci = np.zeros(cap, np.int32); mn = np.zeros(cap, np.float32); mx = np.zeros(cap, np.float32)
    for query, k, tau in (("nn", 10, 0.0), ("within", 1, 150.0)):
        prefix, ids = tdp.candidates(n, None, query=query, knn=k, within_dist=tau)
        for i in range(0, n.n_obj, 211):
            orc.lib().orc_tile_object_box(on.h, i, orc.ptr(box))
            if query == "nn":
                c = orc.lib().orc_query_knn(on.h, orc.ptr(box[:3]), orc.ptr(box[3:6]),
                                            orc.ptr(box[6:]), i, k, orc.ptr(ci), orc.ptr(mn),
                                            orc.ptr(mx), cap)
            else:
                c = orc.lib().orc_query_within(on.h, orc.ptr(box[:3]), orc.ptr(box[3:6]),
                                               orc.ptr(box[6:]), i, tau, orc.ptr(ci),
                                               orc.ptr(mn), orc.ptr(mx), cap)
            assert ids[prefix[i]:prefix[i + 1]].tolist() == ci[:c].tolist(), (query, i)


def test_cli_without_gpu_fails_loudly(small_tiles):
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is present")
    exe = os.path.join(ROOT, "3dpipe_amd", "bin", "tdbase")
    pn, pv = small_tiles
    r = subprocess.run([exe, "join", "--tile1", pn, "--tile2", pv, "-q", "nn", "-k", "3",
                        "--lod", "20", "40", "-g"], capture_output=True, text=True)
    assert r.returncode == 2 and "GPU unavailable" in r.stderr
    assert "ignoring unknown option --lod" in r.stderr
    r = subprocess.run([exe, "join", "-q", "nn"], capture_output=True, text=True)
    assert r.returncode == 1 and "tile1" in r.stderr


# ------------------------------------------------------- operation counts

def _flops_lib():
    import ctypes
    path = os.path.join(ROOT, "oracle", "lib", "libflops.so")
    if not os.path.exists(path):
        pytest.skip("oracle/lib/libflops.so not built")
    lib = ctypes.CDLL(path)
    lib.flop_tridist.restype = ctypes.c_uint64
    lib.flop_tridist.argtypes = [ctypes.c_void_p] * 4
    return lib


def _count(lib, S, T):
    S = np.asarray(S, np.float64)
    T = np.asarray(T, np.float64)
    d = np.zeros(1)
    early = np.zeros(1, np.int32)
    n = lib.flop_tridist(S.ctypes.data, T.ctypes.data, d.ctypes.data, early.ctypes.data)
    return int(n), float(d[0]), bool(early[0])


def test_flop_count_of_the_canonical_full_path():
    """F_TRI (the per-pair algorithmic FP32 work in the throughput figures) is
    the counted cost of TriDist on the full path -- all 9 edge-edge cases
    without an early return, then both vertex-face tests, the second one
    deciding -- plus the 8 operations of the bound update."""
    import importlib.util
    lib = _flops_lib()
    # a small tilted S under a large horizontal T: S's face test finds T's
    # vertices on both sides of its plane, T's face test finds S's top vertex
    S = [1, 1, 3, 2, 1, -1, 1, 2, -1]
    T = [-10, -10, 5, 20, -10, 5, -10, 20, 5]
    n, d, early = _count(lib, S, T)
    assert d == 2.0 and not early
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    assert bench.F_TRI == n + 8
    # an early edge-edge return costs less; no random pair costs far more
    n_edge, d_edge, early_edge = _count(lib, [0, 0, 0, 1, 0, 0, 0, 1, 0],
                                        [0.5, -1, 2, 0.5, 1, 2, 3, 0, 2])
    assert early_edge and n_edge < n
    rng = np.random.default_rng(3)
    counts = [_count(lib, rng.uniform(-5, 5, 9), rng.uniform(-5, 5, 9))[0] for _ in range(500)]
    assert max(counts) < 1.2 * n
