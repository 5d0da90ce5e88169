"""Pins of the oracle's join drivers and voxel bound by brute force and
closed forms (CPU only).

The reference's join drivers (KNNJoin.cpp, WithinJoin.cpp, SpatialJoin.cpp)
need CGAL and cannot be compiled here, so the oracle's restatement of them is
checked against what the mathematics fixes on small tiles:

* within at the single LOD 100: an object pair is reported exactly when its
  voxels' nearest core points are within tau (the filter's object-level
  report, WithinJoin.cpp:164-252) or its all-pairs MeshDist is (the refine
  decision, WithinJoin.cpp:769-792); the reported range is the filter's
  [box min, core-core] aggregate or [d, d];
* kNN at the single LOD 100: the reported objects are the k nearest by
  all-pairs MeshDist, with the filter's or [d, d] ranges, filter reports
  first, each group in candidate-list order (KNNJoin.cpp:133-202, 497-851,
  1524-1811); over the full LOD list 20..100 the same objects are reported;
* the voxel bound (voxel_dist_core, computing_gpu.cu:1606-1631): the
  Euclidean distance between two boxes and the distance between their core
  points, on boxes whose answers are integers.

MeshDist itself comes from the oracle's meshdist, which equals the reference's
compiled MeshDist bit for bit (test_cpu.py); box distances are recomputed here
with numpy in the reference's operation order."""
import numpy as np
import pytest

import orc

LOD100 = 4  # LOD slot of LOD 100


# ---------------------------------------------------------------- voxel bound


def _vb(prec, a, b):
    out = np.zeros(2, np.float32 if prec == "f" else np.float64)
    fn = orc.lib().orc_voxel_bound_f if prec == "f" else orc.lib().orc_voxel_bound_d
    fn(orc.ptr(np.asarray(a, np.float32)), orc.ptr(np.asarray(b, np.float32)), orc.ptr(out))
    return out


@pytest.mark.parametrize("prec", ["f", "d"])
def test_voxel_bound_closed_form(prec):
    """min = Euclidean distance between the boxes (0 when they touch or
    overlap), max = |core1 - core2|; integer answers, so exact in both forms"""
    cases = [
        # box 1 (low, high, core), box 2, min, max
        ([0, 0, 0, 1, 1, 1, 0, 0, 0], [4, 5, 1, 6, 7, 2, 2, 3, 6], 5.0, 7.0),   # gap (3, 4, 0)
        ([4, 5, 1, 6, 7, 2, 2, 3, 6], [0, 0, 0, 1, 1, 1, 0, 0, 0], 5.0, 7.0),   # swapped
        ([0, 0, 0, 2, 2, 2, 1, 1, 1], [2, 0, 0, 3, 1, 1, 3, 4, 7], 0.0, 7.0),   # touching face
        ([0, 0, 0, 2, 2, 2, 1, 1, 1], [2, 2, 2, 3, 3, 3, 3, 3, 2], 0.0, 3.0),   # corner
        ([0, 0, 0, 4, 4, 4, 1, 2, 3], [1, 1, 1, 2, 2, 2, 1, 2, 3], 0.0, 0.0),   # nested, same core
        ([0, 0, 0, 4, 4, 4, 0, 0, 0], [-9, 2, 1, -1, 3, 2, -2, 1, 2], 1.0, 3.0),  # x gap 1
        ([0, 0, 0, 1, 1, 1, 0, 0, 0], [-13, -1, -2, -12, 0, 0, 3, 4, 12], 12.0, 13.0),
        ([0, 0, 0, 1, 1, 1, 0, 0, 0], [4, 5, 13, 5, 6, 14, 0, 0, 0], 13.0, 0.0),  # (3, 4, 12) gap
    ]
    for a, b, mn, mx in cases:
        got = _vb(prec, a, b)
        want = np.array([mn, mx], got.dtype)
        assert np.array_equal(got, want), (a, b, got, want)
        # symmetric in its arguments
        assert np.array_equal(_vb(prec, b, a), got)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_voxel_bound_is_box_distance(prec):
    """random boxes: min equals the distance between the nearest points of
    the two boxes (clamping one box's corner region into the other), max the
    core distance; both within rounding of an independent formulation"""
    rng = np.random.default_rng(5)
    for _ in range(3000):
        a = np.zeros(9, np.float32)
        b = np.zeros(9, np.float32)
        for v in (a, b):
            lo = rng.uniform(-50, 50, 3)
            hi = lo + rng.uniform(0, 20, 3)
            v[:3], v[3:6] = lo, hi
            v[6:] = lo + (hi - lo) * rng.uniform(0, 1, 3)
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
        gap = np.maximum(0, np.maximum(a64[:3] - b64[3:6], b64[:3] - a64[3:6]))
        got = _vb(prec, a, b)
        tol = 1e-5 if prec == "f" else 1e-12
        assert abs(got[0] - np.linalg.norm(gap)) <= tol * (1 + np.linalg.norm(gap))
        dc = np.linalg.norm(a64[6:] - b64[6:])
        assert abs(got[1] - dc) <= tol * (1 + dc)


# ---------------------------------------------------------------- joins


def _aggregates(A, B, prec, self_join):
    """per object pair: all-pairs MeshDist at LOD 100 (inf when a side has no
    triangles there), the minimum voxel box distance and the minimum
    core-core distance over all voxel pairs"""
    dt = np.float32 if prec == "f32" else np.float64
    nA, nB = len(A["obj_vox_off"]) - 1, len(B["obj_vox_off"]) - 1
    M = np.full((nA, nB), np.inf)
    bmin = np.full((nA, nB), np.inf)
    cc = np.full((nA, nB), np.inf)
    recA, recB = A["tri"].reshape(-1, 12), B["tri"].reshape(-1, 12)
    rec = np.concatenate([recA, recB])
    tri9 = np.ascontiguousarray(rec[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]])
    hd = np.ascontiguousarray(rec[:, [3, 7]])
    vA, vB = A["vox_box"].reshape(-1, 9).astype(dt), B["vox_box"].reshape(-1, 9).astype(dt)
    offA, cntA = A["vox_tri_off"].reshape(-1, 5)[:, LOD100], A["vox_tri_cnt"].reshape(-1, 5)[:, LOD100]
    offB, cntB = B["vox_tri_off"].reshape(-1, 5)[:, LOD100], B["vox_tri_cnt"].reshape(-1, 5)[:, LOD100]
    for a in range(nA):
        xa = np.arange(A["obj_vox_off"][a], A["obj_vox_off"][a + 1])
        for b in range(nB):
            if self_join and a == b:
                continue
            xb = np.arange(B["obj_vox_off"][b], B["obj_vox_off"][b + 1])
            X, Y = np.repeat(xa, len(xb)), np.tile(xb, len(xa))
            # the reference's order: axes x, y, z accumulated one at a time
            s = np.zeros(len(X), dt)
            c2 = np.zeros(len(X), dt)
            for i in range(3):
                t1 = vA[X, i] - vB[Y, 3 + i]
                t2 = vA[X, 3 + i] - vB[Y, i]
                s = s + np.where(t2 < 0, t2 * t2, np.where(t1 > 0, t1 * t1, dt(0)))
                dc = vA[X, 6 + i] - vB[Y, 6 + i]
                c2 = c2 + dc * dc
            bmin[a, b] = np.sqrt(s).min()
            cc[a, b] = np.sqrt(c2).min()
            keep = (cntA[X] > 0) & (cntB[Y] > 0)
            if keep.any():
                os_ = np.stack([offA[X[keep]], cntA[X[keep]], len(recA) + offB[Y[keep]],
                                cntB[Y[keep]]], 1).astype(np.uint64)
                M[a, b] = orc.meshdist(tri9, hd, os_, prec)[:, 0].min()
    return M, bmin, cc


@pytest.fixture(scope="module")
def tiny(tdp, tiny_tiles):
    pn, pv = tiny_tiles
    n, v = tdp.Tile.load(pn), tdp.Tile.load(pv)
    return {"on": orc.OracleTile(pn), "ov": orc.OracleTile(pv), "A": n.arrays(), "B": v.arrays()}


@pytest.fixture(scope="module")
def coarse(tdp, tmp_path_factory):
    """the tiny tiles' objects with voxels of about 100 triangles: core points
    lie far from the nearest surface points, so the LOD-20 bounds decide
    object pairs at every Hausdorff level"""
    d = tmp_path_factory.mktemp("coarse")
    n, v = tdp.generate(seed=3, n_vessels=4, nuclei_per_vessel=5, cube=600.0, subdiv=2,
                        vessel_sides=8, rings_min=2, rings_max=2, prototypes=8, voxel_tris=100,
                        place_min=20.0, place_max=120.0)
    pn, pv = str(d / "nuclei.dt"), str(d / "vessels.dt")
    n.save(pn)
    v.save(pv)
    return {"on": orc.OracleTile(pn), "ov": orc.OracleTile(pv), "A": n.arrays(), "B": v.arrays()}


@pytest.fixture(scope="module")
def agg(tiny):
    out = {}
    for prec in ("f32", "f64"):
        for self_join in (False, True):
            B = tiny["A"] if self_join else tiny["B"]
            out[prec, self_join] = _aggregates(tiny["A"], B, prec, self_join)
    return out


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("self_join", [False, True])
def test_oracle_within_equals_bruteforce_at_lod100(tiny, agg, prec, self_join):
    M, bmin, cc = agg[prec, self_join]
    finite = M[np.isfinite(M)]
    # fixed distances and distances that split the object pairs
    taus = [0.0, 5.0, 20.0, 60.0, 150.0] + [float(q) for q in np.quantile(finite, [0.05, 0.3, 0.6])]
    seen = set()
    for tau in taus:
        t, o, mn, mx, _ = orc.join(tiny["on"], None if self_join else tiny["ov"], "within", 1,
                                   tau, lods=(100,), precision=prec)
        T = (np.float32(tau) if prec == "f32" else np.float64(np.float32(tau)))
        want = {(a, b) for a in range(M.shape[0]) for b in range(M.shape[1])
                if not (self_join and a == b) and (cc[a, b] <= T or M[a, b] <= T)}
        got = list(zip(t.tolist(), o.tolist()))
        assert len(got) == len(set(got)), "an object pair reported twice"
        assert set(got) == want, (tau, set(got) ^ want)
        for i, (a, b) in enumerate(got):
            if cc[a, b] <= T:   # decided by the filter
                assert (mn[i], mx[i]) == (bmin[a, b], cc[a, b]), (tau, a, b)
            else:               # decided at LOD 100: exact distance
                assert mn[i] == M[a, b] and mx[i] == M[a, b], (tau, a, b)
        seen.add(len(got))
    # the distances chosen really split the pairs: several different counts
    assert len(seen) >= 4


def _cand_order(tiny, self_join, k):
    """the octree's kNN candidate lists (query_knn_fast order) of every target"""
    A = tiny["A"]
    B = A if self_join else tiny["B"]
    box, tr = B["obj_box"].reshape(-1, 6), B["obj_transform"].reshape(-1, 3)
    low, high = np.ascontiguousarray(box[:, :3]), np.ascontiguousarray(box[:, 3:])
    h = orc.lib().orc_boxset_build(len(box), orc.ptr(low), orc.ptr(high), orc.ptr(tr), 10)
    qb, qt = A["obj_box"].reshape(-1, 6), A["obj_transform"].reshape(-1, 3)
    out = []
    cap = 4096
    for i in range(len(qb)):
        ids = np.zeros(cap, np.int32)
        mn = np.zeros(cap, np.float32)
        mx = np.zeros(cap, np.float32)
        c = orc.lib().orc_query_knn(h, orc.ptr(np.ascontiguousarray(qb[i, :3])),
                                    orc.ptr(np.ascontiguousarray(qb[i, 3:])),
                                    orc.ptr(np.ascontiguousarray(qt[i])), i if self_join else -1,
                                    k, orc.ptr(ids), orc.ptr(mn), orc.ptr(mx), cap)
        out.append(ids[:c].tolist())
    return out


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("self_join", [False, True])
def test_oracle_knn_equals_bruteforce_ranking(tiny, agg, prec, self_join):
    M, bmin, cc = agg[prec, self_join]
    checked = 0
    for K in (1, 2, 3, 5):
        t, o, mn, mx, _ = orc.join(tiny["on"], None if self_join else tiny["ov"], "nn", K, 0,
                                   lods=(100,), precision=prec)
        tf, of, _, _, _ = orc.join(tiny["on"], None if self_join else tiny["ov"], "nn", K, 0,
                                   precision=prec)
        lists = orc.group(t, o, mn, mx)
        full = orc.group(tf, of)
        cands = _cand_order(tiny, self_join, K)
        for a in range(M.shape[0]):
            row = M[a].copy()
            if self_join:
                row[a] = np.inf
            rank = np.argsort(row, kind="stable")
            kk = min(K, int(np.isfinite(row).sum()))
            if kk < len(row) and row[rank[kk - 1]] == row[rank[kk]]:
                continue  # a tie at the k-th place: either object is right
            want = set(rank[:kk].tolist())
            got = lists.get(a, [])
            assert {e[0] for e in got} == want, (K, a)
            assert {e[0] for e in full.get(a, [])} == want, ("LODs 20..100", K, a)
            # ranges: the filter's aggregate, or [d, d] at LOD 100; filter
            # reports come first, each group in candidate-list order
            groups = []
            for ob, lo, hi in got:
                if (lo, hi) == (bmin[a, ob], cc[a, ob]):
                    groups.append(0)
                else:
                    assert lo == M[a, ob] and hi == M[a, ob], (K, a, ob)
                    groups.append(1)
            assert groups == sorted(groups), (K, a)
            pos = {ob: p for p, ob in enumerate(cands[a])}
            for g in (0, 1):
                seq = [pos[e[0]] for e, gg in zip(got, groups) if gg == g]
                assert seq == sorted(seq), (K, a, g)
            checked += 1
    assert checked >= 40


# ---------------------------------------------------------------- kNN status


def _status(ranges, cand_left, fp64):
    r = np.ascontiguousarray(ranges, np.float64).reshape(-1, 2)
    out = np.full(len(r), -1, np.int32)
    orc.lib().orc_knn_status(orc.ptr(r), len(r), int(cand_left), int(fp64), orc.ptr(out))
    return out.tolist()


@pytest.mark.parametrize("fp64", [0, 1])
def test_knn_status_known_answers(fp64):
    """evaluate_knn_kernel (computing_gpu.cu:105-151): sure_i counts the j
    with min_i >= max_j, maybe_i the j without max_i <= min_j; CONFIRMED if
    maybe_i < cand_left, else REMOVE if sure_i >= cand_left, else KEEP. The
    answers were worked out by hand from that rule; the touching ranges make
    >= against > and <= against < visible."""
    C, R, K = 1, 2, 0
    chain = [[1, 2], [2, 3], [3, 4]]
    assert _status(chain, 1, fp64) == [C, R, R]
    assert _status(chain, 2, fp64) == [C, C, R]
    assert _status(chain, 3, fp64) == [C, C, C]
    assert _status([[0, 5], [1, 6], [2, 7], [10, 11]], 2, fp64) == [K, K, K, R]
    assert _status([[0, 5], [1, 6], [2, 7], [10, 11]], 4, fp64) == [C, C, C, C]
    assert _status([[3, 3], [3, 3]], 1, fp64) == [C, C]
    assert _status([[5, 9]], 1, fp64) == [C]
    assert _status([[5, 9]], 0, fp64) == [R]
    assert _status([[0, 4], [4, 8], [4, 8]], 1, fp64) == [C, R, R]
    assert _status([[0, 4], [3, 8], [5, 9]], 1, fp64) == [K, K, R]


# ---------------------------------------------------------------- within, two LODs


def _vp_values(A, B, a, b, prec):
    """every voxel pair of objects (a, b): box distance, core distance,
    MeshDist (d, lo, hi) at LOD 20 and d at LOD 100 (nan where a side is
    empty at that LOD)"""
    dt = np.float32 if prec == "f32" else np.float64
    xa = np.arange(A["obj_vox_off"][a], A["obj_vox_off"][a + 1])
    xb = np.arange(B["obj_vox_off"][b], B["obj_vox_off"][b + 1])
    X, Y = np.repeat(xa, len(xb)), np.tile(xb, len(xa))
    vA, vB = A["vox_box"].reshape(-1, 9).astype(dt), B["vox_box"].reshape(-1, 9).astype(dt)
    s = np.zeros(len(X), dt)
    c2 = np.zeros(len(X), dt)
    for i in range(3):
        t1 = vA[X, i] - vB[Y, 3 + i]
        t2 = vA[X, 3 + i] - vB[Y, i]
        s = s + np.where(t2 < 0, t2 * t2, np.where(t1 > 0, t1 * t1, dt(0)))
        dc = vA[X, 6 + i] - vB[Y, 6 + i]
        c2 = c2 + dc * dc
    recA, recB = A["tri"].reshape(-1, 12), B["tri"].reshape(-1, 12)
    rec = np.concatenate([recA, recB])
    tri9 = np.ascontiguousarray(rec[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]])
    hd = np.ascontiguousarray(rec[:, [3, 7]])
    out = {"bmin": np.sqrt(s), "cc": np.sqrt(c2)}
    for slot, key in ((0, "r20"), (4, "r100")):
        oa, ca = A["vox_tri_off"].reshape(-1, 5)[X, slot], A["vox_tri_cnt"].reshape(-1, 5)[X, slot]
        ob, cb = B["vox_tri_off"].reshape(-1, 5)[Y, slot], B["vox_tri_cnt"].reshape(-1, 5)[Y, slot]
        res = np.full((len(X), 3), np.nan, dt)
        ne = (ca > 0) & (cb > 0)
        if ne.any():
            os_ = np.stack([oa[ne], ca[ne], len(recA) + ob[ne], cb[ne]], 1).astype(np.uint64)
            res[ne] = orc.meshdist(tri9, hd, os_, prec)
        out[key] = res
    return out


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_oracle_within_two_lods_by_level(coarse, prec, level):
    """within over LODs (20, 100): the LOD-20 round folds each voxel pair's
    range with the level's bounds (KNNJoin.cpp:1670-1686,
    WithinJoin.cpp:600-618): level 2 [min(lo), min(hi)] of the per-triangle
    Hausdorff intervals, level 1 d less / plus the objects' Hausdorff /
    proxy values, level 0 only max <- min(max, d); the object range is
    [min of the pairs' min, min of the pairs' max], reported when max <= tau;
    pairs whose min exceeds the object's max are evicted before LOD 100,
    where every pair becomes [d, d]. Restated per object pair in numpy from
    those lines and compared with the oracle's report list and ranges."""
    rounds_seen = set()
    for self_join in (False, True):
        rounds_seen |= _within_two_lods(coarse, prec, level, self_join)
    # reports from the filter, the LOD-20 round and the LOD-100 round
    assert rounds_seen == {0, 1, 2}


def _within_two_lods(tiny, prec, level, self_join):
    A = tiny["A"]
    B = A if self_join else tiny["B"]
    dt = np.float32 if prec == "f32" else np.float64
    hA, hB = A["obj_hd"].reshape(-1, 5, 2), B["obj_hd"].reshape(-1, 5, 2)
    vals = {(a, b): _vp_values(A, B, a, b, prec)
            for a in range(len(A["obj_vox_off"]) - 1) for b in range(len(B["obj_vox_off"]) - 1)
            if not (self_join and a == b)}
    rounds_seen = set()
    for tau in (10.0, 20.0, 30.0, 45.0, 60.0, 90.0, 120.0, 200.0):
        T = dt(np.float32(tau))
        want = {}
        for (a, b), v in vals.items():
            if v["cc"].min() <= T:
                want[a, b] = (v["bmin"].min(), v["cc"].min(), 0)
                continue
            if v["bmin"].min() > T:
                continue
            keep = v["bmin"] <= T
            mn, mx = v["bmin"][keep].copy(), v["cc"][keep].copy()
            r20 = v["r20"][keep]
            ne = ~np.isnan(r20[:, 0])
            d = r20[:, 0]
            if level == 2:
                lo, hi = r20[:, 1], r20[:, 2]
            elif level == 1:
                lo = (d - dt(hA[a, 0, 0])) - dt(hB[b, 0, 0])
                hi = (d + dt(hA[a, 0, 1])) + dt(hB[b, 0, 1])
            else:
                lo, hi = mn, d
            mn = np.where(ne, np.minimum(mn, lo), mn)
            mx = np.where(ne, np.minimum(mx, hi), mx)
            if mx.min() <= T:
                want[a, b] = (mn.min(), mx.min(), 1)
                continue
            surv = ~(mn > mx.min())
            d100 = v["r100"][keep][surv, 0]
            ne = ~np.isnan(d100)
            mn2 = np.where(ne, d100, mn[surv])
            mx2 = np.where(ne, d100, mx[surv])
            if mx2.min() <= T:
                want[a, b] = (mn2.min(), mx2.min(), 2)
        t, o, rmn, rmx, _ = orc.join(tiny["on"], None if self_join else tiny["ov"], "within", 1,
                                     tau, lods=(20, 100), level=level, precision=prec)
        got = {(int(a), int(b)): (rmn[i], rmx[i]) for i, (a, b) in enumerate(zip(t, o))}
        assert len(got) == len(t)
        assert set(got) == set(want), (tau, set(got) ^ set(want))
        for key, (lo, hi, rnd) in want.items():
            assert got[key] == (lo, hi), (tau, key, rnd, got[key], (lo, hi))
            rounds_seen.add(rnd)
    return rounds_seen
