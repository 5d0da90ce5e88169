"""The streaming join (tdb_stream_join) on an MI355X: chunks of targets
decoded on the host, copied into reserved device tiles and joined within a
device budget give exactly the result of one in-memory join of the same
objects -- self-joins and joins against a second source, kNN and within,
with the stages overlapped or in turn."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NUC = dict(seed=4, n_vessels=0, nuclei_per_vessel=600, cube=1500.0, nucleus_r_min=6.0,
           nucleus_r_max=14.0)
MB = 1 << 20


def _same(a, b):
    return (np.array_equal(a.target, b.target) and np.array_equal(a.obj2, b.obj2)
            and np.array_equal(a.rmin.view(np.uint64), b.rmin.view(np.uint64))
            and np.array_equal(a.rmax.view(np.uint64), b.rmax.view(np.uint64))
            and np.array_equal(a.round, b.round))


@pytest.fixture(scope="module")
def nuclei(tdp):
    n, _ = tdp.generate(**NUC)
    return n


@pytest.mark.parametrize("query,k,tau,overlap,share,lazy,workers,precision,tuning", [
    ("nn", 5, 0.0, True, 0.5, True, 0, "mixed", {}),
    ("nn", 5, 0.0, True, 0.5, False, 0, "mixed", {}),
    ("nn", 10, 0.0, False, 0.5, True, 0, "mixed", {}),
    ("nn", 10, 0.0, False, 0.5, False, 0, "mixed", {}),
    ("within", 1, 120.0, True, 0.5, True, 3, "mixed", {}),
    ("within", 1, 120.0, True, 0.5, False, 3, "mixed", {}),
    ("nn", 3, 0.0, True, 0.9, True, 1, "mixed", {}),
    ("nn", 5, 0.0, True, 0.5, True, 2, "mixed", dict(graph=1, drain=1, drain_blocks=512)),
    ("nn", 5, 0.0, True, 0.5, True, 2, "fp64", {}),
    ("within", 1, 90.0, True, 0.5, True, 2, "ref32", {}),
    ("intersect", 1, 0.0, True, 0.5, True, 2, "mixed", {})])
def test_stream_self_join_equals_join(tdp, nuclei, query, k, tau, overlap, share, lazy, workers,
                                      precision, tuning):
    e = tdp.Engine(0)
    if tuning:
        e.set_tuning(**tuning)
    want = e.join(nuclei, None, query=query, knn=k, within_dist=tau, precision=precision)
    src = tdp.Source.generate(**NUC)
    budget = 64 * MB
    got = e.stream_join(src, None, query=query, knn=k, within_dist=tau, precision=precision,
                        hbm_budget=budget, tile_share=share, overlap=overlap, lazy=lazy,
                        workers=workers)
    e.close()
    st = got.stream
    print(query, k, lazy, workers, {x: st[x] for x in (
        "chunks", "splits", "overlap", "peak_device_bytes", "record_bytes_all",
        "record_bytes_shipped", "meta_bytes", "h2d_bytes", "workers")})
    assert st["chunks"] > 4
    assert st["peak_device_bytes"] <= budget
    # as many workers as asked for (4 by default), fewer where a target with
    # its candidates would not fit their tile slots
    assert 1 <= st["workers"] <= (1 if not overlap else (workers or 4))
    assert got.n_targets == want.n_targets == nuclei.n_obj
    assert len(got.target) > 0
    assert _same(got, want)
    # kNN and within candidate lists from the device, intersection lists from
    # the octree
    if query == "intersect":
        assert st["candidate_host_targets"] == -1
    else:
        assert 0 <= st["candidate_host_targets"] < nuclei.n_obj // 50
    # what went over: the voxel and object arrays plus the records shipped;
    # lazily, only those of the voxels some round read
    assert st["h2d_bytes"] == st["meta_bytes"] + st["record_bytes_shipped"]
    if lazy:
        assert 0 < st["record_bytes_shipped"] < st["record_bytes_all"]
    else:
        assert st["record_bytes_shipped"] == st["record_bytes_all"]


@pytest.mark.parametrize("lazy", [True, False])
def test_stream_split_joins_share_the_tile(tdp, nuclei, lazy):
    """budgets whose join buffers force target ranges to be split: the
    ranges joined one after another on the same tile give the same result
    (lazily: the voxels an earlier range put stay held). Tile shares from
    large to small until one splits (a share that leaves too little for a
    one-target join is refused with an error)"""
    e = tdp.Engine(0)
    want = e.join(nuclei, None, query="nn", knn=10, precision="mixed")
    split = 0
    for share in (0.93, 0.9, 0.87, 0.84, 0.8, 0.75, 0.7):
        try:
            got = e.stream_join(tdp.Source.generate(**NUC), None, query="nn", knn=10,
                                precision="mixed", hbm_budget=64 * MB, tile_share=share,
                                lazy=lazy, workers=1)
        except tdp.TdpError as err:
            print(share, "refused:", err)
            continue
        print(share, got.stream)
        assert _same(got, want)
        if got.stream["splits"] > 0:
            split += 1
            break
    e.close()
    assert split, "no tile share split a chunk's join"


def test_stream_against_second_source(tdp):
    """targets from one source, candidates from another (tile source)"""
    n, v = tdp.generate(seed=0, n_vessels=10, nuclei_per_vessel=20, cube=2000.0)
    e = tdp.Engine(0)
    want = e.join(n, v, query="nn", knn=3, precision="fp64")
    got = e.stream_join(tdp.Source.of_tile(n), tdp.Source.of_tile(v), query="nn", knn=3,
                        precision="fp64", hbm_budget=48 * MB, max_chunk_targets=16)
    e.close()
    assert got.stream["chunks"] >= 200 // 16
    assert _same(got, want)


def test_stream_budget_too_small_is_an_error(tdp):
    e = tdp.Engine(0)
    with pytest.raises(tdp.TdpError):
        e.stream_join(tdp.Source.generate(**NUC), None, query="nn", knn=3, hbm_budget=1 * MB)
    e.close()
