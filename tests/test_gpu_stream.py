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


@pytest.mark.parametrize("query,k,tau,overlap,share", [("nn", 5, 0.0, True, 0.5),
                                                       ("nn", 10, 0.0, False, 0.5),
                                                       ("within", 1, 120.0, True, 0.5),
                                                       ("nn", 3, 0.0, True, 0.9)])
def test_stream_self_join_equals_join(tdp, nuclei, query, k, tau, overlap, share):
    e = tdp.Engine(0)
    want = e.join(nuclei, None, query=query, knn=k, within_dist=tau, precision="mixed")
    src = tdp.Source.generate(**NUC)
    budget = 64 * MB
    got = e.stream_join(src, None, query=query, knn=k, within_dist=tau, precision="mixed",
                        hbm_budget=budget, tile_share=share, overlap=overlap)
    e.close()
    st = got.stream
    print(query, k, {x: st[x] for x in ("chunks", "splits", "overlap", "peak_device_bytes")})
    assert st["chunks"] > 4
    assert st["peak_device_bytes"] <= budget
    assert got.n_targets == want.n_targets == nuclei.n_obj
    assert len(got.target) > 0
    assert _same(got, want)


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
