"""The serving loop (tdb_server_*) on an MI355X: queries merged into shared
joins return exactly what each query returns joined alone, for mixed k,
within distances and sizes, in the batched mixed form and in the
one-at-a-time FP64 form."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (size, query, k, tau)
QUERIES = [(7, "nn", 1, 0.0), (30, "nn", 10, 0.0), (3, "nn", 5, 0.0), (45, "nn", 20, 0.0),
           (12, "within", 1, 80.0), (25, "within", 1, 300.0), (18, "nn", 5, 0.0)]


def _query_tile(tdp, i, size):
    # nuclei near the 10 vessels of the resident tile (same vessel seed),
    # placements from their own seed
    n, _ = tdp.generate(seed=0, n_vessels=10, nuclei_per_vessel=max(1, -(-size // 10)),
                        cube=2000.0, max_nuclei=size, nuclei_seed=1000 + i, emit_vessels=0)
    return n


def _same(a, b):
    return (np.array_equal(a.target, b.target) and np.array_equal(a.obj2, b.obj2)
            and np.array_equal(a.rmin.view(np.uint64), b.rmin.view(np.uint64))
            and np.array_equal(a.rmax.view(np.uint64), b.rmax.view(np.uint64))
            and np.array_equal(a.round, b.round))


@pytest.fixture(scope="module")
def vessels(tdp):
    _, v = tdp.generate(seed=0, n_vessels=10, nuclei_per_vessel=1, cube=2000.0)
    return v


@pytest.mark.parametrize("precision,max_batch,overlap,lazy,executors", [
    ("mixed", 64, True, True, 0), ("mixed", 64, True, False, 0), ("mixed", 64, True, True, 3),
    ("mixed", 2, True, True, 1), ("fp64", 1, False, True, 0), ("fp64", 1, False, False, 0),
    ("ref32", 64, True, True, 2), ("mixed", 64, False, True, 0)])
def test_server_matches_joins_alone(tdp, vessels, precision, max_batch, overlap, lazy, executors):
    eng = tdp.Engine(0)
    alone = []
    for i, (size, q, k, tau) in enumerate(QUERIES):
        t = _query_tile(tdp, i, size)
        assert t.n_obj == size
        alone.append(eng.join(t, vessels, query=q, knn=k, within_dist=tau, precision=precision))
    eng.close()
    tiles = [_query_tile(tdp, i, size) for i, (size, _, _, _) in enumerate(QUERIES)]
    srv = tdp.Server(vessels, max_batch=max_batch, overlap=overlap, precision=precision,
                     lazy=lazy, executors=executors)
    # submitted back to back: the queries that wait together share joins
    tickets = [srv.submit(tiles[i], query=q, knn=k, within_dist=tau)
               for i, (size, q, k, tau) in enumerate(QUERIES)]
    for i, tk in enumerate(tickets):
        res, lat = srv.wait(tk)
        assert lat > 0
        assert res.n_targets == QUERIES[i][0]
        assert _same(res, alone[i]), f"query {i} differs from its join alone"
    st = srv.stats()
    srv.close()
    print(st)
    assert st["queries"] == len(QUERIES)
    assert st["executors"] == (1 if not overlap else (executors or 2))
    if max_batch == 1:
        assert st["batches"] == len(QUERIES) and st["max_batch"] == 1
    else:
        # all queries were waiting together: at least two shared a join
        assert st["max_batch"] >= 2 and st["batches"] < len(QUERIES)
    if max_batch > 1 and not overlap:
        # one batch at a time: the queries waiting behind the first form one
        # batch, kNN and within queries together
        assert st["mixed_kind_batches"] >= 1
    # lazily only the records of voxels some round read went over
    if lazy:
        assert 0 < st["h2d_record_bytes"] < st["record_bytes_of_queries"]
    else:
        assert st["h2d_record_bytes"] == st["record_bytes_of_queries"]


def test_server_budget_rejects_oversized_query(tdp, vessels):
    srv = tdp.Server(vessels, hbm_budget=1 << 20)  # below the resident tile itself
    with pytest.raises(tdp.TdpError):
        srv.submit(_query_tile(tdp, 0, 5), query="nn", knn=1)
    srv.close()
