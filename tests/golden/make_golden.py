"""Write the golden vectors in tests/golden from the reference's own sources,
compiled by oracle/Makefile (`make -C oracle ref`) into oracle/_ref/libtdref.so:
  tridist_ref.npz   TriDist (src/geometry/computing_cpu.cpp:403-439) on random
                    and degenerate triangle pairs
  meshdist_ref.npz  MeshDist (computing_cpu.cpp:441-481) per voxel pair
  octree_ref.npz    query_knn_fast candidate lists (src/index/octree.cpp:215-291)
                    on integer-grid boxes (tie-heavy)
Run from the repository root: python tests/golden/make_golden.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import orc  # noqa: E402
from conftest import random_triangles  # noqa: E402


def main():
    ref = orc.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libtdref.so missing: make -C oracle ref")
    rng = np.random.default_rng(2024)
    S, T = random_triangles(rng, 4096)
    d = np.zeros(len(S), np.float32)
    ref.ref_tridist(orc.ptr(S), orc.ptr(T), len(S), orc.ptr(d))
    np.savez_compressed(os.path.join(HERE, "tridist_ref.npz"), S=S, T=T, d=d)

    tri = rng.uniform(-30, 30, (1200, 9)).astype(np.float32)
    hd = rng.uniform(0, 2, (1200, 2)).astype(np.float32)
    P = 200
    os_ = np.stack([rng.integers(0, 1100, P), rng.integers(1, 40, P),
                    rng.integers(0, 1100, P), rng.integers(1, 40, P)], 1).astype(np.uint64)
    os_[:, 1] = np.minimum(os_[:, 1], 1200 - os_[:, 0])
    os_[:, 3] = np.minimum(os_[:, 3], 1200 - os_[:, 2])
    out = np.zeros((P, 3), np.float32)
    ref.ref_meshdist(orc.ptr(tri), orc.ptr(hd), orc.ptr(os_), P, orc.ptr(out))
    np.savez_compressed(os.path.join(HERE, "meshdist_ref.npz"), tri=tri, hd=hd, os=os_, out=out)

    n = 600
    low = np.floor(rng.uniform(0, 30, (n, 3))).astype(np.float32)
    high = (low + np.floor(rng.uniform(1, 6, (n, 3)))).astype(np.float32)
    tr = np.floor((low + high) / 2).astype(np.float32)
    h = ref.ref_boxset_build(n, orc.ptr(low), orc.ptr(high), orc.ptr(tr), 10)
    qi, qk, qp, qids = [], [], [0], []
    cap = 4096
    for i in range(0, n, 5):
        for k in (1, 4, 10):
            ids = np.zeros(cap, np.int32); mn = np.zeros(cap, np.float32); mx = np.zeros(cap, np.float32)
            c = ref.ref_query_knn(h, orc.ptr(low[i]), orc.ptr(high[i]), orc.ptr(tr[i]), i, k,
                                  orc.ptr(ids), orc.ptr(mn), orc.ptr(mx), cap)
            qi.append(i); qk.append(k); qids.extend(ids[:c].tolist()); qp.append(len(qids))
    ref.ref_boxset_free(h)
    np.savez_compressed(os.path.join(HERE, "octree_ref.npz"), low=low, high=high, tr=tr,
                        q_index=np.array(qi, np.int32), q_k=np.array(qk, np.int32),
                        q_prefix=np.array(qp, np.int64), q_ids=np.array(qids, np.int32))
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
