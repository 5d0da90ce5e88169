"""tdp_meshdist_batch_noalloc on an MI355X: on device arrays the caller owns,
with the library's fixed per-GPU scratch, it gives tdp_meshdist_batch's
values bit for bit, on two calls in a row (the second smaller, so that the
scratch holds the first call's contents), in every precision; sizes beyond
the scratch and pairs outside the records are refused."""
import ctypes as c

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """device buffers through the HIP runtime the library already loaded"""

    def __init__(self):
        self.L = c.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [c.POINTER(c.c_void_p), c.c_size_t]
        self.L.hipFree.argtypes = [c.c_void_p]
        self.L.hipMemcpy.argtypes = [c.c_void_p, c.c_void_p, c.c_size_t, c.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = c.c_void_p()
        assert self.L.hipMalloc(c.byref(p), max(a.nbytes, 1)) == 0
        self.held.append(p)
        if a.nbytes:
            assert self.L.hipMemcpy(p, a.ctypes.data_as(c.c_void_p), a.nbytes, 1) == 0
        return p.value

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data_as(c.c_void_p), c.c_void_p(p), out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.held:
            self.L.hipFree(p)
        self.held = []


def _case(tdp, seed, slot, n_pairs, n_vessels=6):
    n, _ = tdp.generate(seed=seed, n_vessels=n_vessels, nuclei_per_vessel=10, cube=800.0)
    a = n.arrays()
    rec = a["tri"].reshape(-1, 12)
    off = a["vox_tri_off"].reshape(-1, 5)[:, slot]
    cnt = a["vox_tri_cnt"].reshape(-1, 5)[:, slot]
    nz = np.flatnonzero(cnt > 0)
    rng = np.random.default_rng(seed)
    v1, v2 = rng.choice(nz, n_pairs), rng.choice(nz, n_pairs)
    os_ = np.stack([off[v1], cnt[v1], off[v2], cnt[v2]], 1).astype(np.uint64)
    return np.ascontiguousarray(rec), os_


def test_noalloc_equals_meshdist_batch(tdp):
    dev = tdp.Device(0)
    hip = _Hip()
    try:
        # two calls in a row per precision, the second with a smaller input
        for precision in ("ref32", "fp64", "mixed"):
            sizes = []
            for seed, slot, n_pairs, n_vessels in ((1, 4, 3000, 6), (2, 1, 700, 3)):
                rec, os_ = _case(tdp, seed, slot, n_pairs, n_vessels)
                sizes.append(len(rec))
                print(precision, seed, slot, len(rec), n_pairs, flush=True)
                want = dev.meshdist(rec, os_, precision)
                d_rec, d_os = hip.put(rec), hip.put(os_)
                d_out = hip.put(np.full_like(want, np.nan))
                dev.meshdist_noalloc(d_rec, len(rec), d_os, len(os_), d_out, precision)
                got = hip.get(d_out, want)
                assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (precision, seed)
                hip.free()
            assert sizes[1] < sizes[0]
        # beyond the scratch: refused before any work
        with pytest.raises(tdp.TdpError, match="more than its scratch"):
            dev.meshdist_noalloc(1, (1 << 18) + 1, 1, 1, 1, "fp64")
        # a pair past the records: refused by the device-side check
        rec, os_ = _case(tdp, 3, 4, 64)
        bad = os_.copy()
        bad[17, 0] = len(rec)  # offset at the end, count > 0
        d_rec, d_os = hip.put(rec), hip.put(bad)
        d_out = hip.put(np.zeros((64, 3), np.float64))
        with pytest.raises(tdp.TdpError, match="outside the records"):
            dev.meshdist_noalloc(d_rec, len(rec), d_os, 64, d_out, "mixed")
        hip.free()
        # and a good call afterwards still matches
        want = dev.meshdist(rec, os_, "mixed")
        d_rec, d_os = hip.put(rec), hip.put(os_)
        d_out = hip.put(np.zeros_like(want))
        dev.meshdist_noalloc(d_rec, len(rec), d_os, len(os_), d_out, "mixed")
        assert np.array_equal(hip.get(d_out, want).view(np.uint8), want.view(np.uint8))
    finally:
        hip.free()
        dev.close()
