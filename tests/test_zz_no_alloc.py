"""los_cl2_sum_noalloc (the C_l sum with no device allocation) against los_cl2_run.

The inputs are placed in device memory with the HIP runtime the library itself links
(libamdhip64), the entry point is called twice in a row (the second time on a smaller
input, so the scratch array holds the first call's values), and each result must equal,
bit for bit, the one los_cl2_create + los_cl2_run + los_cl2_result give on the same
inputs (same kernels, same summation order).
"""
import ctypes as C

import numpy as np
import pytest

from song_amd import lib as L
from tests.test_cl2 import triangle_grids


class _Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _case(seed, nk, n_l, m, rescale):
    rng = np.random.default_rng(seed)
    k = np.sort(rng.uniform(0.002, 0.1, nk))
    pair_off, k3, _ = triangle_grids(k, rng, outside=True)
    pk = 2.1e-9 * (k / 0.05) ** (0.965 - 1)
    rows = n_l * len(m)
    t1 = rng.normal(size=(rows, k3.size))
    t2 = rng.normal(size=(rows, k3.size))
    used = (rng.random((n_l, len(m))) < 0.9).astype(np.int32)
    return k, pk, pair_off, k3, np.array(m, dtype=np.int32), t1, t2, used, rescale


@pytest.mark.gpu
def test_noalloc_equals_run_twice_in_a_row():
    lib = L.load()
    f = lib.los_cl2_sum_noalloc
    f.argtypes = [C.POINTER(L.CCl2Desc), C.c_long, C.c_void_p, C.c_void_p, C.c_char_p]
    f.restype = C.c_int
    hip = _Hip()
    try:
        for seed, nk, n_l, m, rescale in ((3, 24, 40, [0, 1, 2], 1), (4, 9, 7, [0, 2], 0)):
            print(f"case seed {seed} k_size {nk} n_l {n_l} m {m}", flush=True)
            k, pk, po, k3, mm, t1, t2, used, resc = _case(seed, nk, n_l, m, rescale)
            ref = L.Cl2(k, pk, po, k3, mm, t1, t2, used, resc)
            ref.run(1)
            want = ref.result()
            ref.close()
            dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
            dk, dpk = hip.put(k), hip.put(pk)
            dpo, dk3 = hip.put(po.astype(np.int64)), hip.put(k3)
            dm, dt1, dt2 = hip.put(mm), hip.put(t1), hip.put(t2)
            du = hip.put(used)
            dout = hip.put(np.full(n_l * len(m), np.nan))
            desc = L.CCl2Desc(nk, C.cast(dk, dp), C.cast(dpk, dp), C.cast(dpo, lp),
                              C.cast(dk3, dp), n_l, len(m), C.cast(dm, ip), C.cast(dt1, dp),
                              C.cast(dt2, dp), C.cast(du, ip), resc)
            err = C.create_string_buffer(2048)
            assert f(C.byref(desc), k3.size, dout, None, err) == 0, err.value
            assert hip.h.hipDeviceSynchronize() == 0
            got = hip.get(dout, n_l * len(m)).reshape(n_l, len(m))
            assert np.array_equal(got, want), np.max(np.abs(got - want))
        # limits are refused before any device work
        desc.k_size = 200
        assert f(C.byref(desc), k3.size, dout, None, err) == 1
        assert b"k_size" in err.value
    finally:
        hip.free()
