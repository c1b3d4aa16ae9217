"""The stage's output handed to a consumer on the device, without a copy to the host.

los_engine_device_output gives the last run's transfers where they are ([tt][fine k3]);
los_cl2_sum_noalloc sums the C_l of the T rows straight from there. The device output equals
what the host path stores into ptr2->transfer bit for bit, and the C_l equal, bit for bit,
those los_cl2_create / los_cl2_run give on the gathered host copy (same kernels).
"""
import ctypes as C

import numpy as np
import pytest

from song_amd import lib as L

CFG = dict(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=4, l_max_los_p=4, k_size=6,
           k_min=1e-4, k_max=0.05, k3_size=16, tau_size=128, l_size=12, l_max=400,
           dk_fine=1e-3, xx_max=1500.0)


def _hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


@pytest.mark.gpu
def test_cl2_on_engine_output():
    lib = L.load()
    cfg = L.Config(**CFG)
    s = L.Scenario(cfg)
    s.prepare()
    eng = L.Engine(0)
    eng.load(s)
    eng.set_form(L.FORM_FP64)
    eng.run()
    dev, rows, cols = eng.device_output()
    eng.store(s)
    flat = s.gather().reshape(rows, cols)
    hip = _hip()
    bufs = []
    try:
        host = np.empty((rows, cols))
        assert hip.hipMemcpy(host.ctypes.data, dev, host.nbytes, 2) == 0
        assert np.array_equal(host, flat)  # the device output is what the host path stores
        # the C_l of the T rows (every l >= 2 carries m = 0, 1, 2), host path first
        ls, k, _ = s.grids()
        offs = s.pair_offsets()
        pair_off = np.zeros(len(offs) + 1, dtype=np.int64)
        k3 = []
        st = s.s
        for q, (i1, i2) in enumerate(offs):
            n = offs[(i1, i2)][1]
            buf, last = (C.c_double * n)(), C.c_int()
            assert lib.transfer2_get_k3_list(st["ppr"], st["ppr2"], st["ppt2"], st["pbs"],
                                             st["pbs2"], st["ptr2"], i1, i2, buf,
                                             C.byref(last)) == 0
            k3.extend(buf)
            pair_off[q + 1] = pair_off[q] + n
        k3 = np.array(k3)
        m = np.array(cfg.m_list, dtype=np.int32)
        n_l = cfg.l_size
        T = flat[:n_l * m.size].reshape(n_l, m.size, cols)
        pk = 2.1e-9 * (k / 0.05) ** (0.965 - 1)
        used = np.ones((n_l, m.size), dtype=np.int32)
        h = L.Cl2(k, pk, pair_off, k3, m, T, T, used, 1)
        h.run(1)
        want = h.result()
        h.close()

        def put(a):
            a = np.ascontiguousarray(a)
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
            bufs.append(p)
            assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
            return p

        dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        dk, dpk, dpo, dk3 = put(k), put(pk), put(pair_off), put(k3)
        dm, du, dout = put(m), put(used), put(np.zeros(n_l * m.size))
        tptr = C.cast(C.c_void_p(dev), dp)  # T rows: the first n_l x m_size rows of the output
        desc = L.CCl2Desc(k.size, C.cast(dk, dp), C.cast(dpk, dp), C.cast(dpo, lp),
                          C.cast(dk3, dp), n_l, m.size, C.cast(dm, ip), tptr, tptr,
                          C.cast(du, ip), 1)
        f = lib.los_cl2_sum_noalloc
        f.argtypes = [C.POINTER(L.CCl2Desc), C.c_long, C.c_void_p, C.c_void_p, C.c_char_p]
        f.restype = C.c_int
        err = C.create_string_buffer(2048)
        assert f(C.byref(desc), k3.size, dout, lib.los_engine_stream(eng.ptr), err) == 0, err.value
        assert hip.hipDeviceSynchronize() == 0
        got = np.empty(n_l * m.size)
        assert hip.hipMemcpy(got.ctypes.data, dout, got.nbytes, 2) == 0
        assert np.array_equal(got.reshape(n_l, m.size), want)
    finally:
        for p in bufs:
            hip.hipFree(p)
        eng.close()
        s.close()
