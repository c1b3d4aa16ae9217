"""pytest configuration: the `gpu` marker and shared helpers for the LOS tests."""
import ctypes as C
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


ORACLE_PATH = os.path.join(ROOT, "oracle", "liblos_oracle.so")


def load_oracle():
    """The CPU oracle library (test infrastructure; built by `python -m song_amd.build`)."""
    if not os.path.exists(ORACLE_PATH):
        pytest.skip("oracle library not built")
    o = C.CDLL(ORACLE_PATH)
    o.los_ref_transfer2_init.argtypes = [C.c_void_p] * 10
    o.los_ref_transfer2_init.restype = C.c_int
    o.los_ref_transfer2_free.argtypes = [C.c_void_p] * 3
    o.los_ref_transfer2_free.restype = C.c_int
    o.los_ref_get_k3_list.argtypes = [C.c_void_p] * 5 + [C.c_int, C.c_int,
                                                         C.POINTER(C.c_double), C.POINTER(C.c_int)]
    o.los_ref_get_k3_list.restype = C.c_int
    o.los_ref_spline_columns.argtypes = [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double),
                                         C.c_int, C.POINTER(C.c_double)]
    o.los_ref_spline_columns.restype = C.c_int
    o.los_ref_size_indexl_indexm.argtypes = [C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.c_int]
    o.los_ref_size_indexl_indexm.restype = C.c_int
    o.los_ref_offset_indexl_indexm.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int,
                                               C.POINTER(C.c_int), C.c_int]
    o.los_ref_offset_indexl_indexm.restype = C.c_int
    return o


@pytest.fixture(scope="session")
def oracle():
    return load_oracle()


def run_oracle(scen, oracle):
    """Run the oracle's transfer2_init on a scenario and return the flat transfers."""
    rc = oracle.los_ref_transfer2_init(*scen.structs())
    scen._inited = True
    if rc != 0:
        raise RuntimeError("oracle failed: " + scen.error_message())
    out = scen.gather()
    scen.free_transfers(lambda a, b, c: oracle.los_ref_transfer2_free(a, b, c))
    return out


def field_errors(got, ref, tt2_size):
    """|got - ref| / max|ref| over each transfer type's whole (l,m) field."""
    import numpy as np

    g, r = np.asarray(got).reshape(tt2_size, -1), np.asarray(ref).reshape(tt2_size, -1)
    sc = np.abs(r).max(axis=1, keepdims=True)
    return np.divide(np.abs(g - r), sc, out=np.zeros_like(g), where=sc > 0)
