"""ctypes bindings of the song_los C-ABI (include/transfer2.h, include/song_los.h).

The product is the shared library ``song_amd/libsong_los.so`` (HIP kernels for
gfx950 + the C host side of SONG's transfer2 stage). This module only loads it
and declares argument types; it never computes anything itself. Loading fails
loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field, fields
from typing import List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsong_los.so")

SUCCESS, FAILURE = 0, 1
LINEAR, CUBIC = 0, 1
NO_EXTRAP, FLAT_EXTRAP = 0, 1
BESSEL_K3, CLASS_K3 = 0, 1
TAU_SOURCES, TAU_CUSTOM, TAU_BESSEL = 0, 1, 2
FORM_FP64, FORM_FP64_FAST, FORM_FP32C = 0, 1, 2

ERRSZ = 2048
MAX_M = 14


class LosError(RuntimeError):
    pass


class CConfig(C.Structure):
    """Mirror of ``los_config`` in include/song_los.h (field order matters)."""

    _fields_ = [
        ("has_T", C.c_int), ("has_E", C.c_int),
        ("m_list", C.c_int * MAX_M), ("m_size", C.c_int),
        ("l_max_los_t", C.c_int), ("l_max_los_p", C.c_int),
        ("k_size", C.c_int), ("k_min", C.c_double), ("k_max", C.c_double), ("k_log", C.c_int),
        ("k3_size", C.c_int),
        ("tau_size", C.c_int), ("tau_min", C.c_double), ("tau_max", C.c_double),
        ("tau0", C.c_double), ("tau_rec", C.c_double), ("rs_rec", C.c_double),
        ("l_size", C.c_int), ("l_max", C.c_int),
        ("dk_fine", C.c_double),
        ("xx_max", C.c_double), ("xx_step", C.c_double),
        ("J_cut", C.c_double), ("j_cut", C.c_double),
        ("k3_interpolation", C.c_int), ("bessels_interpolation", C.c_int),
        ("tau_sampling", C.c_int), ("time_interpolation", C.c_int),
        ("tau_linstep", C.c_double), ("k3_sampling", C.c_int),
        ("extrapolation", C.c_int), ("extra_osc_left", C.c_double), ("extra_osc_right", C.c_double),
        ("seed", C.c_ulonglong),
        ("tau_peak", C.c_double), ("tau_sigma", C.c_double), ("ramp_amp", C.c_double),
        ("ramp_center", C.c_double), ("ramp_width", C.c_double),
        ("verbose", C.c_int), ("j_build", C.c_int),
    ]


class CK3IntDesc(C.Structure):
    """Mirror of ``los_k3int_desc`` (include/song_los.h)."""
    _fields_ = [("n_pairs", C.c_int), ("pair_off", C.POINTER(C.c_long)),
                ("k3", C.POINTER(C.c_double)), ("n_q", C.c_int),
                ("transfer", C.POINTER(C.c_double)), ("index_L", C.POINTER(C.c_int)),
                ("n_j", C.c_int), ("j_off", C.POINTER(C.c_long)), ("j", C.POINTER(C.c_double)),
                ("x_min", C.POINTER(C.c_double)), ("xx_step", C.c_double), ("n_r", C.c_int),
                ("r", C.POINTER(C.c_double)), ("coef", C.c_double)]


class CCl2Desc(C.Structure):
    """Mirror of ``los_cl2_desc`` (include/song_los.h)."""
    _fields_ = [("k_size", C.c_int), ("k", C.POINTER(C.c_double)),
                ("pk", C.POINTER(C.c_double)), ("pair_off", C.POINTER(C.c_long)),
                ("k3", C.POINTER(C.c_double)), ("n_l", C.c_int), ("m_size", C.c_int),
                ("m", C.POINTER(C.c_int)), ("transfer1", C.POINTER(C.c_double)),
                ("transfer2", C.POINTER(C.c_double)), ("used", C.POINTER(C.c_int)),
                ("rescale", C.c_int)]


def _declare(lib: C.CDLL) -> C.CDLL:
    vp, i, d, ch = C.c_void_p, C.c_int, C.c_double, C.c_char_p
    sig = {
        "los_config_default": (None, [C.POINTER(CConfig)]),
        "los_scenario_create": (i, [C.POINTER(CConfig), C.POINTER(vp), ch]),
        "los_scenario_destroy": (None, [vp]),
        "los_scenario_structs": (None, [vp] + [C.POINTER(vp)] * 10),
        "los_scenario_n_transfers": (C.c_long, [vp]),
        "los_scenario_gather": (i, [vp, C.POINTER(d), C.c_long, ch]),
        "los_scenario_set": (i, [vp, ch, d, ch]),
        "los_scenario_set_string": (i, [vp, ch, ch, ch]),
        "los_scenario_source_ptr": (C.POINTER(d), [vp, i, i, i]),
        "los_scenario_J_row": (C.POINTER(d), [vp, i, i, i, i, C.POINTER(i), C.POINTER(i)]),
        "los_scenario_grids": (i, [vp, C.POINTER(i), C.POINTER(d), C.POINTER(d)]),
        "los_scenario_override_k": (i, [vp, C.POINTER(d), i, ch]),
        "los_scenario_override_k3": (i, [vp, i, i, C.POINTER(d), i, ch]),
        "los_scenario_override_J": (i, [vp, i, i, i, i, i, C.POINTER(d), i, ch]),
        "los_scenario_override_ddJ": (i, [vp, i, i, i, i, C.POINTER(d), i, ch]),
        "los_scenario_scatter": (i, [vp, C.POINTER(d), C.c_long, ch]),
        "los_scenario_k3_sizes": (i, [vp, C.POINTER(i), C.c_long, ch]),
        "los_scenario_store_sources": (i, [vp, ch]),
        "los_bessel2_build": (i, [vp] * 5 + [ch]),
        "los_bessel2_build_gpu": (i, [vp] * 5 + [i, ch]),
        "los_scenario_rebuild_J": (i, [vp, i, i, C.POINTER(d), ch]),
        "los_sph_bessel_array": (None, [i, d, C.POINTER(d)]),
        "los_wigner3j_j1": (i, [i, i, i, i, C.POINTER(i), C.POINTER(i), C.POINTER(d), i]),
        "transfer2_init": (i, [vp] * 10),
        "los_transfer2_prepare": (i, [vp] * 10),
        "transfer2_free": (i, [vp] * 3),
        "transfer2_get_k3_list": (i, [vp] * 6 + [i, i, C.POINTER(d), C.POINTER(i)]),
        "transfer2_load_transfers_from_disk": (i, [vp, vp, i]),
        "transfer2_free_type_level": (i, [vp, vp, i]),
        "los_set_default_form": (i, [i]),
        "los_set_default_device": (i, [i]),
        "los_engine_create": (i, [i, C.POINTER(vp), ch]),
        "los_engine_destroy": (None, [vp]),
        "los_engine_set_form": (i, [vp, i, ch]),
        "los_engine_load": (i, [vp] * 7 + [i, i, ch]),
        "los_engine_load_pairs": (i, [vp] * 7 + [C.POINTER(i), C.POINTER(i), i, ch]),
        "los_engine_autotune": (i, [vp, i, ch]),
        "los_engine_set_pipeline": (i, [vp, i, ch]),
        "los_engine_set_fused_prep": (i, [vp, i, ch]),
        "los_engine_set_shape": (i, [vp, i, i, i, i, i, ch]),
        "los_engine_get_shape": (None, [vp, i] + [C.POINTER(i)] * 4),
        "los_engine_capture": (i, [vp, ch]),
        "los_engine_run": (i, [vp, C.POINTER(C.c_float), ch]),
        "los_server_create": (i, [i, i, d, i] + [vp] * 6 + [C.POINTER(vp), ch]),
        "los_server_destroy": (None, [vp]),
        "los_server_submit": (C.c_long, [vp, C.POINTER(i), C.POINTER(i), i, i, ch]),
        "los_server_job_shape": (i, [vp, C.c_long, C.POINTER(C.c_long), C.POINTER(i),
                                     C.POINTER(i)]),
        "los_server_wait": (i, [vp, C.c_long, d, C.POINTER(d), C.c_long, C.POINTER(d), ch]),
        "los_server_stats": (i, [vp, C.POINTER(d), i]),
        "los_server_hold": (None, [vp, i]),
        "los_engine_last_kernel_ms": (i, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "los_engine_store": (i, [vp, vp, ch]),
        "los_engine_n_elements": (C.c_long, [vp]),
        "los_engine_flops": (i, [vp, C.POINTER(d), C.POINTER(d)]),
        "los_engine_copy_probe": (i, [vp, d, C.POINTER(C.c_float), ch]),
        "los_engine_alg_bytes": (d, [vp]),
        "los_engine_device_output": (i, [vp, C.POINTER(C.c_void_p), C.POINTER(C.c_long),
                                         C.POINTER(C.c_long)]),
        "los_engine_stream": (vp, [vp]),
        "los_k3int_create": (i, [C.POINTER(CK3IntDesc), i, C.POINTER(vp), ch]),
        "los_k3int_run": (i, [vp, i, C.POINTER(C.c_float), ch]),
        "los_k3int_result": (i, [vp, C.POINTER(d), C.c_long, ch]),
        "los_k3int_counts": (i, [vp, C.POINTER(d), i]),
        "los_k3int_destroy": (None, [vp]),
        "los_cl2_create": (i, [C.POINTER(CCl2Desc), i, C.POINTER(vp), ch]),
        "los_cl2_run": (i, [vp, i, C.POINTER(C.c_float), ch]),
        "los_cl2_result": (i, [vp, C.POINTER(d), C.c_long, ch]),
        "los_cl2_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB: Optional[C.CDLL] = None


def load() -> C.CDLL:
    """Load libsong_los.so; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise LosError(
                f"{LIB_PATH} is missing: build it with `python -m song_amd.build` "
                "(the LOS stage has no fallback implementation)")
        _LIB = _declare(C.CDLL(LIB_PATH))
    return _LIB


def _err() -> C.Array:
    return C.create_string_buffer(ERRSZ)


def check(rc: int, err, what: str) -> None:
    if rc != SUCCESS:
        msg = err.value.decode(errors="replace") if err is not None else ""
        raise LosError(f"{what} failed: {msg}")


@dataclass
class Config:
    """Synthetic-input configuration (mirrors ``los_config``)."""

    has_T: int = 1
    has_E: int = 0
    m_list: List[int] = field(default_factory=lambda: [0])
    l_max_los_t: int = 4
    l_max_los_p: int = 4
    k_size: int = 4
    k_min: float = 1e-4
    k_max: float = 0.05
    k_log: int = 1
    k3_size: int = 12
    tau_size: int = 64
    tau_min: float = 200.0
    tau_max: float = 14000.0
    tau0: float = 14200.0
    tau_rec: float = 280.0
    rs_rec: float = 145.0
    l_size: int = 8
    l_max: int = 200
    dk_fine: float = 2e-3
    xx_max: float = 1000.0
    xx_step: float = 0.3
    J_cut: float = 1e-10
    j_cut: float = 1e-12
    k3_interpolation: int = CUBIC
    bessels_interpolation: int = LINEAR
    tau_sampling: int = TAU_SOURCES
    time_interpolation: int = LINEAR
    tau_linstep: float = 0.1
    k3_sampling: int = CLASS_K3
    extrapolation: int = NO_EXTRAP
    extra_osc_left: float = 0.0
    extra_osc_right: float = 0.0
    seed: int = 20260601
    tau_peak: float = 280.0
    tau_sigma: float = 20.0
    ramp_amp: float = 0.1
    ramp_center: float = 7000.0
    ramp_width: float = 2000.0
    verbose: int = 0
    j_build: int = 0  # J tables: 0 host builder, 1 GPU builder (device 0)

    def to_c(self) -> CConfig:
        c = CConfig()
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name == "m_list":
                for q, m in enumerate(v):
                    c.m_list[q] = m
                c.m_size = len(v)
            else:
                setattr(c, f.name, v)
        return c


STRUCT_NAMES = ("ppr", "ppr2", "pba", "pth", "ppt", "ppt2", "pbs", "pbs2", "ptr", "ptr2")


class Scenario:
    """A set of SONG input structures filled by the seeded generator."""

    def __init__(self, cfg: Config):
        self.lib = load()
        self.cfg = cfg
        self._c = cfg.to_c()
        self.ptr = C.c_void_p()
        err = _err()
        check(self.lib.los_scenario_create(C.byref(self._c), C.byref(self.ptr), err), err,
              "los_scenario_create")
        ptrs = [C.c_void_p() for _ in range(10)]
        self.lib.los_scenario_structs(self.ptr, *[C.byref(p) for p in ptrs])
        self.s = dict(zip(STRUCT_NAMES, ptrs))
        self._inited = False
        self._k3_counts = {}

    def structs(self):
        return [self.s[n] for n in STRUCT_NAMES]

    def set(self, key: str, value: float) -> None:
        err = _err()
        check(self.lib.los_scenario_set(self.ptr, key.encode(), float(value), err), err, key)

    def set_string(self, key: str, value: str) -> None:
        err = _err()
        check(self.lib.los_scenario_set_string(self.ptr, key.encode(), value.encode(), err),
              err, key)

    def transfer2_init(self, fn=None) -> None:
        """Run transfer2_init (the product) or another function with its signature."""
        fn = fn or self.lib.transfer2_init
        rc = fn(*self.structs())
        self._inited = True
        if rc != SUCCESS:
            raise LosError("transfer2_init failed: " + self.error_message())

    def prepare(self) -> None:
        """transfer2_init up to the integral: indices, k3 sizes, all k1 levels allocated."""
        rc = self.lib.los_transfer2_prepare(*self.structs())
        self._inited = True
        if rc != SUCCESS:
            raise LosError("los_transfer2_prepare failed: " + self.error_message())

    def lm_l(self):
        """l of each (l,m) offset of the transfer types (l-major, m ascending)."""
        ls = self.grids()[0]
        m = list(self.cfg.m_list)
        return [int(l) for l in ls for mm in m if mm <= l]

    def pair_offsets(self):
        """{(i1, i2): (offset, k3 count)} of each pair inside one transfer type of gather()."""
        sizes = self.k3_sizes()[:, 0]
        out, o, q = {}, 0, 0
        for i1 in range(self.cfg.k_size):
            for i2 in range(i1 + 1):
                out[(i1, i2)] = (o, int(sizes[q]))
                o += int(sizes[q])
                q += 1
        return out

    def error_message(self) -> str:
        # ptr2->error_message lives inside struct transfers2; read it via offset
        return _transfers2_error(self.s["ptr2"])

    def gather(self) -> np.ndarray:
        n = self.lib.los_scenario_n_transfers(self.ptr)
        out = np.empty(n, dtype=np.float64)
        err = _err()
        check(self.lib.los_scenario_gather(self.ptr, out.ctypes.data_as(C.POINTER(C.c_double)),
                                           n, err), err, "gather")
        return out

    def grids(self):
        """(l list, k grid, tau sampling) as generated."""
        ls = np.zeros(self.cfg.l_size, dtype=np.int32)
        ks = np.zeros(self.cfg.k_size)
        ts = np.zeros(self.cfg.tau_size)
        self.lib.los_scenario_grids(self.ptr, ls.ctypes.data_as(C.POINTER(C.c_int)),
                                    ks.ctypes.data_as(C.POINTER(C.c_double)),
                                    ts.ctypes.data_as(C.POINTER(C.c_double)))
        return ls, ks, ts

    def scatter(self, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        err = _err()
        check(self.lib.los_scenario_scatter(self.ptr, v.ctypes.data_as(C.POINTER(C.c_double)),
                                            v.size, err), err, "scatter")

    def k3_sizes(self) -> np.ndarray:
        """(k_size, physical start, physical size) per pair, after the indices."""
        nk = self.cfg.k_size
        out = np.zeros(3 * nk * (nk + 1) // 2, dtype=np.int32)
        err = _err()
        check(self.lib.los_scenario_k3_sizes(self.ptr, out.ctypes.data_as(C.POINTER(C.c_int)),
                                             out.size, err), err, "k3_sizes")
        return out.reshape(-1, 3)

    def store_sources(self) -> None:
        """Write every k1 level of the sources to its sources file and free it."""
        err = _err()
        check(self.lib.los_scenario_store_sources(self.ptr, err), err, "store_sources")

    def override_k(self, k) -> None:
        v = np.ascontiguousarray(k, dtype=np.float64)
        err = _err()
        check(self.lib.los_scenario_override_k(self.ptr, v.ctypes.data_as(C.POINTER(C.c_double)),
                                               v.size, err), err, "override_k")

    def override_k3(self, i1: int, i2: int, k3) -> None:
        v = np.ascontiguousarray(k3, dtype=np.float64)
        err = _err()
        check(self.lib.los_scenario_override_k3(self.ptr, i1, i2,
                                                v.ctypes.data_as(C.POINTER(C.c_double)),
                                                v.size, err), err, "override_k3")
        self._k3_counts[(i1, i2)] = v.size

    def override_J(self, jt: int, L: int, il: int, im: int, xmin: int, values) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        err = _err()
        check(self.lib.los_scenario_override_J(self.ptr, jt, L, il, im, xmin,
                                               v.ctypes.data_as(C.POINTER(C.c_double)), v.size,
                                               err), err, "override_J")

    def override_ddJ(self, jt: int, L: int, il: int, im: int, values) -> None:
        """Second derivatives of one J row (cubic J interpolation), same size as the row."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        err = _err()
        check(self.lib.los_scenario_override_ddJ(self.ptr, jt, L, il, im,
                                                 v.ctypes.data_as(C.POINTER(C.c_double)),
                                                 v.size, err), err, "override_ddJ")

    def source(self, tp: int, i1: int, i2: int) -> np.ndarray:
        """Writable view of ppt2->sources[tp][i1][i2] as [tau][k3]."""
        p = self.lib.los_scenario_source_ptr(self.ptr, tp, i1, i2)
        if not p:
            raise LosError("sources not allocated")
        nk3 = self.k3_count(i1, i2)
        return np.ctypeslib.as_array(p, shape=(self.cfg.tau_size, nk3))

    def k3_count(self, i1: int, i2: int) -> int:
        return self._k3_counts.get((i1, i2), self.cfg.k3_size)

    def J_row(self, jt: int, L: int, il: int, im: int):
        xmin, xs = C.c_int(), C.c_int()
        p = self.lib.los_scenario_J_row(self.ptr, jt, L, il, im, C.byref(xmin), C.byref(xs))
        return xmin.value, np.ctypeslib.as_array(p, shape=(xs.value,))

    def rebuild_J(self, gpu: bool, device: int = 0) -> float:
        """Build the J tables again on the host or on the GPU; returns the seconds taken."""
        sec = C.c_double()
        err = _err()
        check(self.lib.los_scenario_rebuild_J(self.ptr, int(bool(gpu)), device, C.byref(sec), err),
              err, "los_scenario_rebuild_J")
        return sec.value

    def free_transfers(self, fn=None) -> None:
        if self._inited:
            fn = fn or self.lib.transfer2_free
            fn(self.s["ppr2"], self.s["ppt2"], self.s["ptr2"])
            self._inited = False

    def close(self) -> None:
        if self.ptr:
            self.free_transfers()
            self.lib.los_scenario_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """The LOS engine of one HIP device (los_engine_* of include/song_los.h)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.ptr = C.c_void_p()
        err = _err()
        check(self.lib.los_engine_create(device, C.byref(self.ptr), err), err, "los_engine_create")

    def _call(self, name, *args):
        err = _err()
        check(getattr(self.lib, name)(self.ptr, *args, err), err, name)

    def set_form(self, form: int) -> None:
        self._call("los_engine_set_form", form)

    def load(self, scen: "Scenario", k1_begin: int = 0, k1_end: int | None = None) -> None:
        s = scen.s
        k1_end = scen.cfg.k_size if k1_end is None else k1_end
        self._call("los_engine_load", s["ppr"], s["ppr2"], s["ppt2"], s["pbs"], s["pbs2"],
                   s["ptr2"], k1_begin, k1_end)

    def load_pairs(self, scen: "Scenario", pairs) -> None:
        s = scen.s
        k1 = (C.c_int * len(pairs))(*[int(p[0]) for p in pairs])
        k2 = (C.c_int * len(pairs))(*[int(p[1]) for p in pairs])
        self._call("los_engine_load_pairs", s["ppr"], s["ppr2"], s["ppt2"], s["pbs"], s["pbs2"],
                   s["ptr2"], k1, k2, len(pairs))

    def run(self) -> float:
        ms = C.c_float()
        self._call("los_engine_run", C.byref(ms))
        return ms.value

    def last_kernel_ms(self):
        a, b = C.c_float(), C.c_float()
        self.lib.los_engine_last_kernel_ms(self.ptr, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_pipeline(self, on: int) -> None:
        """Pipelined steps (chunked preparation overlapping the LOS kernel): n >= 2 chunks
        from the next load; 0 or 1 off (the default)."""
        self._call("los_engine_set_pipeline", int(on))

    def set_fused_prep(self, on: bool) -> None:
        """Sweep tables straight from the raw sources (default) or through re-laid and
        splined sources in HBM."""
        self._call("los_engine_set_fused_prep", int(bool(on)))

    def autotune(self, reps: int = 2) -> None:
        self._call("los_engine_autotune", reps)

    def set_shape(self, form: int, kind: int, nk: int = 64, ch: int = 0,
                  persistent: int = 0) -> None:
        """Kernel of a form. Plain FP64: kind 0 (workgroup per k3) or 2 (k-lane);
        fast forms: kind 0 stream (nk, ch, persistent), 1 sweep (nk), 2 k-lane."""
        self._call("los_engine_set_shape", form, kind, nk, ch, persistent)

    def get_shape(self, form: int):
        """(kind, nk, ch, persistent) of the form's kernel."""
        v = [C.c_int() for _ in range(4)]
        self.lib.los_engine_get_shape(self.ptr, form, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def capture(self) -> None:
        self._call("los_engine_capture")

    def store(self, scen: "Scenario") -> None:
        self._call("los_engine_store", scen.s["ptr2"])

    def n_elements(self) -> int:
        return self.lib.los_engine_n_elements(self.ptr)

    def device_output(self):
        """(device address, rows, cols) of the last run's transfers, [tt][fine k3]."""
        p, r, c = C.c_void_p(), C.c_long(), C.c_long()
        if self.lib.los_engine_device_output(self.ptr, C.byref(p), C.byref(r), C.byref(c)):
            raise LosError("los_engine_device_output: nothing loaded")
        return p.value, r.value, c.value

    def alg_bytes(self) -> float:
        return self.lib.los_engine_alg_bytes(self.ptr)

    def copy_probe(self, nbytes: float) -> float:
        """ms of one plain device copy of nbytes."""
        ms = C.c_float()
        self._call("los_engine_copy_probe", float(nbytes), C.byref(ms))
        return ms.value

    def flops(self):
        """(live, dense) algorithmic operation counts of one run."""
        a, b = C.c_double(), C.c_double()
        self.lib.los_engine_flops(self.ptr, C.byref(a), C.byref(b))
        return a.value, b.value

    def close(self) -> None:
        if self.ptr:
            self.lib.los_engine_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The LOS stage as a service over one scenario's pairs (los_server_* of
    include/song_los.h)."""

    def __init__(self, scen: "Scenario", form: int, budget_bytes: float, slots: int = 3,
                 device: int = 0):
        self.lib = load()
        self.ptr = C.c_void_p()
        s = scen.s
        err = _err()
        check(self.lib.los_server_create(device, form, float(budget_bytes), slots, s["ppr"],
                                         s["ppr2"], s["ppt2"], s["pbs"], s["pbs2"], s["ptr2"],
                                         C.byref(self.ptr), err), err, "los_server_create")

    def submit(self, pairs, l_max: int) -> int:
        k1 = (C.c_int * len(pairs))(*[int(p[0]) for p in pairs])
        k2 = (C.c_int * len(pairs))(*[int(p[1]) for p in pairs])
        err = _err()
        jid = self.lib.los_server_submit(self.ptr, k1, k2, len(pairs), int(l_max), err)
        if jid < 0:
            raise LosError("los_server_submit failed: " + err.value.decode(errors="replace"))
        return jid

    def shape(self, jid: int):
        n, nlm, ncol = C.c_long(), C.c_int(), C.c_int()
        if self.lib.los_server_job_shape(self.ptr, jid, C.byref(n), C.byref(nlm), C.byref(ncol)):
            raise LosError(f"unknown job {jid}")
        return n.value, nlm.value, ncol.value

    def wait(self, jid: int, timeout_ms: float = -1.0):
        """(result [tt][col], (t_submit, t_admit, t_done)) of a job."""
        n, _, ncol = self.shape(jid)
        out = np.empty(n, dtype=np.float64)
        times = (C.c_double * 3)()
        err = _err()
        rc = self.lib.los_server_wait(self.ptr, jid, float(timeout_ms),
                                      out.ctypes.data_as(C.POINTER(C.c_double)), n, times, err)
        if rc == 2:
            raise TimeoutError(f"job {jid} not done after {timeout_ms} ms")
        if rc != 0:
            raise LosError(err.value.decode(errors="replace"))
        return out.reshape(-1, ncol), tuple(times)

    def hold(self, on: bool) -> None:
        self.lib.los_server_hold(self.ptr, int(bool(on)))

    def stats(self):
        v = (C.c_double * 6)()
        self.lib.los_server_stats(self.ptr, v, 6)
        keys = ("jobs_done", "launches", "max_busy_slots", "peak_admitted_bytes",
                "resident_bytes", "budget_bytes")
        return dict(zip(keys, list(v)))

    def close(self) -> None:
        if self.ptr:
            self.lib.los_server_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class K3Integral:
    """k3 integral of transfer rows against j_L(k3 r) on the GPU (los_k3int_*).

    pair_off (n_pairs+1), k3 (concatenated grids), transfer [n_q][n_k3], index_L (n_q, -1 =
    skipped row), j_off (n_j+1) / j / x_min (j tables from their first stored point), r,
    coef (M3 coefficient). Inputs are uploaded at construction and stay resident.
    """

    def __init__(self, pair_off, k3, transfer, index_L, j_off, j, x_min, xx_step, r,
                 coef=1.0, device: int = 0):
        self.lib = load()
        self._keep = [np.ascontiguousarray(pair_off, dtype=np.int64),
                      np.ascontiguousarray(k3, dtype=np.float64),
                      np.ascontiguousarray(transfer, dtype=np.float64),
                      np.ascontiguousarray(index_L, dtype=np.int32),
                      np.ascontiguousarray(j_off, dtype=np.int64),
                      np.ascontiguousarray(j, dtype=np.float64),
                      np.ascontiguousarray(x_min, dtype=np.float64),
                      np.ascontiguousarray(r, dtype=np.float64)]
        po, k, t, il, jo, jj, xm, rr = self._keep
        self.n_pairs, self.n_q, self.n_r = po.size - 1, il.size, rr.size
        assert t.size == self.n_q * (po[-1] if po.size else 0), "transfer must be [n_q][n_k3]"
        dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        self.desc = CK3IntDesc(self.n_pairs, po.ctypes.data_as(lp), k.ctypes.data_as(dp),
                               self.n_q, t.ctypes.data_as(dp), il.ctypes.data_as(ip),
                               xm.size, jo.ctypes.data_as(lp), jj.ctypes.data_as(dp),
                               xm.ctypes.data_as(dp), float(xx_step), self.n_r,
                               rr.ctypes.data_as(dp), float(coef))
        self.ptr = C.c_void_p()
        err = _err()
        check(self.lib.los_k3int_create(C.byref(self.desc), device, C.byref(self.ptr), err),
              err, "los_k3int_create")

    def run(self, reps: int = 1) -> float:
        """Launch the kernel `reps` times on the resident inputs; HIP-event ms per launch."""
        ms = C.c_float()
        err = _err()
        check(self.lib.los_k3int_run(self.ptr, reps, C.byref(ms), err), err, "los_k3int_run")
        return ms.value

    def result(self) -> np.ndarray:
        out = np.empty((self.n_q, self.n_r, self.n_pairs), dtype=np.float64)
        err = _err()
        check(self.lib.los_k3int_result(self.ptr, out.ctypes.data_as(C.POINTER(C.c_double)),
                                        out.size, err), err, "los_k3int_result")
        return out

    def counts(self):
        v = (C.c_double * 2)()
        self.lib.los_k3int_counts(self.ptr, v, 2)
        return {"live_terms": v[0], "alg_bytes": v[1]}

    def close(self) -> None:
        if self.ptr:
            self.lib.los_k3int_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Cl2:
    """Second-order C_l sum over (k1, k2, k3) on the GPU (los_cl2_*).

    k, pk (primordial spectrum at k), pair_off (k1-major pairs, k2 <= k1), k3 (concatenated
    grids), m (values), transfer1/transfer2 [n_l][m_size][n_k3], used [n_l][m_size], rescale.
    """

    def __init__(self, k, pk, pair_off, k3, m, transfer1, transfer2, used, rescale=0,
                 device: int = 0):
        self.lib = load()
        self._keep = [np.ascontiguousarray(k, dtype=np.float64),
                      np.ascontiguousarray(pk, dtype=np.float64),
                      np.ascontiguousarray(pair_off, dtype=np.int64),
                      np.ascontiguousarray(k3, dtype=np.float64),
                      np.ascontiguousarray(m, dtype=np.int32),
                      np.ascontiguousarray(transfer1, dtype=np.float64),
                      np.ascontiguousarray(transfer2, dtype=np.float64),
                      np.ascontiguousarray(used, dtype=np.int32)]
        kk, pp, po, k3a, mm, t1, t2, uu = self._keep
        self.n_l, self.m_size = uu.shape
        dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        self.desc = CCl2Desc(kk.size, kk.ctypes.data_as(dp), pp.ctypes.data_as(dp),
                             po.ctypes.data_as(lp), k3a.ctypes.data_as(dp), self.n_l,
                             self.m_size, mm.ctypes.data_as(ip), t1.ctypes.data_as(dp),
                             t2.ctypes.data_as(dp), uu.ctypes.data_as(ip), int(rescale))
        self.ptr = C.c_void_p()
        err = _err()
        check(self.lib.los_cl2_create(C.byref(self.desc), device, C.byref(self.ptr), err),
              err, "los_cl2_create")

    def run(self, reps: int = 1) -> float:
        ms = C.c_float()
        err = _err()
        check(self.lib.los_cl2_run(self.ptr, reps, C.byref(ms), err), err, "los_cl2_run")
        return ms.value

    def result(self) -> np.ndarray:
        """cl_m [n_l][m_size]; C_l is the sum over the last axis."""
        out = np.empty((self.n_l, self.m_size), dtype=np.float64)
        err = _err()
        check(self.lib.los_cl2_result(self.ptr, out.ctypes.data_as(C.POINTER(C.c_double)),
                                      out.size, err), err, "los_cl2_result")
        return out

    def close(self) -> None:
        if self.ptr:
            self.lib.los_cl2_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _transfers2_error(ptr2: C.c_void_p) -> str:
    lib = load()
    off = lib_offsets()["transfers2.error_message"]
    buf = C.string_at(ptr2.value + off, ERRSZ)
    return buf.split(b"\0", 1)[0].decode(errors="replace")


_OFFS = None


def lib_offsets():
    global _OFFS
    if _OFFS is None:
        lib = load()
        fn = lib.los_struct_offsets
        fn.restype = C.c_int
        fn.argtypes = [C.POINTER(C.c_long), C.c_int]
        buf = (C.c_long * 16)()
        fn(buf, 16)
        _OFFS = {"transfers2.error_message": buf[0], "transfers2.tt2_size": buf[1],
                 "transfers2.n_transfers": buf[2]}
    return _OFFS
