"""C-ABI surface and host-side behaviour of libsong_los.so, on the CPU (no kernel runs).

* every function declared in include/*.h is exported by the library;
* without a GPU, transfer2_init fails loudly with the HIP message (no CPU fallback);
* SONG's early exits and the checks that fire before any GPU work
  (source/transfer2.c:99-110, :151-157, :1408-1411, :1764-1771);
* transfers written by the oracle in SONG's file format (one raw FP64 file per
  transfer type, source/transfer2.c:2929-2963) load through the library's
  transfer2_load_transfers_from_disk (:705-763);
* sources written in SONG's sources_%03d.dat format
  (source/perturbations2.c:15347-15414, paths :1051-1064) hold the in-memory
  blocks in SONG's order and load back bit for bit (:521-602).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from song_amd import lib as L
from tests.conftest import ROOT, run_oracle

HEADERS = ["include/transfer2.h", "include/song_los.h"]


def _declared_functions():
    names = []
    for h in HEADERS:
        text = open(os.path.join(ROOT, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|long|double|double \*|void \*)\s*\**\s*(\w+)\s*\(",
                             text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_all_declared_symbols_exported():
    lib = L.load()
    names = _declared_functions()
    assert "transfer2_init" in names and "los_engine_run" in names and len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def _hip_devices():
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        return 0
    n = C.c_int(0)
    if hip.hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


SMALL = dict(has_T=1, has_E=1, m_list=[0, 2], l_max_los_t=2, l_max_los_p=2, k_size=3, k3_size=6,
             tau_size=32, l_size=4, l_max=60, dk_fine=2e-3, xx_max=1000.0)


def test_transfer2_init_fails_loudly_without_gpu():
    if _hip_devices() > 0:
        pytest.skip("a GPU is present")
    s = L.Scenario(L.Config(**SMALL))
    with pytest.raises(L.LosError) as exc:
        s.transfer2_init()
    assert "los_engine_create" in str(exc.value)
    s.close()


def test_skip_when_no_cls_requested():
    s = L.Scenario(L.Config(**SMALL))
    s.set("has_cmb_bispectra", 0)
    s.set("has_cmb_spectra", 0)
    s.transfer2_init()  # returns _SUCCESS_ before touching the GPU
    assert L.load().los_scenario_n_transfers(s.ptr) == 0  # has_cls is false
    s._inited = False
    s.close()


@pytest.mark.parametrize("key,value,msg", [
    ("xx_max", 10.0, "xx_max is not large enough"),
    ("q_linstep_song", 0.0, "infinite loop"),
])
def test_errors_before_gpu_work(key, value, msg):
    s = L.Scenario(L.Config(**SMALL))
    s.set(key, value)
    with pytest.raises(L.LosError) as exc:
        s.transfer2_init()
    assert msg in str(exc.value)
    s._inited = False
    s.close()


def test_too_few_coarse_k3_for_cubic():
    cfg = L.Config(**dict(SMALL, k3_size=3))
    s = L.Scenario(cfg)
    with pytest.raises(L.LosError) as exc:
        s.transfer2_init()
    assert "cannot do interpolation in k3" in str(exc.value)
    s._inited = False
    s.close()


def test_load_oracle_written_transfers(oracle, tmp_path):
    cfg = L.Config(**SMALL)
    ref = run_oracle(L.Scenario(cfg), oracle)
    # oracle writes SONG's transfers_%03d.dat files
    w = L.Scenario(cfg)
    w.set("store_transfers_to_disk", 1)
    w.set_string("transfers_dir", str(tmp_path))
    assert oracle.los_ref_transfer2_init(*w.structs()) == 0, w.error_message()
    w._inited = True
    w.free_transfers(lambda a, b, c: oracle.los_ref_transfer2_free(a, b, c))
    w.close()
    files = sorted(os.listdir(tmp_path))
    assert files[0] == "transfers_000.dat"
    # the library reads them back
    lib = L.load()
    r = L.Scenario(cfg)
    r.set("load_transfers_from_disk", 1)
    r.set_string("transfers_dir", str(tmp_path))
    r.transfer2_init()  # indices only, returns before the integral
    tt2 = len(files)
    for tt in range(tt2):
        assert lib.transfer2_load_transfers_from_disk(r.s["ppt2"], r.s["ptr2"], tt) == 0, \
            r.error_message()
    got = r.gather()
    assert np.array_equal(got, ref)
    for tt in range(tt2):
        lib.transfer2_free_type_level(r.s["ppt2"], r.s["ptr2"], tt)
    r.free_transfers()
    r.close()


@pytest.mark.gpu
def test_engine_rejects_unsupported_options():
    """Options the GPU engine does not implement fail with a message, never silently."""
    for key, val, msg in (("tau_sampling", L.TAU_CUSTOM, "tau_sampling"),):
        s = L.Scenario(L.Config(**SMALL))
        s.set(key, val)
        with pytest.raises(L.LosError) as exc:
            s.transfer2_init()
        assert msg in str(exc.value)
        s._inited = False
        s.close()


def _tp2_size(cfg):
    """Source types: (L,m) with m <= L per field, T then E then B (B when m_max > 0)."""
    def n_lm(l_max):
        return sum(1 for L_ in range(l_max + 1) for m in cfg.m_list if m <= L_)
    n = n_lm(cfg.l_max_los_t) if cfg.has_T else 0
    if cfg.has_E:
        n += n_lm(cfg.l_max_los_p) * (2 if max(cfg.m_list) > 0 else 1)
    return n


def test_sources_files_round_trip(tmp_path):
    cfg = L.Config(**SMALL)
    tp2 = _tp2_size(cfg)
    assert tp2 == 12  # T, E, B x (L,m) in {(0,0), (1,0), (2,0), (2,2)}
    ref = L.Scenario(cfg)  # same seed: the sources the writer holds
    w = L.Scenario(cfg)
    with pytest.raises(L.LosError, match="sources_dir"):
        w.store_sources()
    w.set_string("sources_dir", str(tmp_path))
    w.store_sources()  # writes and frees every k1 level
    with pytest.raises(L.LosError):
        w.source(0, 0, 0)
    w.close()
    files = sorted(os.listdir(tmp_path))
    assert files == ["sources_%03d.dat" % i for i in range(cfg.k_size)]
    # file layout: types outermost, then k2 <= k1, then the [tau][k3] block
    for i1 in range(cfg.k_size):
        raw = np.fromfile(tmp_path / files[i1], dtype="<f8")
        want = np.concatenate([ref.source(tp, i1, i2).ravel()
                               for tp in range(tp2) for i2 in range(i1 + 1)])
        assert raw.size == tp2 * cfg.tau_size * cfg.k3_size * (i1 + 1)
        assert np.array_equal(raw, want)
    # the library's perturb2_load_sources_from_disk reads them back bit for bit
    lib = L.load()
    r = L.Scenario(cfg)
    r.set_string("sources_dir", str(tmp_path))
    ppt2 = r.s["ppt2"]
    for i1 in range(cfg.k_size):
        assert lib.perturb2_free_k1_level(ppt2, i1) == 0
        assert lib.perturb2_allocate_k1_level(ppt2, i1) == 0
        assert not np.any(r.source(0, i1, 0))
        assert lib.perturb2_load_sources_from_disk(ppt2, i1) == 0
        for tp in range(tp2):
            for i2 in range(i1 + 1):
                assert np.array_equal(r.source(tp, i1, i2), ref.source(tp, i1, i2))
    # a truncated file is an error, not a short read
    last = tmp_path / files[-1]
    data = last.read_bytes()
    last.write_bytes(data[:-8])
    i1 = cfg.k_size - 1
    assert lib.perturb2_free_k1_level(ppt2, i1) == 0
    assert lib.perturb2_allocate_k1_level(ppt2, i1) == 0
    assert lib.perturb2_load_sources_from_disk(ppt2, i1) == 1
    r.close()
    ref.close()


def test_transfers2_layout_is_songs():
    """struct transfers2 has SONG's fields in SONG's order and types
    (include/transfer2.h:76-238), so its layout equals SONG's for the same CLASS
    constants: the offsets the library was compiled with equal those of a C struct
    declared here field by field from the reference header."""
    import ctypes as C
    FN, LBL, ERR = 256, 64, 2048  # _FILENAMESIZE_, _MAX_LENGTH_LABEL_, _ERRORMSGSIZE_
    p, i, sh, d, lg = C.c_void_p, C.c_int, C.c_short, C.c_double, C.c_long
    fields = [("transfer", p), ("index_tt2_T", i), ("index_tt2_E", i), ("index_tt2_B", i),
              ("tt2_size", i), ("tt2_labels", p), ("has_cls", sh), ("l_size", i), ("l", p),
              ("m", p), ("m_size", i), ("n_transfers", i), ("n_nonzero_transfers_E", i),
              ("n_nonzero_transfers_B", i), ("index_tt2_monopole", p),
              ("index_pt2_monopole", p), ("corresponding_index_l", p),
              ("corresponding_index_m", p), ("lm_array", p), ("k_sampling", i),
              ("tau_sampling", i), ("k_size_k1k2", p), ("k_physical_start_k1k2", p),
              ("k_physical_size_k1k2", p), ("k_min_k1k2", p), ("k_max_k1k2", p),
              ("k3_size_max", i), ("transfers_dir", C.c_char * FN), ("transfers_paths", p),
              ("transfers_files", p), ("transfers_status_file", p),
              ("transfers_status_path", C.c_char * FN), ("tau0", d), ("tau_rec", d),
              ("rs_rec", d), ("transfer2_verbose", sh), ("error_message", C.c_char * ERR),
              ("count_allocated_transfers", lg), ("count_memorised_transfers", lg),
              ("has_allocated_transfers", p), ("stop_at_transfers2", sh)]

    class Song(C.Structure):
        _fields_ = fields

    lib = L.load()
    fn = lib.los_transfers2_layout
    fn.argtypes, fn.restype = [C.POINTER(C.c_long), C.c_int], C.c_int
    buf = (C.c_long * 64)()
    n = fn(buf, 64)
    assert n == len(fields) + 1
    want = [getattr(Song, f).offset for f, _ in fields] + [C.sizeof(Song)]
    assert list(buf[:n]) == want
    assert LBL == 64


def test_hip_sources_never_see_song_structs():
    """Only host C units read SONG's structures (los_transfer2.c fills the flat
    descriptor of los_stage.h for the GPU side): no HIP source, and no header a HIP
    source includes, includes the struct definitions, so recompiling the host unit
    against SONG's headers adapts the library to SONG's layouts."""
    import re
    csrc = os.path.join(os.path.dirname(L.__file__), "csrc")
    inc = re.compile(r'#include\s+"([^"]+)"')
    banned = {"song_los_types.h", "transfer2.h", "los_internal.h"}

    def closure(name, seen):
        if name in seen:
            return
        seen.add(name)
        path = os.path.join(csrc, name)
        if not os.path.exists(path):
            return
        for dep in inc.findall(open(path).read()):
            base = os.path.basename(dep)
            if base in banned:
                raise AssertionError(f"{name} includes {dep}")
            closure(base, seen)

    hips = [f for f in os.listdir(csrc) if f.endswith(".hip")]
    assert len(hips) >= 6
    for h in hips:
        closure(h, set())
