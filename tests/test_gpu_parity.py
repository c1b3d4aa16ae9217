"""GPU parity of the LOS stage against the CPU oracle (bit-exact for the plain FP64 form).

Each case builds the same seeded inputs twice, runs the oracle's transfer2_init on one copy
and the library's transfer2_init (HIP kernels on cuda:0) on the other, and compares every
element of ptr2->transfer. The plain FP64 form keeps the reference's per-element operation
order without FMA contraction, so the bar is equality; the k3 bookkeeping (sizes, grids) is
compared exactly as well.
"""
import numpy as np
import pytest

from song_amd import lib as L
from tests.conftest import run_oracle

CASES = {
    "T_m0": dict(has_T=1, has_E=0, m_list=[0], l_max_los_t=4, k_size=4, k3_size=12, tau_size=80,
                 l_size=10, l_max=300, dk_fine=1e-3, xx_max=1500.0),
    "TEB_m012": dict(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=6, l_max_los_p=6, k_size=5,
                     k3_size=16, tau_size=96, l_size=12, l_max=400, dk_fine=1.5e-3, xx_max=1500.0),
    "E_m02_linear_k3": dict(has_T=0, has_E=1, m_list=[0, 2], l_max_los_p=4, k_size=4, k3_size=10,
                            tau_size=70, l_size=9, l_max=250, dk_fine=2e-3, xx_max=1500.0,
                            k3_interpolation=L.LINEAR),
    "T_flat_extrapolation": dict(has_T=1, m_list=[0, 1], l_max_los_t=3, k_size=4, k3_size=8,
                                 tau_size=64, l_size=8, l_max=200, dk_fine=2e-3, xx_max=1500.0,
                                 extrapolation=L.FLAT_EXTRAP, extra_osc_left=1.5,
                                 extra_osc_right=2.5),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_fp64_bit_exact(name, oracle):
    cfg = L.Config(**CASES[name])
    print(f"case {name}", flush=True)
    ref_s = L.Scenario(cfg)
    ref = run_oracle(ref_s, oracle)
    gpu_s = L.Scenario(cfg)
    gpu_s.transfer2_init()
    got = gpu_s.gather()
    assert got.shape == ref.shape
    assert np.isfinite(got).all()
    nz = np.count_nonzero(ref)
    assert nz > 0.3 * ref.size, "inputs produce too few non-zero transfers to be a test"
    diff = np.flatnonzero(got != ref)
    assert diff.size == 0, (f"{diff.size} of {ref.size} elements differ; first at {diff[:5]}: "
                            f"{got[diff[:5]]} vs {ref[diff[:5]]}")
    gpu_s.close()
    ref_s.close()


@pytest.mark.gpu
def test_closed_form_gpu():
    """J linear in x, sources constant: the library's transfers equal the closed form."""
    from tests.test_oracle_pinning import CF_CFG, closed_form_case, expected_closed_form, \
        _product_k3_list
    s, cfg, fields, tp_of, pair_c, I, lvals = closed_form_case(CF_CFG)
    s.transfer2_init()
    got = s.gather()
    sizes = s.k3_sizes()
    lists, q = {}, 0
    for i1 in range(cfg.k_size):
        for i2 in range(i1 + 1):
            lists[(i1, i2)] = _product_k3_list(s, i1, i2, int(sizes[q, 0]))
            q += 1
    exp = expected_closed_form(s, cfg, fields, pair_c, I, lvals, lists)
    assert got.shape == exp.shape
    assert np.max(np.abs(got - exp)) < 1e-11 * np.max(np.abs(exp))
    s.close()


@pytest.mark.gpu
def test_fp64_bit_exact_wide_grid(oracle):
    """c2-like grid (log k from 1e-5 to 0.1, T+E+B, m=0,1,2, L=0..6) at reduced size."""
    cfg = L.Config(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=6, l_max_los_p=6, k_size=8,
                   k_min=1e-5, k_max=0.1, k3_size=24, tau_size=200, l_size=24, l_max=2000,
                   dk_fine=6e-4, xx_max=1500.0)
    ref = run_oracle(L.Scenario(cfg), oracle)
    s = L.Scenario(cfg)
    s.transfer2_init()
    got = s.gather()
    s.close()
    assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} of {ref.size} differ"


def _field_errors(got, ref, scen):
    sizes = [int(r[0]) for r in scen.k3_sizes()]
    tt2 = got.size // sum(sizes)
    g, r = got.reshape(tt2, -1), ref.reshape(tt2, -1)
    sc = np.abs(r).max(axis=1, keepdims=True)
    return np.divide(np.abs(g - r), sc, out=np.zeros_like(g), where=sc > 0)


def _row_errors(got, ref, scen):
    sizes = [int(r[0]) for r in scen.k3_sizes()]
    tt2 = got.size // sum(sizes)
    err = np.zeros_like(ref)
    o = 0
    for _ in range(tt2):
        for n in sizes:
            r = ref[o:o + n]
            sc = np.max(np.abs(r))
            if sc > 0:
                err[o:o + n] = np.abs(got[o:o + n] - r) / sc
            o += n
    return err


@pytest.mark.gpu
@pytest.mark.parametrize("form,tol_field,tol_row", [(L.FORM_FP64_FAST, 1e-12, 1e-10),
                                                    (L.FORM_FP32C, 1e-5, 1e-4)])
def test_fast_forms_within_tolerance(form, tol_field, tol_row, oracle):
    """FMA FP64 and FP32-compensated forms against the oracle.

    Error of each element relative to the largest |T| of its (l,m) field (all pairs and
    k3 of one transfer type), and, stricter, relative to the largest |T| of its (tt, pair)
    k3 row; single elements are not used as scales because T oscillates in k3."""
    cfg = L.Config(**CASES["TEB_m012"])
    ref_s = L.Scenario(cfg)
    ref = run_oracle(ref_s, oracle)
    lib = L.load()
    s = L.Scenario(cfg)
    assert lib.los_set_default_form(form) == 0
    try:
        s.transfer2_init()
    finally:
        lib.los_set_default_form(L.FORM_FP64)
    got = s.gather()
    field = _field_errors(got, ref, s)
    row = _row_errors(got, ref, s)
    print(f"form {form}: field max {field.max():.3e} rms {np.sqrt(np.mean(field**2)):.3e}; "
          f"row max {row.max():.3e}")
    assert field.max() < tol_field
    assert row.max() < tol_row
    s.close()
    ref_s.close()


@pytest.mark.gpu
def test_fp64_bit_exact_sources_from_disk(oracle, tmp_path):
    """transfer2_init with load_sources_from_disk (source/transfer2.c:329-343): every k1
    level is read from SONG's sources_%03d.dat files that an earlier stage wrote and freed;
    the result equals the oracle's on the in-memory sources bit for bit."""
    cfg = L.Config(**CASES["TEB_m012"])
    print("case sources_from_disk", flush=True)
    ref = run_oracle(L.Scenario(cfg), oracle)
    s = L.Scenario(cfg)
    s.set_string("sources_dir", str(tmp_path))
    s.store_sources()
    s.set("load_sources_from_disk", 1)
    s.transfer2_init()
    got = s.gather()
    assert np.array_equal(got, ref)
    s.close()


@pytest.mark.gpu
def test_store_transfers_files_equal_oracle(oracle, tmp_path):
    """transfer2_init in store_transfers_to_disk mode (source/transfer2.c:549-567,
    :2929-2963): each k1 level is written to SONG's transfers_%03d.dat files and freed;
    the files equal, byte for byte, those the oracle writes from the same inputs."""
    import os
    cfg = L.Config(**CASES["TEB_m012"])
    print("case store_transfers", flush=True)
    dirs = {}
    for who in ("oracle", "gpu"):
        d = tmp_path / who
        d.mkdir()
        s = L.Scenario(cfg)
        s.set("store_transfers_to_disk", 1)
        s.set_string("transfers_dir", str(d))
        if who == "oracle":
            assert oracle.los_ref_transfer2_init(*s.structs()) == 0, s.error_message()
            s._inited = True
            s.free_transfers(lambda a, b, c: oracle.los_ref_transfer2_free(a, b, c))
        else:
            s.transfer2_init()
        s.close()
        dirs[who] = d
    names = sorted(os.listdir(dirs["oracle"]))
    assert names and names == sorted(os.listdir(dirs["gpu"]))
    for n in names:
        a = (dirs["oracle"] / n).read_bytes()
        b = (dirs["gpu"] / n).read_bytes()
        assert len(a) > 0 and a == b, n


@pytest.mark.gpu
def test_small_batches_bit_exact(oracle, monkeypatch):
    """A device budget that holds one k1 level at a time (SONG_LOS_BATCH_BYTES): the k1
    batches of transfer2_init run one after another and the result still equals the
    oracle's bit for bit."""
    cfg = L.Config(**CASES["TEB_m012"])
    ref = run_oracle(L.Scenario(cfg), oracle)
    monkeypatch.setenv("SONG_LOS_BATCH_BYTES", "1")
    print("case small_batches", flush=True)
    s = L.Scenario(cfg)
    s.transfer2_init()
    got = s.gather()
    s.close()
    assert np.array_equal(got, ref)


@pytest.mark.gpu
def test_cubic_J_bit_exact(oracle):
    """Cubic interpolation of J in x (source/transfer2.c:2259-2271, second derivatives from
    the J builder): transfer2_init in the plain form equals the oracle bit for bit, and
    differs from the linear-J result (the cubic branch ran); the fast forms refuse it."""
    kw = dict(CASES["TEB_m012"], bessels_interpolation=L.CUBIC)
    cfg = L.Config(**kw)
    print("case cubic_J", flush=True)
    ref = run_oracle(L.Scenario(cfg), oracle)
    lin = run_oracle(L.Scenario(L.Config(**CASES["TEB_m012"])), oracle)
    s = L.Scenario(cfg)
    s.transfer2_init()
    got = s.gather()
    s.close()
    assert np.array_equal(got, ref)
    assert not np.array_equal(got, lin)
    s = L.Scenario(cfg)
    s.prepare()
    eng = L.Engine(0)
    eng.load(s)
    for form in (L.FORM_FP32C, L.FORM_FP64_FAST):
        eng.set_form(form)
        with pytest.raises(L.LosError, match="cubic"):
            eng.autotune(1)
        with pytest.raises(L.LosError, match="cubic"):
            eng.run()
    eng.close()
    s.close()
