"""Batched form of the LOS engine on cuda:0.

* every tile shape of the FMA-FP64 and FP32-compensated kernels (dispatch and
  persistent work-queue forms) against the oracle;
* explicit pair lists in any order: each pair's k3 row equals the oracle's bit for bit
  (plain FP64 form), whatever else is in the batch;
* uneven triangle-limited k3 counts (8 to several hundred per pair) in one batch;
* hipGraph capture: replays, interleaved with eager runs, equal the eager result.
"""
import numpy as np
import pytest

from song_amd import lib as L
from tests.conftest import field_errors, run_oracle
from tests.test_gpu_parity import CASES

_STREAM = [(0, nk, ch, p) for nk in (2, 4, 8) for ch in (16, 32) for p in (0, 1)]
_KLANE = [(2, 64, b, 0) for b in (1, 2, 4)]
# sweep: ch = 1 is the FP32 pair table (the FP64 forms always read their pair table)
SHAPES = {L.FORM_FP64_FAST: _STREAM + [(1, nk, 0, p) for nk in (1, 2, 4, 8) for p in (0, 1)] +
          _KLANE,
          L.FORM_FP32C: _STREAM + [(1, nk, c, p) for nk in (1, 2, 4, 8) for c in (0, 1)
                                   for p in (0, 1)] + _KLANE}
# the plain form's kernels: per-k3 workgroups, sweep (dispatch and persistent), k-lane
FP64_SHAPES = ([(0, 1, 0, 0)] + [(1, nk, 0, p) for nk in (1, 2, 4, 8) for p in (0, 1)] +
               [(2, 64, b, 0) for b in (1, 2, 4)])

# log k in [1e-5, 0.2], 8 coarse k3 per pair, fine step 5e-4: fine k3 counts from 8
# (narrow triangles, coarse grid only) to about 400 (k1 = k2 = 0.2)
UNEVEN = dict(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=6, l_max_los_p=6, k_size=10,
              k_min=1e-5, k_max=0.2, k3_size=8, tau_size=160, l_size=12, l_max=2000,
              dk_fine=5e-4, xx_max=6000.0)


def _run(eng, scen, form):
    eng.set_form(form)
    eng.run()
    eng.store(scen)
    return scen.gather()


def _tt2(scen, flat):
    return flat.size // int(scen.k3_sizes()[:, 0].sum())


@pytest.mark.gpu
def test_every_tile_shape(oracle):
    cfg = L.Config(**CASES["TEB_m012"])
    ref = run_oracle(L.Scenario(cfg), oracle)
    s = L.Scenario(cfg)
    s.prepare()
    eng = L.Engine(0)
    eng.load(s)
    eng.set_form(L.FORM_FP64)
    for sh in FP64_SHAPES:
        # every kernel of the plain form is bit-exact
        eng.set_shape(L.FORM_FP64, *sh)
        print(f"form 0 shape {sh}", flush=True)
        assert np.array_equal(_run(eng, s, L.FORM_FP64), ref), f"fp64 shape {sh}"
    tt2 = _tt2(s, ref)
    for form, tol in ((L.FORM_FP64_FAST, 1e-12), (L.FORM_FP32C, 1e-5)):
        eng.set_form(form)
        for sh in SHAPES[form]:
            eng.set_shape(form, *sh)
            print(f"form {form} shape {sh}", flush=True)
            err = field_errors(_run(eng, s, form), ref, tt2)
            assert err.max() < tol, f"form {form} shape {sh}: {err.max():.3e}"
    eng.close()
    s.close()


@pytest.mark.gpu
def test_pair_list_any_order(oracle):
    cfg = L.Config(**CASES["TEB_m012"])
    ref = run_oracle(L.Scenario(cfg), oracle)
    s = L.Scenario(cfg)
    s.prepare()
    offs = s.pair_offsets()
    tt2 = _tt2(s, ref)
    per_tt = ref.size // tt2
    pairs = list(offs)
    rng = np.random.default_rng(7)
    eng = L.Engine(0)
    for trial in range(3):
        pick = [pairs[i] for i in rng.permutation(len(pairs))[: 3 + 4 * trial]]
        print(f"pairs {pick}", flush=True)
        eng.load_pairs(s, pick)
        got = _run(eng, s, L.FORM_FP64)
        for p in pick:
            o, n = offs[p]
            for tt in range(tt2):
                a = tt * per_tt + o
                assert np.array_equal(got[a:a + n], ref[a:a + n]), f"pair {p} tt {tt}"
    eng.close()
    s.close()


@pytest.mark.gpu
def test_uneven_k3_counts(oracle):
    cfg = L.Config(**UNEVEN)
    ref = run_oracle(L.Scenario(cfg), oracle)
    s = L.Scenario(cfg)
    s.prepare()
    sizes = s.k3_sizes()[:, 0]
    print(f"fine k3 per pair: min {sizes.min()} max {sizes.max()}", flush=True)
    assert sizes.min() <= 16 and sizes.max() >= 300
    eng = L.Engine(0)
    eng.load(s)
    eng.set_form(L.FORM_FP64)
    for sh in ((0, 1, 0), (1, 2, 0), (1, 2, 0, 1), (2, 64, 2)):
        eng.set_shape(L.FORM_FP64, *sh)
        assert np.array_equal(_run(eng, s, L.FORM_FP64), ref), f"fp64 shape {sh}"
    tt2 = _tt2(s, ref)
    for form, tol in ((L.FORM_FP64_FAST, 1e-12), (L.FORM_FP32C, 1e-5)):
        eng.set_form(form)
        eng.autotune(1)
        print(f"form {form} tuned shape {eng.get_shape(form)}", flush=True)
        err = field_errors(_run(eng, s, form), ref, tt2)
        assert err.max() < tol, f"form {form}: {err.max():.3e}"
        for sh in ((0, 4, 32, 1), (1, 2, 0, 0), (1, 4, 0, 1), (2, 64, 4, 0)) + \
                (((1, 2, 1, 0),) if form == L.FORM_FP32C else ()):
            eng.set_shape(form, *sh)
            err = field_errors(_run(eng, s, form), ref, tt2)
            assert err.max() < tol, f"form {form} shape {sh}: {err.max():.3e}"
    eng.close()
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("form,shape", [(L.FORM_FP64, (0, 1)), (L.FORM_FP64, (2, 64, 2)),
                                        (L.FORM_FP64, (1, 2)), (L.FORM_FP64, (1, 2, 0, 1)),
                                        (L.FORM_FP32C, (0, 4, 32, 1)),
                                        (L.FORM_FP64_FAST, (0, 8, 16, 0)),
                                        (L.FORM_FP32C, (1, 2, 0, 0)), (L.FORM_FP32C, (1, 2, 1, 0)),
                                        (L.FORM_FP32C, (2, 64, 4))])
def test_graph_replay_matches_eager(form, shape):
    cfg = L.Config(**CASES["TEB_m012"])
    s = L.Scenario(cfg)
    s.prepare()
    eng = L.Engine(0)
    eng.load(s)
    eng.set_form(form)
    eng.set_shape(form, *shape)
    eager = _run(eng, s, form)
    eng.capture()
    for rep in range(3):
        got = _run(eng, s, form)
        assert np.array_equal(got, eager), f"replay {rep}"
        # eager work of another form between replays
        other = L.FORM_FP64 if form != L.FORM_FP64 else L.FORM_FP32C
        _run(eng, s, other)
        eng.set_form(form)
    eng.close()
    s.close()


@pytest.mark.gpu
def test_flop_count_matches_brute_force():
    """los_engine_flops: sum over elements and integrals of sum_tau (5 L_live + 2 [L_live > 0])
    with the reference's skip rule, recounted here from the J rows' first live indices."""
    import ctypes as C

    cfg = L.Config(**CASES["TEB_m012"])
    s = L.Scenario(cfg)
    s.prepare()
    eng = L.Engine(0)
    eng.load(s)
    live, dense = eng.flops()
    ls, _, ts = s.grids()
    t0mt = cfg.tau0 - ts
    lib = L.load()
    kf = []
    for (i1, i2), (_, n) in s.pair_offsets().items():
        buf, last = (C.c_double * n)(), C.c_int()
        lib.transfer2_get_k3_list(*[s.s[x] for x in ("ppr", "ppr2", "ppt2", "pbs", "pbs2", "ptr2")],
                                  i1, i2, buf, C.byref(last))
        kf.extend(list(buf))
    tot = dtot = 0
    for k in kf:
        ix = (k * t0mt / cfg.xx_step).astype(np.int64)
        for il, l in enumerate(ls):
            for im, m in enumerate(cfg.m_list):
                if m > l:
                    continue
                calls = [(0, cfg.l_max_los_t), (1, cfg.l_max_los_p)]  # T: TT; E: EE
                if m != 0:
                    calls.append((2, cfg.l_max_los_p))                # E: EB x S_B
                calls += [(1, cfg.l_max_los_p), (2, cfg.l_max_los_p)]  # B: EE, EB
                for jt, lmax in calls:
                    xms = [s.J_row(jt, Lv, il, im)[0] for Lv in range(m, lmax + 1)]
                    n_live = sum((ix >= x).astype(np.int64) for x in xms)
                    tot += 5 * int(n_live.sum()) + 2 * int((n_live > 0).sum())
                    dtot += len(ts) * (5 * len(xms) + 2)
    eng.close()
    s.close()
    assert live == tot and dense == dtot


@pytest.mark.gpu
@pytest.mark.parametrize("form,shape", [(L.FORM_FP64, (1, 2, 0, 0)), (L.FORM_FP64, (1, 4, 0, 0)),
                                        (L.FORM_FP64_FAST, (1, 4, 0, 0)),
                                        (L.FORM_FP32C, (1, 2, 1, 0)), (L.FORM_FP32C, (1, 4, 0, 0))])
def test_pipelined_steps(oracle, form, shape):
    """Chunked steps (preparation of chunk c+1 on a second stream while chunk c's LOS
    kernel runs), eager and captured, equal the oracle as the single-pass steps do."""
    cfg = L.Config(**CASES["TEB_m012"])
    ref = run_oracle(L.Scenario(cfg), oracle)
    s = L.Scenario(cfg)
    s.prepare()
    eng = L.Engine(0)
    eng.set_pipeline(3)
    eng.load(s)
    eng.set_form(form)
    eng.set_shape(form, *shape)
    tt2 = _tt2(s, ref)
    for graph in (False, True):
        if graph:
            eng.capture()
        for rep in range(2):
            got = _run(eng, s, form)
            print(f"form {form} shape {shape} graph {graph} rep {rep}", flush=True)
            if form == L.FORM_FP64:
                assert np.array_equal(got, ref)
            else:
                tol = 1e-12 if form == L.FORM_FP64_FAST else 1e-5
                assert field_errors(got, ref, tt2).max() < tol
    eng.close()
    s.close()



@pytest.mark.gpu
@pytest.mark.parametrize("form", [L.FORM_FP64, L.FORM_FP64_FAST, L.FORM_FP32C])
def test_fused_tables_equal_two_pass_tables(oracle, form):
    """Sweep tables made straight from the raw sources (one kernel, spline in LDS) and
    through the re-laid, splined sources in HBM give the same transfers: bit for bit in
    the plain form (and equal to the oracle), within the form's tolerance otherwise."""
    cfg = L.Config(**UNEVEN)
    ref = run_oracle(L.Scenario(cfg), oracle)
    s = L.Scenario(cfg)
    s.prepare()
    eng = L.Engine(0)
    eng.load(s)
    tt2 = _tt2(s, ref)
    outs = []
    for fused in (True, False):
        eng.set_fused_prep(fused)
        eng.set_form(form)
        eng.set_shape(form, 1, 4, 0, 0)
        print(f"form {form} fused {fused}", flush=True)
        outs.append(_run(eng, s, form))
    if form == L.FORM_FP64:
        assert np.array_equal(outs[0], ref) and np.array_equal(outs[1], ref)
    else:
        tol = 1e-12 if form == L.FORM_FP64_FAST else 1e-5
        for o in outs:
            assert field_errors(o, ref, tt2).max() < tol
    eng.close()
    s.close()
