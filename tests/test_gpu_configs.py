"""GPU parity on the benchmark workloads themselves (c1 to c5), against the CPU oracle.

The inputs are the benchmark's (bench.py WORKLOADS: seeded synthetic sources, J tables on
x step 0.3), at full size:

* c1: T only, m = 0, L = 0..4, 16x16 (k1,k2), 400 tau, 48 l: transfer2_init bit-exact.
* c2: T+E+B, m = 0,1,2, L = 0..6, all 528 pairs, 600 tau, 96 l to 2000: transfer2_init
  bit-exact; through the engine with the autotuned shapes, the plain form bit-exact, the
  FMA-FP64 form within 1e-12 and the FP32-compensated form within 1e-5 of each (l,m)
  field's largest |T| (the row-normalised figure is printed; see
  test_gpu_parity.test_fast_forms_within_tolerance for what it measures).
* c3: 2048 pairs (8 to ~400 fine k3 each) in one captured graph; the oracle on 64 pairs
  that span the fine k3 counts, every pair's rows bit-exact (plain) or within tolerance.
* c4: c2 at l_max 3000 with 128 l; the oracle on 64 pairs spread over the k3 counts; the
  FP32-compensated form's max and RMS error per (l,m) against it.
* c5: serving jobs of 1-64 pairs with l_max in {500, 1000, 2000, 3000} on the c3 grid
  with 128 l to 3000; each job's rows equal the oracle's (plain form, bit for bit), and
  alone equals mixed (FP32-compensated form).
"""
import numpy as np
import pytest

import bench
from song_amd import lib as L
from tests.conftest import oracle_pairs, product_pairs, run_oracle

TOL = {L.FORM_FP64_FAST: 1e-12, L.FORM_FP32C: 1e-5}
pytestmark = pytest.mark.timeout(600)


def _spread(scen, pairs, n):
    """n pairs of `pairs` spread evenly over their fine k3 counts (smallest and largest
    included)."""
    offs = scen.pair_offsets()
    by = sorted(pairs, key=lambda p: (offs[p][1], p))
    idx = np.unique(np.linspace(0, len(by) - 1, n).round().astype(int))
    return [by[i] for i in idx]


def _field_rel(got, ref):
    """{pair: [tt][k3]} dicts -> per-element |got - ref| / max|ref| of the same tt over the
    compared pairs, and the same over each (tt, pair) row, flattened [tt][all k3]."""
    g = np.concatenate([got[p] for p in ref], axis=1)
    r = np.concatenate([ref[p] for p in ref], axis=1)
    d = np.abs(g - r)
    sc = np.abs(r).max(axis=1, keepdims=True)
    field = np.divide(d, sc, out=np.zeros_like(d), where=sc > 0)
    rows = []
    for p in ref:
        rr, dd = ref[p], np.abs(got[p] - ref[p])
        s = np.abs(rr).max(axis=1, keepdims=True)
        rows.append(np.divide(dd, s, out=np.zeros_like(dd), where=s > 0))
    return field, np.concatenate(rows, axis=1)


def _engine_form(eng, scen, form, graph=False):
    eng.set_form(form)
    eng.autotune(1)
    if graph:
        eng.capture()
    eng.run()
    eng.run()
    eng.store(scen)
    print(f"form {form} shape {eng.get_shape(form)}", flush=True)
    return scen.gather()


def _assert_exact(got, ref, what):
    for p in ref:
        bad = np.flatnonzero(got[p] != ref[p])
        assert bad.size == 0, (f"{what}: pair {p}: {bad.size} of {ref[p].size} differ, "
                               f"first {got[p].ravel()[bad[:3]]} vs {ref[p].ravel()[bad[:3]]}")


@pytest.mark.gpu
def test_c1_dropin_bit_exact(oracle):
    cfg = L.Config(**bench.WORKLOADS["c1"])
    print("c1 oracle", flush=True)
    ref = run_oracle(L.Scenario(cfg), oracle)
    s = L.Scenario(cfg)
    print("c1 transfer2_init", flush=True)
    s.transfer2_init()
    got = s.gather()
    s.close()
    assert got.shape == ref.shape and np.count_nonzero(ref) > 0.3 * ref.size
    assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} of {ref.size} differ"


@pytest.fixture(scope="module")
def c2_ref(oracle):
    cfg = L.Config(**bench.WORKLOADS["c2"])
    print("c2 oracle (528 pairs)", flush=True)
    return cfg, run_oracle(L.Scenario(cfg), oracle)


@pytest.mark.gpu
def test_c2_dropin_bit_exact(c2_ref):
    cfg, ref = c2_ref
    s = L.Scenario(cfg)
    print("c2 transfer2_init", flush=True)
    s.transfer2_init()
    got = s.gather()
    s.close()
    assert got.shape == ref.shape and np.isfinite(got).all()
    diff = np.flatnonzero(got != ref)
    assert diff.size == 0, f"{diff.size} of {ref.size} differ, first at {diff[:5]}"


@pytest.mark.gpu
def test_c2_forms_against_oracle(c2_ref):
    cfg, ref = c2_ref
    s = L.Scenario(cfg)
    s.prepare()
    eng = L.Engine(0)
    eng.load(s)
    pairs = list(s.pair_offsets())
    want = product_pairs(s, ref, pairs)
    got = _engine_form(eng, s, L.FORM_FP64)
    _assert_exact(product_pairs(s, got, pairs), want, "c2 fp64")
    for form in (L.FORM_FP64_FAST, L.FORM_FP32C):
        got = product_pairs(s, _engine_form(eng, s, form), pairs)
        field, row = _field_rel(got, want)
        print(f"c2 form {form}: field max {field.max():.3e} rms {np.sqrt(np.mean(field**2)):.3e}"
              f"; row max {row.max():.3e}", flush=True)
        assert field.max() < TOL[form]
    eng.close()
    s.close()


@pytest.mark.gpu
def test_c3_batched_graph(oracle):
    cfg = L.Config(**bench.WORKLOADS["c3"])
    s = L.Scenario(cfg)
    s.prepare()
    pairs = bench.select_pairs(cfg, bench.PAIRS["c3"])
    sample = _spread(s, pairs, 64)
    offs = s.pair_offsets()
    ks = [offs[p][1] for p in sample]
    print(f"c3 oracle on {len(sample)} pairs, fine k3 {min(ks)}..{max(ks)}", flush=True)
    assert min(ks) <= 16 and max(ks) >= 300
    want = oracle_pairs(s, oracle, sample)
    eng = L.Engine(0)
    eng.load_pairs(s, pairs)
    got = product_pairs(s, _engine_form(eng, s, L.FORM_FP64, graph=True), sample)
    _assert_exact(got, want, "c3 fp64 graph")
    for form in (L.FORM_FP64_FAST, L.FORM_FP32C):
        got = product_pairs(s, _engine_form(eng, s, form, graph=True), sample)
        field, row = _field_rel(got, want)
        print(f"c3 form {form}: field max {field.max():.3e}; row max {row.max():.3e}", flush=True)
        assert field.max() < TOL[form]
    eng.close()
    s.close()


@pytest.mark.gpu
def test_c4_fp32c_per_lm(oracle):
    cfg = L.Config(**bench.WORKLOADS["c4"])
    s = L.Scenario(cfg)
    s.prepare()
    pairs = list(s.pair_offsets())
    sample = _spread(s, pairs, 64)
    print(f"c4 oracle on {len(sample)} pairs", flush=True)
    want = oracle_pairs(s, oracle, sample)
    eng = L.Engine(0)
    eng.load(s)
    _assert_exact(product_pairs(s, _engine_form(eng, s, L.FORM_FP64), sample), want, "c4 fp64")
    got = product_pairs(s, _engine_form(eng, s, L.FORM_FP32C), sample)
    field, row = _field_rel(got, want)
    per_max = field.max(axis=1)
    per_rms = np.sqrt(np.mean(field ** 2, axis=1))
    worst = int(np.argmax(per_max))
    print(f"c4 fp32c per (l,m): max of max {per_max.max():.3e} (tt {worst}), "
          f"max of rms {per_rms.max():.3e}, median max {np.median(per_max):.3e}; "
          f"row max {row.max():.3e}", flush=True)
    assert per_max.max() < TOL[L.FORM_FP32C]
    assert per_rms.max() < TOL[L.FORM_FP32C] / 10
    eng.close()
    s.close()


@pytest.mark.gpu
def test_c5_serving_jobs(oracle):
    cfg = L.Config(**bench.WORKLOADS["c5"])
    s = L.Scenario(cfg)
    s.prepare()
    allp = list(s.pair_offsets())
    rng = np.random.default_rng(11)
    jobs = []
    for lmax in (500, 1000, 2000, 3000, 3000):
        k = int(rng.integers(1, 65))
        jobs.append(([allp[i] for i in rng.choice(len(allp), k, replace=False)], lmax))
    jobs.append(([allp[-1]], 2000))  # the largest pair (k1 = k2 = 0.2) alone
    union = sorted({p for pp, _ in jobs for p in pp})
    print(f"c5: {len(jobs)} jobs, oracle on {len(union)} pairs", flush=True)
    want = oracle_pairs(s, oracle, union)
    lm_l = s.lm_l()
    n_lm_all = len(lm_l)

    def expected(pp, lmax):
        n_lm = max(i for i, l in enumerate(lm_l) if l <= lmax) + 1
        rows = [np.concatenate([want[p][f * n_lm_all + q] for p in pp])
                for f in range(3) for q in range(n_lm)]
        return np.array(rows)

    srv = L.Server(s, L.FORM_FP64, budget_bytes=8 * 2.0 ** 30, slots=3)
    srv.hold(True)
    ids = [srv.submit(pp, lm) for pp, lm in jobs]
    srv.hold(False)
    for (pp, lm), i in zip(jobs, ids):
        got, _ = srv.wait(i, 120000)
        assert np.array_equal(got, expected(pp, lm)), f"job of {len(pp)} pairs l_max {lm}"
    srv.close()
    # field scale: the largest |T| of each transfer type over every pair of the jobs
    fscale = np.max(np.abs(np.concatenate([want[p] for p in union], axis=1)), axis=1)
    srv = L.Server(s, L.FORM_FP32C, budget_bytes=8 * 2.0 ** 30, slots=3)
    alone = [srv.wait(srv.submit(pp, lm), 120000)[0] for pp, lm in jobs]
    srv.hold(True)
    ids = [srv.submit(pp, lm) for pp, lm in jobs]
    srv.hold(False)
    for (pp, lm), i, a in zip(jobs, ids, alone):
        got, _ = srv.wait(i, 120000)
        assert np.array_equal(got, a)
        exp = expected(pp, lm)
        n_lm = exp.shape[0] // 3
        sc = np.array([fscale[f * n_lm_all + q] for f in range(3) for q in range(n_lm)])
        assert np.all(np.abs(got - exp) <= TOL[L.FORM_FP32C] * sc[:, None])
    srv.close()
    s.close()
