"""The LOS stage as a service (los_server_*) on cuda:0.

* a job's result equals the oracle's transfers of its pairs and (l,m) offsets
  (plain FP64 form: bit for bit);
* each job's result is the same alone and in a mix of concurrently admitted jobs
  (bit for bit, all forms: tiles never mix jobs and the kernel shape is fixed);
* admission respects the HBM budget, and a job larger than the budget is refused;
* jobs admitted while others run overlap on several streams.
"""
import numpy as np
import pytest

from song_amd import lib as L
from tests.conftest import run_oracle

CFG = dict(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=4, l_max_los_p=4, k_size=8,
           k_min=1e-4, k_max=0.05, k3_size=12, tau_size=96, l_size=16, l_max=600,
           dk_fine=2e-3, xx_max=1500.0)


def _jobs(cfg, n, seed):
    rng = np.random.default_rng(seed)
    allp = [(i1, i2) for i1 in range(cfg["k_size"]) for i2 in range(i1 + 1)]
    lmaxes = [40, 150, 400, 600]
    out = []
    for _ in range(n):
        k = int(rng.integers(1, 9))
        pick = [allp[i] for i in rng.choice(len(allp), k, replace=False)]
        out.append((pick, int(rng.choice(lmaxes))))
    return out


def _expected(scen, ref, pairs, l_max):
    """Rows of the oracle's flat transfers for a job: [field x lm with l <= l_max][cols]."""
    offs = scen.pair_offsets()
    ls = scen.grids()[0]
    per_tt = ref.size // (3 * len(scen.lm_l()))
    lm_l = scen.lm_l()
    n_lm = max(i for i, l in enumerate(lm_l) if l <= l_max) + 1
    rows = []
    for field in range(3):
        for lm in range(n_lm):
            tt = field * len(lm_l) + lm
            rows.append(np.concatenate([ref[tt * per_tt + offs[p][0]:
                                            tt * per_tt + offs[p][0] + offs[p][1]]
                                        for p in pairs]))
    del ls
    return np.array(rows)


@pytest.fixture(scope="module")
def scen_ref(oracle):
    cfg = L.Config(**CFG)
    ref = run_oracle(L.Scenario(cfg), oracle)
    s = L.Scenario(cfg)
    s.prepare()
    yield s, ref
    s.close()


@pytest.mark.gpu
def test_job_results_match_oracle(scen_ref):
    s, ref = scen_ref
    srv = L.Server(s, L.FORM_FP64, budget_bytes=1 << 30, slots=2)
    for pairs, l_max in _jobs(CFG, 6, 1):
        print(f"job {pairs} l_max {l_max}", flush=True)
        got, _ = srv.wait(srv.submit(pairs, l_max), 60000)
        assert np.array_equal(got, _expected(s, ref, pairs, l_max))
    srv.close()


@pytest.mark.gpu
@pytest.mark.parametrize("form", [L.FORM_FP64, L.FORM_FP32C, L.FORM_FP64_FAST])
def test_alone_equals_mix(scen_ref, form):
    s, _ = scen_ref
    jobs = _jobs(CFG, 24, 2)
    srv = L.Server(s, form, budget_bytes=1 << 30, slots=3)
    alone = [srv.wait(srv.submit(p, lm), 60000)[0] for p, lm in jobs]
    assert srv.stats()["launches"] == len(jobs)
    srv.hold(True)  # queue every job, then admit them together: merged launches
    ids = [srv.submit(p, lm) for p, lm in jobs]
    srv.hold(False)
    mixed = [srv.wait(i, 60000)[0] for i in ids]
    st = srv.stats()
    print(f"form {form}: {st}", flush=True)
    assert st["launches"] <= len(jobs) + 3  # the mix ran in at most one launch per slot
    for a, b in zip(alone, mixed):
        assert np.array_equal(a, b)
    srv.close()


@pytest.mark.gpu
def test_budget_admission(scen_ref):
    s, _ = scen_ref
    jobs = _jobs(CFG, 16, 3)
    # a budget that holds only a couple of jobs at a time
    probe = L.Server(s, L.FORM_FP32C, budget_bytes=1 << 30, slots=1)
    sizes = []
    for p, lm in jobs:
        jid = probe.submit(p, lm)
        sizes.append(8 * probe.shape(jid)[0])
        probe.wait(jid, 60000)
    probe.close()
    budget = 2.5 * max(sizes) + 3 * 4096 * 4
    srv = L.Server(s, L.FORM_FP32C, budget_bytes=budget, slots=4)
    ids = [srv.submit(p, lm) for p, lm in jobs]
    for i in ids:
        srv.wait(i, 60000)
    st = srv.stats()
    print(st, flush=True)
    assert 0 < st["peak_admitted_bytes"] <= budget
    with pytest.raises(L.LosError):
        allp = [(i1, i2) for i1 in range(CFG["k_size"]) for i2 in range(i1 + 1)]
        srv.submit(allp, 600)  # needs more than the whole budget
    srv.close()


OVERLAP = dict(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=6, l_max_los_p=6, k_size=16,
               k_min=1e-5, k_max=0.1, k3_size=60, tau_size=600, l_size=96, l_max=2000,
               dk_fine=1.5e-4, xx_max=1500.0)


@pytest.mark.gpu
def test_jobs_arriving_during_a_launch(oracle):
    """Continuous batching: small jobs submitted while a long launch runs are admitted to
    other slots and finish before it (their admission precedes its completion, and at
    least two slots were busy at once); each equals its result alone and the oracle's rows
    bit for bit (plain FP64 form)."""
    import time
    from tests.conftest import oracle_pairs

    cfg = L.Config(**OVERLAP)
    s = L.Scenario(cfg)
    s.prepare()
    allp = list(s.pair_offsets())
    small = [([allp[5]], 500), ([allp[40], allp[7]], 1000), ([allp[-1]], 2000)]
    want = oracle_pairs(s, oracle, sorted({p for pp, _ in small for p in pp}))
    lm_l = s.lm_l()
    n_all = len(lm_l)

    def expected(pp, lmax):
        n_lm = max(i for i, l in enumerate(lm_l) if l <= lmax) + 1
        return np.array([np.concatenate([want[p][f * n_all + q] for p in pp])
                         for f in range(3) for q in range(n_lm)])

    srv = L.Server(s, L.FORM_FP64, budget_bytes=16 * 2.0 ** 30, slots=3)
    alone = [srv.wait(srv.submit(pp, lm), 120000)[0] for pp, lm in small]
    big = srv.submit(allp, 2000)
    time.sleep(0.005)  # the long job is admitted and running
    ids = [srv.submit(pp, lm) for pp, lm in small]
    res = [srv.wait(i, 120000) for i in ids]
    _, (_, _, big_done) = srv.wait(big, 120000)
    st = srv.stats()
    admits = [t[1] for _, t in res]
    print(f"big done at {big_done:.4f}; small admitted at {admits}; stats {st}", flush=True)
    assert st["max_busy_slots"] >= 2
    assert min(admits) < big_done
    for (pp, lm), a, (got, _) in zip(small, alone, res):
        assert np.array_equal(got, a)
        assert np.array_equal(got, expected(pp, lm))
    srv.close()
    s.close()
