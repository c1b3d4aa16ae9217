"""Closed forms that pin the oracle's k3 interpolation, J skip, flat extrapolation and time
grids (CPU), each with planted-mistake companions.

Every case runs the oracle's transfer2_init on inputs for which the mathematics fixes the
answer, independently of the oracle's code:

* sources quadratic in k3 and constant in tau, cubic k3 interpolation: the clamped spline
  with end slopes from the parabola through the three end points reproduces a quadratic
  exactly, so each fine k3 sees P(k3) (source/transfer2.c:2639-2706); with sources linear
  in k3, linear interpolation is exact (:2695-2699);
* J rows linear in x stored from a per-row first live index index_xmin_J > 0: the
  reference skips the steps with index_x < index_xmin_J (:2232). With x decreasing in tau,
  the live steps are a prefix t <= c, and 1/2 sum_t dtau_t f_t over that prefix is the
  exact integral of the linear integrand over [tau_0, tau_c] plus (tau_{c+1} - tau_c)
  f(tau_c) / 2 (the last live step keeps its full trapezoid weight);
* flat k3 extrapolation (:2740-2758) with sources that vary in k3: points left of the
  physical range take the value at the first physical point, points right of it the value
  at the last;
* custom and bessel tau sampling (:2366-2585, :2783-2908): the grid is sorted, holds every
  sampling time of the sources, starts and ends at theirs, has no gap wider than
  tau_step_max and adds ceil(gap / tau_step_max) - 1 points inside each gap of the sources'
  sampling; sum dtau / 2 = tau_max - tau_min; with J constant and sources linear in tau,
  linear and cubic time interpolation and the trapezoid rule are exact.

Each companion builds the oracle again with one deliberate mistake (LOS_PLANT = n in
oracle/los_ref.c: swapped spline weights, a wrong h^2/6 term, <= in the J skip, swapped
extrapolation fills, swapped linear weights in k3 or in time) and checks that the same
closed form then fails.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

from song_amd import lib as L
from tests.conftest import ROOT, load_oracle

BASE = dict(has_T=1, has_E=1, m_list=[0, 2], l_max_los_t=2, l_max_los_p=2, k_size=3,
            k_min=2e-3, k_max=1e-2, k3_size=7, tau_size=40, tau_min=200.0, tau_max=14000.0,
            tau0=14200.0, l_size=4, l_max=40, dk_fine=1e-3, xx_max=200.0, xx_step=0.25)
U = np.array([0.0, 0.08, 0.21, 0.4, 0.57, 0.81, 1.0])  # non-uniform coarse k3 positions


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    """Oracle builds with deliberate mistake n (compiled on first use)."""
    d = tmp_path_factory.mktemp("planted")
    built = {}

    def get(n):
        if n not in built:
            out = str(d / f"liblos_oracle_plant{n}.so")
            subprocess.run(["gcc", "-O1", "-std=c99", "-D_GNU_SOURCE", "-fopenmp", "-fPIC",
                            "-shared", "-DDEBUG", f"-DLOS_PLANT={n}", "-Wno-format-truncation",
                            "-o", out, os.path.join(ROOT, "oracle", "los_ref.c"), "-lm"],
                           check=True)
            built[n] = load_oracle(out)
        return built[n]

    return get


def _pair_grid(k1, k2, kmin, kmax):
    """A non-uniform coarse k3 grid inside the pair's triangle (the generator's bounds)."""
    lo = max(abs(k1 - k2) + 1e-10, (k1 + k2) / 100.0, kmin)
    hi = min(k1 + k2 - 1e-10, kmax)
    return lo + (hi - lo) * U


class Case:
    """Inputs with J rows a + b x stored from a chosen first index and sources
    c(key) P(k3) Q(tau); expected transfers from the closed forms."""

    def __init__(self, cfg_kw, seed, P, Q=None, xmin_max=0, b_scale=0.01):
        self.cfg = L.Config(**cfg_kw)
        cfg = self.cfg
        self.s = s = L.Scenario(cfg)
        rng = np.random.default_rng(seed)
        self.lvals, self.kvals, self.tau = s.grids()
        self.P, self.Q = P, Q or (lambda t: np.ones_like(t))
        m = cfg.m_list
        self.jtypes = ["TT", "EE", "EB"] if cfg.has_E else ["TT"]
        L_size = max(cfg.l_max_los_t, cfg.l_max_los_p if cfg.has_E else 0, max(m)) + 1
        xx_size = int(cfg.xx_max / cfg.xx_step + 1)
        xx = np.arange(xx_size) * (cfg.xx_max / (xx_size - 1))
        self.rows = {}
        for jt_i, jt in enumerate(self.jtypes):
            for Lq in range(L_size):
                for il, l in enumerate(self.lvals):
                    for im, mm in enumerate(m):
                        if mm > min(Lq, l):
                            continue
                        a, b = rng.normal(), rng.normal() * b_scale
                        xm = int(rng.integers(0, xmin_max + 1))
                        self.rows[(jt, Lq, il, im)] = (a, b, xm)
                        s.override_J(jt_i, Lq, il, im, xm, (a + b * xx)[xm:])
        self.fields = ["T"] + (["E", "B"] if cfg.has_E else [])
        tp_of, tp = {}, 0
        for f in self.fields:
            Lm = cfg.l_max_los_t if f == "T" else cfg.l_max_los_p
            for Lq in range(Lm + 1):
                for mm in m:
                    if mm <= Lq:
                        tp_of[(f, Lq, mm)] = tp
                        tp += 1
        self.c = {}
        q = self.Q(self.tau)
        for i1 in range(cfg.k_size):
            for i2 in range(i1 + 1):
                kc = _pair_grid(self.kvals[i1], self.kvals[i2], self.kvals[0], self.kvals[-1])
                s.override_k3(i1, i2, kc)
                for key, t in tp_of.items():
                    c = rng.normal()
                    self.c[(i1, i2) + key] = c
                    s.source(t, i1, i2)[:, :] = c * np.outer(q, P(kc))

    def run(self, oracle):
        s = self.s
        rc = oracle.los_ref_transfer2_init(*s.structs())
        s._inited = True
        assert rc == 0, s.error_message()
        got = s.gather()
        sizes = s.k3_sizes()
        lists, q = {}, 0
        for i1 in range(self.cfg.k_size):
            for i2 in range(i1 + 1):
                n, first, nphys = (int(v) for v in sizes[q])
                buf, last = np.zeros(n + 4), C.c_int()
                st = s.s
                assert oracle.los_ref_get_k3_list(
                    st["ppr"], st["ppr2"], st["ppt2"], st["pbs2"], st["ptr2"], i1, i2,
                    buf.ctypes.data_as(C.POINTER(C.c_double)), C.byref(last)) == 0
                lists[(i1, i2)] = (buf[:n], first, nphys)
                q += 1
        s.free_transfers(lambda a, b, c: oracle.los_ref_transfer2_free(a, b, c))
        return got, lists

    def integral(self, row, k):
        """1/2 sum_t dtau_t J(k (tau0 - tau_t)) over the live steps, J = a + b x, closed form."""
        a, b, xm = self.rows[row]
        cfg, tau = self.cfg, self.tau
        T0 = cfg.tau0
        ix = (k * (T0 - tau) / cfg.xx_step).astype(np.int64)
        live = ix >= xm
        if not live.any():
            return 0.0
        c = int(np.flatnonzero(live)[-1])
        assert live[:c + 1].all(), "live steps must be a prefix"

        def F(t1, t2):  # integral of a + b k (T0 - tau) over [t1, t2]
            return a * (t2 - t1) + b * k * (T0 * (t2 - t1) - (t2 ** 2 - t1 ** 2) / 2)

        if c == tau.size - 1:
            return F(tau[0], tau[-1])
        return F(tau[0], tau[c]) + 0.5 * (tau[c + 1] - tau[c]) * (a + b * k * (T0 - tau[c]))

    def expected(self, lists, flat=False):
        cfg = self.cfg
        m = cfg.m_list
        out = []
        for f in self.fields:
            Lm = cfg.l_max_los_t if f == "T" else cfg.l_max_los_p
            for il, l in enumerate(self.lvals):
                for im, mm in enumerate(m):
                    if mm > l:
                        continue
                    for i1 in range(cfg.k_size):
                        for i2 in range(i1 + 1):
                            ks, first, nphys = lists[(i1, i2)]
                            for ik, k in enumerate(ks):
                                # flat extrapolation: the value at the nearest physical point
                                ke = ks[min(max(ik, first), first + nphys - 1)] if flat else k
                                pk = self.P(np.array([ke]))[0]
                                Ls = [Lq for Lq in range(mm, Lm + 1)]
                                c = lambda F, Lq: self.c[(i1, i2, F, Lq, mm)] * pk

                                def I(jt, Lq):
                                    return self.integral((jt, Lq, il, im), k)

                                if f == "T":
                                    v = sum(c("T", Lq) * I("TT", Lq) for Lq in Ls)
                                elif f == "E":
                                    v = sum(c("E", Lq) * I("EE", Lq) for Lq in Ls)
                                    if mm != 0:
                                        v += sum(c("B", Lq) * I("EB", Lq) for Lq in Ls)
                                else:
                                    v = sum(c("B", Lq) * I("EE", Lq) for Lq in Ls) - \
                                        sum(c("E", Lq) * I("EB", Lq) for Lq in Ls)
                                out.append(v / (8.0 * (2 * l + 1)))
        return np.array(out)

    def error(self, oracle, flat=False):
        got, lists = self.run(oracle)
        exp = self.expected(lists, flat)
        assert got.shape == exp.shape
        return float(np.max(np.abs(got - exp)) / np.max(np.abs(exp)))

    def close(self):
        self.s.close()


def quad(k):
    return 1.0 + 40.0 * k + 3000.0 * k * k


def lin(k):
    return 1.0 - 60.0 * k


CASES = {
    # name: (config changes, P, first-index spread of the J rows, flat, mistakes it must catch)
    "cubic_k3_quadratic_sources": (dict(), quad, 0, False, (1, 2)),
    "linear_k3_linear_sources": (dict(k3_interpolation=L.LINEAR), lin, 0, False, (5,)),
    "J_live_start": (dict(), quad, 300, False, (3,)),
    # negative oscillation counts extend every pair's grid to the whole k range
    "flat_extrapolation": (dict(extrapolation=L.FLAT_EXTRAP, extra_osc_left=-1,
                                extra_osc_right=-1), quad, 0, True, (4,)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_closed_form(name, oracle):
    kw, P, xm, flat, _ = CASES[name]
    c = Case(dict(BASE, **kw), seed=31, P=P, xmin_max=xm)
    if xm:
        # the cut must fall inside the time range for many (k3, row) combinations
        assert sum(1 for r in c.rows.values() if r[2] > 0) > 0.9 * len(c.rows)
    if flat:
        _, lists = c.run(oracle)
        ext = [(f, len(ks) - f - n) for ks, f, n in lists.values()]
        assert sum(a for a, _ in ext) >= 8 and sum(b for _, b in ext) >= 8, ext
        c.close()
        c = Case(dict(BASE, **kw), seed=31, P=P, xmin_max=xm)
    err = c.error(oracle, flat)
    c.close()
    print(f"{name}: max |oracle - closed form| / max = {err:.2e}")
    assert err < 1e-11


@pytest.mark.parametrize("name", list(CASES))
def test_closed_form_catches_planted_mistakes(name, planted):
    kw, P, xm, flat, mistakes = CASES[name]
    for n in mistakes:
        c = Case(dict(BASE, **kw), seed=31, P=P, xmin_max=xm)
        err = c.error(planted(n), flat)
        c.close()
        print(f"{name}: mistake {n}: {err:.2e}")
        assert err > 1e-6, f"mistake {n} not detected"


def test_J_live_start_cut_lands_on_grid_steps(oracle):
    """The J-skip case exercises the boundary itself: for many (k3, row) pairs some
    time step has index_x exactly equal to index_xmin_J (the step a <= test would drop)."""
    c = Case(dict(BASE), seed=31, P=quad, xmin_max=300)
    _, lists = c.run(oracle)
    hits = 0
    for (ks, _, _) in lists.values():
        for k in ks:
            ix = (k * (c.cfg.tau0 - c.tau) / c.cfg.xx_step).astype(np.int64)
            hits += sum(1 for r in c.rows.values() if r[2] > 0 and np.any(ix == r[2]))
    c.close()
    assert hits > 100


# ---------------------------------------------------------------------------
# custom and bessel time sampling
# ---------------------------------------------------------------------------
def _time_grid(oracle, s, k, cap=20000):
    f = oracle.los_ref_time_grid
    f.argtypes = [C.c_void_p] * 4 + [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double), C.c_int]
    f.restype = C.c_int
    tau, dt = np.zeros(cap), np.zeros(cap)
    left = np.zeros(cap, dtype=np.int32)
    st = s.s
    # the oracle's bookkeeping first (it copies tau0 into transfers2)
    oracle.los_ref_indices.argtypes = [C.c_void_p] * 8
    assert oracle.los_ref_indices(st["ppr"], st["ppr2"], st["pba"], st["pth"], st["ppt2"],
                                  st["pbs"], st["pbs2"], st["ptr2"]) == 0
    s._inited = True
    n = f(st["ppr2"], st["ppt2"], st["pbs2"], st["ptr2"], k,
          tau.ctypes.data_as(C.POINTER(C.c_double)), left.ctypes.data_as(C.POINTER(C.c_int)),
          dt.ctypes.data_as(C.POINTER(C.c_double)), cap)
    s.free_transfers(lambda a, b, c: oracle.los_ref_transfer2_free(a, b, c))
    assert n > 0, s.error_message()
    return tau[:n], left[:n], dt[:n]


@pytest.mark.parametrize("sampling", [L.TAU_CUSTOM, L.TAU_BESSEL])
@pytest.mark.parametrize("k", [2e-3, 7.3e-3, 2.1e-2])
def test_time_grid_properties(oracle, sampling, k):
    cfg = L.Config(**dict(BASE, tau_sampling=sampling, tau_linstep=0.4, tau_size=25))
    s = L.Scenario(cfg)
    src = s.grids()[2]
    tau, left, dt = _time_grid(oracle, s, k)
    s.close()
    step = (2 * math.pi / k * 0.4) if sampling == L.TAU_CUSTOM else cfg.xx_step / k
    assert np.all(np.diff(tau) > 0)
    assert tau[0] == src[0] and tau[-1] == src[-1]
    assert np.all(np.isin(src, tau))
    assert np.max(np.diff(tau)) <= step * (1 + 1e-12)
    for i in range(src.size - 1):
        inside = np.count_nonzero((tau > src[i]) & (tau < src[i + 1]))
        assert inside == max(math.ceil((src[i + 1] - src[i]) / step - 1e-9) - 1, 0)
    # index_tau_left: the reference stores, for a point added inside (src[i-1], src[i]),
    # i, the index of the next sampling time (source/transfer2.c:2493-2508), and for
    # src[i] itself i; the last index is moved back by one (:2515-2518). Interpolation
    # in time then extends the piece of [src[i], src[i+1]] leftwards, which is exact for
    # the linear sources below.
    want = np.minimum(np.searchsorted(src, tau, side="left"), src.size - 2)
    assert np.array_equal(left, want)
    assert abs(0.5 * dt.sum() - (src[-1] - src[0])) < 1e-9 * (src[-1] - src[0])
    # brute-force weights of the trapezoid on this grid
    w = np.empty_like(tau)
    w[0], w[-1] = tau[1] - tau[0], tau[-1] - tau[-2]
    w[1:-1] = tau[2:] - tau[:-2]
    assert np.array_equal(w, dt)


def _custom_case(oracle, time_interp, q=2e-4):
    """J constant per row (b = 0), sources c (1 + q tau): the integrand is linear in tau, so
    any time grid integrates it exactly and linear or cubic time interpolation is exact."""
    kw = dict(BASE, tau_sampling=L.TAU_CUSTOM, tau_linstep=0.4, tau_size=25,
              time_interpolation=time_interp, has_E=0, m_list=[0])
    c = Case(kw, seed=17, P=lambda k: np.ones_like(k), Q=lambda t: 1.0 + q * t, b_scale=0.0)
    got, lists = c.run(oracle)
    t0, tN = c.tau[0], c.tau[-1]
    cfg = c.cfg
    out = []
    for il, l in enumerate(c.lvals):
        for i1 in range(cfg.k_size):
            for i2 in range(i1 + 1):
                for k in lists[(i1, i2)][0]:
                    v = sum(c.c[(i1, i2, "T", Lq, 0)] * c.rows[("TT", Lq, il, 0)][0] *
                            ((tN - t0) + q * (tN ** 2 - t0 ** 2) / 2)
                            for Lq in range(cfg.l_max_los_t + 1))
                    out.append(v / (8.0 * (2 * l + 1)))
    exp = np.array(out)
    c.close()
    return float(np.max(np.abs(got - exp)) / np.max(np.abs(exp)))


@pytest.mark.parametrize("interp", [L.LINEAR, L.CUBIC])
def test_custom_sampling_closed_form(oracle, interp, planted):
    assert _custom_case(oracle, interp) < 1e-11
    if interp == L.LINEAR:
        assert _custom_case(planted(6), interp) > 1e-6


# ---------------------------------------------------------------------------
# cubic interpolation of J in x
# ---------------------------------------------------------------------------
def _cubic_j_case(oracle, seed=23, i0=36):
    """J rows a + b x + g x^2 + h x^3 with their exact second derivatives 2 g + 6 h x, sources
    that vanish except at one time step tau_i0 (constant in k3): the cubic form of
    source/transfer2.c:2262-2268 is the cubic through the two nodes with the given second
    derivatives, so it reproduces a cubic exactly, and each transfer is
    dtau_i0 / 2 sum_L c J_L(k (tau0 - tau_i0)) / (8 (2l + 1)) with E = EE + [m != 0] EB and
    B = EE - EB as in source/transfer2.c:1885-2061."""
    cfg = L.Config(**dict(BASE, bessels_interpolation=L.CUBIC))
    s = L.Scenario(cfg)
    rng = np.random.default_rng(seed)
    lvals, kvals, tau = s.grids()
    m = cfg.m_list
    xx_size = int(cfg.xx_max / cfg.xx_step + 1)
    xx = np.arange(xx_size) * (cfg.xx_max / (xx_size - 1))
    coef = {}
    for jt_i, jt in enumerate(["TT", "EE", "EB"]):
        for Lq in range(max(cfg.l_max_los_t, cfg.l_max_los_p) + 1):
            for il, l in enumerate(lvals):
                for im, mm in enumerate(m):
                    if mm > min(Lq, l):
                        continue
                    # near the observer (x of a few units) the cubic term dominates and the
                    # second derivatives differ between neighbouring nodes
                    a, b = rng.normal() * 1e-2, rng.normal() * 1e-2
                    g, h = rng.normal() * 1e-2, rng.normal() * 1e-2
                    coef[(jt, Lq, il, im)] = (a, b, g, h)
                    s.override_J(jt_i, Lq, il, im, 0, a + b * xx + g * xx * xx + h * xx ** 3)
                    s.override_ddJ(jt_i, Lq, il, im, 2 * g + 6 * h * xx)
    keys = [(f, Lq, mm) for f, Lm in (("T", cfg.l_max_los_t), ("E", cfg.l_max_los_p),
                                       ("B", cfg.l_max_los_p))
            for Lq in range(Lm + 1) for mm in m if mm <= Lq]
    c = {}
    for i1 in range(cfg.k_size):
        for i2 in range(i1 + 1):
            for t, key in enumerate(keys):
                c[(i1, i2) + key] = rng.normal()
                src = s.source(t, i1, i2)
                src[:, :] = 0.0
                src[i0, :] = c[(i1, i2) + key]
    rc = oracle.los_ref_transfer2_init(*s.structs())
    s._inited = True
    assert rc == 0, s.error_message()
    got = s.gather()
    sizes = s.k3_sizes()
    lists, q = {}, 0
    for i1 in range(cfg.k_size):
        for i2 in range(i1 + 1):
            n = int(sizes[q, 0])
            buf, last = np.zeros(n + 4), C.c_int()
            st = s.s
            assert oracle.los_ref_get_k3_list(st["ppr"], st["ppr2"], st["ppt2"], st["pbs2"],
                                              st["ptr2"], i1, i2,
                                              buf.ctypes.data_as(C.POINTER(C.c_double)),
                                              C.byref(last)) == 0
            lists[(i1, i2)] = buf[:n]
            q += 1
    s.free_transfers(lambda a, b, c_: oracle.los_ref_transfer2_free(a, b, c_))
    s.close()
    dt = tau[i0 + 1] - tau[i0 - 1]
    d0 = cfg.tau0 - tau[i0]

    def J(jt, Lq, il, im, k):
        a, b, g, h = coef[(jt, Lq, il, im)]
        x = k * d0
        return a + b * x + g * x * x + h * x ** 3

    out = []
    for f in ("T", "E", "B"):
        Lm = cfg.l_max_los_t if f == "T" else cfg.l_max_los_p
        for il, l in enumerate(lvals):
            for im, mm in enumerate(m):
                if mm > l:
                    continue
                for i1 in range(cfg.k_size):
                    for i2 in range(i1 + 1):
                        for k in lists[(i1, i2)]:
                            Ls = range(mm, Lm + 1)
                            cc = lambda F, Lq: c[(i1, i2, F, Lq, mm)]
                            if f == "T":
                                v = sum(cc("T", Lq) * J("TT", Lq, il, im, k) for Lq in Ls)
                            elif f == "E":
                                v = sum(cc("E", Lq) * J("EE", Lq, il, im, k) for Lq in Ls)
                                if mm != 0:
                                    v += sum(cc("B", Lq) * J("EB", Lq, il, im, k) for Lq in Ls)
                            else:
                                v = sum(cc("B", Lq) * J("EE", Lq, il, im, k) for Lq in Ls) - \
                                    sum(cc("E", Lq) * J("EB", Lq, il, im, k) for Lq in Ls)
                            out.append(0.5 * dt * v / (8.0 * (2 * l + 1)))
    exp = np.array(out)
    assert got.shape == exp.shape
    return float(np.max(np.abs(got - exp)) / np.max(np.abs(exp)))


def test_cubic_J_closed_form(oracle, planted):
    good, bad = _cubic_j_case(oracle), _cubic_j_case(planted(7))
    print(f"cubic J: oracle {good:.2e}, swapped weights {bad:.2e}")
    assert good < 1e-11
    assert bad > 1e-9  # 1e8 times the correct case
