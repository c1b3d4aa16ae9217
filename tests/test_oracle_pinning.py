"""Pins of the CPU oracle (and of the library's host bookkeeping) on the CPU.

The reference cannot be built here (its CLASS submodule is empty) and ships no
golden data, so each oracle function is pinned by something it cannot simply
restate:
  * k3 grids: hand-derived dyadic fixtures (tests/golden/k3_grid_kat.json) and
    the reference's DEBUG invariants on seeded grids;
  * (l,m) offsets: hand-enumerated fixtures (tests/golden/lm_offsets.json);
  * k3 spline: exact reproduction of quadratics and agreement with scipy's
    clamped CubicSpline given the same end slopes;
  * LOS integral: closed forms for J linear in x and sources constant in tau,
    where linear interpolation and the trapezoid rule are exact, which fixes
    the J weights, the trapezoid weights, the 1/2, the 1/(8(2l+1)) rescaling,
    the source and transfer (l,m) indexing and the E = EE+EB, B = EE-EB signs.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from song_amd import lib as L
from tests.conftest import run_oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------------------
# (l,m) offsets
# ---------------------------------------------------------------------------
def test_lm_offsets_fixture(oracle):
    lib = L.load()
    lib.los_offset_indexl_indexm.restype = C.c_int
    lib.los_size_indexl_indexm.restype = C.c_int
    gold = json.load(open(os.path.join(GOLD, "lm_offsets.json")))
    for case in gold["cases"]:
        lv = np.array(case["l"], dtype=np.int32)
        mv = np.array(case["m"], dtype=np.int32)
        for impl_size, impl_off in (
                (oracle.los_ref_size_indexl_indexm, oracle.los_ref_offset_indexl_indexm),
                (lib.los_size_indexl_indexm, lib.los_offset_indexl_indexm)):
            assert impl_size(_ip(lv), lv.size, _ip(mv), mv.size) == case["size"]
            for l, m, off in case["offsets"]:
                assert impl_off(l, m, _ip(lv), lv.size, _ip(mv), mv.size) == off
    for lvl, mvl, l, m in gold["absent"]:
        lv = np.array(lvl, dtype=np.int32)
        mv = np.array(mvl, dtype=np.int32)
        assert oracle.los_ref_offset_indexl_indexm(l, m, _ip(lv), lv.size, _ip(mv), mv.size) == -1


def test_lm_offsets_bijection(oracle):
    rng = np.random.default_rng(7)
    for _ in range(20):
        lv = np.unique(rng.integers(0, 60, size=12)).astype(np.int32)
        mv = np.unique(rng.integers(0, 6, size=3)).astype(np.int32)
        n = oracle.los_ref_size_indexl_indexm(_ip(lv), lv.size, _ip(mv), mv.size)
        offs = [oracle.los_ref_offset_indexl_indexm(int(l), int(m), _ip(lv), lv.size, _ip(mv), mv.size)
                for l in lv for m in mv if m <= l]
        assert sorted(offs) == list(range(n))
        assert n == sum(1 for l in lv for m in mv if m <= l)


# ---------------------------------------------------------------------------
# k3 grids
# ---------------------------------------------------------------------------
def _kat_scenario(case, gold):
    cfg = L.Config(has_T=1, has_E=0, m_list=[0], l_max_los_t=2, k_size=3, k_min=0.25, k_max=2.0,
                   k3_size=5, tau_size=16, tau_min=0.5, tau_max=3.5, tau0=gold["tau0"], tau_rec=1.0,
                   l_size=3, l_max=12, xx_max=gold["xx_max"], xx_step=gold["xx_step"],
                   k3_sampling=L.BESSEL_K3, extrapolation=case["extrapolation"],
                   extra_osc_left=case["extra_osc"][0], extra_osc_right=case["extra_osc"][1])
    s = L.Scenario(cfg)
    s.override_k(gold["k"])
    for key, k3 in gold["other_pairs_k3"].items():
        i1, i2 = map(int, key.split(","))
        s.override_k3(i1, i2, k3)
    s.override_k3(*gold["pair"], case["k3_coarse"])
    return s


def _oracle_k3_list(oracle, s, i1, i2, n):
    out = np.zeros(n + 4)
    last = C.c_int()
    st = s.s
    rc = oracle.los_ref_get_k3_list(st["ppr"], st["ppr2"], st["ppt2"], st["pbs2"], st["ptr2"], i1, i2,
                                    _dp(out), C.byref(last))
    assert rc == 0, s.error_message()
    return out[:n]


def _product_k3_list(s, i1, i2, n):
    lib = L.load()
    out = np.zeros(n + 4)
    last = C.c_int()
    st = s.s
    rc = lib.transfer2_get_k3_list(st["ppr"], st["ppr2"], st["ppt2"], st["pbs"], st["pbs2"],
                                   st["ptr2"], i1, i2, _dp(out), C.byref(last))
    assert rc == 0, s.error_message()
    return out[:n]


def _oracle_indices(oracle, s):
    oracle.los_ref_indices.argtypes = [C.c_void_p] * 8
    oracle.los_ref_indices.restype = C.c_int
    st = s.s
    rc = oracle.los_ref_indices(st["ppr"], st["ppr2"], st["pba"], st["pth"], st["ppt2"], st["pbs"],
                                st["pbs2"], st["ptr2"])
    s._inited = True
    assert rc == 0, s.error_message()


def _free_oracle(oracle, s):
    s.free_transfers(lambda a, b, c: oracle.los_ref_transfer2_free(a, b, c))


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_k3_grid_fixtures(which, oracle):
    gold = json.load(open(os.path.join(GOLD, "k3_grid_kat.json")))
    for case in gold["cases"]:
        s = _kat_scenario(case, gold)
        if which == "oracle":
            _oracle_indices(oracle, s)
        else:
            lib = L.load()
            assert lib.los_transfer2_prepare(*s.structs()) == 0, s.error_message()
            s._inited = True
        sizes = s.k3_sizes()
        i1, i2 = gold["pair"]
        row = sizes[i1 * (i1 + 1) // 2 + i2]
        assert tuple(row) == (case["k_size"], case["physical_start"], case["physical_size"]), case["name"]
        getter = _oracle_k3_list if which == "oracle" else (lambda o, *a: _product_k3_list(*a))
        got = getter(oracle, s, i1, i2, case["k_size"])
        assert got.tolist() == case["k3"], case["name"]
        if which == "oracle":
            _free_oracle(oracle, s)
        s.close()


BOOK_CASES = {
    "class_sampling": dict(k_size=6, k3_size=20, dk_fine=7e-4, k_max=0.05),
    "bessel_sampling": dict(k_size=5, k3_size=9, k3_sampling=L.BESSEL_K3, xx_max=900.0, k_max=0.05),
    "extrapolation": dict(k_size=5, k3_size=12, dk_fine=9e-4, k_max=0.05,
                          extrapolation=L.FLAT_EXTRAP, extra_osc_left=2.0, extra_osc_right=3.0),
    "extrapolation_to_ends": dict(k_size=5, k3_size=12, dk_fine=9e-4, k_max=0.05,
                                  extrapolation=L.FLAT_EXTRAP, extra_osc_left=-1, extra_osc_right=-1),
    "wide_log_grid": dict(k_size=12, k_min=1e-5, k_max=0.1, k3_size=30, dk_fine=1.5e-4,
                          xx_max=1600.0),
}


@pytest.mark.parametrize("name", list(BOOK_CASES))
def test_k3_bookkeeping_product_equals_oracle(name, oracle):
    """Seeded (non-dyadic) grids: the library's host bookkeeping and the oracle agree
    bit for bit, and the reference's DEBUG invariants hold (both check them)."""
    cfg = L.Config(**BOOK_CASES[name])
    so, sp = L.Scenario(cfg), L.Scenario(cfg)
    _oracle_indices(oracle, so)
    assert L.load().los_transfer2_prepare(*sp.structs()) == 0, sp.error_message()
    sp._inited = True
    a, b = so.k3_sizes(), sp.k3_sizes()
    assert np.array_equal(a, b)
    q = 0
    for i1 in range(cfg.k_size):
        for i2 in range(i1 + 1):
            n = int(a[q, 0])
            ko = _oracle_k3_list(oracle, so, i1, i2, n)
            kp = _product_k3_list(sp, i1, i2, n)
            assert np.array_equal(ko, kp), (i1, i2)
            assert np.all(np.diff(ko) > 0)
            q += 1
    _free_oracle(oracle, so)
    so.close()
    sp.close()


# ---------------------------------------------------------------------------
# k3 spline
# ---------------------------------------------------------------------------
def _spline(oracle, x, y):
    n = x.size
    y2 = np.ascontiguousarray(y.reshape(-1, n))
    dd = np.zeros_like(y2)
    assert oracle.los_ref_spline_columns(_dp(x), n, _dp(y2), y2.shape[0], _dp(dd)) == 0
    return dd


def _eval(x, y, dd, xs):
    i = np.clip(np.searchsorted(x, xs) - 1, 0, x.size - 2)
    h = x[i + 1] - x[i]
    b = (xs - x[i]) / h
    a = 1 - b
    return a * y[i] + b * y[i + 1] + ((a ** 3 - a) * dd[i] + (b ** 3 - b) * dd[i + 1]) * h * h / 6.0


def test_spline_exact_on_quadratics(oracle):
    rng = np.random.default_rng(3)
    for n in (4, 7, 60):
        x = np.sort(rng.uniform(0.01, 0.2, n))
        c0, c1, c2 = rng.normal(size=3)
        y = c0 + c1 * x + c2 * x ** 2
        dd = _spline(oracle, x, y)[0]
        scale = abs(c2) * 2
        assert np.max(np.abs(dd - 2 * c2)) < 1e-8 * max(scale, 1)
        xs = rng.uniform(x[0], x[-1], 50)
        assert np.max(np.abs(_eval(x, y, dd, xs) - (c0 + c1 * xs + c2 * xs ** 2))) < 1e-12


def test_spline_matches_scipy_clamped(oracle):
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(11)
    x = np.sort(rng.uniform(0.0, 1.0, 23))
    y = rng.normal(size=(5, x.size))
    dd = _spline(oracle, x, y)
    for r in range(y.shape[0]):
        yy = y[r]
        d0 = ((x[2] - x[0]) ** 2 * (yy[1] - yy[0]) - (x[1] - x[0]) ** 2 * (yy[2] - yy[0])) / \
             ((x[2] - x[0]) * (x[1] - x[0]) * (x[2] - x[1]))
        e, b, a = -1, -2, -3
        dn = ((x[e] - x[a]) ** 2 * (yy[e] - yy[b]) - (x[e] - x[b]) ** 2 * (yy[e] - yy[a])) / \
             ((x[e] - x[a]) * (x[e] - x[b]) * (x[b] - x[a]))
        cs = CubicSpline(x, yy, bc_type=((1, d0), (1, dn)))
        ref = cs(x, 2)
        assert np.max(np.abs(dd[r] - ref)) < 1e-9 * np.max(np.abs(ref))


# ---------------------------------------------------------------------------
# LOS integral closed forms
# ---------------------------------------------------------------------------
CF_CFG = dict(has_T=1, has_E=1, m_list=[0, 2], l_max_los_t=2, l_max_los_p=2, k_size=3,
              k_min=2e-3, k_max=1e-2, k3_size=6, tau_size=40, tau_min=200.0, tau_max=14000.0,
              tau0=14200.0, l_size=4, l_max=40, dk_fine=1e-3, xx_max=200.0, xx_step=0.25)


def _lm_pairs(lvals, mvals):
    return [(l, m) for l in lvals for m in mvals if m <= l]


def closed_form_case(cfg_kw, seed=5):
    """Build a scenario with J = alpha + beta*x (every row, stored from x=0) and sources
    constant in (k3, tau); return the scenario and the exact transfers."""
    cfg = L.Config(**cfg_kw)
    s = L.Scenario(cfg)
    rng = np.random.default_rng(seed)
    lvals, kvals, tau = s.grids()
    mvals = cfg.m_list
    Lmax_t, Lmax_p = cfg.l_max_los_t, cfg.l_max_los_p
    # J types present: TT (T), EE and EB (E with some m > 0)
    jtypes = []
    if cfg.has_T:
        jtypes.append("TT")
    if cfg.has_E:
        jtypes.append("EE")
        if max(mvals) > 0:
            jtypes.append("EB")
    L_size = max(Lmax_t if cfg.has_T else 0, Lmax_p if cfg.has_E else 0, max(mvals)) + 1
    xx_size = int(cfg.xx_max / cfg.xx_step + 1)
    xx = np.arange(xx_size) * (cfg.xx_max / (xx_size - 1))
    coef = {}
    for jt_i, jt in enumerate(jtypes):
        for Lq in range(L_size):
            a, b = rng.normal(), rng.normal() * 0.01
            coef[(jt, Lq)] = (a, b)
            for il, l in enumerate(lvals):
                for im, m in enumerate(mvals):
                    if m > min(Lq, l):
                        continue
                    s.override_J(jt_i, Lq, il, im, 0, a + b * xx)
    # source types: field blocks T, E, B; inside each, (L,m) l-major, m ascending
    fields = (["T"] if cfg.has_T else []) + (["E"] if cfg.has_E else []) + \
             (["B"] if (cfg.has_E and max(mvals) > 0) else [])
    tp_of = {}
    tp = 0
    for f in fields:
        Lm = Lmax_t if f == "T" else Lmax_p
        for (Lq, m) in _lm_pairs(range(Lm + 1), mvals):
            tp_of[(f, Lq, m)] = tp
            tp += 1
    nk = cfg.k_size
    pair_c = {}
    for i1 in range(nk):
        for i2 in range(i1 + 1):
            c = {key: rng.normal() for key in tp_of}
            pair_c[(i1, i2)] = c
            for key, t in tp_of.items():
                src = s.source(t, i1, i2)
                src[:, :] = c[key]
    t0, tN = tau[0], tau[-1]
    D = tN - t0
    T0 = cfg.tau0

    def I(jt, Lq, k):
        a, b = coef[(jt, Lq)]
        return a * D + b * k * (T0 * D - (tN ** 2 - t0 ** 2) / 2)

    return s, cfg, fields, tp_of, pair_c, I, lvals


def expected_closed_form(s, cfg, fields, pair_c, I, lvals, k3_lists):
    mvals = cfg.m_list
    lm = _lm_pairs(lvals, mvals)
    out = []
    nk = cfg.k_size
    for f in fields:
        Lm = cfg.l_max_los_t if f == "T" else cfg.l_max_los_p
        for (l, m) in lm:
            for i1 in range(nk):
                for i2 in range(i1 + 1):
                    c = pair_c[(i1, i2)]
                    for k in k3_lists[(i1, i2)]:
                        Ls = [Lq for Lq in range(Lm + 1) if Lq >= m]
                        if f == "T":
                            v = sum(c[("T", Lq, m)] * I("TT", Lq, k) for Lq in Ls)
                        elif f == "E":
                            v = sum(c[("E", Lq, m)] * I("EE", Lq, k) for Lq in Ls)
                            if m != 0:
                                v += sum(c[("B", Lq, m)] * I("EB", Lq, k) for Lq in Ls)
                        else:
                            v = sum(c[("B", Lq, m)] * I("EE", Lq, k) for Lq in Ls) - \
                                sum(c[("E", Lq, m)] * I("EB", Lq, k) for Lq in Ls)
                        out.append(v / (8.0 * (2 * l + 1)))
    return np.array(out)


def k3_lists_of(run_and_list, cfg):
    return run_and_list


def test_closed_form_oracle(oracle):
    s, cfg, fields, tp_of, pair_c, I, lvals = closed_form_case(CF_CFG)
    st = s.s
    assert oracle.los_ref_transfer2_init(*s.structs()) == 0, s.error_message()
    s._inited = True
    got = s.gather()
    sizes = s.k3_sizes()
    lists = {}
    q = 0
    for i1 in range(cfg.k_size):
        for i2 in range(i1 + 1):
            lists[(i1, i2)] = _oracle_k3_list(oracle, s, i1, i2, int(sizes[q, 0]))
            q += 1
    _free_oracle(oracle, s)
    exp = expected_closed_form(s, cfg, fields, pair_c, I, lvals, lists)
    assert got.shape == exp.shape
    scale = np.max(np.abs(exp))
    assert np.max(np.abs(got - exp)) < 1e-11 * scale
    s.close()


def test_closed_form_detects_planted_sign_error(oracle):
    """The closed form separates B = EE - EB from EE + EB: flipping the sign of the
    mixing term in the expectation must break the agreement."""
    s, cfg, fields, tp_of, pair_c, I, lvals = closed_form_case(CF_CFG, seed=9)
    assert oracle.los_ref_transfer2_init(*s.structs()) == 0
    s._inited = True
    got = s.gather()
    sizes = s.k3_sizes()
    lists, q = {}, 0
    for i1 in range(cfg.k_size):
        for i2 in range(i1 + 1):
            lists[(i1, i2)] = _oracle_k3_list(oracle, s, i1, i2, int(sizes[q, 0]))
            q += 1
    _free_oracle(oracle, s)

    def I_flip(jt, Lq, k):
        return -I(jt, Lq, k) if jt == "EB" else I(jt, Lq, k)

    wrong = expected_closed_form(s, cfg, fields, pair_c, I_flip, lvals, lists)
    assert np.max(np.abs(got - wrong)) > 1e-6 * np.max(np.abs(wrong))
    s.close()
