"""The projection tables J_Llm(x) (input of the LOS stage) and the special functions behind them.

SONG builds them in its bessel2 module (source/bessel2.c) from CLASS's bessel_j and drc3jj,
which are absent from the reference tree; the library's host builder (los_bessel2.c) is pinned
here against independent implementations and closed forms:
* j_l(x) against scipy.special.spherical_jn, on both sides of the turning point and at x = 0;
* Wigner 3j symbols against sympy's exact wigner_3j;
* J_{0,l,0}(x) = (2l+1) j_l(x) and J_{1,l,0}(x) = l j_{l-1}(x) - (l+1) j_{l+1}(x), the
  closed forms of the TT rows with L = 0, 1 (source/bessel2.c:1543-1555 with the 3j sums),
  up to the j_l1 cut of the reference (terms with |j_l1| < j_cut are zero);
* index_xmin_J brackets the first grid point where |J| reaches J_cut.
"""
import ctypes as C

import numpy as np
import pytest

from song_amd import lib as L

SMALL = dict(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=3, l_max_los_p=3, k_size=3,
             k3_size=6, tau_size=16, l_size=8, l_max=300, xx_max=600.0)


def _jl_array(lib, lmax, x):
    out = (C.c_double * (lmax + 1))()
    lib.los_sph_bessel_array(lmax, x, out)
    return np.array(out[:])


@pytest.mark.parametrize("x", [0.0, 0.3, 2.5, 37.1, 150.0, 600.0, 2999.7])
def test_spherical_bessel_matches_scipy(x):
    from scipy.special import spherical_jn
    lib = L.load()
    lmax = 400
    got = _jl_array(lib, lmax, x)
    ls = np.arange(lmax + 1)
    ref = spherical_jn(ls, x)
    # absolute error against the function's scale; values below 1e-250 underflow in scipy
    scale = max(np.max(np.abs(ref)), 1e-300)
    live = np.abs(ref) > 1e-200
    assert np.all(np.abs(got[live] - ref[live]) <= 1e-12 * scale + 1e-13 * np.abs(ref[live])), x
    assert np.all(np.abs(got[~live]) < 1e-190)


@pytest.mark.parametrize("j2,j3,m2,m3", [(3, 5, 0, 0), (6, 4, -2, 2), (10, 7, 1, -1),
                                         (2, 2, 0, 0), (6, 30, 2, -2), (40, 50, 3, -1),
                                         (70, 45, 0, 0)])
def test_wigner3j_matches_sympy(j2, j3, m2, m3):
    from sympy.physics.wigner import wigner_3j
    lib = L.load()
    out = (C.c_double * 160)()
    a, b = C.c_int(), C.c_int()
    n = lib.los_wigner3j_j1(j2, j3, m2, m3, C.byref(a), C.byref(b), out, 160)
    assert n == b.value - a.value + 1
    m1 = -m2 - m3
    for q in range(n):
        j1 = a.value + q
        ref = float(wigner_3j(j1, j2, j3, m1, m2, m3))
        assert abs(out[q] - ref) <= 1e-13 * max(1.0, abs(ref)), (j1, out[q], ref)


def _grid(cfg):
    n = int(cfg.xx_max / cfg.xx_step + 1)
    return np.arange(n) * (cfg.xx_max / (n - 1))


def test_tt_rows_closed_forms():
    from scipy.special import spherical_jn
    cfg = L.Config(**SMALL)
    s = L.Scenario(cfg)
    lvals = s.grids()[0]
    xx = _grid(cfg)
    for il, l in enumerate(lvals):
        for Lq in (0, 1):
            xmin, row = s.J_row(0, Lq, il, 0)  # J type 0 = TT, m = 0
            x = xx[xmin:]
            if Lq == 0:
                ref = (2 * l + 1) * spherical_jn(l, x)
            else:
                ref = l * spherical_jn(l - 1, x) - (l + 1) * spherical_jn(l + 1, x)
            scale = np.max(np.abs(ref))
            # terms with |j_l1| below j_cut are zero by construction (source/bessel2.c:1669-1815):
            # at most (sum of |coefficients|) j_cut = (2l+1) j_cut near the start of the row
            tol = 1e-11 * scale + (2 * l + 1) * cfg.j_cut
            assert np.max(np.abs(row - ref)) <= tol, (l, Lq)
    s.close()


def test_index_xmin_brackets_the_cut():
    cfg = L.Config(**SMALL)
    s = L.Scenario(cfg)
    lvals = s.grids()[0]
    n = int(cfg.xx_max / cfg.xx_step + 1)
    checked = 0
    for jt in range(3):
        for Lq in range(4):
            for il, l in enumerate(lvals):
                for im, m in enumerate(cfg.m_list):
                    if m > min(Lq, l):
                        continue
                    xmin, row = s.J_row(jt, Lq, il, im)
                    if xmin == n - 1 or xmin == 0:
                        continue
                    # the row is stored from the point before |J| first reaches the cut
                    assert abs(row[0]) < cfg.J_cut and abs(row[1]) >= cfg.J_cut, (jt, Lq, l, m)
                    checked += 1
    assert checked > 20
    s.close()


def test_wigner3j_rejects_short_buffer():
    """A j1 range longer than the caller's buffer is refused, not written past its end."""
    lib = L.load()
    out = (C.c_double * 8)()
    a, b = C.c_int(), C.c_int()
    assert lib.los_wigner3j_j1(10, 12, 0, 0, C.byref(a), C.byref(b), out, 8) == -1
