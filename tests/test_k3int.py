"""k3 integral of the transfer functions (the bispectrum module's consumer of this stage).

Reference: source/bispectra2.c:1182-1410 (loop over l3, k1, k2, r; trapezoid measure
:1283-1287; M3 coefficient and factor 2 :1355-1361) with bessel2_convolution
(source/bessel2.c:499-605, linear interpolation, g = NULL). The reference ships no golden
vectors for it, so the oracle's restatement (oracle/los_ref.c, los_ref_k3_integral) is
pinned on the CPU by
* a closed form: a j table that is linear in x (j(x) = x, which linear interpolation
  reproduces) and a transfer row that is non-zero at one k3 give
  I = 2 coef 0.5 k^2 (k r) T dk exactly, which fixes k^2, the trapezoid measure at the
  ends and inside, the 0.5 and 2, the M3 coefficient, a against 1 - a (the point is
  not at a cell midpoint) and the x < x_min skip;
* quadrature: j_2 from scipy on a fine x grid and a smooth T(k3) against twice
  scipy.integrate.quad of k^2 j_2(k r) T(k) dk (the trapezoid sum with the measure
  k[i+1] - k[i-1] and the 0.5 is the integral; the reference then doubles it);
and the GPU kernel (los_k3int_*) is compared with it bit for bit on ragged grids.
"""
import ctypes as C

import numpy as np
import pytest

from song_amd import lib as L
from tests.conftest import load_oracle


def _oracle_fn():
    o = load_oracle()
    f = o.los_ref_k3_integral
    dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
    f.argtypes = [C.c_int, lp, dp, C.c_int, dp, ip, lp, dp, dp, C.c_double, C.c_int, dp,
                  C.c_int, C.c_double, dp]
    f.restype = C.c_int
    return f


def oracle_k3int(pair_off, k3, transfer, index_L, j_off, j, x_min, step, r, abs_M3, coef):
    f = _oracle_fn()
    arrs = [np.ascontiguousarray(a, dtype=t) for a, t in
            ((pair_off, np.int64), (k3, np.float64), (transfer, np.float64),
             (index_L, np.int32), (j_off, np.int64), (j, np.float64), (x_min, np.float64),
             (r, np.float64))]
    po, k, t, il, jo, jj, xm, rr = arrs
    n_pairs, n_q, n_r = po.size - 1, il.size, rr.size
    out = np.zeros((n_q, n_r, n_pairs))
    dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
    rc = f(n_pairs, po.ctypes.data_as(lp), k.ctypes.data_as(dp), n_q, t.ctypes.data_as(dp),
           il.ctypes.data_as(ip), jo.ctypes.data_as(lp), jj.ctypes.data_as(dp),
           xm.ctypes.data_as(dp), step, n_r, rr.ctypes.data_as(dp), abs_M3, coef,
           out.ctypes.data_as(dp))
    assert rc == 0
    return out


def j_tables(ls, x_max, step, cut=1e-12):
    """j_l on the grid x = i step, each stored from the point before the first |j_l| >= cut
    (the index_xmin_l1 rule of source/bessel2.c:1669-1815); returns j_off, j, x_min."""
    from scipy.special import spherical_jn
    n = int(x_max / step) + 1
    xx = np.arange(n) * step
    off, vals, xmin = [0], [], []
    for l in ls:
        jl = spherical_jn(l, xx)
        i0 = 0 if l == 0 else max(int(np.argmax(np.abs(jl) >= cut)) - 1, 0)
        vals.append(jl[i0:])
        xmin.append(xx[i0])
        off.append(off[-1] + n - i0)
    return np.array(off), np.concatenate(vals), np.array(xmin)


def test_oracle_closed_form_linear_table():
    step = 0.25
    x_min = 2.0
    j = x_min + step * np.arange(4000)          # j(x) = x from x_min on
    k3 = np.array([0.010, 0.013, 0.0175, 0.021, 0.030])
    dk = np.array([0.003, 0.0075, 0.008, 0.0125, 0.009])
    r = np.array([50.0, 150.0, 333.3])
    for i in range(k3.size):
        T = np.zeros(k3.size)
        T[i] = 1.7
        for abs_M3, coef in ((0, 1.0), (2, 0.37)):
            got = oracle_k3int([0, k3.size], k3, T[None, :], [0], [0, j.size], j, [x_min],
                               step, r, abs_M3, coef)
            for ir, rv in enumerate(r):
                x = k3[i] * rv
                want = 0.0 if x < x_min else 2 * coef * 0.5 * k3[i] ** 2 * x * T[i] * dk[i]
                assert got[0, ir, 0] == pytest.approx(want, rel=1e-12, abs=1e-300), (i, ir)
    # k3 * 50 < x_min = 2 for the first three points: those terms are skipped
    T = np.ones(k3.size)
    got = oracle_k3int([0, k3.size], k3, T[None, :], [0], [0, j.size], j, [x_min], step,
                       r[:1], 0, 1.0)
    want = sum(k3[i] ** 2 * (k3[i] * 50.0) * dk[i] for i in range(k3.size) if k3[i] * 50 >= 2)
    assert got[0, 0, 0] == pytest.approx(want, rel=1e-12)


def test_oracle_against_quadrature():
    from scipy.integrate import quad
    from scipy.special import spherical_jn
    step = 0.002
    kmin, kmax, n = 0.002, 0.05, 3001
    k3 = np.linspace(kmin, kmax, n)
    k0, w = 0.02, 0.008
    T = np.exp(-((k3 - k0) / w) ** 2)
    r = np.array([40.0, 120.0, 300.0])
    j_off, j, x_min = j_tables([2], kmax * r.max() + 1.0, step)
    got = oracle_k3int([0, n], k3, T[None, :], [0], j_off, j, x_min, step, r, 0, 1.0)
    for ir, rv in enumerate(r):
        want, _ = quad(lambda k: k * k * spherical_jn(2, k * rv) * np.exp(-((k - k0) / w) ** 2),
                       kmin, kmax, limit=400, epsabs=0, epsrel=1e-12)
        # the trapezoid sum is the integral; :1361 doubles it (T(k1,k2,k3) + T(k2,k1,k3))
        assert got[0, ir, 0] == pytest.approx(2 * want, rel=2e-5), (rv, got[0, ir, 0], want)


def test_create_rejects_bad_inputs_before_gpu_work():
    """Checks that fire on the host, before any device call (no GPU needed)."""
    j_off, j, x_min = [0, 100], np.zeros(100), [0.0]
    with pytest.raises(L.LosError, match="less than two"):
        L.K3Integral([0, 1], [0.01], np.ones((1, 1)), [0], j_off, j, x_min, 0.3, [10.0])
    with pytest.raises(L.LosError, match="beyond the j table"):
        L.K3Integral([0, 2], [0.01, 0.1], np.ones((1, 2)), [0], j_off, j, x_min, 0.3, [400.0])
    with pytest.raises(L.LosError, match="negative"):
        L.K3Integral([0, 2], [0.01, 0.1], np.ones((1, 2)), [0], j_off, j, x_min, 0.3, [-1.0])


def ragged_case(seed, n_pairs=29, n_r=150, abs_M3=1, coef=1.3):
    rng = np.random.default_rng(seed)
    counts = rng.integers(2, 400, n_pairs)
    counts[0], counts[1] = 2, 399
    pair_off = np.concatenate([[0], np.cumsum(counts)])
    k3 = np.concatenate([np.sort(rng.uniform(1e-4, 0.2, c)) for c in counts])
    ls = [0, 2, 5, 40, 700]
    rows = [0, 1, -1, 2, 3, 4, 1, 5]  # table 5: j_l1 negligible over the whole x grid
    T = rng.standard_normal((len(rows), k3.size))
    T[rng.random(T.shape) < 0.1] = 0.0
    r = np.sort(rng.uniform(0.0, 14000.0, n_r))
    step = 0.3
    j_off, j, x_min = j_tables(ls, 0.2 * 14000.0 + 2.0, step)
    # source/bessel2.c:1756-1759, :1789-1792: a table whose values are all below the cut holds one zero
    # at x_min_l1 = x_max; every term of its rows is skipped
    j_off = np.append(j_off, j_off[-1] + 1)
    j = np.append(j, 0.0)
    x_min = np.append(x_min, 15000.0)
    return dict(pair_off=pair_off, k3=k3, transfer=T, index_L=np.array(rows), j_off=j_off,
                j=j, x_min=x_min, step=step, r=r, abs_M3=abs_M3, coef=coef)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,abs_M3,coef,n_r", [(1, 0, 1.0, 150), (2, 1, 1.3, 64),
                                                  (3, 2, 0.41, 1)])
def test_k3int_bit_exact(seed, abs_M3, coef, n_r):
    print(f"case k3int seed={seed} M3={abs_M3} n_r={n_r}", flush=True)
    c = ragged_case(seed, n_r=n_r, abs_M3=abs_M3, coef=coef)
    ref = oracle_k3int(c["pair_off"], c["k3"], c["transfer"], c["index_L"], c["j_off"], c["j"],
                       c["x_min"], c["step"], c["r"], abs_M3, coef)
    # the kernel always applies coef; the reference only for M3 != 0 (coef 1 for M3 = 0)
    h = L.K3Integral(c["pair_off"], c["k3"], c["transfer"], c["index_L"], c["j_off"], c["j"],
                     c["x_min"], c["step"], c["r"], coef if abs_M3 else 1.0)
    h.run(2)
    got = h.result()
    h.close()
    assert np.count_nonzero(ref) > 0.5 * ref.size
    assert not np.any(got[2]) and not np.any(ref[2])  # skipped row
    assert not np.any(got[7]) and not np.any(ref[7])  # negligible table
    diff = np.flatnonzero(got != ref)
    assert diff.size == 0, f"{diff.size} of {ref.size} differ; first {diff[:5]}"


@pytest.mark.gpu
def test_k3int_on_gpu_transfers(oracle):
    """The chain: transfers from the LOS stage on the GPU, their k3 grids from
    transfer2_get_k3_list, integrated against j_l3(k3 r) for every l3 of the T field (M3 = 0)."""
    cfg = L.Config(has_T=1, has_E=0, m_list=[0], l_max_los_t=4, k_size=5, k3_size=12,
                   tau_size=80, l_size=10, l_max=300, dk_fine=1e-3, xx_max=1500.0)
    s = L.Scenario(cfg)
    s.transfer2_init()
    flat = s.gather()
    sizes = s.k3_sizes()[:, 0]
    pair_off = np.concatenate([[0], np.cumsum(sizes)])
    from tests.test_oracle_pinning import _product_k3_list
    k3 = np.empty(pair_off[-1])
    q = 0
    for i1 in range(cfg.k_size):
        for i2 in range(i1 + 1):
            k3[pair_off[q]:pair_off[q + 1]] = _product_k3_list(s, i1, i2, int(sizes[q]))
            q += 1
    assert np.all(np.diff(k3[pair_off[0]:pair_off[1]]) > 0)
    n_l = cfg.l_size
    T = flat[: n_l * pair_off[-1]].reshape(n_l, pair_off[-1])  # T rows: l-major, pairs, k3
    ls = [int(v) for v in s.grids()[0][:n_l]]
    r = np.linspace(13000.0, 14500.0, 90)
    j_off, j, x_min = j_tables(ls, float(k3.max() * r.max()) + 2.0, 0.3)
    ref = oracle_k3int(pair_off, k3, T, np.arange(n_l), j_off, j, x_min, 0.3, r, 0, 1.0)
    h = L.K3Integral(pair_off, k3, T, np.arange(n_l), j_off, j, x_min, 0.3, r, 1.0)
    h.run(1)
    got = h.result()
    h.close()
    s.close()
    assert np.count_nonzero(ref) > 0.5 * ref.size
    assert np.array_equal(got, ref)
