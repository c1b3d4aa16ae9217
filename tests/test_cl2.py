"""Second-order C_l: the (k1, k2, k3) sum of spectra2_cls (source/spectra2.c:241-590).

The reference ships no golden vectors for it. The oracle's restatement (oracle/los_ref.c,
los_ref_cl2) is pinned on the CPU by
* a closed form: with P(k) = 1/k (pk = k^2 / (2 pi^2)) and T1 = T2 = 1/sqrt(2 k2 k3) every
  pair's k3 sum is sum dk3 / (2 k2) = 1 exactly when the triangle-edge corrections
  (:486-500) are right (the corrected half steps add up to the triangle's width 2 k2), the
  k2 half steps add up to k1 - k_min, the k1 ones integrate k1 - k_min exactly, so
  C = (k_max - k_min)^2 / (4 pi^4): this fixes every step rule, the 2 and the 2/pi/2/pi/2/pi/2/pi
  constant, and P = 2 pi^2 / k^3 pk;
* the rescaling (1 - cos^2)^m on a fine k3 grid against scipy quadrature of the same
  integrand over the triangle, and m = 0 rescaled equal to not rescaled;
and the GPU kernel (los_cl2_*) is compared with it to 1e-12 of the sum of |terms| (its sum
order differs from the reference's sequential one).
"""
import ctypes as C

import numpy as np
import pytest

from song_amd import lib as L
from tests.conftest import load_oracle


def oracle_cl2(k, pk, pair_off, k3, m, t1, t2, used, rescale):
    o = load_oracle()
    f = o.los_ref_cl2
    dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
    f.argtypes = [C.c_int, dp, dp, lp, dp, C.c_int, C.c_int, ip, dp, dp, ip, C.c_int, dp, dp]
    f.restype = C.c_int
    a = [np.ascontiguousarray(x, dtype=t) for x, t in
         ((k, np.float64), (pk, np.float64), (pair_off, np.int64), (k3, np.float64),
          (m, np.int32), (t1, np.float64), (t2, np.float64), (used, np.int32))]
    kk, pp, po, k3a, mm, a1, a2, uu = a
    n_l, m_size = uu.shape
    cl = np.zeros(n_l)
    cl_m = np.zeros((n_l, m_size))
    assert f(kk.size, kk.ctypes.data_as(dp), pp.ctypes.data_as(dp), po.ctypes.data_as(lp),
             k3a.ctypes.data_as(dp), n_l, m_size, mm.ctypes.data_as(ip), a1.ctypes.data_as(dp),
             a2.ctypes.data_as(dp), uu.ctypes.data_as(ip), int(rescale), cl.ctypes.data_as(dp),
             cl_m.ctypes.data_as(dp)) == 0
    return cl, cl_m


def triangle_grids(k, rng, n_lo=2, n_hi=30, outside=False, eps_frac=0.013):
    """Per pair (k1-major, k2 <= k1) a k3 grid strictly inside [k1-k2, k1+k2]; with
    `outside`, some pairs get extra points beyond the triangle (the extrapolated range)."""
    grids = []
    for i1 in range(k.size):
        for i2 in range(i1 + 1):
            k1, k2 = k[i1], k[i2]
            n = int(rng.integers(n_lo, n_hi + 1))
            lo, hi = k1 - k2, k1 + k2
            eps = eps_frac * (hi - lo)
            g = np.linspace(lo + eps, hi - 0.7 * eps, n)
            if outside and rng.random() < 0.5:
                g = np.concatenate([g, [hi + 0.1 * k2, hi + 0.3 * k2]])
                if lo > 0.2 * k2:
                    g = np.concatenate([[lo - 0.1 * k2], g])
            grids.append(g)
    off = np.concatenate([[0], np.cumsum([g.size for g in grids])])
    return off, np.concatenate(grids), grids


def test_oracle_closed_form():
    rng = np.random.default_rng(5)
    k = np.sort(rng.uniform(0.002, 0.1, 9))
    pair_off, k3, grids = triangle_grids(k, rng)
    T = np.concatenate([1 / np.sqrt(2 * k[i2] * g) for (i1, i2), g in
                        zip([(a, b) for a in range(k.size) for b in range(a + 1)], grids)])
    pk = k ** 2 / (2 * np.pi ** 2)
    cl, cl_m = oracle_cl2(k, pk, pair_off, k3, [0], T[None, None, :], T[None, None, :],
                          [[1]], 0)
    want = (k[-1] - k[0]) ** 2 / (4 * np.pi ** 4)
    assert cl[0] == pytest.approx(want, rel=1e-12)
    assert cl_m[0, 0] == cl[0]
    # m = 0 is unchanged by the rescaling; a skipped row is zero
    cl_r, _ = oracle_cl2(k, pk, pair_off, k3, [0], T[None, None, :], T[None, None, :], [[1]], 1)
    assert cl_r[0] == pytest.approx(want, rel=1e-12)
    cl0, cl0_m = oracle_cl2(k, pk, pair_off, k3, [0], T[None, None, :], T[None, None, :],
                            [[0]], 0)
    assert cl0[0] == 0 and cl0_m[0, 0] == 0


def test_oracle_rescale_against_quadrature():
    from scipy.integrate import quad
    k = np.array([0.01, 0.013, 0.02])
    rng = np.random.default_rng(1)
    # points close to the edges: the end corrections are rectangles, whose error is
    # second order in the distance to the edge
    pair_off, k3, grids = triangle_grids(k, rng, 3000, 3000, eps_frac=1e-5)
    pairs = [(a, b) for a in range(k.size) for b in range(a + 1)]
    T = np.concatenate([1 / np.sqrt(2 * k[i2] * g) for (i1, i2), g in zip(pairs, grids)])
    pk = k ** 2 / (2 * np.pi ** 2)
    for m in (1, 2):
        cl, _ = oracle_cl2(k, pk, pair_off, k3, [m], T[None, None, :], T[None, None, :],
                           [[1]], 1)
        want = 0.0
        for i1 in range(k.size):
            s1 = ((k[1] - k[0]) / 2 if i1 == 0 else (k[-1] - k[-2]) / 2 if i1 == k.size - 1
                  else (k[i1 + 1] - k[i1 - 1]) / 2)
            for i2 in range(1, i1 + 1) if i1 > 0 else []:
                s2 = ((k[1] - k[0]) / 2 if i2 == 0 else (k[i1] - k[i1 - 1]) / 2 if i2 == i1
                      else (k[i2 + 1] - k[i2 - 1]) / 2)
                k1, k2 = k[i1], k[i2]
                inner, _ = quad(lambda q: (1 - ((q * q + k1 * k1 - k2 * k2) / (2 * q * k1)) ** 2)
                                ** m / (2 * k2), k1 - k2, k1 + k2, epsabs=0, epsrel=1e-12)
                want += 2 * s1 * s2 * inner / (4 * np.pi ** 4)
            if i1 > 0:  # i2 = 0 term
                k1, k2 = k[i1], k[0]
                s2 = (k[1] - k[0]) / 2
                inner, _ = quad(lambda q: (1 - ((q * q + k1 * k1 - k2 * k2) / (2 * q * k1)) ** 2)
                                ** m / (2 * k2), k1 - k2, k1 + k2, epsabs=0, epsrel=1e-12)
                want += 2 * s1 * s2 * inner / (4 * np.pi ** 4)
        assert cl[0] == pytest.approx(want, rel=1e-6), (m, cl[0], want)


def test_create_rejects_bad_inputs_before_gpu_work():
    k = np.array([0.01, 0.02])
    po = np.array([0, 2, 3, 5])  # second pair has one point
    k3 = np.array([0.001, 0.015, 0.01, 0.01, 0.03])
    t = np.ones((1, 1, 5))
    with pytest.raises(L.LosError, match="less than two"):
        L.Cl2(k, [1.0, 1.0], po, k3, [0], t, t, [[1]])
    po_ok = np.array([0, 2, 4, 6])
    k3_ok = np.array([0.001, 0.015, 0.01, 0.02, 0.01, 0.03])
    t6 = np.ones((1, 1, 6))
    with pytest.raises(L.LosError, match="non-negative"):
        L.Cl2(k, [1.0, 1.0], po_ok, k3_ok, [-1], t6, t6, [[1]])


def random_case(seed, k_size=10, m=(0, 1, 2), n_l=5):
    rng = np.random.default_rng(seed)
    k = np.exp(np.sort(rng.uniform(np.log(1e-4), np.log(0.1), k_size)))
    pair_off, k3, _ = triangle_grids(k, rng, 2, 60, outside=True)
    pk = 2.1e-9 * (k / 0.05) ** (0.965 - 1)
    t1 = rng.standard_normal((n_l, len(m), k3.size))
    t2 = rng.standard_normal((n_l, len(m), k3.size))
    used = np.ones((n_l, len(m)), dtype=np.int32)
    used[1, 0] = 0
    return k, pk, pair_off, k3, list(m), t1, t2, used


@pytest.mark.gpu
@pytest.mark.parametrize("seed,rescale,same", [(1, 0, True), (2, 1, False), (3, 1, True)])
def test_cl2_matches_oracle(seed, rescale, same):
    print(f"case cl2 seed={seed} rescale={rescale}", flush=True)
    k, pk, po, k3, m, t1, t2, used = random_case(seed)
    if same:
        t2 = t1
    _, ref = oracle_cl2(k, pk, po, k3, m, t1, t2, used, rescale)
    _, scale = oracle_cl2(k, pk, po, k3, m, np.abs(t1), np.abs(t2), used, rescale)
    h = L.Cl2(k, pk, po, k3, m, t1, t2, used, rescale)
    h.run(2)
    got = h.result()
    h.close()
    assert np.all(got[1, 0] == 0) and np.all(ref[1, 0] == 0)
    assert np.all(scale[used == 1] > 0)
    err = np.abs(got - ref)
    assert np.all(err <= 1e-12 * scale), float(np.max(err / np.where(scale > 0, scale, 1)))


@pytest.mark.gpu
def test_cl2_on_gpu_transfers():
    """TT spectrum from the LOS stage's transfers (T field, m = 0, 1, 2)."""
    from tests.test_oracle_pinning import _product_k3_list
    cfg = L.Config(has_T=1, has_E=0, m_list=[0, 1, 2], l_max_los_t=4, k_size=6, k3_size=12,
                   tau_size=80, l_size=8, l_max=300, dk_fine=1e-3, xx_max=1500.0)
    s = L.Scenario(cfg)
    s.transfer2_init()
    flat = s.gather()
    sizes = s.k3_sizes()[:, 0]
    pair_off = np.concatenate([[0], np.cumsum(sizes)])
    n_k3 = int(pair_off[-1])
    k3 = np.empty(n_k3)
    q = 0
    for i1 in range(cfg.k_size):
        for i2 in range(i1 + 1):
            k3[pair_off[q]:pair_off[q + 1]] = _product_k3_list(s, i1, i2, int(sizes[q]))
            q += 1
    k = s.grids()[1][:cfg.k_size]
    n_lm = flat.size // n_k3
    rows = flat.reshape(n_lm, n_k3)  # T field, (l,m) l-major, m ascending (all l >= m here)
    T = rows.reshape(cfg.l_size, 3, n_k3)
    pk = 2.1e-9 * (k / 0.05) ** (0.965 - 1)
    used = np.ones((cfg.l_size, 3), dtype=np.int32)
    _, ref = oracle_cl2(k, pk, pair_off, k3, [0, 1, 2], T, T, used, 1)
    h = L.Cl2(k, pk, pair_off, k3, [0, 1, 2], T, T, used, 1)
    h.run(1)
    got = h.result()
    h.close()
    s.close()
    assert np.all(ref > 0)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)
