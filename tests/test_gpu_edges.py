"""Edge cases of the LOS stage on cuda:0, each against the CPU oracle.

* the largest coarse k3 grid the stage accepts (256 points per pair, the LDS tile of the
  source re-layout above 64 KiB) and an odd count next to it;
* the smallest coarse grid (4 points: the spline's end formulas use three of them);
* sources up to L = 7 (the largest multipole the kernels are built for);
* six azimuthal modes m = 0..5 (the widest table-kernel workgroup);
* an empty pair list: load, autotune, capture, run and store do nothing and succeed.
The plain FP64 form must be bit-identical; the fast forms run their sweep shapes to the
tolerances of the parity tests.
"""
import numpy as np
import pytest

from song_amd import lib as L
from tests.conftest import field_errors, run_oracle

BASE = dict(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=3, l_max_los_p=3, k_size=3,
            k3_size=12, tau_size=48, l_size=8, l_max=300, dk_fine=2e-3, xx_max=1200.0)
EDGES = {
    "coarse_256": dict(BASE, k3_size=256, dk_fine=4e-4),
    "coarse_255": dict(BASE, k3_size=255, dk_fine=4e-4),
    "coarse_4": dict(BASE, k3_size=4),
    "L_7": dict(BASE, l_max_los_t=7, l_max_los_p=7, m_list=[0, 2]),
    "m_0_to_5": dict(BASE, m_list=[0, 1, 2, 3, 4, 5], l_max_los_t=5, l_max_los_p=5),
}


def _check(cfg_kw, oracle):
    cfg = L.Config(**cfg_kw)
    ref = run_oracle(L.Scenario(cfg), oracle)
    s = L.Scenario(cfg)
    s.prepare()
    eng = L.Engine(0)
    eng.load(s)
    tt2 = ref.size // int(s.k3_sizes()[:, 0].sum())
    for form, shapes, tol in ((L.FORM_FP64, [(0, 1, 0, 0), (1, 2, 0, 0), (2, 64, 1, 0)], 0.0),
                              (L.FORM_FP64_FAST, [(1, 4, 0, 0)], 1e-12),
                              (L.FORM_FP32C, [(1, 4, 1, 0), (1, 2, 0, 0)], 1e-5)):
        eng.set_form(form)
        for sh in shapes:
            eng.set_shape(form, *sh)
            eng.run()
            eng.store(s)
            got = s.gather()
            print(f"form {form} shape {sh}", flush=True)
            if tol == 0.0:
                assert np.array_equal(got, ref), f"form {form} shape {sh}"
            else:
                assert field_errors(got, ref, tt2).max() < tol, f"form {form} shape {sh}"
    eng.close()
    s.close()
    assert np.count_nonzero(ref) > 0.2 * ref.size


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(EDGES))
def test_edge_case(name, oracle):
    print(f"case {name}", flush=True)
    _check(EDGES[name], oracle)


@pytest.mark.gpu
def test_transfer2_init_at_largest_coarse_grid(oracle):
    """The drop-in entry point itself on the 256-point coarse grid."""
    cfg = L.Config(**EDGES["coarse_256"])
    ref = run_oracle(L.Scenario(cfg), oracle)
    s = L.Scenario(cfg)
    s.transfer2_init()
    assert np.array_equal(s.gather(), ref)
    s.close()


@pytest.mark.gpu
def test_empty_pair_list():
    cfg = L.Config(**BASE)
    s = L.Scenario(cfg)
    s.prepare()
    eng = L.Engine(0)
    eng.load_pairs(s, [])
    assert eng.n_elements() == 0
    for form in (L.FORM_FP64, L.FORM_FP32C):
        eng.set_form(form)
        eng.autotune(1)
        eng.capture()
        assert eng.run() == 0.0
        eng.store(s)
    eng.close()
    s.close()
