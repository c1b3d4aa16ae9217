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
