"""run_wave with N chosen by get_N / special_get_N (N=None) and with
oversample, against the oracle evaluated at the same N: bit-exact start
bins, FP64 waveforms within 1e-9 relative (1e-8 with a third body)."""

import numpy as np
import pytest

from oracle import fastgb as F
from tests.helpers import YR, draw_params, rel_err

pytestmark = pytest.mark.gpu

DT = 15.0


@pytest.mark.parametrize("oversample", [1, 2, 4])
def test_run_wave_N_none_matches_oracle(oversample):
    from gbgpu_amd import GBGPU
    from gbgpu_amd.utils.utility import get_N

    rng = np.random.default_rng(810 + oversample)
    p = draw_params(24, rng, f_lo=1e-3, f_hi=8e-3)
    print(f"run_wave N=None oversample={oversample}", flush=True)
    gb = GBGPU()
    gb.run_wave(*p, N=None, T=YR, dt=DT, oversample=oversample)
    N = int(gb.N)
    assert N == int(np.max(get_N(p[0], p[1], F.adjust_T(YR, DT), oversample=oversample)))
    assert N == 64 * oversample  # 1-8 mHz over 1 yr: the 64-sample band, quiet sources
    o = F.run_wave(*p, N=N, T=YR, dt=DT)
    assert np.array_equal(gb.start_inds.cpu().numpy(), o["start_inds"])
    assert rel_err(gb.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2)).max() <= 1e-9


def test_third_body_N_none_matches_oracle():
    """A 0.02-yr outer orbit over 1 yr needs 100 samples per year: N doubles
    from the 64-sample band to 128."""
    from gbgpu_amd import GBGPUThirdBody

    rng = np.random.default_rng(820)
    B = 6
    p = draw_params(B, rng, f_lo=1e-3, f_hi=4e-3)
    third = [rng.uniform(0, 30, B), rng.uniform(0, 2 * np.pi, B), np.zeros(B), np.full(B, 0.02), rng.uniform(0, 1, B)]
    gb = GBGPUThirdBody()
    gb.run_wave(*p, *third, N=None, T=YR, dt=DT)
    assert int(gb.N) == 128
    o = F.run_wave(*p, *third, N=128, T=YR, dt=DT)
    assert np.array_equal(gb.start_inds.cpu().numpy(), o["start_inds"])
    assert rel_err(gb.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2)).max() <= 1e-8
