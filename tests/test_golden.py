"""The committed golden vectors (tests/golden/fastgb_small.npz, written by
tests/golden/make_fixtures.py) against the CPU oracle: a regression pin of
the oracle, whose correctness is established by tests/test_oracle.py and
tests/test_oracle_physics.py.  The HIP path is compared with the same
vectors in tests/test_gpu_golden.py."""

from pathlib import Path

import numpy as np

from oracle import fastgb as F
from tests.helpers import rel_err

G = np.load(Path(__file__).resolve().parent / "golden" / "fastgb_small.npz")


def test_golden_circular_and_tdi2():
    dt, T = float(G["dt"]), float(G["T_in"])
    o = F.run_wave(*G["circ_params"], N=int(G["circ_N"]), T=T, dt=dt)
    assert np.array_equal(o["start_inds"], G["circ_start"])
    assert rel_err(o["XYZf"], G["circ_XYZf"]) < 1e-12
    assert rel_err(o["AETf"], G["circ_AETf"]) < 1e-12
    o = F.run_wave(*G["tdi2_params"], N=int(G["tdi2_N"]), T=T, dt=dt, tdi2=True)
    assert np.array_equal(o["start_inds"], G["tdi2_start"])
    assert rel_err(o["XYZf"], G["tdi2_XYZf"]) < 1e-12


def test_golden_third_body():
    dt, T = float(G["dt"]), float(G["T_in"])
    o = F.run_wave(*G["third_params"], N=int(G["third_N"]), T=T, dt=dt)
    assert np.array_equal(o["start_inds"], G["third_start"])
    assert rel_err(o["XYZf"], G["third_XYZf"]) < 1e-12


def test_golden_likelihood_window():
    dt, T = float(G["dt"]), float(G["T_in"])
    s_lo = int(G["ll_start_freq_ind"])
    o = F.run_wave(*G["ll_params"], N=int(G["ll_N"]), T=T, dt=dt)
    d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"] - s_lo, G["ll_A"], G["ll_E"],
                             G["ll_S"], G["ll_S"], o["df"])
    assert np.allclose(d_h, G["ll_d_h"], rtol=1e-12, atol=0)
    assert np.allclose(h_h, G["ll_h_h"], rtol=1e-12, atol=0)
    ll = F.loglike(float(G["ll_d_d"]), d_h, h_h)
    assert np.allclose(ll, G["ll_loglike"], rtol=1e-13, atol=0)
