"""The FP64 matrix-core transform (GBGPU.transform = "mfma_f64" with the
FP64 policy, N = 256 without a third body): the length-256 transform as two
16 x 16 complex matrix products with FP64 products and sums, instead of the
LDS radix-4 transform.

Bars: the FP64 waveform bar (1e-9 relative plus the carrier-phase floor
5e-16 x 2 pi f T) against the oracle, as for the LDS transform; against the
LDS transform itself 1e-12 relative (the two differ by rounding only);
log-likelihoods by tests.helpers.ll_bar against the oracle and within
1e-9 + 1e-12 (|<h|h>| + |<d|h>|) of the LDS transform's; the
allocation-free entry equals gbgpu_wave_ll bit for bit."""

import numpy as np
import pytest

from oracle import fastgb as F
from tests.helpers import YR, draw_params, ll_bar, make_data, rel_err

pytestmark = pytest.mark.gpu

DT = 15.0


def fp64_tol(p, T):
    fmax = np.abs(p[1]) + np.abs(p[2]) * T + 0.5 * np.abs(p[3]) * T**2
    return 1e-9 + 5e-16 * 2 * np.pi * fmax * T


def test_run_wave_mfma_f64_matches_oracle_and_lds():
    from gbgpu_amd import GBGPU

    rng = np.random.default_rng(1700)
    B = 96
    # every series tier and the range-reduced path beyond it
    p = draw_params(B, rng, f_lo=1e-4, f_hi=4e-2, log_f=True)
    print("run_wave fp64 N=256 transform=mfma_f64", flush=True)
    gb = GBGPU(precision="fp64", transform="mfma_f64")
    gb.run_wave(*p, N=256, T=4 * YR, dt=DT)
    o = F.run_wave(*p, N=256, T=4 * YR, dt=DT)
    assert np.array_equal(gb.start_inds.cpu().numpy(), o["start_inds"])
    tol = fp64_tol(p, 4 * YR)
    err = rel_err(gb.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2))
    err_aet = rel_err(gb.AETf.cpu().numpy(), o["AETf"], axis=(1, 2))
    print(f"  max error / bar XYZ {np.max(err / tol):.3e} AET {np.max(err_aet / tol):.3e}")
    assert np.all(err <= tol), np.max(err / tol)
    assert np.all(err_aet <= tol), np.max(err_aet / tol)
    gl = GBGPU(precision="fp64")
    gl.run_wave(*p, N=256, T=4 * YR, dt=DT)
    d = rel_err(gb.XYZf.cpu().numpy(), gl.XYZf.cpu().numpy(), axis=(1, 2))
    print(f"  max relative difference from the LDS transform {d.max():.3e}")
    assert d.max() <= 1e-12, d.max()


def test_get_ll_mfma_f64_matches_oracle_lds_and_noalloc():
    import torch

    from gbgpu_amd import GBGPU, _capi

    rng = np.random.default_rng(1800)
    B = 64
    p = draw_params(B, rng)
    A, E, SA, SE, df, L = make_data(p[:, :10], 256, YR, DT, rng)
    q = p.copy()
    q[1] *= 1 + 1e-7 * rng.standard_normal(B)
    q[4] += 0.1
    print("get_ll fp64 N=256 transform=mfma_f64", flush=True)
    gb = GBGPU(precision="fp64", transform="mfma_f64")
    gb.d_d = 0.0
    ll = gb.get_ll(q, [A, E], [SA, SE], N=256, T=YR, dt=DT)
    o = F.run_wave(*q, N=256, T=YR, dt=DT)
    d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"], A, E, SA, SE, df)
    ll_o = F.loglike(0.0, d_h, h_h)
    assert np.all(np.abs(ll - ll_o) <= ll_bar(d_h, h_h)), np.max(np.abs(ll - ll_o) / ll_bar(d_h, h_h))
    gl = GBGPU(precision="fp64")
    gl.d_d = 0.0
    ll_lds = gl.get_ll(q, [A, E], [SA, SE], N=256, T=YR, dt=DT)
    bar = 1e-9 + 1e-12 * (np.abs(h_h) + np.abs(d_h))
    assert np.all(np.abs(ll - ll_lds) <= bar), np.max(np.abs(ll - ll_lds) / bar)
    # the allocation-free entry with the same transform: bit for bit
    lib = _capi.load()
    pt, N, Tadj = gb._prepare(*q[:9], q[9:], 256, YR, DT, 1)
    cfg, _tab = gb._cfg(N, Tadj, False, "fp64", _capi.TRANSFORMS["mfma_f64"])
    dev = lambda x, dt: torch.as_tensor(x, dtype=dt, device="cuda")  # noqa: E731
    At, Et, SAt, SEt = dev(A, torch.complex128), dev(E, torch.complex128), dev(SA, torch.float64), \
        dev(SE, torch.float64)
    dat = _capi.Data(A=At.data_ptr(), E=Et.data_ptr(), psdA=SAt.data_ptr(), psdE=SEt.data_ptr(), data_index=None,
                     noise_index=None, data_length=L, start_freq_ind=0, df=df)
    out = {}
    for name in ("gbgpu_wave_ll", "gbgpu_wave_ll_noalloc"):
        r = torch.empty(B, dtype=torch.float64, device="cuda")
        _capi.check(getattr(lib, name)(r.data_ptr(), None, None, pt.data_ptr(), 9, B, cfg, dat, 0.0, 0,
                                       _capi.stream_handle()), name)
        torch.cuda.synchronize()
        out[name] = r.cpu().numpy()
    assert np.array_equal(out["gbgpu_wave_ll"], out["gbgpu_wave_ll_noalloc"])
    assert np.allclose(out["gbgpu_wave_ll"], ll, rtol=0, atol=1e-9)
