"""GPU parity at N = 4096 and 8192 (the large-N kernels: one 1024-thread
workgroup per binary, channels transformed in passes that fit in LDS), against
the CPU FP64 oracle on identical seeded inputs.

Bars, as for the other sizes (tests/test_gpu_parity.py):
  * start bins: bit-exact;
  * FP64 waveforms: per-binary relative L2 error <= 1e-9 plus the carrier
    phase's FP64 floor 5e-16 * 2 pi f_max T (fp64_tol);
  * FP64 log-likelihoods: tests.helpers.ll_bar (absolute, on h_h - 2 Re d_h);
  * mixed: waveforms 1e-6 relative; logL 1e-6 (|h_h| + |d_h|) + 1e-6.
"""

import numpy as np
import pytest

from oracle import fastgb as F
from tests.helpers import YR, draw_params, ll_bar, make_data, rel_err

pytestmark = pytest.mark.gpu

DT = 15.0
BIG = [4096, 8192]


def _gb(**kw):
    from gbgpu_amd import GBGPU

    return GBGPU(**kw)


def fp64_tol(p, T):
    fmax = np.abs(p[1]) + np.abs(p[2]) * T + 0.5 * np.abs(p[3]) * T**2
    return 1e-9 + 5e-16 * 2 * np.pi * fmax * T


def _third(B, rng, e2=0.3):
    """A2, varpi, e2, P2 (yr), T2 (yr)."""
    return [rng.uniform(0, 100, B), rng.uniform(0, 2 * np.pi, B), np.full(B, e2), 10 ** rng.uniform(0, 1, B),
            rng.uniform(0, 1, B)]


@pytest.mark.parametrize("N", BIG)
@pytest.mark.parametrize("tdi2", [False, True])
def test_run_wave_large_N_fp64(N, tdi2):
    """0.5 mHz to 0.3 Hz at 4 yr: the low-frequency series tiers and the
    reduced sine/cosine of the high-frequency links, X, Y, Z through three
    channel passes (FP64, N = 8192) or two (N = 4096)."""
    rng = np.random.default_rng(4000 + N + tdi2)
    B = 10
    p = draw_params(B, rng, f_lo=5e-4, f_hi=0.3, log_f=True)
    # above 20 mHz, slow the chirp so that the signal stays in the N-bin band
    hi = p[1] > 0.02
    p[2, hi] *= 1e-4
    p[3] = 11.0 / 3.0 * p[2] ** 2 / p[1]
    print(f"run_wave fp64 N={N} tdi2={tdi2}", flush=True)
    gb = _gb()
    gb.run_wave(*p, N=N, T=4 * YR, dt=DT, tdi2=tdi2)
    o = F.run_wave(*p, N=N, T=4 * YR, dt=DT, tdi2=tdi2)
    assert np.array_equal(gb.start_inds.cpu().numpy(), o["start_inds"])
    tol = fp64_tol(p, 4 * YR)
    err = rel_err(gb.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2))
    assert np.all(err <= tol), np.max(err / tol)
    aet = gb.AETf.cpu().numpy()
    err = rel_err(aet[:, :2], o["AETf"][:, :2], axis=(1, 2))
    assert np.all(err <= tol), np.max(err / tol)
    # T cancels to O((f/f*)^2) of X: compared on the A/E scale per binary
    scale = np.abs(o["AETf"]).max(axis=(1, 2))
    assert np.all(np.abs(aet[:, 2] - o["AETf"][:, 2]).max(axis=1) <= tol * scale)


@pytest.mark.parametrize("N", BIG)
def test_run_wave_large_N_without_xyz_output(N):
    """gbgpu_run_wave with XYZf = NULL keeps X and Y between channel passes in
    the AETf buffer itself; its A, E, T equal the call that also writes XYZf
    bit for bit."""
    import torch

    from gbgpu_amd import _capi

    rng = np.random.default_rng(4100 + N)
    p = draw_params(6, rng, f_lo=1e-3, f_hi=2e-2, log_f=True)
    gb = _gb()
    pt, N, Tadj = gb._prepare(*p, (), N, YR, DT, 1)
    B = pt.shape[1]
    cfg, _tab = gb._cfg(N, Tadj, False, "fp64")
    lib = _capi.load()
    out = {}
    for with_xyz in (True, False):
        xyz = torch.empty((B, 3, N), dtype=torch.complex128, device="cuda") if with_xyz else None
        aet = torch.full((B, 3, N), np.nan, dtype=torch.complex128, device="cuda")
        st = torch.empty(B, dtype=torch.int64, device="cuda")
        _capi.check(lib.gbgpu_run_wave(_capi.ptr(xyz), aet.data_ptr(), st.data_ptr(), pt.data_ptr(), pt.shape[0], B,
                                       cfg, _capi.stream_handle()), "gbgpu_run_wave")
        out[with_xyz] = (aet.cpu().numpy(), st.cpu().numpy())
    assert np.array_equal(out[True][1], out[False][1])
    assert np.array_equal(out[True][0], out[False][0])


@pytest.mark.parametrize("N", BIG)
def test_get_ll_large_N_fp64_multi_stream(N):
    """Fused likelihood at N = 4096 / 8192 on 2-D data with per-binary
    data_index / noise_index and a window offset (start_freq_ind), including
    windows that run over either end of the stream."""
    import torch

    rng = np.random.default_rng(4200 + N)
    B, T = 12, YR
    p = draw_params(B, rng, f_lo=1.0e-3, f_hi=1.2e-3)
    df = 1 / F.adjust_T(T, DT)
    k_lo = int(np.rint(1.0e-3 / df))
    start_freq_ind = k_lo - N // 4
    L = int(np.rint(1.2e-3 / df)) - start_freq_ind + N // 4
    S = 2
    dA = (rng.standard_normal((S, L)) + 1j * rng.standard_normal((S, L))) * 1e-20
    dE = (rng.standard_normal((S, L)) + 1j * rng.standard_normal((S, L))) * 1e-20
    from gbgpu_amd.noise import A1TDISens

    f = (np.arange(L) + start_freq_ind) * df
    SA = np.stack([A1TDISens.get_Sn(f) * (1 + 0.2 * s) for s in range(S)])
    SE = SA * 1.1
    di = rng.integers(0, S, B).astype(np.int32)
    ni = rng.integers(0, S, B).astype(np.int32)
    print(f"get_ll fp64 N={N} L={L}", flush=True)
    gb = _gb()
    gb.d_d = 0.0
    ll = gb.get_ll(p, [torch.from_numpy(dA).cuda(), torch.from_numpy(dE).cuda()], [SA, SE], N=N, T=T, dt=DT,
                   start_freq_ind=start_freq_ind, data_index=di, noise_index=ni)
    o = F.run_wave(*p, N=N, T=T, dt=DT)
    d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"] - start_freq_ind, dA, dE, SA, SE, df,
                             di, ni, L)
    ll_o = F.loglike(0.0, d_h, h_h)
    assert np.all(np.abs(ll - ll_o) <= ll_bar(d_h, h_h)), np.max(np.abs(ll - ll_o) / ll_bar(d_h, h_h))
    assert np.allclose(gb.d_h.cpu().numpy(), d_h, rtol=1e-9, atol=1e-9 * np.abs(d_h).max())


@pytest.mark.parametrize("N", BIG)
def test_get_ll_large_N_with_injection_and_phase_marginalisation(N):
    """Signal-dominated data (the binaries themselves injected, 1 yr): logL
    and the phase-marginalised logL against the oracle."""
    rng = np.random.default_rng(4300 + N)
    B = 8
    p = draw_params(B, rng, f_lo=2e-3, f_hi=4e-3)
    p[0] *= 30.0
    A, E, SA, SE, df, L = make_data(p, N, YR, DT, rng)
    gb = _gb()
    gb.d_d = float(4 * df * (np.sum(np.abs(A) ** 2 / SA) + np.sum(np.abs(E) ** 2 / SE)))
    o = F.run_wave(*p, N=N, T=YR, dt=DT)
    d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"], A, E, SA, SE, df)
    for marg in (False, True):
        print(f"get_ll fp64 N={N} phase_marginalize={marg}", flush=True)
        ll = gb.get_ll(p, [A, E], [SA, SE], N=N, T=YR, dt=DT, phase_marginalize=marg)
        ll_o = F.loglike(gb.d_d, d_h, h_h, phase_marginalize=marg)
        assert np.all(np.abs(ll - ll_o) <= ll_bar(d_h, h_h)), np.max(np.abs(ll - ll_o) / ll_bar(d_h, h_h))


@pytest.mark.parametrize("N", BIG)
def test_large_N_mixed_precision(N):
    """Mixed policy at N = 4096 / 8192 (FP32 transform of length N): A/E
    waveforms within 1e-6 of the FP64 oracle, T marked NaN, logL within the
    mixed bar."""
    rng = np.random.default_rng(4400 + N)
    B = 8
    p = draw_params(B, rng, f_lo=2e-3, f_hi=4e-3)
    p[0] *= 10.0
    A, E, SA, SE, df, L = make_data(p, 256, YR, DT, rng)
    o = F.run_wave(*p, N=N, T=YR, dt=DT)
    print(f"mixed N={N}", flush=True)
    gb = _gb(precision="mixed")
    gb.run_wave(*p, N=N, T=YR, dt=DT)
    assert np.array_equal(gb.start_inds.cpu().numpy(), o["start_inds"])
    aet = gb.AETf.cpu().numpy()
    assert rel_err(aet[:, :2], o["AETf"][:, :2], axis=(1, 2)).max() <= 1e-6
    assert rel_err(gb.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2)).max() <= 1e-6
    assert np.all(np.isnan(aet[:, 2]))
    gb.d_d = float(4 * df * (np.sum(np.abs(A) ** 2 / SA) + np.sum(np.abs(E) ** 2 / SE)))
    ll = gb.get_ll(p, [A, E], [SA, SE], N=N, T=YR, dt=DT)
    d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"], A, E, SA, SE, df)
    ll_o = F.loglike(gb.d_d, d_h, h_h)
    bound = 1e-6 * (np.abs(h_h) + np.abs(d_h)) + 1e-6
    assert np.all(np.abs(ll - ll_o) <= bound), np.max(np.abs(ll - ll_o) / bound)


@pytest.mark.parametrize("N", BIG)
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_run_wave_ae_large_N(N, precision):
    """The A/E-only entry (global templates) at N = 4096 / 8192 equals
    run_wave's A and E (1e-13 of the largest bin in FP64; the mixed bar in
    mixed), same start bins."""
    import torch

    from gbgpu_amd import _capi

    rng = np.random.default_rng(4500 + N)
    p = draw_params(6, rng, f_lo=1e-3, f_hi=4e-2, log_f=True)
    gb = _gb(precision=precision)
    gb.run_wave(*p, N=N, T=YR, dt=DT)
    pt, N, Tadj = gb._prepare(*p, (), N, YR, DT, 1)
    B = pt.shape[1]
    cfg, _tab = gb._cfg(N, Tadj, False, precision)
    ae = torch.empty((B, 2, N), dtype=torch.complex128, device="cuda")
    st = torch.empty(B, dtype=torch.int64, device="cuda")
    _capi.check(_capi.load().gbgpu_run_wave_ae(ae.data_ptr(), st.data_ptr(), pt.data_ptr(), pt.shape[0], B, cfg,
                                               _capi.stream_handle()), "gbgpu_run_wave_ae")
    ae = ae.cpu().numpy()
    ref = gb.AETf[:, :2].cpu().numpy()
    assert np.array_equal(st.cpu().numpy(), gb.start_inds.cpu().numpy())
    if precision == "fp64":
        assert np.abs(ae - ref).max() <= 1e-13 * np.abs(ref).max()
    else:
        o = F.run_wave(*p, N=N, T=YR, dt=DT)
        assert rel_err(ae, o["AETf"][:, :2], axis=(1, 2)).max() <= 1e-6


@pytest.mark.parametrize("N", BIG)
def test_third_body_large_N(N):
    """Eccentric third body at N = 4096 / 8192: the pre-pass spans 16
    wavefronts of 4 or 8 consecutive samples per thread; FP64 run_wave and
    get_ll against the oracle."""
    from gbgpu_amd import GBGPUThirdBody

    rng = np.random.default_rng(4600 + N)
    B = 6
    p = draw_params(B, rng, f_lo=1e-3, f_hi=4e-3)
    third = _third(B, rng)
    print(f"third body N={N}", flush=True)
    gb = GBGPUThirdBody()
    gb.run_wave(*p, *third, N=N, T=YR, dt=DT)
    o = F.run_wave(*p, *third, N=N, T=YR, dt=DT)
    assert np.array_equal(gb.start_inds.cpu().numpy(), o["start_inds"])
    assert rel_err(gb.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2)).max() <= 1e-8
    A, E, SA, SE, df, L = make_data(p, 256, YR, DT, rng)
    gb.d_d = float(4 * df * (np.sum(np.abs(A) ** 2 / SA) + np.sum(np.abs(E) ** 2 / SE)))
    ll = gb.get_ll(np.vstack([p, np.array(third)]), [A, E], [SA, SE], N=N, T=YR, dt=DT)
    d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"], A, E, SA, SE, df)
    ll_o = F.loglike(gb.d_d, d_h, h_h)
    # the FP64 third-body waveform bar (1e-8) carried through the inner products
    assert np.all(np.abs(ll - ll_o) <= ll_bar(d_h, h_h, rel=1e-8))


def test_auto_N_reaches_the_large_kernels():
    """run_wave(N=None) above 0.3 Hz: get_N's band table gives 1024 m, i.e.
    4096 at 4 yr and 8192 beyond 4 yr (m = 8), and the templates match the
    oracle at that N (monochromatic sources, so the band holds the signal)."""
    from gbgpu_amd.utils.utility import get_N

    rng = np.random.default_rng(4700)
    p = draw_params(3, rng, f_lo=0.31, f_hi=0.4)
    p[2] = p[3] = 0.0
    for yrs, want in ((4, 4096), (6, 8192)):
        T = F.adjust_T(yrs * YR, DT)
        assert np.all(get_N(p[0], p[1], T) == want)
        print(f"auto N={want} T={yrs} yr", flush=True)
        gb = _gb()
        gb.run_wave(*p, N=None, T=yrs * YR, dt=DT)
        assert gb.N == want
        o = F.run_wave(*p, N=want, T=yrs * YR, dt=DT)
        assert np.array_equal(gb.start_inds.cpu().numpy(), o["start_inds"])
        err = rel_err(gb.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2))
        assert np.all(err <= fp64_tol(p, T)), np.max(err / fp64_tol(p, T))


def test_noalloc_large_N_matches_wave_ll():
    """The allocation-free entry at N = 4096 and 8192 (twiddle tables in its
    fixed scratch) is bit-identical to gbgpu_wave_ll, on two consecutive
    calls."""
    import torch

    from gbgpu_amd import _capi

    lib = _capi.load()
    rng = np.random.default_rng(4800)
    inj = draw_params(16, rng)
    A, E, SA, SE, df, L = make_data(inj, 256, YR, DT, rng)
    gb = _gb()
    d_d = float(4 * df * (np.sum(np.abs(A) ** 2 / SA) + np.sum(np.abs(E) ** 2 / SE)))
    Tadj = F.adjust_T(YR, DT)
    for B, N, prec in ((300, 8192, "fp64"), (100, 4096, "mixed")):
        print(f"noalloc N={N} B={B} {prec}", flush=True)
        p = torch.from_numpy(draw_params(B, rng)).to("cuda")
        dA, dE, S_A, S_E, di, ni, Ld = gb._prep_data([A, E], [SA, SE], None, None, B)
        dat = _capi.Data(A=dA.data_ptr(), E=dE.data_ptr(), psdA=S_A.data_ptr(), psdE=S_E.data_ptr(),
                         data_index=di.data_ptr(), noise_index=ni.data_ptr(), data_length=int(Ld),
                         start_freq_ind=0, df=float(1 / Tadj))
        cfg, _tab = gb._cfg(N, Tadj, False, prec)
        res = []
        for fn in (lib.gbgpu_wave_ll, lib.gbgpu_wave_ll_noalloc):
            ll = torch.full((B,), np.nan, dtype=torch.float64, device="cuda")
            _capi.check(fn(ll.data_ptr(), None, None, p.data_ptr(), 9, B, cfg, dat, d_d, 0, _capi.stream_handle()),
                        "wave_ll")
            res.append(ll.cpu().numpy())
        assert np.all(np.isfinite(res[1]))
        assert np.array_equal(res[0], res[1])
