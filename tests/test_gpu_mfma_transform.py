"""The matrix-core transforms of the mixed N = 256 path (GBGPU.transform =
"mfma_f32" / "mfma_bf16x3" / "mfma_bf16x6") against the FP64 oracle.

Bars: per-binary relative L2 waveform error <= 1e-6 and logL error <=
1e-6 (|<h|h>| + |<d|h>|) + 1e-6 -- the mixed-precision bars of the LDS
transform -- for the FP32 and three-piece bf16 variants; the two-piece bf16
split keeps about 16 significant bits, and is held to 3e-5 on waveforms and
3e-5 (|<h|h>| + |<d|h>|) + 1e-6 on logL.  Every variant also matches its own
batched and allocation-free entry points bit for bit."""

import numpy as np
import pytest

from oracle import fastgb as F
from tests.helpers import YR, draw_params, make_data, rel_err

pytestmark = pytest.mark.gpu

DT = 15.0
WAVE_BAR = {"lds": 1e-6, "mfma_f32": 1e-6, "mfma_bf16x6": 1e-6, "mfma_bf16x3": 3e-5}
TRANSFORMS = ["mfma_f32", "mfma_bf16x3", "mfma_bf16x6"]


@pytest.mark.parametrize("transform", TRANSFORMS)
def test_run_wave_mfma_matches_oracle(transform):
    from gbgpu_amd import GBGPU

    rng = np.random.default_rng(1300)
    B = 96
    # every series tier and the range-reduced path beyond it
    p = draw_params(B, rng, f_lo=1e-4, f_hi=4e-2, log_f=True)
    print(f"run_wave mixed N=256 transform={transform}", flush=True)
    gb = GBGPU(precision="mixed", transform=transform)
    gb.run_wave(*p, N=256, T=4 * YR, dt=DT)
    o = F.run_wave(*p, N=256, T=4 * YR, dt=DT)
    assert np.array_equal(gb.start_inds.cpu().numpy(), o["start_inds"])
    err = rel_err(gb.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2))
    err_ae = rel_err(gb.AETf.cpu().numpy()[:, :2], o["AETf"][:, :2], axis=(1, 2))
    print(f"  max per-binary error XYZ {err.max():.3e} AE {err_ae.max():.3e} (median {np.median(err):.3e})")
    assert err.max() <= WAVE_BAR[transform], err.max()
    assert err_ae.max() <= WAVE_BAR[transform], err_ae.max()
    assert np.all(np.isnan(gb.AETf.cpu().numpy()[:, 2]))
    # the LDS transform of the same policy for comparison: same start bins
    gl = GBGPU(precision="mixed")
    gl.run_wave(*p, N=256, T=4 * YR, dt=DT)
    assert rel_err(gb.XYZf.cpu().numpy(), gl.XYZf.cpu().numpy(), axis=(1, 2)).max() <= 2 * WAVE_BAR[transform]


@pytest.mark.parametrize("transform", TRANSFORMS)
def test_get_ll_mfma_matches_oracle(transform):
    from gbgpu_amd import GBGPU

    rng = np.random.default_rng(1400)
    B = 64
    p = draw_params(B, rng)
    A, E, SA, SE, df, L = make_data(p[:, :10], 256, YR, DT, rng)
    q = p.copy()
    q[1] *= 1 + 1e-7 * rng.standard_normal(B)
    q[4] += 0.1
    print(f"get_ll mixed N=256 transform={transform}", flush=True)
    gb = GBGPU(precision="mixed", transform=transform)
    gb.d_d = float(4 * df * (np.sum(np.abs(A) ** 2 / SA) + np.sum(np.abs(E) ** 2 / SE)))
    ll = gb.get_ll(q, [A, E], [SA, SE], N=256, T=YR, dt=DT)
    o = F.run_wave(*q, N=256, T=YR, dt=DT)
    d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"], A, E, SA, SE, df)
    ll_o = F.loglike(gb.d_d, d_h, h_h)
    rel = 3e-5 if transform == "mfma_bf16x3" else 1e-6
    bound = rel * (np.abs(h_h) + np.abs(d_h)) + 1e-6
    print(f"  max |dlogL| / bound {np.max(np.abs(ll - ll_o) / bound):.3e}")
    assert np.all(np.abs(ll - ll_o) <= bound), np.max(np.abs(ll - ll_o) / bound)
    llm = gb.get_ll(q, [A, E], [SA, SE], N=256, T=YR, dt=DT, phase_marginalize=True)
    llm_o = F.loglike(gb.d_d, d_h, h_h, phase_marginalize=True)
    assert np.all(np.abs(llm - llm_o) <= bound)


@pytest.mark.parametrize("transform", TRANSFORMS)
def test_batch_and_noalloc_with_mfma_match_uniform_call(transform):
    """The batched N = 256 bucket and the allocation-free entry point run the
    same transform as get_ll: identical loglike bit for bit; the other
    buckets keep the LDS transform (identical to a batch without it)."""
    import torch

    from gbgpu_amd import GBGPU, _capi
    from gbgpu_amd.batch import BatchedLogLike

    rng = np.random.default_rng(1500)
    B = 600
    p = draw_params(B, rng, f_lo=5e-4, f_hi=8e-3, log_f=True)
    N = np.where(np.arange(B) % 3 == 0, 128, 256).astype(np.int32)
    T = 4 * YR
    A, E, SA, SE, df, L = make_data(p[:, :20], 256, T, DT, rng)
    gb = GBGPU(precision="mixed", transform=transform)
    gb.d_d = float(4 * df * (np.sum(np.abs(A) ** 2 / SA) + np.sum(np.abs(E) ** 2 / SE)))
    print(f"batch + noalloc transform={transform}", flush=True)
    sel = N == 256
    ref256 = gb.get_ll(p[:, sel], [A, E], [SA, SE], N=256, T=T, dt=DT)
    bl = BatchedLogLike(gb, [A, E], [SA, SE], T, DT, precision="mixed", strategy="buckets", guard_tol=None)
    ll = bl(p, N=N).cpu().numpy()
    assert np.array_equal(ll[sel], ref256)
    bl0 = BatchedLogLike(gb, [A, E], [SA, SE], T, DT, precision="mixed", strategy="buckets", guard_tol=None,
                         transform="lds")
    ll0 = bl0(p, N=N).cpu().numpy()
    assert np.array_equal(ll[~sel], ll0[~sel])
    # allocation-free entry point, same transform
    lib = _capi.load()
    Tadj = int(T / DT) * DT
    pt = torch.from_numpy(np.ascontiguousarray(p[:, sel])).to("cuda")
    Bs = pt.shape[1]
    dA, dE, S_A, S_E, di, ni, Ld = gb._prep_data([A, E], [SA, SE], None, None, Bs)
    dat = _capi.Data(A=dA.data_ptr(), E=dE.data_ptr(), psdA=S_A.data_ptr(), psdE=S_E.data_ptr(),
                     data_index=di.data_ptr(), noise_index=ni.data_ptr(), data_length=int(Ld), start_freq_ind=0,
                     df=float(1 / Tadj))
    cfg, _tab = gb._cfg(256, Tadj, False, "mixed", _capi.TRANSFORMS[transform])
    out = torch.full((Bs,), np.nan, dtype=torch.float64, device="cuda")
    rc = lib.gbgpu_wave_ll_noalloc(out.data_ptr(), None, None, pt.data_ptr(), 9, Bs, cfg, dat, gb.d_d, 0,
                                   _capi.stream_handle())
    _capi.check(rc, "gbgpu_wave_ll_noalloc")
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref256)


def test_transform_falls_back_to_lds_where_unsupported():
    """GBGPU applies the matrix-core transform to mixed N = 256 circular
    calls only: an FP64 call or another N with transform set gives exactly
    the LDS-transform result."""
    from gbgpu_amd import GBGPU

    rng = np.random.default_rng(1600)
    p = draw_params(16, rng)
    for prec, N in (("fp64", 256), ("mixed", 128)):
        a, b = GBGPU(precision=prec, transform="mfma_f32"), GBGPU(precision=prec)
        a.run_wave(*p, N=N, T=YR, dt=DT)
        b.run_wave(*p, N=N, T=YR, dt=DT)
        assert np.array_equal(a.XYZf.cpu().numpy(), b.XYZf.cpu().numpy())
