"""GPU parity of the uneven-N batch (gbgpu_wave_ll_batched, via get_ll with an
array N and BatchedLogLike) beyond circular / TDI1 / one stream / N <= 2048:
N from 64 to 8192 in one call, TDI2, 2-D data with per-binary data_index and
noise_index, and third-body rows.

Bars:
  * the batch equals one standalone get_ll call per N bit for bit (the same
    per-binary arithmetic: the bucket launches run the uniform kernels' tiles);
  * FP64 against the oracle: tests.helpers.ll_bar (absolute, on
    h_h - 2 Re d_h; the third-body waveform bar 1e-8 carried through);
  * mixed against the oracle: 1e-6 (|h_h| + |d_h|) + 1e-6.
"""

import numpy as np
import pytest

from oracle import fastgb as F
from tests.helpers import YR, draw_params, ll_bar, make_data

pytestmark = pytest.mark.gpu

DT = 15.0
N_ALL = np.array([64, 128, 256, 512, 1024, 2048, 4096, 8192])


def _streams(rng, T=YR):
    """Two full-length A/E streams (noise + a few injected binaries) and their
    PSDs (the second scaled), as 2-D arrays."""
    inj = draw_params(8, rng)
    A, E, SA, SE, df, L = make_data(inj, 256, T, DT, rng)
    A2 = 0.7 * A + 0.3 * np.roll(A, 17)
    E2 = 0.7 * E + 0.3 * np.roll(E, 17)
    return (np.stack([A, A2]), np.stack([E, E2]), np.stack([SA, 1.3 * SA]), np.stack([SE, 1.1 * SE]), df, L)


def _case(rng, B, third):
    p = draw_params(B, rng, f_lo=1e-3, f_hi=4e-3)
    p[0] *= 10.0
    if third:  # A2, varpi, e2, P2 (yr), T2 (yr)
        p = np.vstack([p, rng.uniform(0, 100, B), rng.uniform(0, 2 * np.pi, B), rng.choice([0.0, 0.3], B),
                       10 ** rng.uniform(0, 1, B), rng.uniform(0, 1, B)])
    N = rng.choice(N_ALL, B)
    N[: len(N_ALL)] = N_ALL  # every bucket populated
    return p, N


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("third", [False, True])
def test_uneven_batch_tdi2_streams_third_body(precision, third):
    from gbgpu_amd import GBGPU, GBGPUThirdBody

    rng = np.random.default_rng(5000 + 2 * third + (precision == "mixed"))
    B = 64
    p, N = _case(rng, B, third)
    A, E, SA, SE, df, L = _streams(rng)
    di = rng.integers(0, 2, B).astype(np.int32)
    ni = rng.integers(0, 2, B).astype(np.int32)
    gb = GBGPUThirdBody(precision=precision) if third else GBGPU(precision=precision)
    gb.d_d = 123.0
    print(f"uneven batch {precision} third={third} N={sorted(set(N.tolist()))}", flush=True)
    kw = dict(T=YR, dt=DT, tdi2=True)
    ll = gb.get_ll(p, [A, E], [SA, SE], N=N, data_index=di, noise_index=ni, **kw)
    ref = np.empty(B)
    for n in np.unique(N):
        sel = N == n
        ref[sel] = gb.get_ll(p[:, sel], [A, E], [SA, SE], N=int(n), data_index=di[sel], noise_index=ni[sel], **kw)
    assert np.array_equal(ll, ref)  # identical arithmetic per binary
    # oracle on two binaries of every bucket
    for n in np.unique(N):
        sel = np.flatnonzero(N == n)[:2]
        o = F.run_wave(*p[:, sel], N=int(n), T=YR, dt=DT, tdi2=True)
        d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"], A, E, SA, SE, df, di[sel],
                                 ni[sel], L)
        ll_o = F.loglike(gb.d_d, d_h, h_h)
        if precision == "mixed":
            bar = 1e-6 * (np.abs(h_h) + np.abs(d_h)) + 1e-6
        else:
            bar = ll_bar(d_h, h_h, rel=1e-8 if third else 1e-9)
        assert np.all(np.abs(ll[sel] - ll_o) <= bar), (n, np.abs(ll[sel] - ll_o) / bar)


def test_uneven_batch_2d_streams_without_indices_use_stream_0():
    """2-D data without data_index: every binary reads stream 0 and a window
    past data_length is clipped there (the uniform get_ll's semantics), also
    at N = 8192 near the top of the stream."""
    from gbgpu_amd import GBGPU

    rng = np.random.default_rng(5100)
    A, E, SA, SE, df, L = _streams(rng)
    B = 16
    p, N = _case(rng, B, False)
    # two binaries whose 8192-bin windows run past the last bin
    p[1, :2] = (L - 1000 + np.array([0.0, 500.0])) * df
    N[:2] = 8192
    gb = GBGPU()
    gb.d_d = 0.0
    ll = gb.get_ll(p, [A, E], [SA, SE], N=N, T=YR, dt=DT)
    ref = np.empty(B)
    for n in np.unique(N):
        sel = N == n
        ref[sel] = gb.get_ll(p[:, sel], [A[0], E[0]], [SA[0], SE[0]], N=int(n), T=YR, dt=DT)
    assert np.array_equal(ll, ref)
    o = F.run_wave(*p[:, :2], N=8192, T=YR, dt=DT)
    d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"], A[0], E[0], SA[0], SE[0], df,
                             data_length=L)
    assert np.all(np.abs(ll[:2] - F.loglike(0.0, d_h, h_h)) <= ll_bar(d_h, h_h))


def test_guarded_mixed_batch_with_large_N_and_third_body():
    """Guarded mixed batch (routing at set-up + re-evaluation) over every
    bucket up to 8192 with third-body rows and noise_index: every logL
    within 1e-3 of the FP64 batch, and the binaries evaluated in FP64 equal
    it bit for bit."""
    from gbgpu_amd import GBGPUThirdBody
    from gbgpu_amd.batch import BatchedLogLike

    rng = np.random.default_rng(5200)
    A, E, SA, SE, df, L = _streams(rng)
    B = 96
    p, N = _case(rng, B, True)
    p[0, :20] *= 300.0  # loud: routed to FP64 at set-up or re-evaluated
    ni = rng.integers(0, 2, B).astype(np.int32)
    di = rng.integers(0, 2, B).astype(np.int32)
    gb = GBGPUThirdBody()
    gb.d_d = 0.0
    res = {}
    for prec, guard in (("fp64", None), ("mixed", 1e-3)):
        bl = BatchedLogLike(gb, [A, E], [SA, SE], YR, DT, precision=prec, guard_tol=guard, data_index=di,
                            noise_index=ni)
        res[prec] = bl(bl.kernel_params(p), N=N).cpu().numpy()
        if prec == "mixed":
            counts = bl.guard_counts()
    assert np.all(np.abs(res["mixed"] - res["fp64"]) <= 1e-3), np.abs(res["mixed"] - res["fp64"]).max()
    assert sum(counts.values()) >= 20
    same = res["mixed"] == res["fp64"]
    assert same.sum() >= 20


def test_captured_loop_with_third_body_tdi2_and_streams():
    """The hipGraph-captured sampler loop over a widened batch replays the
    eager loop bit for bit, with eager work between replays."""
    import torch

    from gbgpu_amd import GBGPUThirdBody
    from gbgpu_amd.batch import BatchedLogLike

    rng = np.random.default_rng(5300)
    A, E, SA, SE, df, L = _streams(rng)
    B = 48
    p, _ = _case(rng, B, True)
    di = rng.integers(0, 2, B).astype(np.int32)
    gb = GBGPUThirdBody()
    gb.d_d = 0.0
    bl = BatchedLogLike(gb, [A, E], [SA, SE], YR, DT, precision="fp64", guard_tol=None, tdi2=True, data_index=di,
                        noise_index=di, N_levels=(128, 512, 4096), f_edges=(2e-3, 3.5e-3))
    kp = bl.kernel_params(p)
    eager = bl.run_loop(kp, 5, sigma=1e-3, seed=9).cpu().numpy()
    assert np.all(np.isfinite(eager))
    g, out = bl.capture_loop(kp, 5, sigma=1e-3, seed=9)
    for _ in range(3):
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), eager)
        bl(kp)  # eager work between replays
        torch.cuda.synchronize()


def test_bucket_capacity_bounds_do_not_change_results():
    """bucket_capacity only selects kernels: no bounds (the strided kernel
    alone), the exact bucket sizes (one pass per workgroup), halved sizes and
    one binary per bucket (nearly every binary past its bound, taken by the
    strided kernel after the single pass) give the same logL bit for bit, in
    FP64 and in guarded mixed (whose FP64 passes take the same bounds)."""
    from gbgpu_amd import GBGPU, _capi
    from gbgpu_amd.batch import BatchedLogLike

    rng = np.random.default_rng(5400)
    A, E, SA, SE, df, L = _streams(rng)
    B = 300
    p = draw_params(B, rng, f_lo=1e-3, f_hi=4e-3)
    p[0, :40] *= 300.0  # loud: routed to FP64 or re-evaluated in the mixed batch
    N = rng.choice([64, 128, 256, 512, 1024, 2048], B).astype(np.int32)
    counts = np.bincount(np.log2(N).astype(int) - 5, minlength=_capi.NBUCKET)
    gb = GBGPU()
    gb.d_d = 0.0
    for prec, guard in (("fp64", None), ("mixed", 1e-3)):
        bl = BatchedLogLike(gb, [A[0], E[0]], [SA[0], SE[0]], YR, DT, precision=prec, guard_tol=guard)
        kp = bl.kernel_params(p)
        print(f"capacity bounds {prec}: exact", flush=True)
        ref = bl(kp, N=N).cpu().numpy()
        assert list(bl.cfg.bucket_capacity) == [int(c) for c in counts]
        assert np.all(np.isfinite(ref))
        for name, caps in (("none", np.zeros_like(counts)), ("halved", counts // 2),
                           ("one", np.minimum(counts, 1))):

            def fixed(Nn, slack=False, caps=caps):
                for k in range(_capi.NBUCKET):
                    bl.cfg.bucket_capacity[k] = int(caps[k])

            bl._set_capacity = fixed
            print(f"capacity bounds {prec}: {name}", flush=True)
            got = bl(kp, N=N).cpu().numpy()
            assert np.array_equal(got, ref), (prec, name, np.abs(got - ref).max())
