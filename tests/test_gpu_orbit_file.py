"""The HIP path with orbits read from sampled positions (an orbit file,
interpolated linearly) against the oracle run with the same samples
(numpy ``interp``): FP64 waveforms within 1e-9 relative per binary, FP64
log-likelihoods within tests.helpers.ll_bar, the third body within 1e-8,
and the uneven-N batch identical to per-call get_ll.  The samples are the
analytic orbits plus a slow 1% breathing of the constellation, so a path
that ignored the samples and used the analytic orbits fails by O(1)."""

import numpy as np
import pytest

from oracle import fastgb as F
from tests.helpers import YR, draw_params, ll_bar, make_data, psd_AE, rel_err

pytestmark = pytest.mark.gpu

DT = 15.0


def orbit_samples(span, dt=1e4):
    from gbgpu_amd import EqualArmlengthOrbits

    a = EqualArmlengthOrbits()
    t = np.arange(int(np.floor(span / dt)) + 1) * dt
    x = np.stack([a.get_pos(t, s) for s in (1, 2, 3)], axis=1)
    for s in range(3):
        ph = 2 * np.pi * t / (0.37 * YR) + 2.1 * s
        x[:, s] += 0.01 * a.armlength * np.stack([np.cos(ph), np.sin(ph), 0.5 * np.cos(0.7 * ph)], axis=-1)
    return t, x


@pytest.mark.parametrize("N", [128, 512])
def test_run_wave_with_orbit_file_matches_oracle(N):
    from gbgpu_amd import GBGPU, InterpolatedOrbits

    t, x = orbit_samples(1.1 * YR)
    rng = np.random.default_rng(900 + N)
    p = draw_params(32, rng, f_lo=3e-4, f_hi=2e-2, log_f=True)
    print(f"orbit file run_wave N={N}", flush=True)
    gb = GBGPU(orbits=InterpolatedOrbits(t, x))
    gb.run_wave(*p, N=N, T=YR, dt=DT)
    o = F.run_wave(*p, N=N, T=YR, dt=DT, orbits=F.TableOrbits(t, x))
    assert np.array_equal(gb.start_inds.cpu().numpy(), o["start_inds"])
    assert rel_err(gb.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2)).max() <= 1e-9
    assert rel_err(gb.AETf.cpu().numpy()[:, :2], o["AETf"][:, :2], axis=(1, 2)).max() <= 1e-9
    # the samples, not the analytic orbits, drive the kernels
    oa = F.run_wave(*p, N=N, T=YR, dt=DT)
    assert rel_err(o["XYZf"], oa["XYZf"], axis=(1, 2)).min() > 1e-2


def test_get_ll_and_batch_with_orbit_file_match_oracle():
    from gbgpu_amd import GBGPU, InterpolatedOrbits
    from gbgpu_amd.batch import BatchedLogLike

    T = 4 * YR
    t, x = orbit_samples(4.05 * YR)
    rng = np.random.default_rng(950)
    B = 60
    p = draw_params(B, rng)
    Tadj = F.adjust_T(T, DT)
    df = 1.0 / Tadj
    L = int(Tadj / DT) // 2 + 1
    SA, SE = psd_AE(L, df)
    A = (rng.standard_normal(L) + 1j * rng.standard_normal(L)) * np.sqrt(SA / (4 * df))
    E = (rng.standard_normal(L) + 1j * rng.standard_normal(L)) * np.sqrt(SE / (4 * df))
    orb = F.TableOrbits(t, x)
    o = F.run_wave(*p[:, :5], N=256, T=T, dt=DT, orbits=orb)
    for k in range(5):  # inject five of the sources, drawn with the file orbits
        s = int(o["start_inds"][k])
        A[s : s + 256] += o["AETf"][k, 0]
        E[s : s + 256] += o["AETf"][k, 1]
    N = np.where(np.arange(B) % 3 == 0, 512, 256).astype(np.int32)
    print("orbit file get_ll + batch, 4 yr", flush=True)
    gb = GBGPU(orbits=InterpolatedOrbits(t, x))
    gb.d_d = float(4 * df * (np.sum(np.abs(A) ** 2 / SA) + np.sum(np.abs(E) ** 2 / SE)))
    ll = gb.get_ll(p, [A, E], [SA, SE], N=N, T=T, dt=DT)
    for n in (256, 512):
        sel = N == n
        oo = F.run_wave(*p[:, sel], N=n, T=T, dt=DT, orbits=orb)
        d_h, h_h = F.inner_terms(oo["AETf"][:, 0], oo["AETf"][:, 1], oo["start_inds"], A, E, SA, SE, df)
        ll_o = F.loglike(gb.d_d, d_h, h_h)
        assert np.all(np.abs(ll[sel] - ll_o) <= ll_bar(d_h, h_h)), np.abs(ll[sel] - ll_o).max()
    bl = BatchedLogLike(gb, [A, E], [SA, SE], T, DT, precision="fp64", strategy="buckets", guard_tol=None)
    assert np.array_equal(bl(p, N=N).cpu().numpy(), ll)


def test_third_body_with_orbit_file_matches_oracle():
    from gbgpu_amd import GBGPUThirdBody, InterpolatedOrbits

    t, x = orbit_samples(1.1 * YR)
    rng = np.random.default_rng(975)
    B = 12
    p = draw_params(B, rng)
    third = np.array([rng.uniform(10, 80, B), rng.uniform(0, 2 * np.pi, B), rng.uniform(0, 0.6, B),
                      rng.uniform(1.5, 6.0, B), rng.uniform(-1, 1, B)])
    print("orbit file third body N=256", flush=True)
    g3 = GBGPUThirdBody(orbits=InterpolatedOrbits(t, x))
    g3.run_wave(*p, *third, N=256, T=YR, dt=DT)
    o = F.run_wave(*p, *third, N=256, T=YR, dt=DT, orbits=F.TableOrbits(t, x))
    assert np.array_equal(g3.start_inds.cpu().numpy(), o["start_inds"])
    assert rel_err(g3.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2)).max() <= 1e-8
