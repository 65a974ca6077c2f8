"""Constants (notebook known answers) and the analytic orbit model (closed forms)."""

import numpy as np
import pytest

from gbgpu_amd.orbits import EqualArmlengthOrbits, orbit_table
from gbgpu_amd.utils import utility as U
from oracle import fastgb as F

YR = F.YRSID_SI


def test_notebook_get_amplitude_and_fdot():
    # examples/GBGPU_tutorial.ipynb cells 30 and 32 (printed to 9 digits)
    amp = U.get_amplitude(np.full(3, 0.2), np.full(3, 0.1), np.full(3, 0.1), np.full(3, 6.0))
    fdot = U.get_fdot(np.full(3, 0.001), m1=np.full(3, 0.2), m2=np.full(3, 0.1))
    assert np.allclose(amp, 6.37562267e-23, rtol=2e-9, atol=0)
    assert np.allclose(fdot, 1.73157515e-19, rtol=2e-9, atol=0)
    # oracle agrees
    assert np.allclose(F.get_amplitude(0.2, 0.1, 0.1, 6.0), 6.37562267e-23, rtol=2e-9, atol=0)
    mc = U.get_chirp_mass(0.2, 0.1)
    assert np.allclose(F.get_fdot(0.001, mc), 1.73157515e-19, rtol=2e-9, atol=0)
    # chirp-mass inversion round trip
    assert np.allclose(U.get_chirp_mass_from_f_fdot(0.001, fdot), mc, rtol=1e-12)


def test_get_fGW_and_eta():
    f0 = np.array([1e-3, 2e-3])
    t = np.array([[0.0, 10.0], [5.0, 7.0]])
    f = U.get_fGW(f0, np.array([1e-15, 2e-15]), np.array([1e-25, 0.0]), t)
    assert np.allclose(f[0, 1], 1e-3 + 1e-14 + 0.5e-25 * 100)
    assert U.get_eta(1.0, 1.0) == 0.25


@pytest.mark.parametrize("orb_cls", [EqualArmlengthOrbits, F.AnalyticOrbits])
def test_orbit_closed_forms(orb_cls):
    orb = orb_cls()
    t = np.linspace(0, 3 * YR, 4001)
    P = [orb.get_pos(t, np.full(t.shape, s)) for s in (1, 2, 3)]
    L = orb.armlength
    # arm lengths equal L to first order in ec: |L_ij/L - 1| <= few * ec
    for i, j in ((0, 1), (0, 2), (1, 2)):
        Lij = np.linalg.norm(P[j] - P[i], axis=-1)
        assert np.abs(Lij / L - 1).max() < 5 * F.EC
    # guiding centre on the 1 AU circle, 1/fm period
    c = (P[0] + P[1] + P[2]) / 3
    assert np.allclose(np.linalg.norm(c, axis=-1), F.AU_SI, rtol=3 * F.EC**2)
    per = 1 / F.FM
    for s in (1, 2, 3):
        assert np.allclose(orb.get_pos(np.array([0.0]), s), orb.get_pos(np.array([per]), s), rtol=0, atol=1e-3)
    # constellation plane inclined 60 degrees to the ecliptic
    nrm = np.cross(P[1] - P[0], P[2] - P[0])
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    tilt = np.degrees(np.arccos(np.abs(nrm[:, 2])))
    assert np.allclose(tilt, 60.0, atol=0.5)


def test_product_and_oracle_orbits_agree():
    t = np.linspace(0, YR, 257)
    a, b = EqualArmlengthOrbits(), F.AnalyticOrbits()
    for s in (1, 2, 3):
        assert np.allclose(a.get_pos(t, s), b.get_pos(t, s), rtol=1e-14, atol=1e-3)


def test_orbit_table_layout():
    orb = EqualArmlengthOrbits()
    T, N = YR, 64
    tab = orbit_table(orb, T, N)
    assert tab.shape == (22, N) and tab.dtype == np.float64 and tab.flags.c_contiguous
    tm = np.linspace(0, T, N, endpoint=False)
    assert np.array_equal(tab[0], tm)
    P = [orb.get_pos(tm, s) for s in (1, 2, 3)]
    assert np.allclose(tab[1:4].T, (P[0] + P[1] + P[2]) / 3, rtol=1e-15, atol=0)
    assert np.allclose(tab[1:4].T + tab[7:10].T, P[1], rtol=1e-15, atol=1e-4)
    assert np.abs(tab[4:13]).max() < 2e9  # offsets ~ L / sqrt(3)
    r13 = (P[2] - P[0]) / orb.armlength
    assert np.allclose(tab[16:19].T, r13, rtol=0, atol=1e-15)


def test_orbit_table_refuses_a_stretched_constellation():
    """The kernels' phase-offset series assume |P_s - Pc| <= L/sqrt(3).
    Moving spacecraft 1 to Pc + f (P_1 - Pc) puts it (2f + 1)/3 L/sqrt(3) from
    the new centroid: f = 2.5 (2.0) is refused, f = 1.1 (1.07) is accepted
    within the 1.3 margin."""

    class Stretched(EqualArmlengthOrbits):
        def __init__(self, factor):
            super().__init__()
            self.factor = factor

        def get_pos(self, t, sc):
            P = [super(Stretched, self).get_pos(t, k) for k in (1, 2, 3)]
            Pc = (P[0] + P[1] + P[2]) / 3
            s = int(np.atleast_1d(sc)[0])
            if s == 1:
                return Pc + self.factor * (P[0] - Pc)
            return P[s - 1]

    orbit_table(Stretched(1.1), YR, 32)
    with pytest.raises(ValueError, match="constellation centre"):
        orbit_table(Stretched(2.5), YR, 32)


def test_get_N_bands_without_sensitivity():
    f0 = np.array([5e-4, 2e-3, 0.02, 0.05, 0.2])
    amp = np.full(5, 1e-30)
    N = U.get_N(amp, f0, YR, sensitivity=False)
    # 1 yr -> mult 1: 32, 64, 256, 256 (overlapping band applied last), 256
    assert list(N) == [32, 64, 256, 256, 256]
    assert list(U.get_N(amp, f0, 4 * YR, oversample=2, sensitivity=False)) == [256, 512, 2048, 2048, 2048]


def test_get_N_notebook_cell34():
    # examples/GBGPU_tutorial.ipynb cell 34: weak 1 mHz source over 1 yr -> 64
    amp = U.get_amplitude(0.2, 0.1, 0.1, 6.0)
    assert list(U.get_N(np.full(3, amp), np.full(3, 0.001), YR)) == [64, 64, 64]


def test_baseline_N_rule():
    f = np.array([0.5e-3, 1e-3, 2.9e-3, 3e-3, 5.9e-3, 6e-3, 9.9e-3, 10e-3, 20e-3])
    assert list(U.baseline_N(f)) == [128, 256, 256, 512, 512, 1024, 1024, 2048, 2048]


def test_get_N_band_table_known_answers():
    """The reference's band table at 1 yr (m = 1), boundaries included; the
    overlapping 10-300 mHz band wins from 30 to 300 mHz (it is assigned last
    there, src/gbgpu/utils/utility.py:236-242)."""
    f0 = np.array([5e-4, 1e-3, 9.99e-3, 0.01, 0.029, 0.03, 0.05, 0.0999, 0.1, 0.2, 0.299, 0.3, 0.5])
    assert list(U.band_N(f0, YR)) == [32, 64, 64, 256, 256, 256, 256, 256, 256, 256, 256, 1024, 1024]
    assert list(U.band_N(np.array([2e-3]), 3 * YR)) == [256]  # m = 4 up to 4 yr
    assert list(U.band_N(np.array([2e-3]), 5 * YR)) == [512]  # m = 8 beyond


def test_get_N_snr_rule_and_its_deviations():
    from gbgpu_amd.noise import X1TDISens

    f0 = np.array([2e-3])
    s = np.sin(f0 / U.FSTAR)
    Sm = np.sqrt(X1TDISens.get_Sn(f0)) / (4 * s * s)
    per_amp = 2 * np.sqrt(YR / Sm)  # snr_N = floor(amp * per_amp)
    # snr_N == band_N exactly: the reference returns 0 there, this rule returns N
    amp_eq = (64 + 0.5) / per_amp
    assert U.snr_N(amp_eq, f0, YR)[0] == 64 and U.band_N(f0, YR)[0] == 64
    assert U.get_N(amp_eq, f0, YR)[0] == 64
    # a loud source: snr_N = 300 -> rounded up to 512
    amp_300 = 300.5 / per_amp
    assert U.snr_N(amp_300, f0, YR)[0] == 300
    assert U.get_N(amp_300, f0, YR)[0] == 512
    # capped at the reference's 8192, then oversampled once
    assert U.get_N(1e6 / per_amp, f0, YR)[0] == 8192
    assert U.get_N(1e6 / per_amp, f0, YR, oversample=2)[0] == 16384


def test_third_body_N_oversampled_once():
    """A 0.1-yr outer orbit over 4 yr needs 2/P2 = 20 samples per year: the
    4-yr base N = 256 (2 mHz) gives 64 per year, already enough; a 0.01-yr
    orbit needs 200 per year -> 1024.  oversample = 2 doubles the result once
    (the reference's doubling branch applies it twice)."""
    from gbgpu_amd import GBGPUThirdBody

    gb = GBGPUThirdBody()
    amp, f0 = np.array([1e-30]), np.array([2e-3])
    assert gb.special_get_N(amp, f0, 4 * YR, 0.0, 0.0, 0.0, np.array([0.1]), 0.0)[0] == 256
    assert gb.special_get_N(amp, f0, 4 * YR, 0.0, 0.0, 0.0, np.array([0.01]), 0.0)[0] == 1024
    assert gb.special_get_N(amp, f0, 4 * YR, 0.0, 0.0, 0.0, np.array([0.01]), 0.0, oversample=2)[0] == 2048


def test_unsupported_N_rejected_before_device_work():
    """The HIP kernels take powers of two in [32, 8192] (get_N's cap); other
    N (given, or chosen by get_N with oversampling) are refused with a
    ValueError before anything touches a device."""
    from gbgpu_amd import GBGPU

    gb = GBGPU()
    p = [np.array([x]) for x in (1e-22, 2e-3, 1e-18, 0.0, 0.1, 0.2, 0.3, 0.4, 0.5)]
    for N in (100, 16, 6144, 16384):
        with pytest.raises(ValueError, match="powers of two in \\[32, 8192\\]"):
            gb.run_wave(*p, N=N, T=YR, dt=15.0)
    p[1] = np.array([0.35])  # f0 >= 0.3 Hz over 8 yr: band N = 1024 * 8 = 8192, oversampled twice
    with pytest.raises(ValueError, match="powers of two"):
        gb.run_wave(*p, N=None, T=8 * YR, dt=1.0, oversample=2)
