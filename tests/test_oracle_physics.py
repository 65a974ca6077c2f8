"""Pins the oracle's slow part, transform conventions, eccentric third-body
phase and information matrix against computations that share none of its
code: the direct time-domain plane-wave response (tests/td_response.py),
numpy's real FFT of a sampled time series, an independent Kepler solve and
the closed forms of linear parameters.

Each check is chosen so that a plausible slip in the oracle fails it by a
wide margin (measured once on the CPU by mutating oracle/fastgb.py):
  * light-travel phase om k.P/c with the wrong sign (fastgb.py:246):
    single-link error 1.9-2.0, TDI spectrum error 1.3;
  * TDI delay factor exp(+i wL) instead of exp(-i wL) (fastgb.py:259):
    TDI spectrum error 1.2;
  * a link evaluated at the wrong spacecraft, or the transfer-function
    argument with (1 - k.r): single-link error 0.16;
  * a wrong beta in the eccentric true anomaly (fastgb.py:127): the phase
    differs by O(e) rad.
"""

import numpy as np
import pytest

from oracle import fastgb as F
from tests import td_response as TD
from tests.helpers import YR

FSTAR = TD.C_SI / (F.LARM * 2 * np.pi)

# a chirping 8 mHz source (f/f* = 0.42, so the TDI delay and transfer
# functions are far from their long-wavelength limits); pi fdot T^2 = 6 rad
SRC = np.array([1e-21, 8e-3, 2e-15, 0.0, 0.3, 0.7, 0.4, 1.1, 0.5])
SRC[3] = 11 / 3 * SRC[2] ** 2 / SRC[1]


def _frame(p):
    """The (B = 1) source-frame arrays run_wave hands to the slow part."""
    amp, f0, fdot, fddot, phi0, iota, psi, lam, beta = [np.array([x]) for x in p]
    th = np.pi / 2 - beta
    ci = np.cos(iota)
    Ap, Ac = amp * (1 + ci * ci), -2 * amp * ci
    DP = Ap * np.cos(2 * psi) - 1j * Ac * np.sin(2 * psi)
    DC = -Ap * np.sin(2 * psi) - 1j * Ac * np.cos(2 * psi)
    u = np.array([np.cos(th) * np.cos(lam), np.cos(th) * np.sin(lam), -np.sin(th)]).T
    v = np.array([np.sin(lam), -np.cos(lam), np.zeros(1)]).T
    k = np.array([-np.sin(th) * np.cos(lam), -np.sin(th) * np.sin(lam), -np.cos(th)]).T
    ep = v[:, :, None] * v[:, None, :] - u[:, :, None] * u[:, None, :]
    ec = u[:, :, None] * v[:, None, :] + v[:, :, None] * u[:, None, :]
    return f0, fdot, fddot, phi0, k, DP, DC, ep, ec


def test_single_link_responses_match_time_domain():
    """Each of the six slow-part links G_ij (receiver i, emitter j) is the
    complex envelope of the exact one-arm response:
    y(i <- j)(t) = Re[-2i (f_i/f*) G_ij(t) exp(i Omega_q t)], f_i the
    instantaneous frequency at spacecraft i.  Checked sample by sample at
    the FastGB level (nominal arm length in r_ij, instantaneous-frequency
    transfer function): 1e-3 of the peak response."""
    orb = F.AnalyticOrbits()
    T = F.adjust_T(YR, 15.0)
    N = 256
    tm = np.arange(N) * T / N
    P = [orb.get_pos(tm, np.full(N, s)) for s in (1, 2, 3)]
    f0, fdot, fddot, phi0, k, DP, DC, ep, ec = _frame(SRC)
    Gs, q = F._construct_slow_part(T, orb.armlength, P, tm, f0, fdot, fddot, FSTAR, phi0, k, DP, DC, ep, ec)
    Omq = 2 * np.pi * q[0] / T
    for ij in ("12", "23", "31", "21", "32", "13"):
        r, s = int(ij[0]), int(ij[1])
        y = TD.y_link(SRC, orb, tm, r, s)
        xi = tm - (orb.get_pos(tm, np.full(N, r)) @ k[0]) / TD.C_SI
        f = f0[0] + fdot[0] * xi + 0.5 * fddot[0] * xi**2
        y_fastgb = np.real(-2j * (f / FSTAR) * Gs[ij][0] * np.exp(1j * Omq * tm))
        err = np.max(np.abs(y_fastgb - y)) / np.max(np.abs(y))
        assert err < 1e-3, (ij, err)


def test_tdi_x_spectrum_matches_time_domain():
    """The oracle's X spectrum against numpy's real FFT (times dt) of the
    directly evaluated TDI X(t), sampled every 30 s for a year, on the
    template's N = 512 bins start .. start + N - 1.  This fixes together
    the slow part, the TDI delays and combination, the transform's sign
    convention, the fftshift, the 0.5 T/N scale and the start bin.  The
    remaining difference is FastGB's N-point transform of a non-periodic
    envelope (it falls as N^-1/2: 2.5e-2 at N = 256, 8.8e-3 at 2048)."""
    orb = F.AnalyticOrbits()
    dt = 30.0
    T = F.adjust_T(YR, dt)
    N = 512
    o = F.run_wave(*[np.array([x]) for x in SRC], N=N, T=T, dt=dt)
    t = np.arange(int(T / dt)) * dt
    Xf = np.fft.rfft(TD.tdi_x(SRC, orb, t)) * dt
    s0 = int(o["start_inds"][0])
    ref = o["XYZf"][0, 0]
    err = np.linalg.norm(Xf[s0 : s0 + N] - ref) / np.linalg.norm(ref)
    assert err < 3e-2, err
    # the power the template misses lies outside its window: < 1e-3
    full = np.sum(np.abs(Xf) ** 2)
    assert np.sum(np.abs(Xf[s0 : s0 + N]) ** 2) > (1 - 1e-3) * full


@pytest.mark.parametrize("m0", [5, -7])
def test_transform_of_a_pure_tone(monkeypatch, m0):
    """A slow part that is one pure tone, exp(2 pi i m0 t/T) on link 21
    alone, is the real signal X(t) = Re[K exp(2 pi i (q + m0) t/T)] with
    K = 4 wL sin(wL) exp(-i wL) for fdot = fddot = 0.  numpy's real FFT of
    that signal (times dt) is K T/2 in bin q + m0 and zero elsewhere, so
    run_wave must return exactly that in X at index N/2 + m0 with
    start = q - N/2, and Y = -exp(-i wL) X (Y's slow part is -G21 e^{-iwL}),
    Z = 0.  Fixes the transform's sign convention, fftshift, the 0.5 T/N
    scale and the start offset independently of the slow part."""
    dt = 15.0
    T = F.adjust_T(YR, dt)
    N = 128
    q = 250_000
    f0 = (q + 0.25) / T  # rint(f0 T) = q

    def tone_slow_part(T_, L_, P_, tm, f0_, fdot_, fddot_, fstar, *rest, hooks=None):
        z = np.zeros((1, N), np.complex128)
        Gs = {ij: z for ij in ("12", "23", "31", "13", "32")}
        Gs["21"] = np.exp(2j * np.pi * m0 * tm / T_)[None, :]
        return Gs, np.rint(f0_ * T_)

    monkeypatch.setattr(F, "_construct_slow_part", tone_slow_part)
    p = [np.array([x]) for x in (2e-22, f0, 0.0, 0.0, 0.0, 0.3, 0.2, 0.1, 0.4)]
    o = F.run_wave(*p, N=N, T=T, dt=dt)
    wL = f0 / FSTAR
    K = 4 * wL * np.sin(wL) * np.exp(-1j * wL)
    # the real signal and its FFT (no oracle code involved)
    t = np.arange(int(T / dt)) * dt
    Xt = np.real(K * np.exp(2j * np.pi * (q + m0) * t / T))
    Xf = np.fft.rfft(Xt) * dt
    assert int(o["start_inds"][0]) == q - N // 2
    win = Xf[q - N // 2 : q + N // 2]
    assert np.abs(win[N // 2 + m0] - K * T / 2) < 1e-9 * abs(K * T / 2)
    X = o["XYZf"][0, 0]
    assert np.abs(X - win).max() < 1e-9 * abs(K * T / 2)
    assert np.abs(o["XYZf"][0, 1] + np.exp(-1j * wL) * X).max() < 1e-12 * abs(K * T / 2)
    assert np.abs(o["XYZf"][0, 2]).max() == 0.0


def _kepler_E(M, e):
    """Eccentric anomaly by bisection on E - e sin E = M (monotone in E)."""
    lo = M - 1.0 - e
    hi = M + 1.0 + e
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        f = mid - e * np.sin(mid) - M
        lo = np.where(f < 0, mid, lo)
        hi = np.where(f < 0, hi, mid)
    return 0.5 * (lo + hi)


@pytest.mark.parametrize("e", [0.05, 0.3, 0.6, 0.9])
def test_third_body_phase_is_the_true_anomaly(e):
    """get_phi returns the true anomaly of the outer orbit, tan(nu/2) =
    sqrt((1+e)/(1-e)) tan(E/2), with E from Kepler's equation at mean anomaly
    n2 (t - T2) (independent bisection here, not the Mikkola solver)."""
    n2 = np.array([2 * np.pi / (2.3 * F.YRSID_SI)])
    T2 = np.array([0.4 * F.YRSID_SI])
    t = np.linspace(-3 * F.YRSID_SI, 5 * F.YRSID_SI, 3001)[None, None, :]
    phi = F.get_phi(t, T2, np.array([e]), n2)[0, 0]
    M = n2[0] * (t[0, 0] - T2[0])
    E = _kepler_E(M, e)
    nu = 2 * np.arctan2(np.sqrt(1 + e) * np.sin(E / 2), np.sqrt(1 - e) * np.cos(E / 2))
    assert np.abs(np.angle(np.exp(1j * (phi - nu)))).max() < 1e-9


def test_information_matrix_closed_forms():
    """With multiplicative steps, d h / d ln(amp) = h and
    d h / d ln(phi0) = i phi0 h exactly, so F_amp,amp = <h|h>,
    F_phi0,phi0 = phi0^2 <h|h> and F_amp,phi0 = Re<h| i phi0 h> = 0, with
    <x|y> = 4 df sum Re(conj x y)/S over A and E (white PSD here)."""
    B = 2
    p = np.array([np.full(B, v) for v in (1e-22, 3e-3, 1e-16, 0.0, 0.8, 0.6, 0.3, 1.4, -0.2)])
    p[3] = 11 / 3 * p[2] ** 2 / p[1]
    p[:, 1] *= [1.0, 1.3, 1.0, 1.0, 1.7, 0.5, 2.0, 0.3, -2.0]
    S0 = 1e-40
    N, T, dt = 256, YR, 15.0
    info = F.information_matrix(p, lambda f: np.full_like(f, S0), eps=1e-9, inds=np.array([0, 4]), N=N, T=T, dt=dt)
    o = F.run_wave(*p, N=N, T=T, dt=dt)
    h = o["AETf"][:, :2]
    hh = 4 * o["df"] * np.sum(np.abs(h) ** 2, axis=(1, 2)) / S0
    assert np.allclose(info[:, 0, 0], hh, rtol=1e-6)
    assert np.allclose(info[:, 1, 1], p[4] ** 2 * hh, rtol=1e-5)
    assert np.all(np.abs(info[:, 0, 1]) < 1e-5 * np.sqrt(info[:, 0, 0] * info[:, 1, 1]))
