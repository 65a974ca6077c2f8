"""The oracle's derived-waveform hooks (oracle.fastgb.run_wave(hooks=...))
act where the reference's _construct_slow_part calls shift_frequency and
add_to_argS (src/gbgpu/gbgpu.py:453, :505).

The reference's own third-body waveform is built on exactly these two hooks
(src/gbgpu/thirdbody.py:155-226), so expressing the oracle's third-body
formulas as hooks on a plain binary must reproduce its built-in third-body
waveform; identity hooks must reproduce the plain one.  A hook applied at
the wrong place (e.g. the frequency shift before the light-time delay, or
argS after the om k.P/c term is subtracted) fails these."""

import numpy as np

from oracle import fastgb as F
from tests.helpers import YR, draw_params

DT = 15.0


def third_body_hooks(A2, varpi, e2, P2, T2):
    """The oracle's third-body effect as (shift_frequency, add_to_argS)."""
    args = F.third_prepare(A2, varpi, e2, P2, T2)

    def shift(fi, xi):
        return fi * (1 + F.get_vLOS(xi, *args) / F.C_SI)

    def add(argS, f0, fdot, fddot, xi):
        term = np.zeros_like(xi)
        term[:, :, 1:] = np.cumsum(F.parab_step_ET(f0, fdot, fddot, *args, xi[:, :, 1:], xi[:, :, :-1]), axis=-1)
        return argS + term

    return shift, add


def test_identity_hooks_reproduce_the_plain_waveform():
    p = draw_params(5, np.random.default_rng(1))
    o = F.run_wave(*p, N=128, T=YR, dt=DT)
    h = F.run_wave(*p, N=128, T=YR, dt=DT, hooks=(lambda fi, xi: fi, lambda a, f0, fd, fdd, xi: a))
    assert np.array_equal(o["XYZf"], h["XYZf"])


def test_third_body_as_hooks_reproduces_the_third_body_waveform():
    rng = np.random.default_rng(2)
    B = 6
    p = draw_params(B, rng)
    third = [rng.uniform(0, 100, B), rng.uniform(0, 2 * np.pi, B), np.full(B, 0.3), 10 ** rng.uniform(0, 1, B),
             rng.uniform(0, 1, B)]
    o = F.run_wave(*p, *third, N=256, T=YR, dt=DT)
    h = F.run_wave(*p, N=256, T=YR, dt=DT, hooks=third_body_hooks(*third))
    assert np.array_equal(o["start_inds"], h["start_inds"])
    np.testing.assert_allclose(h["XYZf"], o["XYZf"], rtol=0, atol=1e-14 * np.abs(o["XYZf"]).max())
    # and the effect is far above the tolerance: the Doppler phase of an
    # A2 <= 100 m/s orbit moves the waveform by a few per cent over a year
    plain = F.run_wave(*p, N=256, T=YR, dt=DT)
    assert np.abs(plain["XYZf"] - o["XYZf"]).max() > 1e-3 * np.abs(o["XYZf"]).max()
