"""Analytic LISA noise model: first-generation TDI X and A/E sensitivities.

The reference takes these from lisatools (``A1TDISens``/``X1TDISens``,
src/gbgpu/gbgpu.py:14,1028, src/gbgpu/utils/utility.py:249), which is not
vendored.  This module states the standard two-component model instead:
test-mass acceleration noise 3e-15 m s^-2 Hz^-1/2 with its 0.4 mHz and
8 mHz shaping, optical-metrology noise 15e-12 m Hz^-1/2 with its 2 mHz
shaping, arm length 2.5e9 m, and the equal-arm TDI-1 transfer functions
  S_X = 16 sin^2 x [S_oms + (3 + cos 2x) S_acc]
  S_A = S_E = 8 sin^2 x [2 (3 + 2 cos x + cos 2x) S_acc + (2 + cos x) S_oms]
with x = 2 pi f L / c and both noises as fractional-frequency PSDs.
It only generates synthetic data and supplies default PSDs; it is not on
the waveform/likelihood hot path.
"""

from __future__ import annotations

import numpy as np

from .utils.constants import C_SI, LARM_SI

ACC_NOISE = 3e-15  # m s^-2 Hz^-1/2
OMS_NOISE = 15e-12  # m Hz^-1/2


def _components(f, acc=ACC_NOISE, oms=OMS_NOISE):
    f = np.asarray(f, dtype=np.float64)
    Sa_a = acc**2 * (1.0 + (0.4e-3 / f) ** 2) * (1.0 + (f / 8e-3) ** 4)
    Sa_nu = Sa_a / (2.0 * np.pi * f * C_SI) ** 2
    Soms_d = oms**2 * (1.0 + (2e-3 / f) ** 4)
    Soms_nu = Soms_d * (2.0 * np.pi * f / C_SI) ** 2
    return Sa_nu, Soms_nu


class X1TDISens:
    """First-generation TDI X-channel PSD (fractional frequency, 1/Hz)."""

    @staticmethod
    def get_Sn(f, armlength=LARM_SI, **kwargs):
        Sa, So = _components(f, **kwargs)
        x = 2.0 * np.pi * np.asarray(f) * armlength / C_SI
        return 16.0 * np.sin(x) ** 2 * (So + (3.0 + np.cos(2 * x)) * Sa)


class A1TDISens:
    """First-generation TDI A (= E) channel PSD (fractional frequency, 1/Hz)."""

    @staticmethod
    def get_Sn(f, armlength=LARM_SI, **kwargs):
        Sa, So = _components(f, **kwargs)
        x = 2.0 * np.pi * np.asarray(f) * armlength / C_SI
        return 8.0 * np.sin(x) ** 2 * (2.0 * (3.0 + 2.0 * np.cos(x) + np.cos(2 * x)) * Sa + (2.0 + np.cos(x)) * So)


E1TDISens = A1TDISens


def noise_realisation(psd, df, rng):
    """Complex Gaussian noise with E|n|^2 = psd / (4 df) (one-sided PSD)."""
    sigma = np.sqrt(psd / (4.0 * df) / 2.0)
    return sigma * (rng.standard_normal(psd.shape) + 1j * rng.standard_normal(psd.shape))
