"""Galactic binaries with a third body (src/gbgpu/thirdbody.py:22-535).

The line-of-sight Doppler modulation (``shift_frequency``) and its phase
integral (``add_to_argS``) run inside the HIP waveform kernels: the Mikkola
Kepler inverse, true anomaly, v_LOS and the cumulative trapezoid over the N
samples (an LDS prefix scan) are evaluated per binary per spacecraft.  The
host-side hooks here prepare the five extra parameter rows and choose N.
"""

from __future__ import annotations

import numpy as np

from .gbgpu import GBGPUBase
from .utils.constants import YRSID_SI
from .utils.utility import get_N


class GBGPUThirdBody(GBGPUBase):
    """Extra ``run_wave`` args: ``A2, varpi, e2, P2, T2`` (A2 in m/s, P2 and T2 in years)."""

    _kernel_variant = "third"

    def prepare_additional_args(self, A2, varpi, e2, P2, T2):
        """src/gbgpu/thirdbody.py:39-80: returns (A2, varpi, e2, n2, T2[s])."""
        A2 = np.atleast_1d(np.asarray(A2, dtype=np.float64)).copy()
        varpi = np.atleast_1d(np.asarray(varpi, dtype=np.float64)).copy()
        e2 = np.atleast_1d(np.asarray(e2, dtype=np.float64)).copy()
        P2 = np.atleast_1d(np.asarray(P2, dtype=np.float64)).copy()
        T2 = np.atleast_1d(np.asarray(T2, dtype=np.float64)).copy()
        self.P2 = P2
        n2 = 2 * np.pi / (P2 * YRSID_SI)
        T2 = T2 * YRSID_SI
        return (A2, varpi, e2, n2, T2)

    def special_get_N(self, amp, f0, T, A2, varpi, e2, P2, T2, oversample=1):
        """N per binary that also resolves the third-body orbit (the job of
        src/gbgpu/thirdbody.py:82-153): the circular-binary N (get_N, without
        oversampling), doubled while the sampling rate N / (T in years) is
        below two samples per outer period (2 / P2 per year), then
        oversampled once."""
        amp = np.atleast_1d(amp)
        f0 = np.atleast_1d(f0)
        P2 = np.atleast_1d(np.asarray(P2, dtype=np.float64))
        assert len(amp) == len(f0)
        assert len(P2) == len(f0)
        N = get_N(amp, f0, T, oversample=1).astype(np.int64)
        yrs = T / YRSID_SI
        need = 2.0 / P2  # samples per year that resolve the outer orbit
        while np.any(N / yrs < need):
            N = np.where(N / yrs < need, 2 * N, N)
        return (N * int(oversample)).astype(int)

    def shift_frequency(self, fi, xi, A2, varpi, e2, n2, T2):
        """Evaluated in-kernel: f_i *= 1 + v_LOS(xi)/c (src/gbgpu/thirdbody.py:155-180)."""
        raise NotImplementedError("evaluated inside the HIP waveform kernel")

    def add_to_argS(self, argS, f0, fdot, fddot, xi, A2, varpi, e2, n2, T2):
        """Evaluated in-kernel: cumulative trapezoid of 2 pi v_LOS f_GW / c
        (src/gbgpu/thirdbody.py:182-226)."""
        raise NotImplementedError("evaluated inside the HIP waveform kernel")
