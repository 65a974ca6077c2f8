"""Utility functions with the reference's names and semantics
(src/gbgpu/utils/utility.py).  Host-side numpy; none of these is on the GPU
hot path except ``AET`` (applied in-kernel) and ``get_N`` (host N choice).
"""

from __future__ import annotations

import numpy as np

from .constants import C_SI, G_SI, MSUN_SI, PC_SI, YRSID_SI, FSTAR


def AET(X, Y, Z):
    """A, E, T from X, Y, Z (src/gbgpu/utils/utility.py:24-38)."""
    return (
        (Z - X) / np.sqrt(2.0),
        (X - 2.0 * Y + Z) / np.sqrt(6.0),
        (X + Y + Z) / np.sqrt(3.0),
    )


def get_fGW(f0, fdot, fddot, t, xp=None):
    """f(t) = f0 + fdot t + fddot t^2 / 2 (src/gbgpu/utils/utility.py:41-81)."""
    if xp is None:
        xp = np
    tdim = t.ndim
    dim_diff = tdim - 1
    if dim_diff > 0:
        assert (t.shape[0] == len(f0)) or (t.shape[0] == 1)
        dims_to_expand = tuple(range(1, dim_diff + 1))
        f0 = xp.expand_dims(f0, dims_to_expand)
        fdot = xp.expand_dims(fdot, dims_to_expand)
        fddot = xp.expand_dims(fddot, dims_to_expand)
    return f0 + fdot * t + 0.5 * fddot * t * t


def get_chirp_mass(m1, m2):
    """src/gbgpu/utils/utility.py:84-96."""
    return (m1 * m2) ** (3 / 5) / (m1 + m2) ** (1 / 5)


def get_eta(m1, m2):
    """src/gbgpu/utils/utility.py:99-111."""
    return (m1 * m2) / (m1 + m2) ** 2


def get_amplitude(m1, m2, f, d):
    """Strain amplitude, d in kpc (src/gbgpu/utils/utility.py:114-130)."""
    Mc = get_chirp_mass(m1, m2) * MSUN_SI
    d = d * 1e3 * PC_SI
    return 2 * (G_SI * Mc) ** (5.0 / 3.0) / (C_SI**4 * d) * (np.pi * f) ** (2.0 / 3.0)


def get_fdot(f, m1=None, m2=None, Mc=None):
    """Quadrupole fdot (src/gbgpu/utils/utility.py:133-169)."""
    if m1 is None and m2 is None and Mc is None:
        raise ValueError("Must provide either m1 & m2 or Mc.")
    elif m1 is not None or m2 is not None:
        assert m1 is not None and m2 is not None
        Mc = get_chirp_mass(m1, m2) * MSUN_SI
    else:
        Mc = Mc * MSUN_SI
    return (96.0 / 5.0) * np.pi ** (8 / 3) * (G_SI * Mc / C_SI**3) ** (5 / 3) * f ** (11 / 3)


def get_chirp_mass_from_f_fdot(f, fdot):
    """Inverse of get_fdot (src/gbgpu/utils/utility.py:172-194)."""
    Mc_SI = (5.0 / 96.0 * np.pi ** (-8 / 3) * (G_SI / C_SI**3) ** (-5 / 3) * f ** (-11 / 3) * fdot) ** (3 / 5)
    return Mc_SI / MSUN_SI


def band_N(f0, Tobs):
    """Sample count from frequency and observation time alone: the band
    table of src/gbgpu/utils/utility.py:217-242, which sets N per binary as
    32 m, 64 m for 1 <= f0 < 10 mHz, 256 m for 10 mHz <= f0 < 300 mHz,
    512 m for 30 <= f0 < 100 mHz and 1024 m for f0 >= 100 mHz, with
    m = 1, 2, 4 or 8 for Tobs up to 1, 2, 4 or more years.  The 256 m band
    overlaps the 512 m and 1024 m ones; the reference assigns it after them,
    so it wins from 30 to 300 mHz, and the same precedence applies here."""
    f0 = np.atleast_1d(np.asarray(f0, dtype=np.float64))
    yrs = Tobs / YRSID_SI
    m = 1 if yrs <= 1.0 else 2 if yrs <= 2.0 else 4 if yrs <= 4.0 else 8
    n = np.select(
        [(f0 >= 0.001) & (f0 < 0.01), (f0 >= 0.01) & (f0 < 0.3), (f0 >= 0.03) & (f0 < 0.1), f0 >= 0.1],
        [64, 256, 512, 1024],
        default=32,
    )
    return (n * m).astype(np.int64)


def snr_N(amp, f0, Tobs):
    """Sample count a loud source needs: 2 A sqrt(T / S_m(f0)), with
    S_m = sqrt(S_X) / (4 sin^2(f0/f*)) the Michelson-response noise of
    src/gbgpu/utils/utility.py:246-259 (S_X: this package's analytic X-channel
    sensitivity, gbgpu_amd.noise.X1TDISens, in place of lisatools').  Rounded
    down to an integer as the reference rounds it."""
    from ..noise import X1TDISens

    amp = np.atleast_1d(np.asarray(amp, dtype=np.float64))
    f0 = np.atleast_1d(np.asarray(f0, dtype=np.float64))
    s = np.sin(f0 / FSTAR)
    Sm = np.sqrt(X1TDISens.get_Sn(f0)) / (4.0 * s * s)
    return np.floor(2.0 * amp * np.sqrt(Tobs / Sm)).astype(np.int64)


N_MAX_REFERENCE = 8192


def get_N(amp, f0, Tobs, oversample=1, sensitivity=True):
    """Slow-part sample count per binary (the job of
    src/gbgpu/utils/utility.py:197-273), as this package's own rule:

        N = next_pow2(min(max(band_N, snr_N), 8192)) * oversample

    band_N is the reference's frequency/observation-time table (band_N
    above); snr_N raises N for loud sources (snr_N above; skipped with
    ``sensitivity=False``).  It differs from the reference in three ways,
    each deliberate and tested (tests/test_constants_orbits.py):
      * N is rounded up to a power of two (the transforms here are radix-4/2;
        the reference's numpy FFT takes any length);
      * when snr_N equals band_N exactly the result is that N (the
        reference's ``M*(M>N) + N*(M<N)`` returns 0 there);
      * ``oversample`` multiplies once (the third-body variant of the
        reference multiplies twice when it doubles N, thirdbody.py:82-153).
    """
    n = band_N(f0, Tobs)
    if sensitivity:
        n = np.maximum(n, snr_N(amp, f0, Tobs))
    n = np.minimum(n, N_MAX_REFERENCE)
    return (next_pow2(n) * int(oversample)).astype(int)


def next_pow2(n):
    """Smallest power of two >= n (elementwise)."""
    n = np.maximum(np.asarray(n, dtype=np.int64), 1)
    return (1 << np.ceil(np.log2(n)).astype(np.int64)).astype(np.int64)


def baseline_N(f0):
    """N by frequency at Tobs = 4 yr for the uneven-batch workloads:
    128 below 1 mHz, 256 below 3 mHz, 512 below 6 mHz, 1024 below 10 mHz,
    2048 above."""
    f0 = np.asarray(f0)
    N = np.full(f0.shape, 2048, dtype=np.int64)
    N[f0 < 10e-3] = 1024
    N[f0 < 6e-3] = 512
    N[f0 < 3e-3] = 256
    N[f0 < 1e-3] = 128
    return N
