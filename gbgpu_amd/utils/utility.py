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


def get_N(amp, f0, Tobs, oversample=1, sensitivity=True):
    """Slow-part sample count (src/gbgpu/utils/utility.py:197-273).

    Same T multiplier and frequency bands as the reference (including its
    overlapping 10-300 mHz band, applied in the same order).  The SNR bump
    uses this package's analytic X-channel sensitivity
    (``gbgpu_amd.noise.X1TDISens``) in place of lisatools'; pass
    ``sensitivity=False`` to skip it.  Capped at 8192 like the reference.
    """
    amp = np.atleast_1d(amp)
    f0 = np.atleast_1d(f0)
    mult = 8
    if (Tobs / YRSID_SI) <= 1.0:
        mult = 1
    elif (Tobs / YRSID_SI) <= 2.0:
        mult = 2
    elif (Tobs / YRSID_SI) <= 4.0:
        mult = 4
    elif (Tobs / YRSID_SI) <= 8.0:
        mult = 8
    mult = np.full_like(f0, mult, dtype=np.int32)
    N = 32 * mult
    N[f0 >= 0.1] = 1024 * mult[f0 >= 0.1]
    N[(f0 >= 0.03) & (f0 < 0.1)] = 512 * mult[(f0 >= 0.03) & (f0 < 0.1)]
    N[(f0 >= 0.01) & (f0 < 0.3)] = 256 * mult[(f0 >= 0.01) & (f0 < 0.3)]
    N[(f0 >= 0.001) & (f0 < 0.01)] = 64 * mult[(f0 >= 0.001) & (f0 < 0.01)]
    if sensitivity:
        from ..noise import X1TDISens

        fonfs = f0 / FSTAR
        SnX = np.sqrt(X1TDISens.get_Sn(f0))
        Sm = SnX / (4.0 * np.sin(fonfs) * np.sin(fonfs))
        Acut = amp * np.sqrt(Tobs / Sm)
        M = (2.0 ** (np.log(Acut) / np.log(2.0) + 1.0)).astype(int)
        M = M * (M > N) + N * (M < N)
        N = M * (M > N) + N * (M < N)
    else:
        M = N
    M = np.asarray(M).copy()
    M[M > 8192] = 8192
    N = M
    return (N * oversample).astype(int)


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
