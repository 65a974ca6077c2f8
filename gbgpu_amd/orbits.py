"""Spacecraft orbits for the FastGB slow part.

The reference obtains positions from ``lisatools.detector.EqualArmlengthOrbits``
(src/gbgpu/gbgpu.py:86-101, :390-397), which interpolates an orbit file that
is not vendored.  This module supplies the analytic rigid-rotating
constellation of Cornish & Rubbo (Phys. Rev. D 67, 022001, 2003, eq. 56),
parametrised by the constants in src/gbgpu/cutils/LISA.h:12-27
(eccentricity ``ec``, ``kappa``, ``lambda``, modulation frequency ``fm``).

``InterpolatedOrbits`` takes sampled positions (an orbit file) instead and
interpolates them linearly.  Any object with the same small interface (``armlength``, ``t_base``,
``get_pos(t, sc)``, ``configure(...)``) can be passed to ``GBGPU(orbits=...)``:
the GPU kernels only read a per-(T, N) table of positions built on the host
from ``get_pos``, so the device code does not depend on the orbit model.
"""

from __future__ import annotations

import numpy as np

from .utils.constants import (
    AU_SI,
    LARM_SI,
    LISA_EC,
    LISA_FM,
    LISA_KAPPA,
    LISA_LAMBDA,
    YRSID_SI,
)


class Orbits:
    """Minimal orbit interface used by the FastGB path."""

    armlength: float
    t_base: np.ndarray

    def configure(self, linear_interp_setup: bool = False, **kwargs) -> None:
        """Kept for interface parity with lisatools' ``Orbits.configure``."""
        self.linear_interp_setup = linear_interp_setup

    def get_pos(self, t, sc):  # pragma: no cover - interface
        raise NotImplementedError


class EqualArmlengthOrbits(Orbits):
    """Analytic equal-arm (to first order in ``ec``) rigid-rotating orbits.

    Args:
        armlength: nominal arm length in metres (LISA.h ``Larm``).
        ec: orbital eccentricity; ``None`` keeps ``armlength / (2*sqrt(3)*AU)``
            which equals LISA.h's 0.0048241852 for 2.5e9 m.
        kappa, lam: initial guiding-centre phase and constellation orientation.
        fm: orbital modulation frequency in Hz.
        t_max: last time (s) for which the orbits are declared valid; ``run_wave``
            refuses observation times beyond it, as the reference does against
            its orbit file (src/gbgpu/gbgpu.py:181-184).
    """

    def __init__(
        self,
        armlength: float = LARM_SI,
        ec: float | None = LISA_EC,
        kappa: float = LISA_KAPPA,
        lam: float = LISA_LAMBDA,
        fm: float = LISA_FM,
        t_max: float = 10.0 * YRSID_SI,
    ):
        self.armlength = float(armlength)
        self.ec = float(ec) if ec is not None else armlength / (2.0 * np.sqrt(3.0) * AU_SI)
        self.kappa = float(kappa)
        self.lam = float(lam)
        self.fm = float(fm)
        self.t_base = np.array([0.0, float(t_max)])
        self.linear_interp_setup = False

    def get_pos(self, t, sc):
        """Position (m) of spacecraft ``sc`` (1, 2 or 3) at times ``t``.

        ``sc`` may be a scalar or an array broadcastable against ``t``; the
        result has shape ``t.shape + (3,)`` like lisatools' ``get_pos``.
        """
        t = np.asarray(t, dtype=np.float64)
        sc = np.asarray(sc)
        alpha = 2.0 * np.pi * self.fm * t + self.kappa
        beta = 2.0 * np.pi * (sc.astype(np.float64) - 1.0) / 3.0 + self.lam
        sa, ca = np.sin(alpha), np.cos(alpha)
        sb, cb = np.sin(beta), np.cos(beta)
        a = AU_SI
        e = self.ec
        x = a * ca + a * e * (sa * ca * sb - (1.0 + sa * sa) * cb)
        y = a * sa + a * e * (sa * ca * cb - (1.0 + ca * ca) * sb)
        z = -np.sqrt(3.0) * a * e * (ca * cb + sa * sb)
        return np.stack(np.broadcast_arrays(x, y, z), axis=-1)


class InterpolatedOrbits(Orbits):
    """Orbits given as sampled spacecraft positions, interpolated linearly.

    This is the role of the orbit files behind lisatools' ``Orbits``
    classes, which the reference configures with ``linear_interp_setup=True``
    and queries through ``get_pos(t, sc)`` (src/gbgpu/gbgpu.py:86-101,
    :390-397).  Positions between samples are the straight-line
    interpolation of the two neighbouring samples, per spacecraft and
    coordinate.

    Args:
        t: sample times in seconds, strictly increasing, shape (K,), K >= 2.
        x: positions in metres, shape (K, 3, 3): time, spacecraft 1..3, x/y/z.
        armlength: nominal arm length in metres used for the unit vectors
            r_ij = (P_j - P_i)/armlength and the TDI delays.

    On a uniform grid (every spacing within 1e-9 of the first) a position is
    found by index arithmetic; otherwise by binary search.  ``configure(
    linear_interp_setup=True, dt=...)`` resamples the samples onto a uniform
    grid of step ``dt`` first.  Times outside ``[t[0], t[-1]]`` raise
    ``ValueError``: the table is never extrapolated.
    """

    def __init__(self, t, x, armlength: float = LARM_SI):
        t = np.ascontiguousarray(t, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        if t.ndim != 1 or len(t) < 2:
            raise ValueError("t must be a 1-D array of at least two sample times")
        if x.shape != (len(t), 3, 3):
            raise ValueError(f"x must have shape (len(t), 3, 3) = ({len(t)}, 3, 3), got {x.shape}")
        if not np.all(np.diff(t) > 0):
            raise ValueError("sample times must be strictly increasing")
        if not np.all(np.isfinite(x)):
            raise ValueError("positions must be finite")
        self.armlength = float(armlength)
        self.linear_interp_setup = False
        self._set_samples(t, x)

    def _set_samples(self, t, x):
        self.t_base = t
        self._x = x
        d = np.diff(t)
        self.dt_base = float(d[0]) if np.all(np.abs(d - d[0]) <= 1e-9 * d[0]) else None

    @classmethod
    def from_file(cls, path, armlength: float | None = None):
        """Loads a ``.npz`` orbit file holding ``t`` (K,) and ``x`` (K, 3, 3)
        and, optionally, ``armlength``; an explicit ``armlength`` argument
        wins over the file's.  The file is read without unpickling."""
        with np.load(path, allow_pickle=False) as f:
            t, x = f["t"], f["x"]
            L = float(f["armlength"]) if "armlength" in f.files else LARM_SI
        return cls(t, x, armlength=L if armlength is None else armlength)

    def save(self, path) -> None:
        """Writes the samples in the layout ``from_file`` reads."""
        np.savez(path, t=self.t_base, x=self._x, armlength=np.float64(self.armlength))

    def configure(self, linear_interp_setup: bool = False, dt: float | None = None, **kwargs) -> None:
        """With ``linear_interp_setup`` and a ``dt``, resamples the positions
        onto the uniform grid t[0] + k dt (k = 0 .. floor((t[-1]-t[0])/dt))
        by the same linear interpolation; without ``dt`` the samples are
        kept as given."""
        self.linear_interp_setup = linear_interp_setup
        if linear_interp_setup and dt is not None:
            if not dt > 0:
                raise ValueError("dt must be positive")
            t0 = self.t_base[0]
            K = int(np.floor((self.t_base[-1] - t0) / dt + 1e-9)) + 1
            if K < 2:
                raise ValueError("dt is longer than the span of the samples")
            tn = t0 + np.arange(K) * float(dt)
            tn[-1] = min(tn[-1], self.t_base[-1])
            xn = np.stack([self.get_pos(tn, s) for s in (1, 2, 3)], axis=1)
            self._set_samples(tn, xn)

    def get_pos(self, t, sc):
        """Position (m) of spacecraft ``sc`` (1, 2 or 3) at times ``t``;
        ``sc`` is a scalar or an array broadcastable against ``t``.  Result
        shape ``broadcast(t, sc).shape + (3,)``."""
        t = np.asarray(t, dtype=np.float64)
        sc = np.asarray(sc)
        if not np.all(np.isin(sc, (1, 2, 3))):
            raise ValueError("spacecraft index must be 1, 2 or 3")
        t, sc = np.broadcast_arrays(t, sc)
        tb = self.t_base
        if t.size and (t.min() < tb[0] or t.max() > tb[-1]):
            raise ValueError(f"times outside the orbit samples [{tb[0]}, {tb[-1]}]")
        K = len(tb)
        if self.dt_base is not None:
            u = (t - tb[0]) / self.dt_base
            i = np.clip(np.floor(u).astype(np.int64), 0, K - 2)
            w = np.clip(u - i, 0.0, 1.0)
        else:
            i = np.clip(np.searchsorted(tb, t, side="right") - 1, 0, K - 2)
            w = (t - tb[i]) / (tb[i + 1] - tb[i])
        s = sc.astype(np.int64) - 1
        w = w[..., None]
        return self._x[i, s] * (1.0 - w) + self._x[i + 1, s] * w


def orbit_table(orbits: Orbits, T: float, N: int) -> np.ndarray:
    """Host-side per-(T, N) table read by the GPU kernels.

    Rows (structure of arrays, each of length N, float64):
      0      t_n = n*T/N                             (src/gbgpu/gbgpu.py:269)
      1..3   Pc = (P1 + P2 + P3)/3, the guiding centre
      4..12  dP_s = P_s - Pc for s = 1, 2, 3 (x, y, z) (src/gbgpu/gbgpu.py:390-397)
      13..21 r12, r13, r23 = (Pj - Pi)/armlength     (src/gbgpu/gbgpu.py:424-427)
    Spacecraft positions are carried as centre + offset so that k.P_s/c
    splits into a large common delay (~500 s, kept in FP64) and small
    per-spacecraft offsets (< 5 s).
    """
    tm = np.linspace(0.0, T, num=N, endpoint=False)
    P = [np.asarray(orbits.get_pos(tm, np.full(N, s, dtype=np.int32)), dtype=np.float64) for s in (1, 2, 3)]
    L = orbits.armlength
    r12 = (P[1] - P[0]) / L
    r13 = (P[2] - P[0]) / L
    r23 = (P[2] - P[1]) / L
    Pc = (P[0] + P[1] + P[2]) / 3.0
    # The kernels' short series for the per-spacecraft phase offsets assume
    # |P_s - Pc| <= L/sqrt(3) (an equilateral constellation); their error
    # stays below 1e-16 up to 30% beyond that.
    dmax = max(float(np.max(np.linalg.norm(p - Pc, axis=1))) for p in P)
    if dmax > 1.3 * L / np.sqrt(3.0):
        raise ValueError(
            f"spacecraft sit up to {dmax:.4g} m from the constellation centre, beyond 1.3 armlength/sqrt(3)"
        )
    rows = [tm, Pc[:, 0], Pc[:, 1], Pc[:, 2]]
    for p in P:
        d = p - Pc
        rows += [d[:, 0], d[:, 1], d[:, 2]]
    for r in (r12, r13, r23):
        rows += [r[:, 0], r[:, 1], r[:, 2]]
    return np.ascontiguousarray(np.stack(rows, axis=0))


ORBIT_TABLE_ROWS = 22
