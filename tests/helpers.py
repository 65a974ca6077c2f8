"""Seeded workload generators shared by the tests (parameter draws follow the
benchmark configurations: amp log-uniform, chirp-mass fdot, isotropic sky)."""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from oracle import fastgb as F

ROOT = Path(__file__).resolve().parent.parent
YR = F.YRSID_SI


def draw_params(B, rng, f_lo=1e-3, f_hi=4e-3, log_f=False, mc=(0.2, 1.0)):
    """(9, B) parameters: amp, f0, fdot, fddot, phi0, iota, psi, lam, beta."""
    amp = 10 ** rng.uniform(-23, -21, B)
    f0 = 10 ** rng.uniform(np.log10(f_lo), np.log10(f_hi), B) if log_f else rng.uniform(f_lo, f_hi, B)
    Mc = rng.uniform(*mc, B)
    fdot = F.get_fdot(f0, Mc)
    fddot = 11.0 / 3.0 * fdot**2 / f0
    phi0 = rng.uniform(0, 2 * np.pi, B)
    iota = np.arccos(rng.uniform(-1, 1, B))
    psi = rng.uniform(0, np.pi, B)
    lam = rng.uniform(0, 2 * np.pi, B)
    beta = np.arcsin(rng.uniform(-1, 1, B))
    return np.array([amp, f0, fdot, fddot, phi0, iota, psi, lam, beta])


def psd_AE(L, df):
    from gbgpu_amd.noise import A1TDISens

    f = np.arange(L) * df
    f[0] = df  # avoid f = 0
    S = A1TDISens.get_Sn(f)
    return S, S.copy()


def make_data(params, N, T, dt, rng, L=None, noise=True):
    """A/E data = noise + the sum of the given binaries (oracle waveforms)."""
    Tadj = F.adjust_T(T, dt)
    df = 1.0 / Tadj
    if L is None:
        L = int(Tadj / dt) // 2 + 1
    SA, SE = psd_AE(L, df)
    o = F.run_wave(*params, N=N, T=T, dt=dt)
    fill = F.fill_global(1, L, o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"], np.zeros(params.shape[1], int))
    A, E = fill[0, 0].copy(), fill[0, 1].copy()
    if noise:
        from gbgpu_amd.noise import noise_realisation

        A += noise_realisation(SA, df, rng)
        E += noise_realisation(SE, df, rng)
    return A, E, SA, SE, df, L


def oracle_c():
    lib = ctypes.CDLL(str(ROOT / "oracle" / "liboracle_c.so"))
    P = ctypes.c_void_p
    lib.oracle_get_ll.argtypes = [P] * 8 + [ctypes.c_double, P, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int]
    lib.oracle_fill_global.argtypes = [P] * 5 + [ctypes.c_int, ctypes.c_int, P, ctypes.c_int]
    return lib


def rel_err(a, b, axis=None):
    """||a - b|| / ||b|| (per binary when axis is given)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if axis is None:
        return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())
    num = np.sqrt(np.sum(np.abs(a - b) ** 2, axis=axis))
    den = np.sqrt(np.sum(np.abs(b) ** 2, axis=axis))
    return num / den


def ll_bar(d_h, h_h, rel=1e-9, atol=1e-6):
    """FP64 log-likelihood bar, absolute, on the terms a template contributes:
    logL = -(d_d + h_h - 2 Re d_h)/2, and d_d is an input, so
    |dlogL| <= (atol + rel (|h_h| + 2 |d_h|)) / 2 per binary -- rel is the
    FP64 waveform bar (a relative waveform error e moves h_h by 2 e h_h and
    d_h by e |d_h|)."""
    return 0.5 * (atol + rel * (np.abs(h_h) + 2 * np.abs(d_h)))
