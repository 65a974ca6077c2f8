"""Direct time-domain LISA response of a plane gravitational wave, written
from the physics of the single-link Doppler response, independently of the
FastGB approximation the oracle restates.  Test infrastructure: the oracle's
frequency-domain slow part is checked against it.

Conventions (stated here so that the checks pin them):
  * the wave travels along k = -(sin th cos lam, sin th sin lam, cos th),
    th = pi/2 - beta, and reaches position x at time t with retarded time
    xi = t - k.x / c;
  * strain h_ab(xi) = Re[(DP e+_ab + DC ex_ab) exp(i Psi(xi))], with
    Psi(xi) = phi0 + 2 pi f0 xi + pi fdot xi^2 + pi/3 fddot xi^3, the
    polarisation amplitudes DP, DC and the tensors e+ = v v - u u,
    ex = u v + v u of the source frame (u, v, k);
  * a photon emitted by spacecraft s at t_e and received by spacecraft r at
    t travels along n = (P_r(t) - P_s(t_e)) / |...|, with t_e from the light
    time |P_r(t) - P_s(t_e)| / c; its fractional frequency shift is
    y(r <- s)(t) = n.[h(xi_e) - h(xi_r)].n / (2 (1 - k.n))
    (the one-arm Doppler response of Estabrook & Wahlquist 1975, in the form
    of Armstrong, Estabrook & Tinto 1999);
  * delays are by the nominal light time L/c of the constellation.
"""

from __future__ import annotations

import numpy as np

C_SI = 299792458.0


def source_frame(amp, iota, psi, lam, beta):
    """Polarisation tensor H = DP e+ + DC ex (3x3 complex) and propagation k."""
    th = np.pi / 2 - beta
    ci = np.cos(iota)
    Ap = amp * (1 + ci * ci)
    Ac = -2 * amp * ci
    DP = Ap * np.cos(2 * psi) - 1j * Ac * np.sin(2 * psi)
    DC = -Ap * np.sin(2 * psi) - 1j * Ac * np.cos(2 * psi)
    u = np.array([np.cos(th) * np.cos(lam), np.cos(th) * np.sin(lam), -np.sin(th)])
    v = np.array([np.sin(lam), -np.cos(lam), 0.0])
    k = np.array([-np.sin(th) * np.cos(lam), -np.sin(th) * np.sin(lam), -np.cos(th)])
    ep = np.outer(v, v) - np.outer(u, u)
    ec = np.outer(u, v) + np.outer(v, u)
    return DP * ep + DC * ec, k


def phase(p, xi):
    """Psi(xi) of a source p = (amp, f0, fdot, fddot, phi0, ...)."""
    _, f0, fdot, fddot, phi0 = p[:5]
    return phi0 + 2 * np.pi * f0 * xi + np.pi * fdot * xi**2 + np.pi / 3 * fddot * xi**3


def y_link(p, orbits, t, r, s):
    """y(r <- s)(t) for source p (9 parameters) and an orbit model with
    get_pos(t, sc) -> (..., 3) metres."""
    H, k = source_frame(p[0], p[5], p[6], p[7], p[8])
    Pr = orbits.get_pos(t, np.full(np.shape(t), r))
    te = t - orbits.armlength / C_SI
    for _ in range(3):  # light-time fixed point (converges to ~1e-12 s)
        Ps = orbits.get_pos(te, np.full(np.shape(t), s))
        te = t - np.linalg.norm(Pr - Ps, axis=-1) / C_SI
    Ps = orbits.get_pos(te, np.full(np.shape(t), s))
    d = Pr - Ps
    n = d / np.linalg.norm(d, axis=-1)[..., None]
    nHn = np.einsum("...i,ij,...j->...", n, H, n)
    h_e = np.real(nHn * np.exp(1j * phase(p, te - (Ps @ k) / C_SI)))
    h_r = np.real(nHn * np.exp(1j * phase(p, t - (Pr @ k) / C_SI)))
    return (h_e - h_r) / (2 * (1 - n @ k))


def tdi_x(p, orbits, t):
    """First-generation TDI X(t) at spacecraft 1 for equal, constant arms:
    X = (1 - D^2) [y(2<-1) + D y(1<-2) - y(3<-1) - D y(1<-3)], D the delay by
    L/c.  (The FastGB slow part labels its six links by receiving spacecraft
    and combines them in this order.)"""
    D = orbits.armlength / C_SI

    def arms(tt):
        return (y_link(p, orbits, tt, 2, 1) + y_link(p, orbits, tt - D, 1, 2)
                - y_link(p, orbits, tt, 3, 1) - y_link(p, orbits, tt - D, 1, 3))

    return arms(t) - arms(t - 2 * D)
