"""numpy FP64 restatement of the reference's FastGB path (the oracle).

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.  Every function cites the
reference lines it restates.  Arrays follow the reference's shapes:
parameters (B,), slow part (B, 3, N), outputs (B, 3, N) complex128.
"""

from __future__ import annotations

import numpy as np

# constants (values of lisatools.utils.constants as pinned by the notebook's
# get_amplitude/get_fdot outputs; LISA.h:12-27 for the constellation)
C_SI = 299792458.0
G_SI = 6.67408e-11
MSUN_SI = 1.98848e30
PC_SI = 3.085677581491367e16
YRSID_SI = 31558149.763545603
AU_SI = 149597870700.0
LARM = 2.5e9
EC = 0.0048241852
KAPPA = 0.0
LAMBDA = 0.0
FM = 3.168753575e-8


class AnalyticOrbits:
    """Cornish & Rubbo (2003) rigid-rotating constellation, LISA.h constants.

    Oracle-side statement of the orbit model the product tables are built
    from; pinned by closed-form checks in tests/test_constants_orbits.py.
    """

    def __init__(self, armlength=LARM, ec=EC, kappa=KAPPA, lam=LAMBDA, fm=FM, t_max=10 * YRSID_SI):
        self.armlength = armlength
        self.ec, self.kappa, self.lam, self.fm = ec, kappa, lam, fm
        self.t_base = np.array([0.0, t_max])

    def get_pos(self, t, sc):
        t = np.asarray(t, dtype=np.float64)
        sc = np.asarray(sc, dtype=np.float64)
        alpha = 2 * np.pi * self.fm * t + self.kappa
        beta = 2 * np.pi * (sc - 1) / 3 + self.lam
        sa, ca, sb, cb = np.sin(alpha), np.cos(alpha), np.sin(beta), np.cos(beta)
        x = AU_SI * ca + AU_SI * self.ec * (sa * ca * sb - (1 + sa * sa) * cb)
        y = AU_SI * sa + AU_SI * self.ec * (sa * ca * cb - (1 + ca * ca) * sb)
        z = -np.sqrt(3.0) * AU_SI * self.ec * np.cos(alpha - beta)
        return np.stack(np.broadcast_arrays(x, y, z), axis=-1)


class TableOrbits:
    """Orbits from sampled positions t (K,), x (K, 3, 3), linearly
    interpolated per spacecraft and coordinate with numpy's ``interp``: the
    oracle-side statement of an orbit file read with lisatools'
    ``linear_interp_setup`` (src/gbgpu/gbgpu.py:86-101, :390-397)."""

    def __init__(self, t, x, armlength=LARM):
        self.t = np.asarray(t, dtype=np.float64)
        self.x = np.asarray(x, dtype=np.float64)
        self.armlength = armlength
        self.t_base = self.t

    def get_pos(self, t, sc):
        t = np.asarray(t, dtype=np.float64)
        t, sc = np.broadcast_arrays(t, np.asarray(sc))
        out = np.empty(t.shape + (3,))
        for s in (1, 2, 3):
            m = sc == s
            for c in range(3):
                out[..., c][m] = np.interp(t[m], self.t, self.x[:, s - 1, c])
        return out


# ---------------------------------------------------------------------------
# utilities (src/gbgpu/utils/utility.py)
# ---------------------------------------------------------------------------
def AET(X, Y, Z):
    """src/gbgpu/utils/utility.py:24-38."""
    return (Z - X) / np.sqrt(2.0), (X - 2.0 * Y + Z) / np.sqrt(6.0), (X + Y + Z) / np.sqrt(3.0)


def get_fdot(f, Mc):
    """src/gbgpu/utils/utility.py:133-169 (chirp-mass branch), Mc in Msun."""
    Mc = Mc * MSUN_SI
    return (96.0 / 5.0) * np.pi ** (8 / 3) * (G_SI * Mc / C_SI**3) ** (5 / 3) * f ** (11 / 3)


def get_amplitude(m1, m2, f, d):
    """src/gbgpu/utils/utility.py:114-130."""
    Mc = (m1 * m2) ** (3 / 5) / (m1 + m2) ** (1 / 5) * MSUN_SI
    d = d * 1e3 * PC_SI
    return 2 * (G_SI * Mc) ** (5.0 / 3.0) / (C_SI**4 * d) * (np.pi * f) ** (2.0 / 3.0)


# ---------------------------------------------------------------------------
# third body (src/gbgpu/thirdbody.py)
# ---------------------------------------------------------------------------
def third_prepare(A2, varpi, e2, P2, T2):
    """src/gbgpu/thirdbody.py:39-80: n2 = 2 pi/(P2 yr), T2 in seconds."""
    A2, varpi, e2, P2, T2 = (np.atleast_1d(np.asarray(x, dtype=np.float64)).copy() for x in (A2, varpi, e2, P2, T2))
    n2 = 2 * np.pi / (P2 * YRSID_SI)
    T2 = T2 * YRSID_SI
    return A2, varpi, e2, n2, T2


def get_u(l, e):
    """Mikkola Kepler inverse, src/gbgpu/thirdbody.py:228-320."""
    l = np.array(l, dtype=np.float64, copy=True)
    e = np.broadcast_to(e, l.shape).astype(np.float64)
    neg = l < 0
    l = l * np.sign(l)
    over2pi = l > 2 * np.pi
    mult = np.floor(l[over2pi] / (2 * np.pi))
    l[over2pi] -= mult * 2.0 * np.pi
    overpi = l > np.pi
    l[overpi] = 2.0 * np.pi - l[overpi]
    alpha = (1.0 - e) / (4.0 * e + 0.5)
    beta = 0.5 * l / (4.0 * e + 0.5)
    z = np.sqrt(beta * beta + alpha * alpha * alpha)
    z = np.where(neg, beta - z, beta + z)
    z = np.where(z < 0, -1.0, 1.0) * np.abs(z) ** (1 / 3)
    s = z - alpha / z
    w = s - 0.078 * s**5 / (1.0 + e)
    u0 = l + e * (3.0 * w - 4.0 * w**3)
    f = u0 - e * np.sin(u0) - l
    f1 = 1.0 - e * np.cos(u0)
    f2 = u0 - f - l
    f3 = 1.0 - f1
    f4 = -f2
    f2 = f2 * 0.5
    f3 = f3 * 0.166666666666667
    f4 = f4 * 0.0416666666666667
    u1 = -f / f1
    u2 = -f / (f1 + f2 * u1)
    u3 = -f / (f1 + f2 * u2 + f3 * u2 * u2)
    u4 = -f / (f1 + f2 * u3 + f3 * u3 * u3 + f4 * u3 * u3 * u3)
    u = u0 + u4
    u[overpi] = 2.0 * np.pi - u[overpi]
    u[over2pi] = 2.0 * np.pi * mult + u[over2pi]
    u[neg] *= -1
    return u


def get_phi(t, T2, e2, n2):
    """src/gbgpu/thirdbody.py:322-350 (t is (B, 3, K))."""
    u = get_u(n2[:, None, None] * (t - T2[:, None, None]), e2[:, None, None])
    adjust = e2 > 1e-6
    if np.any(adjust):
        ea = e2[adjust]
        beta = (1.0 - np.sqrt(1.0 - ea * ea)) / ea
        u[adjust] += 2.0 * np.arctan2(
            beta[:, None, None] * np.sin(u[adjust]),
            1.0 - beta[:, None, None] * np.cos(u[adjust]),
        )
    return u


def get_vLOS(t, A2, varpi, e2, n2, T2):
    """src/gbgpu/thirdbody.py:352-380."""
    phi2 = get_phi(t, T2, e2, n2)
    return A2[:, None, None] * (np.sin(phi2 + varpi[:, None, None]) + e2[:, None, None] * np.sin(varpi[:, None, None]))


def parab_step_ET(f0, fdot, fddot, A2, varpi, e2, n2, T2, t0, t0_old):
    """Trapezoid step of the third-body phase, src/gbgpu/thirdbody.py:382-428."""
    dtt = t0 - t0_old
    fgw_old = f0[:, None, None] + fdot[:, None, None] * t0_old + 0.5 * fddot[:, None, None] * t0_old * t0_old
    fgw = f0[:, None, None] + fdot[:, None, None] * t0 + 0.5 * fddot[:, None, None] * t0 * t0
    g1 = get_vLOS(t0_old, A2, varpi, e2, n2, T2) * fgw_old
    g3 = get_vLOS(t0, A2, varpi, e2, n2, T2) * fgw
    return dtt * (g1 + g3) / 2.0 * np.pi * 2 / C_SI


# ---------------------------------------------------------------------------
# waveform (src/gbgpu/gbgpu.py:108-540)
# ---------------------------------------------------------------------------
def adjust_T(T, dt):
    """src/gbgpu/gbgpu.py:177-179."""
    N_obs = int(T / dt)
    return N_obs * dt


def run_wave(amp, f0, fdot, fddot, phi0, iota, psi, lam, beta, *third, N, T, dt, orbits=None, tdi2=False,
             hooks=None):
    """Restates GBGPUBase.run_wave (src/gbgpu/gbgpu.py:108-316).

    ``third`` is empty or (A2, varpi, e2, P2, T2) as GBGPUThirdBody takes them.
    ``hooks``: a derived waveform's (shift_frequency, add_to_argS), numpy
    callables (fi, xi) -> fi and (argS, f0, fdot, fddot, xi) -> argS on
    (B, 3, N) arrays (their extra arguments bound), applied where the
    reference's _construct_slow_part calls them (src/gbgpu/gbgpu.py:453, :505).
    Returns dict with XYZf, AETf (B, 3, N) complex128, start_inds (B,) int64, df, T.
    """
    if orbits is None:
        orbits = AnalyticOrbits()
    T = adjust_T(T, dt)
    amp, f0, fdot, fddot, phi0, iota, psi, lam, beta = (
        np.atleast_1d(np.asarray(x, dtype=np.float64)) for x in (amp, f0, fdot, fddot, phi0, iota, psi, lam, beta)
    )
    add_args = third_prepare(*third) if len(third) else ()
    theta = np.pi / 2 - beta
    df = 1 / T
    cosiota = np.cos(iota)
    L = orbits.armlength
    fstar = C_SI / (L * 2 * np.pi)
    cosps, sinps = np.cos(2.0 * psi), np.sin(2.0 * psi)
    Aplus = amp * (1.0 + cosiota * cosiota)
    Across = -2.0 * amp * cosiota
    DP = Aplus * cosps - 1.0j * Across * sinps
    DC = -Aplus * sinps - 1.0j * Across * cosps
    sinth, costh = np.sin(theta), np.cos(theta)
    sinph, cosph = np.sin(lam), np.cos(lam)
    u = np.array([costh * cosph, costh * sinph, -sinth]).T  # (B,3)
    v = np.array([sinph, -cosph, np.zeros_like(cosph)]).T
    k = np.array([-sinth * cosph, -sinth * sinph, -costh]).T
    eplus = v[:, :, None] * v[:, None, :] - u[:, :, None] * u[:, None, :]
    ecross = u[:, :, None] * v[:, None, :] + v[:, :, None] * u[:, None, :]
    tm = np.linspace(0, T, num=N, endpoint=False)
    P = [np.asarray(orbits.get_pos(tm, np.full(N, s, dtype=np.int32)), dtype=np.float64) for s in (1, 2, 3)]

    Gs, q = _construct_slow_part(T, L, P, tm, f0, fdot, fddot, fstar, phi0, k, DP, DC, eplus, ecross, *add_args,
                                 hooks=hooks)
    XYZf, f_min = _computeXYZ(T, Gs, f0, fdot, fddot, fstar, amp, q, tm)
    start_inds = np.round(f_min / df).astype(np.int64)
    fctr = 0.5 * T / N
    if tdi2:
        omegaL = 2 * np.pi * f0 * (L / C_SI)
        fctr = fctr * (2.0j * np.sin(2 * omegaL) * np.exp(-2j * omegaL))
        XYZf = XYZf * fctr[:, None, None]
    else:
        XYZf = XYZf * fctr
    AETf = np.stack(AET(XYZf[:, 0], XYZf[:, 1], XYZf[:, 2]), axis=1)
    return dict(XYZf=XYZf, AETf=AETf, start_inds=start_inds, df=df, T=T, N=N)


def _construct_slow_part(T, L, P, tm, f0, fdot, fddot, fstar, phi0, k, DP, DC, eplus, ecross, *add_args, hooks=None):
    """src/gbgpu/gbgpu.py:399-540."""
    P1, P2, P3 = P
    r = {}
    r["12"] = (P2 - P1) / L
    r["13"] = (P3 - P1) / L
    r["23"] = (P3 - P2) / L
    r["31"] = -r["13"]
    kdotr = {}
    for ij in ["12", "13", "23"]:
        kdotr[ij] = k @ r[ij].T  # (B, N)
        kdotr[ij[-1] + ij[0]] = -kdotr[ij]
    kdotP = np.stack([k @ P1.T, k @ P2.T, k @ P3.T], axis=1) / C_SI  # (B,3,N)
    xi = tm - kdotP
    fi = f0[:, None, None] + fdot[:, None, None] * xi + 1 / 2.0 * fddot[:, None, None] * xi**2
    if add_args:
        fi = fi * (1 + get_vLOS(xi, *add_args) / C_SI)  # thirdbody.py:179
    if hooks is not None:
        fi = hooks[0](fi, xi)  # gbgpu.py:453
    fonfs = fi / fstar
    q = np.rint(f0 * T)
    dfq = 2.0 * np.pi * (q / T)
    om = 2.0 * np.pi * f0
    A = {}
    for ij in ["12", "23", "31"]:
        rr = r[ij]  # (N,3)
        ep = np.einsum("bij,nj->bin", eplus, rr)
        ec = np.einsum("bij,nj->bin", ecross, rr)
        A[ij] = (ep * rr.T[None] * DP[:, None, None] + ec * rr.T[None] * DC[:, None, None]).sum(axis=1)
    argS = (
        phi0[:, None, None]
        + (om[:, None, None] - dfq[:, None, None]) * tm[None, None, :]
        + np.pi * fdot[:, None, None] * (xi**2)
        + 1 / 3 * np.pi * fddot[:, None, None] * (xi**3)
    )
    if add_args:
        # thirdbody.py:182-226: cumulative trapezoid of v_LOS * f_GW
        A2, varpi, e2, n2, T2 = add_args
        term = np.zeros_like(xi)
        term[:, :, 1:] = np.cumsum(parab_step_ET(f0, fdot, fddot, A2, varpi, e2, n2, T2, xi[:, :, 1:], xi[:, :, :-1]), axis=-1)
        argS = argS + term
    if hooks is not None:
        argS = hooks[1](argS, f0, fdot, fddot, xi)  # gbgpu.py:505
    kdotP = om[:, None, None] * kdotP - argS
    Gs = {}
    for ij, ij_sym, s in [("12", "12", 0), ("23", "23", 1), ("31", "31", 2), ("21", "12", 1), ("32", "23", 2), ("13", "31", 0)]:
        arg_ij = 0.5 * fonfs[:, s, :] * (1 + kdotr[ij])
        Gs[ij] = 0.25 * np.sin(arg_ij) / arg_ij * np.exp(-1.0j * (arg_ij + kdotP[:, s])) * A[ij_sym]
    return Gs, q


def _computeXYZ(T, Gs, f0, fdot, fddot, fstar, ampl, q, tm):
    """src/gbgpu/gbgpu.py:348-388."""
    f = f0[:, None] + fdot[:, None] * tm[None, :] + 1 / 2 * fddot[:, None] * tm[None, :] ** 2
    omL = f / fstar
    SomL = np.sin(omL)
    fctr = np.exp(-1.0j * omL)
    fctr2 = 4.0 * omL * SomL * fctr / ampl[:, None]
    Xsl = Gs["21"] - Gs["31"] + (Gs["12"] - Gs["13"]) * fctr
    Ysl = Gs["32"] - Gs["12"] + (Gs["23"] - Gs["21"]) * fctr
    Zsl = Gs["13"] - Gs["23"] + (Gs["31"] - Gs["32"]) * fctr
    XYZsl = fctr2[:, None, :] * np.stack([Xsl, Ysl, Zsl], axis=1)
    XYZf_slow = ampl[:, None, None] * np.fft.fft(XYZsl, axis=-1)
    M = XYZf_slow.shape[2]
    XYZf = np.fft.fftshift(XYZf_slow, axes=-1)
    fmin = (q - M / 2) / T
    return XYZf, fmin


def freqs(start_inds, N, df):
    """src/gbgpu/gbgpu.py:572-578."""
    return (np.arange(N)[None, :] + start_inds[:, None]) * df


# ---------------------------------------------------------------------------
# likelihood / global template (src/gbgpu/cutils/gbgpu_utils.cu, gbgpu.py)
# ---------------------------------------------------------------------------
def inner_terms(A, E, start, dataA, dataE, psdA, psdE, df, data_index=None, noise_index=None, data_length=None):
    """Brute-force d_h, h_h (gbgpu_utils.cu:125-185) with (B, N) templates.

    Data and PSD are flat arrays of S streams of ``data_length`` bins.  Bins
    outside [0, data_length) contribute nothing (the reference reads out of
    bounds there).
    """
    B, N = A.shape
    if data_length is None:
        data_length = dataA.shape[-1]
    dataA, dataE = dataA.reshape(-1), dataE.reshape(-1)
    psdA, psdE = psdA.reshape(-1), psdE.reshape(-1)
    di = np.zeros(B, np.int64) if data_index is None else np.asarray(data_index, np.int64)
    ni = np.zeros(B, np.int64) if noise_index is None else np.asarray(noise_index, np.int64)
    j = np.asarray(start, np.int64)[:, None] + np.arange(N)[None, :]
    valid = (j >= 0) & (j < data_length)
    jc = np.clip(j, 0, data_length - 1)
    dA = np.where(valid, dataA[di[:, None] * data_length + jc], 0)
    dE = np.where(valid, dataE[di[:, None] * data_length + jc], 0)
    SA = psdA[ni[:, None] * data_length + jc]
    SE = psdE[ni[:, None] * data_length + jc]
    Av = np.where(valid, A, 0)
    Ev = np.where(valid, E, 0)
    d_h = 4.0 * df * (np.sum(np.conj(dA) * Av / SA, axis=1) + np.sum(np.conj(dE) * Ev / SE, axis=1))
    h_h = 4.0 * df * (np.sum(np.conj(Av) * Av / SA, axis=1) + np.sum(np.conj(Ev) * Ev / SE, axis=1))
    return d_h, h_h


def loglike(d_d, d_h, h_h, phase_marginalize=False):
    """src/gbgpu/gbgpu.py:714-728."""
    if phase_marginalize:
        d_h = np.abs(d_h)
    return -1.0 / 2.0 * (d_d + h_h - 2 * d_h).real


def fill_global(num_groups, data_length, A, E, start, group_index):
    """Serial scatter in binary order (gbgpu_utils.cu:55-101), (B, N) templates."""
    out = np.zeros((num_groups, 2, data_length), dtype=np.complex128)
    B, N = A.shape
    for b in range(B):
        g = int(group_index[b])
        s = int(start[b])
        lo, hi = max(0, s), min(data_length, s + N)
        if hi <= lo:
            continue
        out[g, 0, lo:hi] += A[b, lo - s : hi - s]
        out[g, 1, lo:hi] += E[b, lo - s : hi - s]
    return out


def inner(x, y, start_x, start_y, psdA, psdE, df, data_length):
    """<x|y> = 4 df sum conj(x) y / S over A and E on overlapping windows.

    x, y: tuples (A, E) of 1-D windows starting at start_x, start_y.
    """
    out = 0j
    for c, psd in ((0, psdA), (1, psdE)):
        xa = np.zeros(data_length, np.complex128)
        ya = np.zeros(data_length, np.complex128)
        for arr, st, dst in ((x[c], start_x, xa), (y[c], start_y, ya)):
            n = len(arr)
            lo, hi = max(0, st), min(data_length, st + n)
            if hi > lo:
                dst[lo:hi] = arr[lo - st : hi - st]
        out += 4.0 * df * np.sum(np.conj(xa) * ya / psd)
    return out


def swap_ll_diff_bruteforce(resA, resE, psdA, psdE, df, hr, sr, ha, sa):
    """Delta logL of replacing h_r by h_a against residual r (brute force).

    logL(x) = -<x|x>/2; returns logL(r + h_r - h_a) - logL(r) by forming the
    new residual explicitly.
    """
    L = len(resA)
    newA, newE = resA.copy(), resE.copy()
    for (h, s, sign) in ((hr, sr, +1.0), (ha, sa, -1.0)):
        n = len(h[0])
        lo, hi = max(0, s), min(L, s + n)
        if hi > lo:
            newA[lo:hi] += sign * h[0][lo - s : hi - s]
            newE[lo:hi] += sign * h[1][lo - s : hi - s]

    def ll(a, e):
        return -0.5 * (4 * df * (np.sum(np.abs(a) ** 2 / psdA) + np.sum(np.abs(e) ** 2 / psdE)))

    return ll(newA, newE) - ll(resA, resE)


# ---------------------------------------------------------------------------
# information matrix (src/gbgpu/gbgpu.py:948-1138)
# ---------------------------------------------------------------------------
def information_matrix(params, psd_func, eps=1e-9, inds=None, N=1024, T=None, dt=None, easy_central_difference=False, orbits=None):
    """5-point (or central) multiplicative-step FD Fisher, gbgpu.py:1052-1132."""
    params = np.atleast_2d(np.asarray(params, dtype=np.float64))
    num_params, num_bins = params.shape
    if inds is None:
        inds = np.arange(num_params)
    nd = len(inds)
    dh = np.zeros((num_bins, nd, 2, N), np.complex128)

    def wave(p):
        o = run_wave(*p, N=N, T=T, dt=dt, orbits=orbits)
        return o, o["AETf"][:, :2]

    for i, ind in enumerate(inds):
        def step(m):
            p = np.ones_like(params)
            p[ind] += m * eps
            return p * params

        _, up1 = wave(step(1))
        _, dn1 = wave(step(-1))
        if easy_central_difference:
            dh[:, i] = (up1 - dn1) / (2 * eps)
        else:
            _, up2 = wave(step(2))
            o, dn2 = wave(step(-2))
            dh[:, i] = (-up2 + dn2 + 8 * (up1 - dn1)) / (12 * eps)
    o = run_wave(*params, N=N, T=T, dt=dt, orbits=orbits)
    fr = freqs(o["start_inds"], N, o["df"])
    psd = psd_func(fr)
    info = np.zeros((num_bins, nd, nd))
    for i in range(nd):
        for j in range(i, nd):
            ip = 4 * o["df"] * np.sum((dh[:, i].conj() * dh[:, j]).real / psd[:, None, :], axis=(1, 2))
            info[:, i, j] = ip
            info[:, j, i] = ip
    return info
