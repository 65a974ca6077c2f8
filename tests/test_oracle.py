"""Pins the FP64 oracle (oracle/fastgb.py) with checks a plausible mistake
fails: invariants of the FastGB response, closed forms, a library DFT, brute
force sums, and the reference's own compiled native loops (oracle/_ref)."""

import ctypes
from pathlib import Path

import numpy as np
import pytest

from oracle import fastgb as F
from tests.helpers import YR, draw_params, oracle_c, rel_err

DT = 15.0
N = 128


def _wave(p, **kw):
    kw.setdefault("N", N)
    kw.setdefault("T", YR)
    kw.setdefault("dt", DT)
    return F.run_wave(*p, **kw)


@pytest.fixture(scope="module")
def base():
    p = draw_params(6, np.random.default_rng(1))
    return p, _wave(p)


def test_reference_smoke_cases_finite():
    # the reference's own tests (tests/test_gbgpu.py, tests/test_thirdbody.py) assert finiteness
    B = 10
    p = np.array([np.full(B, v) for v in (1e-22, 2e-3, 1e-18, 0.0, 0.1, 0.2, 0.3, 0.4, 0.5)])
    o = F.run_wave(*p, N=256, T=0.95 * YR, dt=15.0)
    assert np.all(np.isfinite(o["XYZf"])) and np.all(np.isfinite(o["AETf"]))
    o3 = F.run_wave(*p, np.full(B, 19.5), np.zeros(B), np.full(B, 0.3), np.full(B, 0.6), np.zeros(B), N=256, T=YR,
                    dt=15.0)
    assert np.all(np.isfinite(o3["XYZf"]))


def test_phase_covariance(base):
    # argS carries +phi0 and G ~ exp(-i(... - argS)), so h(phi0 + d) = e^{+i d} h(phi0)
    p, o = base
    q = p.copy()
    q[4] += 0.7
    assert rel_err(_wave(q)["XYZf"], np.exp(0.7j) * o["XYZf"]) < 1e-13


def test_amplitude_linearity(base):
    p, o = base
    q = p.copy()
    q[0] *= 3.5
    assert rel_err(_wave(q)["XYZf"], 3.5 * o["XYZf"]) < 1e-14


def test_polarisation_symmetries(base):
    p, o = base
    q = p.copy()
    q[6] += np.pi / 2
    assert rel_err(_wave(q)["XYZf"], -o["XYZf"]) < 1e-13
    q = p.copy()
    q[6] += np.pi
    assert rel_err(_wave(q)["XYZf"], o["XYZf"]) < 1e-13
    q = p.copy()
    q[7] += 2 * np.pi
    assert rel_err(_wave(q)["XYZf"], o["XYZf"]) < 1e-12


def test_cyclic_spacecraft_relabelling(base):
    # rotating the constellation by 2pi/3 relabels spacecraft 1->3, 2->1, 3->2:
    # the new X is the old Y, Y -> Z, Z -> X
    p, o = base
    rot = F.AnalyticOrbits(lam=F.LAMBDA + 2 * np.pi / 3)
    r = _wave(p, orbits=rot)["XYZf"]
    assert rel_err(r[:, 0], o["XYZf"][:, 1]) < 1e-12
    assert rel_err(r[:, 1], o["XYZf"][:, 2]) < 1e-12
    assert rel_err(r[:, 2], o["XYZf"][:, 0]) < 1e-12


def test_aet_closed_form(base):
    _, o = base
    X, Y, Z = o["XYZf"][:, 0], o["XYZf"][:, 1], o["XYZf"][:, 2]
    A, E, T = o["AETf"][:, 0], o["AETf"][:, 1], o["AETf"][:, 2]
    assert np.allclose(A * np.sqrt(2), Z - X)
    assert np.allclose(E * np.sqrt(6), X - 2 * Y + Z)
    assert np.allclose(T * np.sqrt(3), X + Y + Z)


def test_start_bins_closed_form_and_rounding():
    T = F.adjust_T(YR, DT)
    ks = np.array([40000.5, 40001.5, 40002.49, 40002.51, 99999.0])
    p = draw_params(len(ks), np.random.default_rng(2))
    p[1] = ks / T
    o = _wave(p)
    assert np.array_equal(o["start_inds"], np.rint(ks).astype(np.int64) - N // 2)
    assert list(np.rint(ks).astype(int)[:2]) == [40000, 40002]  # half to even


def test_long_wavelength_limit():
    """For f << f*, the TDI response scales as f^2 at fixed strain
    (fctr2 = 4 omL sin omL ~ 4 omL^2, transfer functions -> 1), so by
    Parseval halving f0 quarters the L2 norm of X, Y, Z up to O(f/f*)."""
    p = draw_params(4, np.random.default_rng(3))
    p[2] = p[3] = 0.0
    p[1] = 1e-4
    a = np.linalg.norm(_wave(p)["XYZf"], axis=2)
    p[1] = 0.5e-4
    b = np.linalg.norm(_wave(p)["XYZf"], axis=2)
    assert np.allclose(a / b, 4.0, rtol=0.01)


def test_inner_terms_bruteforce_small():
    rng = np.random.default_rng(4)
    B, M, L = 5, 8, 30
    A = rng.standard_normal((B, M)) + 1j * rng.standard_normal((B, M))
    E = rng.standard_normal((B, M)) + 1j * rng.standard_normal((B, M))
    dA = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    dE = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    SA = rng.uniform(1, 2, L)
    SE = rng.uniform(1, 2, L)
    st = np.array([-3, 0, 10, 25, 29])
    d_h, h_h = F.inner_terms(A, E, st, dA, dE, SA, SE, 0.5)
    for b in range(B):
        dh = hh = 0
        for i in range(M):
            j = st[b] + i
            if 0 <= j < L:
                dh += np.conj(dA[j]) * A[b, i] / SA[j] + np.conj(dE[j]) * E[b, i] / SE[j]
                hh += abs(A[b, i]) ** 2 / SA[j] + abs(E[b, i]) ** 2 / SE[j]
        assert np.isclose(d_h[b], 2 * dh) and np.isclose(h_h[b], 2 * hh)


def test_c_oracle_matches_numpy():
    rng = np.random.default_rng(6)
    B, M, L = 40, 16, 300
    A = rng.standard_normal((B, M)) + 1j * rng.standard_normal((B, M))
    E = rng.standard_normal((B, M)) + 1j * rng.standard_normal((B, M))
    dA = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    dE = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    SA = rng.uniform(1, 2, L)
    SE = rng.uniform(1, 2, L)
    st = rng.integers(-20, L, B).astype(np.int32)
    At, Et = np.ascontiguousarray(A.T), np.ascontiguousarray(E.T)
    c = oracle_c()
    dh = np.zeros(B, np.complex128)
    hh = np.zeros(B, np.complex128)
    z = np.zeros(B, np.int32)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    c.oracle_get_ll(P(dh), P(hh), P(At), P(Et), P(dA), P(dE), P(SA), P(SE), 0.25, P(st), M, B, P(z), P(z), L)
    d_h, h_h = F.inner_terms(A, E, st, dA, dE, SA, SE, 0.25)
    assert np.allclose(dh, d_h, rtol=1e-13) and np.allclose(hh, h_h, rtol=1e-13)
    g = rng.integers(0, 2, B).astype(np.int32)
    gA = np.zeros((2, L), np.complex128)
    gE = np.zeros((2, L), np.complex128)
    c.oracle_fill_global(P(gA), P(gE), P(At), P(Et), P(st), M, B, P(g), L)
    ref = F.fill_global(2, L, A, E, st, g)
    assert np.allclose(gA, ref[:, 0], rtol=1e-14, atol=1e-14) and np.allclose(gE, ref[:, 1], rtol=1e-14, atol=1e-14)


REF_LIB = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libgbgpu_ref.so"


@pytest.mark.skipif(not REF_LIB.exists(), reason="reference build (oracle/_ref) not present")
def test_c_oracle_matches_reference_build():
    """cref.c against the reference's own gbgpu_utils.cu compiled for the host
    (in-range windows only: the reference reads out of bounds otherwise)."""
    rng = np.random.default_rng(8)
    B, M, L, S = 50, 32, 500, 2
    At = rng.standard_normal((M, B)) + 1j * rng.standard_normal((M, B))
    Et = rng.standard_normal((M, B)) + 1j * rng.standard_normal((M, B))
    dA = rng.standard_normal((S, L)) + 1j * rng.standard_normal((S, L))
    dE = rng.standard_normal((S, L)) + 1j * rng.standard_normal((S, L))
    SA = rng.uniform(1, 2, (S, L))
    SE = rng.uniform(1, 2, (S, L))
    st = rng.integers(0, L - M, B).astype(np.int32)
    di = rng.integers(0, S, B).astype(np.int32)
    ni = rng.integers(0, S, B).astype(np.int32)
    ref = ctypes.CDLL(str(REF_LIB))
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    ref.ref_get_ll.argtypes = [ctypes.c_void_p] * 8 + [ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    ref.ref_fill_global.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    c = oracle_c()
    out = []
    for fn in (ref.ref_get_ll, c.oracle_get_ll):
        dh = np.zeros(B, np.complex128)
        hh = np.zeros(B, np.complex128)
        fn(P(dh), P(hh), P(At), P(Et), P(dA), P(dE), P(SA), P(SE), 0.3, P(st), M, B, P(di), P(ni), L)
        out.append((dh, hh))
    assert np.allclose(out[0][0], out[1][0], rtol=1e-14) and np.allclose(out[0][1], out[1][1], rtol=1e-14)
    st2 = rng.integers(-M, L + 5, B).astype(np.int32)  # fill_global clips, so any start is fine
    g = rng.integers(0, 3, B).astype(np.int32)
    res = []
    for fn in (ref.ref_fill_global, c.oracle_fill_global):
        gA = np.zeros((3, L), np.complex128)
        gE = np.zeros((3, L), np.complex128)
        fn(P(gA), P(gE), P(At), P(Et), P(st2), M, B, P(g), L)
        res.append((gA, gE))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_swap_algebra_matches_bruteforce():
    rng = np.random.default_rng(10)
    L = 400
    r = (rng.standard_normal(L) + 1j * rng.standard_normal(L), rng.standard_normal(L) + 1j * rng.standard_normal(L))
    S = (rng.uniform(1, 2, L), rng.uniform(1, 2, L))
    df = 0.01
    hr = (rng.standard_normal(50) + 1j * rng.standard_normal(50), rng.standard_normal(50) + 1j * rng.standard_normal(50))
    ha = (rng.standard_normal(50) + 1j * rng.standard_normal(50), rng.standard_normal(50) + 1j * rng.standard_normal(50))
    sr, sa = 100, 120
    brute = F.swap_ll_diff_bruteforce(r[0], r[1], S[0], S[1], df, hr, sr, ha, sa)
    ip = lambda x, sx, y, sy: F.inner(x, y, sx, sy, S[0], S[1], df, L)
    rr = (r[0], r[1])
    alg = (ip(rr, 0, ha, sa).real - ip(rr, 0, hr, sr).real
           - 0.5 * (ip(ha, sa, ha, sa).real + ip(hr, sr, hr, sr).real - 2 * ip(hr, sr, ha, sa).real))
    assert np.isclose(alg, brute, rtol=1e-10)


def test_kepler_solver_residual():
    rng = np.random.default_rng(12)
    l = rng.uniform(-3 * np.pi, 3 * np.pi, 4000)
    e = rng.uniform(0, 0.9, 4000)
    u = F.get_u(l, e)
    assert np.abs(u - e * np.sin(u) - l).max() < 1e-12


def test_third_body_circular_closed_form():
    """e2 = 0: v_LOS = A2 sin(n2 (t - T2) + varpi) and, for fdot = fddot = 0, the
    phase integral is 2 pi f0 A2/(c n2) [cos(phi(0)) - cos(phi(t))] (trapezoid
    error O(dt^2))."""
    t = np.linspace(0, YR, 4097)[None, None, :]
    A2, varpi, e2, P2, T2 = (np.array([x]) for x in (30.0, 0.4, 0.0, 1.7, 0.2))
    A2, varpi, e2, n2, T2s = F.third_prepare(A2, varpi, e2, P2, T2)
    v = F.get_vLOS(t, A2, varpi, e2, n2, T2s)
    ph = n2[0] * (t - T2s[0]) + varpi[0]
    assert np.allclose(v, A2[0] * np.sin(ph), rtol=1e-12, atol=1e-12)
    f0 = np.array([3e-3])
    steps = F.parab_step_ET(f0, np.zeros(1), np.zeros(1), A2, varpi, e2, n2, T2s, t[:, :, 1:], t[:, :, :-1])
    integ = np.concatenate([[0.0], np.cumsum(steps[0, 0])])
    exact = 2 * np.pi * f0[0] * A2[0] / (F.C_SI * n2[0]) * (np.cos(ph[0, 0, 0]) - np.cos(ph[0, 0]))
    assert np.abs(integ - exact).max() < 1e-4 * np.abs(exact).max()


def test_information_matrix_runs_and_is_symmetric():
    from gbgpu_amd.noise import A1TDISens

    B = 2
    p = np.array([np.full(B, v) for v in (1e-22, 2e-3, 1e-14, 0.0, 0.1, 0.2, 0.3, 0.4, 0.5)])
    inds = np.array([0, 1, 2, 4, 5, 6, 7, 8])
    info = F.information_matrix(p, A1TDISens.get_Sn, eps=1e-9, inds=inds, N=256, T=0.95 * YR, dt=15.0)
    assert info.shape == (B, 8, 8)
    assert np.allclose(info, np.transpose(info, (0, 2, 1)))
    assert not np.any(np.isnan(np.linalg.pinv(info)))
