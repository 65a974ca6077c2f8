"""GPU checks of the derivative kernels (adjoint gradient, forward-tangent
Fisher matrices) against finite differences of the FP64 HIP likelihood and
waveforms, which the parity suite pins to the oracle.

Bars (stated here, checked below):
  * gradient: |g - g_FD| <= 1e-5 * max(|g_FD|, sqrt(F_kk)) per component,
    g_FD a Richardson-extrapolated central difference (steps h, h/2, h/4) of
    gb.get_ll in FP64, with steps that move the phase by ~1e-3 rad
    ("natural" steps).  The literal 1e-7 relative / 1e-9 absolute (angles)
    steps of the differentiation workload are held to the same 1e-5, with
    the differences taken of the template terms -(h_h - 2 Re d_h)/2 (d_d set
    to 0: the gradient does not depend on it, and the rounding of a large
    d_d would swamp differences over such small steps);
  * Fisher: |F - F_FD| <= 1e-5 * sqrt(F_ii F_jj), F_FD the reference's 5-point
    stencil (eps = 1e-9, multiplicative steps) over HIP run_wave outputs,
    fddot excluded as in the reference's own test (its 1e-9 step moves the
    phase by ~1e-16 rad);
  * the Fisher kernel's gradient equals the adjoint one to 1e-9.
"""

import numpy as np
import pytest

from tests.helpers import YR, draw_params, make_data

pytestmark = pytest.mark.gpu

DT = 15.0
T1 = YR


def _gb():
    from gbgpu_amd import GBGPU

    return GBGPU()


def _safe_params(B, rng, f_lo=1e-3, f_hi=4e-3):
    p = draw_params(B, rng, f_lo, f_hi)
    # keep cos(iota), sin(beta) away from +-1 so that every stencil point is valid
    p[5] = np.arccos(np.clip(np.cos(p[5]), -0.9, 0.9))
    p[8] = np.arcsin(np.clip(np.sin(p[8]), -0.9, 0.9))
    # loud enough for derivatives well above the FP64 noise of the likelihood
    p[0] = 10 ** rng.uniform(-22, -21, B)
    return p


def _to_K(p):
    return np.array([np.log(p[0]), p[1], p[2], p[3], p[4], np.cos(p[5]), p[6], p[7], np.sin(p[8])])


def _from_K(K):
    return np.array([np.exp(K[0]), K[1], K[2], K[3], K[4], np.arccos(K[5]), K[6], K[7], np.arcsin(K[8])])


def _offset(p, rng, T):
    """A point near the injection so that every gradient component is O(1 sigma) or larger."""
    q = p.copy()
    B = p.shape[1]
    q[0] *= 1.0 + 0.05 * rng.standard_normal(B)
    q[1] += 0.3 / T * rng.standard_normal(B)
    q[2] += 0.3 / T**2 * rng.standard_normal(B)
    q[4] += 0.2 * rng.standard_normal(B)
    q[5] = np.arccos(np.clip(np.cos(q[5]) + 0.02 * rng.standard_normal(B), -0.9, 0.9))
    q[6] += 0.05 * rng.standard_normal(B)
    q[7] += 0.01 * rng.standard_normal(B)
    q[8] = np.arcsin(np.clip(np.sin(q[8]) + 0.01 * rng.standard_normal(B), -0.9, 0.9))
    return q


def _natural_steps(K, T, f0):
    B = K.shape[1]
    h = np.empty_like(K)
    h[0] = 1e-4
    h[1] = 1e-3 / (2 * np.pi * T)
    h[2] = 1e-3 / (np.pi * T**2)
    h[3] = 1e-3 / (np.pi / 3 * T**3)
    h[4] = 1e-3
    h[5] = 1e-4
    h[6] = 1e-3
    h[7] = 1e-3 / (2 * np.pi * f0 * 500.0)
    h[8] = 1e-4
    return h + 0 * np.zeros(B)


def _fd_grad(ll_of_raw, K, h, tie=False):
    """Richardson-extrapolated central differences of ll_of_raw along each K coordinate."""
    g = np.zeros((K.shape[1], 9))
    for k in range(9):
        if tie and k == 3:
            continue
        D = []
        for m in (1.0, 0.5, 0.25):
            vals = []
            for sgn in (1, -1):
                Kp = K.copy()
                Kp[k] = K[k] + sgn * m * h[k]
                if tie:
                    Kp[3] = 11.0 / 3.0 * Kp[2] ** 2 / Kp[1]
                vals.append(ll_of_raw(_from_K(Kp)))
            D.append((vals[0] - vals[1]) / (2 * m * h[k]))
        r12 = (4 * D[1] - D[0]) / 3
        r23 = (4 * D[2] - D[1]) / 3
        g[:, k] = (16 * r23 - r12) / 15
    return g


def _setup(seed, B=6, N=256, T=T1, f_lo=1e-3, f_hi=4e-3, tdi2=False):
    rng = np.random.default_rng(seed)
    p = _safe_params(B, rng, f_lo, f_hi)
    A, E, SA, SE, df, L = make_data(p, N, T, DT, rng)
    gb = _gb()
    gb.d_d = 4 * df * np.sum(np.abs(A) ** 2 / SA + np.abs(E) ** 2 / SE)
    q = _offset(p, rng, T)
    return gb, q, [A, E], [SA, SE], N


def _check_grad(g, g_fd, F, skip=()):
    scale = np.maximum(np.abs(g_fd), np.sqrt(np.abs(np.diagonal(F, axis1=1, axis2=2))))
    err = np.abs(g - g_fd) / scale
    for k in skip:
        err[:, k] = 0.0
    print("max rel err per component:", np.array2string(err.max(axis=0), precision=2))
    assert err.max() <= 1e-5, err


@pytest.mark.parametrize("N", [128, 256, 1024])
def test_adjoint_gradient_matches_fd(N):
    gb, q, data, psd, N = _setup(300 + N, N=N)
    print(f"adjoint grad vs FD N={N}", flush=True)
    kw = dict(N=N, T=T1, dt=DT)
    ll, g = gb.get_ll_grad(q, data, psd, **kw)
    ll_ref = gb.get_ll(q, data, psd, **kw)
    assert np.allclose(ll, ll_ref, rtol=1e-12, atol=1e-9)
    F = gb.fisher_matrix(q, psd, **kw) if N <= 1024 else None
    K = _to_K(q)
    g_fd = _fd_grad(lambda raw: gb.get_ll(raw, data, psd, **kw), K, _natural_steps(K, T1, q[1]))
    _check_grad(g, g_fd, F)


def test_adjoint_gradient_high_frequency_tiers():
    """f0 in 10-30 mHz (below the 33 mHz Nyquist frequency of dt = 15 s):
    series tiers 1-2 and the reduced sine/cosine path of the transfer
    derivative; N = 2048 exercises the 256-thread binary group."""
    gb, q, data, psd, N = _setup(311, B=4, N=2048, f_lo=1e-2, f_hi=3e-2)
    print("adjoint grad high-f N=2048", flush=True)
    kw = dict(N=N, T=T1, dt=DT)
    ll, g = gb.get_ll_grad(q, data, psd, **kw)
    F = gb.fisher_matrix(q, psd, N=1024, T=T1, dt=DT)  # scale only
    K = _to_K(q)
    g_fd = _fd_grad(lambda raw: gb.get_ll(raw, data, psd, **kw), K, _natural_steps(K, T1, q[1]))
    _check_grad(g, g_fd, F)


def test_adjoint_gradient_phase_marginalized():
    gb, q, data, psd, N = _setup(320)
    print("adjoint grad phase-marginalised", flush=True)
    kw = dict(N=N, T=T1, dt=DT)
    ll, g = gb.get_ll_grad(q, data, psd, phase_marginalize=True, **kw)
    assert np.allclose(ll, gb.get_ll(q, data, psd, phase_marginalize=True, **kw), rtol=1e-12, atol=1e-9)
    F = gb.fisher_matrix(q, psd, **kw)
    K = _to_K(q)
    g_fd = _fd_grad(lambda raw: gb.get_ll(raw, data, psd, phase_marginalize=True, **kw), K,
                    _natural_steps(K, T1, q[1]))
    _check_grad(g, g_fd, F)
    # the marginalised likelihood does not depend on phi0
    assert np.all(np.abs(g[:, 4]) <= 1e-8 * np.sqrt(F[:, 4, 4]))


def test_adjoint_gradient_tdi2():
    gb, q, data, psd, N = _setup(330)
    print("adjoint grad TDI2", flush=True)
    kw = dict(N=N, T=T1, dt=DT, tdi2=True)
    ll, g = gb.get_ll_grad(q, data, psd, **kw)
    F = gb.fisher_matrix(q, psd, **kw)
    K = _to_K(q)
    g_fd = _fd_grad(lambda raw: gb.get_ll(raw, data, psd, **kw), K, _natural_steps(K, T1, q[1]))
    _check_grad(g, g_fd, F)


def test_tied_gradient_autograd():
    """torch.autograd through SamplingLogLike (fddot tied) against
    Richardson-extrapolated central differences.

    Natural steps and the literal workload steps (1e-7 relative; 1e-9
    absolute for phi0, psi, lambda; 1e-7 absolute for cos iota, sin beta) are
    both held to the 1e-5 bar."""
    import torch

    from gbgpu_amd.derivatives import FastGBLikelihood, raw_to_sampling

    gb, q, data, psd, N = _setup(340, N=256)
    q[3] = 11.0 / 3.0 * q[2] ** 2 / q[1]
    print("autograd tied gradient", flush=True)
    like = FastGBLikelihood(gb, data, psd, N=N, T=T1, dt=DT)
    theta = raw_to_sampling(torch.from_numpy(q).cuda()).clone().requires_grad_(True)
    ll = like.sampling(theta)
    ll.sum().backward()
    g8 = theta.grad.T.cpu().numpy()  # (B, 8)
    kw = dict(N=N, T=T1, dt=DT)
    K = _to_K(q)
    F = gb.fisher_matrix(q, psd, basis="sampling", tie_fddot=True, **kw)
    cols = [0, 1, 2, 4, 5, 6, 7, 8]
    ll_of = lambda raw: gb.get_ll(raw, data, psd, **kw)  # noqa: E731
    g_nat = _fd_grad(ll_of, K, _natural_steps(K, T1, q[1]), tie=True)[:, cols]
    _check_grad(g8, g_nat, F)
    h = np.empty_like(K)
    h[0] = 1e-7
    h[1] = 1e-7 * np.abs(K[1])
    h[2] = 1e-7 * np.abs(K[2])
    h[3] = 0.0
    h[4] = h[6] = h[7] = 1e-9
    h[5] = h[8] = 1e-7
    # the gradient does not depend on d_d; the FP64 rounding of
    # d_d + h_h - 2 d_h does, and over these small steps it dominates the
    # difference, so it is taken of the template terms alone (d_d = 0)
    d_d = gb.d_d
    gb.d_d = 0.0
    g_lit = _fd_grad(ll_of, K, h, tie=True)[:, cols]
    gb.d_d = d_d
    scale = np.maximum(np.abs(g_lit), np.sqrt(np.diagonal(F, axis1=1, axis2=2)))
    err = np.abs(g8 - g_lit) / scale
    print("literal-step FD max rel err per component:", np.array2string(err.max(axis=0), precision=2))
    assert err.max() <= 1e-5
    # raw-basis autograd: chain rule of the same kernel output
    pr = torch.from_numpy(q).cuda().requires_grad_(True)
    like(pr).sum().backward()
    _, g_raw = gb.get_ll_grad(q, data, psd, basis="raw", **kw)
    assert np.allclose(pr.grad.T.cpu().numpy(), g_raw, rtol=1e-12, atol=0)


@pytest.mark.parametrize("N", [256, 1024])
def test_fisher_matches_reference_stencil(N):
    rng = np.random.default_rng(400 + N)
    p = _safe_params(5, rng)
    # the multiplicative stencil steps x by x*eps: keep the angles away from 0,
    # where its absolute step (and so its resolution) vanishes
    p[4] = rng.uniform(0.5, 2 * np.pi, 5)
    p[6] = rng.uniform(0.5, np.pi - 0.5, 5)
    p[7] = rng.uniform(0.5, 2 * np.pi, 5)
    p[8] = np.arcsin(rng.choice([-1.0, 1.0], 5) * rng.uniform(0.2, 0.9, 5))
    gb = _gb()
    print(f"Fisher exact vs FD N={N}", flush=True)
    inds = np.array([0, 1, 2, 4, 5, 6, 7, 8])
    kw = dict(N=N, T=T1, dt=DT)
    F_ex = gb.information_matrix(p, inds=inds, method="exact", **kw)
    F_fd = gb.information_matrix(p, inds=inds, method="fd", **kw)
    d = np.sqrt(np.abs(np.diagonal(F_fd, axis1=1, axis2=2)))
    err = np.abs(F_ex - F_fd) / (d[:, :, None] * d[:, None, :])
    print("max normalised Fisher err:", err.max())
    assert err.max() <= 1e-5
    assert np.allclose(F_ex, np.transpose(F_ex, (0, 2, 1)), rtol=1e-12, atol=0)
    assert np.all(np.linalg.eigvalsh(F_ex) > -1e-8 * np.abs(F_ex).max())


def test_fisher_transforms_and_gradient_consistency():
    """Fisher kernel gradient == adjoint gradient; parameter_transforms as the reference."""
    gb, q, data, psd, N = _setup(410)
    print("Fisher gradient consistency", flush=True)
    kw = dict(N=N, T=T1, dt=DT)
    _, g = gb.get_ll_grad(q, data, psd, **kw)
    from gbgpu_amd.derivatives import FastGBLikelihood

    like = FastGBLikelihood(gb, data, psd, N=N, T=T1, dt=DT)
    F, g2 = like.fisher(q, with_grad=True)
    F = F.cpu().numpy()
    g2 = g2.cpu().numpy()
    scale = np.maximum(np.abs(g), np.sqrt(np.diagonal(F, axis1=1, axis2=2)))
    assert np.max(np.abs(g - g2) / scale) <= 1e-9
    # transforms: x0 = ln amp mapped by exp; the multiplicative stencil then gives
    # x0 dh/dx0 = ln(amp) * h
    x = q.copy()
    x[0] = np.log(q[0])
    tr = {0: np.exp}
    inds = np.array([0, 1, 2, 4, 5, 6, 7, 8])
    # eps = 1e-7: the 1e-9 stencil's rounding noise (~1e-7 |h| per derivative)
    # reaches 1e-5 of sqrt(F_ii) for weakly constrained parameters; at 1e-7
    # the 5-point truncation stays below 1e-6 for these binaries
    F_ex = gb.information_matrix(x, eps=1e-7, parameter_transforms=tr, inds=inds, method="exact", **kw)
    F_fd = gb.information_matrix(x, eps=1e-7, parameter_transforms=tr, inds=inds, method="fd", **kw)
    d = np.sqrt(np.abs(np.diagonal(F_fd, axis1=1, axis2=2)))
    assert np.max(np.abs(F_ex - F_fd) / (d[:, :, None] * d[:, None, :])) <= 1e-5
