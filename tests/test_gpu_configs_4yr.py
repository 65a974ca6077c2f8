"""The differentiation and global-template/swap workloads at their own
observation time, 4 yr (Tobs = 126,230,400 s, dt = 15 s: 4,207,681 bins per
stream), where the frequency-derivative sensitivities grow as T^2 and T^3.

Bars:
  * gradient (N = 256): |g - g_FD| <= 1e-5 max(|g_FD|, sqrt(F_kk)) against
    Richardson-extrapolated central differences of the FP64 likelihood, both
    with steps moving the phase by ~1e-3 rad and with the literal 1e-7
    relative / 1e-9 absolute steps (differences of the template terms, d_d
    set to 0, whose rounding would otherwise swamp such small steps);
  * Fisher (N = 256): |F - F_FD| <= 1e-5 sqrt(F_ii F_jj) against the
    reference's 5-point multiplicative stencil;
  * fill: the N = 128-1024 frequency mix summed into one full-length stream
    is bit-identical on repeat, within 1e-12 (of the largest bin) of the
    serial C restatement summing the same HIP templates, and within 1e-9 of
    the oracle's templates summed by the oracle;
  * swap: within the FP64 waveform bar carried through the inner products of
    brute force logL(r + h_r - h_a) - logL(r), circular and third-body
    (circular outer orbit, P2 = 1-10 yr) proposals.
"""

import ctypes

import numpy as np
import pytest

from oracle import fastgb as F
from tests.helpers import YR, draw_params, make_data, oracle_c
from tests.test_gpu_derivatives import _check_grad, _fd_grad, _natural_steps, _safe_params, _to_K

pytestmark = pytest.mark.gpu

DT = 15.0
T4 = 4 * 31_557_600.0  # 4 Julian years, as the workloads define Tobs: 4,207,681 bins at dt = 15 s


def _data_on_gpu(seed, B, N, f_lo=1e-3, f_hi=6e-3):
    import torch

    from gbgpu_amd import GBGPU

    rng = np.random.default_rng(seed)
    p = _safe_params(B, rng, f_lo, f_hi)
    A, E, SA, SE, df, L = make_data(p, N, T4, DT, rng)
    assert L == 4_207_681
    gb = GBGPU()
    gb.d_d = 4 * df * np.sum(np.abs(A) ** 2 / SA + np.abs(E) ** 2 / SE)
    data = [torch.from_numpy(A).cuda(), torch.from_numpy(E).cuda()]
    psd = [torch.from_numpy(SA).cuda(), torch.from_numpy(SE).cuda()]
    return gb, p, data, psd, rng


def _near(p, rng):
    """A point near the injection (every gradient component O(1 sigma) or more at 4 yr)."""
    q = p.copy()
    B = p.shape[1]
    q[0] *= 1.0 + 0.05 * rng.standard_normal(B)
    q[1] += 0.3 / T4 * rng.standard_normal(B)
    q[2] += 0.3 / T4**2 * rng.standard_normal(B)
    q[3] = 11.0 / 3.0 * q[2] ** 2 / q[1]
    q[4] += 0.2 * rng.standard_normal(B)
    q[5] = np.arccos(np.clip(np.cos(q[5]) + 0.02 * rng.standard_normal(B), -0.9, 0.9))
    q[6] += 0.05 * rng.standard_normal(B)
    q[7] += 0.003 * rng.standard_normal(B)
    q[8] = np.arcsin(np.clip(np.sin(q[8]) + 0.003 * rng.standard_normal(B), -0.9, 0.9))
    return q


def test_gradient_4yr_matches_fd():
    gb, p, data, psd, rng = _data_on_gpu(3001, 6, 256)
    q = _near(p, rng)
    print("adjoint gradient 4 yr N=256", flush=True)
    kw = dict(N=256, T=T4, dt=DT)
    ll, g = gb.get_ll_grad(q, data, psd, **kw)
    assert np.allclose(ll, gb.get_ll(q, data, psd, **kw), rtol=1e-12, atol=1e-9)
    Fm = gb.fisher_matrix(q, psd, **kw)
    K = _to_K(q)
    ll_of = lambda raw: gb.get_ll(raw, data, psd, **kw)  # noqa: E731
    _check_grad(g, _fd_grad(ll_of, K, _natural_steps(K, T4, q[1])), Fm)
    # the literal steps of the workload, in its 8-parameter basis with fddot
    # tied to 11/3 fdot^2 / f0: 1e-7 relative, 1e-9 absolute for phi0, psi,
    # lambda, 1e-7 absolute for cos iota and sin beta.  The gradient does not
    # depend on d_d, but the FP64 rounding of d_d + h_h - 2 d_h does: at
    # |d_d| ~ 8e6 (4,207,681 bins per channel) its ulp, 2e-9, over a 2e-9
    # step would swamp the difference, so these differences are taken with
    # d_d = 0, i.e. of the template terms -(h_h - 2 Re d_h)/2 alone.
    gb.d_d = 0.0
    h = np.empty_like(K)
    h[0] = 1e-7
    h[1] = 1e-7 * np.abs(K[1])
    h[2] = 1e-7 * np.abs(K[2])
    h[3] = 0.0
    h[4] = h[6] = h[7] = 1e-9
    h[5] = h[8] = 1e-7
    cols = [0, 1, 2, 4, 5, 6, 7, 8]
    g_lit = _fd_grad(ll_of, K, h, tie=True)[:, cols]
    # the kernel's gradient in the tied basis, by the chain rule through fddot
    gt = g.copy()
    gt[:, 1] += g[:, 3] * (-q[3] / q[1])
    gt[:, 2] += g[:, 3] * (2 * q[3] / q[2])
    scale = np.maximum(np.abs(g_lit), np.sqrt(np.diagonal(Fm, axis1=1, axis2=2))[:, cols])
    err = np.abs(gt[:, cols] - g_lit) / scale
    print("literal-step FD max rel err per component:", np.array2string(err.max(axis=0), precision=2))
    assert err.max() <= 1e-5


def test_fisher_4yr_matches_reference_stencil():
    from gbgpu_amd import GBGPU

    rng = np.random.default_rng(3002)
    p = _safe_params(4, rng, 1e-3, 6e-3)
    p[3] = 11.0 / 3.0 * p[2] ** 2 / p[1]
    p[4] = rng.uniform(0.5, 2 * np.pi, 4)
    p[6] = rng.uniform(0.5, np.pi - 0.5, 4)
    p[7] = rng.uniform(0.5, 2 * np.pi, 4)
    p[8] = np.arcsin(rng.choice([-1.0, 1.0], 4) * rng.uniform(0.2, 0.9, 4))
    gb = GBGPU()
    print("Fisher 4 yr exact vs FD", flush=True)
    inds = np.array([0, 1, 2, 4, 5, 6, 7, 8])
    kw = dict(N=256, T=T4, dt=DT)
    F_ex = gb.information_matrix(p, inds=inds, method="exact", **kw)
    F_fd = gb.information_matrix(p, inds=inds, method="fd", **kw)
    d = np.sqrt(np.abs(np.diagonal(F_fd, axis1=1, axis2=2)))
    err = np.abs(F_ex - F_fd) / (d[:, :, None] * d[:, None, :])
    print("max normalised Fisher err:", err.max())
    assert err.max() <= 1e-5


def _global_mix(B, seed):
    """The global-template population: f0 log-uniform 0.1-10 mHz, N = 128
    below 1 mHz, 256 below 3, 512 below 6, 1024 above."""
    rng = np.random.default_rng(seed)
    p = draw_params(B, rng, f_lo=1e-4, f_hi=1e-2, log_f=True)
    N = np.select([p[1] < 1e-3, p[1] < 3e-3, p[1] < 6e-3], [128, 256, 512], 1024)
    return p, N


def test_fill_4yr_frequency_mix_full_stream():
    import torch

    from gbgpu_amd import GBGPU, _capi

    p, Nb = _global_mix(3000, 3003)
    L = 4_207_681
    gb = GBGPU()
    lib = _capi.load()
    c = oracle_c()
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    print("fill 4 yr frequency mix", flush=True)
    runs = []
    gA = np.zeros(L, np.complex128)
    gE = np.zeros(L, np.complex128)
    oA = np.zeros(L, np.complex128)
    for rep in range(2):
        tA = torch.zeros(L, dtype=torch.complex128, device="cuda")
        tE = torch.zeros(L, dtype=torch.complex128, device="cuda")
        for n in (128, 256, 512, 1024):
            sel = Nb == n
            gb.run_wave(*p[:, sel], N=n, T=T4, dt=DT)
            AET = gb.AETf
            st = gb.start_inds
            g0 = torch.zeros(int(sel.sum()), dtype=torch.int32, device="cuda")
            rc = lib.gbgpu_fill_global_rows(tA.data_ptr(), tE.data_ptr(), AET[:, 0].data_ptr(), AET[:, 1].data_ptr(),
                                            AET.stride(0), st.data_ptr(), 0, n, int(sel.sum()), g0.data_ptr(), L, 1,
                                            _capi.stream_handle())
            _capi.check(rc, "gbgpu_fill_global_rows")
            if rep == 0:
                # the serial C restatement summing the same HIP templates, (M, B)-major
                At = np.ascontiguousarray(AET[:, 0].cpu().numpy().T)
                Et = np.ascontiguousarray(AET[:, 1].cpu().numpy().T)
                s32 = st.cpu().numpy().astype(np.int32)
                c.oracle_fill_global(P(gA), P(gE), P(At), P(Et), P(s32), n, int(sel.sum()),
                                     P(np.zeros(int(sel.sum()), np.int32)), L)
                o = F.run_wave(*p[:, sel], N=n, T=T4, dt=DT)
                oA += F.fill_global(1, L, o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"],
                                    np.zeros(int(sel.sum()), int))[0, 0]
        torch.cuda.synchronize()
        runs.append((tA.cpu().numpy(), tE.cpu().numpy()))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    hA, hE = runs[0]
    scale = np.abs(gA).max()
    assert np.abs(hA - gA).max() <= 1e-12 * scale and np.abs(hE - gE).max() <= 1e-12 * scale
    # overlapping windows: many bins receive several templates
    cover = np.zeros(L + 1, np.int64)
    for n in (128, 256, 512, 1024):
        st = np.rint(p[1, Nb == n] * F.adjust_T(T4, DT)).astype(np.int64) - n // 2
        np.add.at(cover, np.clip(st, 0, L), 1)
        np.add.at(cover, np.clip(st + n, 0, L), -1)
    assert np.count_nonzero(np.cumsum(cover)[:L] >= 2) > 50_000
    assert np.linalg.norm(hA - oA) <= 1e-9 * np.linalg.norm(oA)


@pytest.mark.parametrize("variant", ["circular", "third_body"])
def test_swap_4yr_full_stream(variant):
    from gbgpu_amd import GBGPU, GBGPUThirdBody

    rng = np.random.default_rng(3004 + (variant == "third_body"))
    P, N = 40, 512
    p = draw_params(P, rng, f_lo=3e-3, f_hi=3.05e-3)  # overlapping windows
    A, E, SA, SE, df, L = make_data(p, N, T4, DT, rng)
    q = p.copy()
    q[1] *= 1 + 5e-9 * rng.standard_normal(P)
    q[0] *= 1.1
    q[4] += 0.3
    if variant == "third_body":
        zeros = np.zeros(P)
        rem = np.concatenate([p, [zeros, zeros, zeros, np.ones(P), zeros]])  # A2 = 0: the plain binary
        add = np.concatenate([q, [rng.uniform(0, 100, P), rng.uniform(0, 2 * np.pi, P), zeros,
                                  10 ** rng.uniform(0, 1, P), zeros]])
        gb = GBGPUThirdBody()
    else:
        rem, add = p, q
        gb = GBGPU()
    print(f"swap 4 yr {variant}", flush=True)
    dll, terms = gb.swap_likelihood_difference(rem, add, [A, E], [SA, SE], N=N, T=T4, dt=DT, return_terms=True)
    o_r = F.run_wave(*rem, N=N, T=T4, dt=DT)
    o_a = F.run_wave(*add, N=N, T=T4, dt=DT)
    ref = np.array([
        F.swap_ll_diff_bruteforce(A, E, SA, SE, df, (o_r["AETf"][i, 0], o_r["AETf"][i, 1]), o_r["start_inds"][i],
                                  (o_a["AETf"][i, 0], o_a["AETf"][i, 1]), o_a["start_inds"][i])
        for i in range(P)
    ])
    # FP64 waveform bar at 4 yr: 1e-9 plus the carrier-phase floor
    # 5e-16 * 2 pi f T (1.2e-9 at 3 mHz); 1e-8 with a third body
    wbar = 1e-8 if variant == "third_body" else 3e-9
    scale = np.abs(terms).sum(axis=1)
    assert np.all(np.abs(dll - ref) <= 4 * wbar * scale + 1e-6), np.max(np.abs(dll - ref) / scale)
