"""Derived waveforms on the GPU: subclasses of GBGPUBase defined outside the
package, with their own shift_frequency / add_to_argS hooks written against
``self.xp`` (torch on this backend), as the reference's extension API takes
them (src/gbgpu/gbgpu.py:1140-1229; docs/source/user/derivedwaves.rst).

Bars: FP64 waveform bar (1e-9 relative plus the carrier-phase floor; 1e-8
for the third-body formulas, as for GBGPUThirdBody) against the oracle run
with numpy versions of the same hooks; log-likelihoods by tests.helpers.ll_bar.
"""

import numpy as np
import pytest

from oracle import fastgb as F
from tests.helpers import YR, draw_params, ll_bar, make_data, rel_err

pytestmark = pytest.mark.gpu

DT = 15.0


def _classes():
    import torch

    from gbgpu_amd import GBGPUBase
    from gbgpu_amd.utils.constants import C_SI, YRSID_SI
    from gbgpu_amd.utils.utility import get_N

    class DriftShift(GBGPUBase):
        """A toy modification: fi -> fi (1 + delta), argS -> argS + alpha xi^2,
        with per-binary (delta, alpha)."""

        def prepare_additional_args(self, delta, alpha):
            xp = self.xp
            return (xp.as_tensor(np.atleast_1d(delta), dtype=xp.float64, device="cuda"),
                    xp.as_tensor(np.atleast_1d(alpha), dtype=xp.float64, device="cuda"))

        def special_get_N(self, amp, f0, T, delta, alpha, oversample=1):
            return get_N(amp, f0, T, oversample=oversample)

        def shift_frequency(self, fi, xi, delta, alpha):
            return fi * (1.0 + delta[:, None, None])

        def add_to_argS(self, argS, f0, fdot, fddot, xi, delta, alpha):
            return argS + alpha[:, None, None] * xi * xi

    class TorchThirdBody(GBGPUBase):
        """The circular third body (src/gbgpu/thirdbody.py:155-226 with
        e2 = 0) written as hooks in torch."""

        def prepare_additional_args(self, A2, varpi, P2, T2):
            xp = self.xp
            t = lambda x: xp.as_tensor(np.atleast_1d(x), dtype=xp.float64, device="cuda")  # noqa: E731
            return t(A2), t(varpi), 2 * np.pi / (t(P2) * YRSID_SI), t(T2) * YRSID_SI

        def special_get_N(self, amp, f0, T, A2, varpi, P2, T2, oversample=1):
            return get_N(amp, f0, T, oversample=oversample)

        @staticmethod
        def _vlos(xi, A2, varpi, n2, T2):
            return A2[:, None, None] * torch.sin(n2[:, None, None] * (xi - T2[:, None, None]) + varpi[:, None, None])

        def shift_frequency(self, fi, xi, A2, varpi, n2, T2):
            return fi * (1.0 + self._vlos(xi, A2, varpi, n2, T2) / C_SI)

        def add_to_argS(self, argS, f0, fdot, fddot, xi, A2, varpi, n2, T2):
            xp = self.xp
            fgw = f0[:, None, None] + fdot[:, None, None] * xi + 0.5 * fddot[:, None, None] * xi * xi
            g = self._vlos(xi, A2, varpi, n2, T2) * fgw
            step = (xi[..., 1:] - xi[..., :-1]) * (g[..., :-1] + g[..., 1:]) / 2.0 * np.pi * 2 / C_SI
            term = xp.zeros_like(xi)
            term[..., 1:] = xp.cumsum(step, dim=-1)
            return argS + term

    return DriftShift, TorchThirdBody


def _drift_hooks(delta, alpha):
    return (lambda fi, xi: fi * (1.0 + delta[:, None, None]),
            lambda argS, f0, fdot, fddot, xi: argS + alpha[:, None, None] * xi * xi)


def fp64_tol(p, T):
    fmax = np.abs(p[1]) + np.abs(p[2]) * T + 0.5 * np.abs(p[3]) * T**2
    return 1e-9 + 5e-16 * 2 * np.pi * fmax * T


@pytest.mark.parametrize("N", [64, 256, 2048])
def test_derived_run_wave_matches_oracle_hooks(N):
    DriftShift, _ = _classes()
    rng = np.random.default_rng(6000 + N)
    B = 12
    p = draw_params(B, rng, f_lo=1e-3, f_hi=2e-2 if N == 2048 else 4e-3)
    delta = rng.uniform(-1e-4, 1e-4, B)
    alpha = rng.uniform(-1e-14, 1e-14, B)
    print(f"derived run_wave N={N}", flush=True)
    gb = DriftShift()
    assert gb._user_hooks
    gb.run_wave(*p, delta, alpha, N=N, T=YR, dt=DT)
    o = F.run_wave(*p, N=N, T=YR, dt=DT, hooks=_drift_hooks(delta, alpha))
    assert np.array_equal(gb.start_inds.cpu().numpy(), o["start_inds"])
    err = rel_err(gb.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2))
    assert np.all(err <= fp64_tol(p, YR)), np.max(err / fp64_tol(p, YR))
    # the hooks change the waveform far beyond the bar
    plain = F.run_wave(*p, N=N, T=YR, dt=DT)
    assert np.min(rel_err(plain["XYZf"], o["XYZf"], axis=(1, 2))) > 1e-3


def test_derived_get_ll_and_global_template_match_oracle():
    import torch

    DriftShift, _ = _classes()
    rng = np.random.default_rng(6100)
    B, N = 16, 256
    p = draw_params(B, rng)
    p[0] *= 10.0
    delta = rng.uniform(-1e-4, 1e-4, B)
    alpha = rng.uniform(-1e-14, 1e-14, B)
    A, E, SA, SE, df, L = make_data(p, N, YR, DT, rng)
    gb = DriftShift()
    gb.d_d = float(4 * df * (np.sum(np.abs(A) ** 2 / SA) + np.sum(np.abs(E) ** 2 / SE)))
    print("derived get_ll", flush=True)
    ll = gb.get_ll(np.vstack([p, delta, alpha]), [A, E], [SA, SE], N=N, T=YR, dt=DT, tdi2=True)
    o = F.run_wave(*p, N=N, T=YR, dt=DT, hooks=_drift_hooks(delta, alpha), tdi2=True)
    d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"], A, E, SA, SE, df)
    ll_o = F.loglike(gb.d_d, d_h, h_h)
    assert np.all(np.abs(ll - ll_o) <= ll_bar(d_h, h_h)), np.max(np.abs(ll - ll_o) / ll_bar(d_h, h_h))
    # templates after get_ll are the hooked ones
    assert rel_err(gb.AETf.cpu().numpy()[:, :2], o["AETf"][:, :2], axis=(1, 2)).max() <= 1e-8
    # global template of the derived waveform
    print("derived generate_global_template", flush=True)
    templates = torch.zeros((1, 2, L), dtype=torch.complex128, device="cuda")
    gb.generate_global_template(np.vstack([p, delta, alpha]).T, np.zeros(B, dtype=np.int32), templates, N=N, T=YR,
                                dt=DT, tdi2=True)
    ref = F.fill_global(1, L, o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"], np.zeros(B, int))
    got = templates.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max()


def test_torch_third_body_hooks_match_the_kernel_third_body():
    """A circular third body written as torch hooks equals the oracle's and
    the kernels' built-in third body (GBGPUThirdBody) to the third-body bar."""
    from gbgpu_amd import GBGPUThirdBody

    _, TorchThirdBody = _classes()
    rng = np.random.default_rng(6200)
    B, N = 8, 512
    p = draw_params(B, rng)
    A2, varpi, P2, T2 = rng.uniform(0, 100, B), rng.uniform(0, 2 * np.pi, B), 10 ** rng.uniform(0, 1, B), \
        rng.uniform(0, 1, B)
    print("torch third body hooks", flush=True)
    gb = TorchThirdBody()
    gb.run_wave(*p, A2, varpi, P2, T2, N=N, T=YR, dt=DT)
    o = F.run_wave(*p, A2, varpi, np.zeros(B), P2, T2, N=N, T=YR, dt=DT)
    assert rel_err(gb.XYZf.cpu().numpy(), o["XYZf"], axis=(1, 2)).max() <= 1e-8
    kb = GBGPUThirdBody()
    kb.run_wave(*p, A2, varpi, np.zeros(B), P2, T2, N=N, T=YR, dt=DT)
    assert rel_err(gb.XYZf.cpu().numpy(), kb.XYZf.cpu().numpy(), axis=(1, 2)).max() <= 1e-8


def test_derived_waveforms_refuse_what_the_kernels_cannot_hook():
    DriftShift, _ = _classes()
    gb = DriftShift()
    rng = np.random.default_rng(6300)
    p = draw_params(2, rng)
    ext = np.vstack([p, np.zeros(2), np.zeros(2)])
    A = np.zeros(1000, dtype=np.complex128)
    S = np.ones(1000)
    gb.d_d = 0.0
    with pytest.raises(NotImplementedError, match="swap_likelihood_difference"):
        gb.swap_likelihood_difference(ext, ext, [A, A], [S, S], N=64, T=YR, dt=DT)
    with pytest.raises(NotImplementedError, match="one N per binary"):
        gb.get_ll(ext, [A, A], [S, S], N=np.array([64, 128]), T=YR, dt=DT)
    with pytest.raises(ValueError, match="FP64"):
        gb.run_wave(*p, np.zeros(2), np.zeros(2), N=64, T=YR, dt=DT, precision="mixed")
