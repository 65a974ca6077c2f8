"""Gradients and Fisher matrices of the FastGB likelihood.

The kernels (``gbgpu_wave_ll_grad``: adjoint; ``gbgpu_fisher``: forward
tangents) return derivatives in the *sampling basis*

    K = (ln amp, f0, fdot, fddot, phi0, cos iota, psi, lambda, sin beta)

with f0, fdot, fddot independent.  This module maps them to

* the raw waveform parameters (amp, f0, fdot, fddot, phi0, iota, psi, lam,
  beta) -- the coordinates of ``run_wave`` and of the reference's
  ``information_matrix`` (src/gbgpu/gbgpu.py:957-1138);
* the 8-parameter sampling basis with fddot tied to 11/3 fdot^2/f0;

and wraps the likelihood in ``torch.autograd.Function`` objects whose
backward pass is the adjoint kernel.
"""

from __future__ import annotations

import numpy as np

from . import _capi

SAMPLING_NAMES = ("ln_amp", "f0", "fdot", "fddot", "phi0", "cos_iota", "psi", "lam", "sin_beta")
TIED_NAMES = ("ln_amp", "f0", "fdot", "phi0", "cos_iota", "psi", "lam", "sin_beta")
_TIED_ROWS = (0, 1, 2, 4, 5, 6, 7, 8)


def _torch():
    import torch

    return torch


def basis_jacobian(params, basis="sampling", tie_fddot=False):
    """J (B, 9, M) with J[b, k, j] = d K_k / d x_j for the requested basis x.

    ``params`` is (9, B) raw parameters (torch or numpy).  basis "sampling"
    (x = K, or its tied 8-parameter form) or "raw" (x = the raw waveform
    parameters; with ``tie_fddot`` the fddot column is dropped and its
    dependence on f0 and fdot folded in).  Gradients map as g_x = J^T g_K and
    Fisher matrices as F_x = J^T F_K J.
    """
    torch = _torch()
    p = params if torch.is_tensor(params) else torch.from_numpy(np.asarray(params, dtype=np.float64))
    p = p.to(torch.float64)
    B = p.shape[1]
    J = torch.zeros((B, 9, 9), dtype=torch.float64, device=p.device)
    idx = torch.arange(9, device=p.device)
    J[:, idx, idx] = 1.0
    if basis == "raw":
        J[:, 0, 0] = 1.0 / p[0]               # d ln amp / d amp
        J[:, 5, 5] = -torch.sin(p[5])          # d cos iota / d iota
        J[:, 8, 8] = torch.cos(p[8])           # d sin beta / d beta
    elif basis != "sampling":
        raise ValueError("basis must be 'sampling' or 'raw'")
    if tie_fddot:
        f0, fdot = p[1], p[2]
        # fddot = 11/3 fdot^2 / f0
        J[:, 3, 1] = -11.0 / 3.0 * fdot ** 2 / f0 ** 2
        J[:, 3, 2] = 22.0 / 3.0 * fdot / f0
        J = J[:, :, list(_TIED_ROWS)]
    return J


def sampling_to_raw(theta):
    """(8, B) tied sampling parameters -> (9, B) raw parameters."""
    torch = _torch()
    lnA, f0, fdot, phi0, c, psi, lam, s = theta
    fddot = 11.0 / 3.0 * fdot ** 2 / f0
    return torch.stack([torch.exp(lnA), f0, fdot, fddot, phi0, torch.arccos(c), psi, lam, torch.arcsin(s)])


def raw_to_sampling(params):
    """(9, B) raw parameters -> (8, B) tied sampling parameters (fddot dropped)."""
    torch = _torch()
    p = params
    return torch.stack([torch.log(p[0]), p[1], p[2], p[4], torch.cos(p[5]), p[6], p[7], torch.sin(p[8])])


class FastGBLikelihood:
    """Resident FP64 likelihood of independent binaries with derivatives.

    logL_b = -1/2 (d_d + <h_b|h_b> - 2 Re<d|h_b>) for every column b of the
    parameter batch (as ``GBGPU.get_ll``), its adjoint gradient and its
    Fisher matrix.

    Args:
        gb: a ``GBGPU`` instance (orbits, device, ``d_d``).
        data, psd: [A, E] 1-D streams (numpy or torch).
        N: samples per waveform (power of two; Fisher matrices need N <= 1024).
        T, dt: observation time and cadence (T is reduced to int(T/dt)*dt).
    """

    def __init__(self, gb, data, psd, N, T, dt=10.0, start_freq_ind=0, phase_marginalize=False, tdi2=False):
        if gb.d_d is None:
            raise ValueError("gb.d_d must be set before building a likelihood")
        self.gb = gb
        self.T = int(T / dt) * dt
        self.N = int(N)
        self.phase_marginalize = bool(phase_marginalize)
        self.A = gb._to_dev(data[0], np.complex128).reshape(-1)
        self.E = gb._to_dev(data[1], np.complex128).reshape(-1)
        self.SA = gb._to_dev(psd[0], np.float64).reshape(-1)
        self.SE = gb._to_dev(psd[1], np.float64).reshape(-1)
        self.cfg, self._tab = gb._cfg(self.N, self.T, tdi2, "fp64")
        self.data = _capi.Data(A=self.A.data_ptr(), E=self.E.data_ptr(), psdA=self.SA.data_ptr(),
                               psdE=self.SE.data_ptr(), data_index=None, noise_index=None,
                               data_length=int(self.A.numel()), start_freq_ind=int(start_freq_ind),
                               df=float(1.0 / self.T))
        self.d_d = float(np.real(gb.d_d))

    def _params(self, params):
        torch = _torch()
        p = params if torch.is_tensor(params) else torch.from_numpy(np.ascontiguousarray(params, np.float64))
        p = p.detach().to(self.gb._device(), torch.float64).contiguous()
        if p.ndim != 2 or p.shape[0] != 9:
            raise ValueError("params must be (9, num_bin)")
        return p

    def ll_and_grad(self, params):
        """(loglike (B,), grad (B, 9) in the sampling basis), device tensors."""
        torch = _torch()
        p = self._params(params)
        B = p.shape[1]
        ll = torch.empty(B, dtype=torch.float64, device=p.device)
        g = torch.empty((B, 9), dtype=torch.float64, device=p.device)
        rc = _capi.load().gbgpu_wave_ll_grad(ll.data_ptr(), g.data_ptr(), None, None, p.data_ptr(), 9, B, self.cfg,
                                             self.data, self.d_d, int(self.phase_marginalize),
                                             _capi.stream_handle())
        _capi.check(rc, "gbgpu_wave_ll_grad")
        return ll, g

    def fisher(self, params, with_grad=False):
        """Fisher matrices (B, 9, 9) in the sampling basis (and the gradient)."""
        torch = _torch()
        p = self._params(params)
        B = p.shape[1]
        F = torch.empty((B, 9, 9), dtype=torch.float64, device=p.device)
        g = torch.empty((B, 9), dtype=torch.float64, device=p.device) if with_grad else None
        rc = _capi.load().gbgpu_fisher(F.data_ptr(), _capi.ptr(g), None, p.data_ptr(), 9, B, self.cfg, self.data,
                                       self.d_d, _capi.stream_handle())
        _capi.check(rc, "gbgpu_fisher")
        return (F, g) if with_grad else F

    def __call__(self, params):
        """Differentiable log-likelihood of (9, B) raw parameters."""
        return WaveLogLike.apply(params, self)

    def sampling(self, theta):
        """Differentiable log-likelihood of (8, B) tied sampling parameters."""
        return SamplingLogLike.apply(theta, self)


def _function_base():
    return _torch().autograd.Function


class WaveLogLike(_function_base()):
    """logL(raw params) with the adjoint kernel as backward pass."""

    @staticmethod
    def forward(ctx, params, like):
        ll, g = like.ll_and_grad(params)
        J = basis_jacobian(params.detach(), "raw")
        g_raw = (J.transpose(1, 2) @ g.unsqueeze(-1)).squeeze(-1)  # (B, 9)
        ctx.save_for_backward(g_raw.to(params.device))
        return ll.to(params.device)

    @staticmethod
    def backward(ctx, gout):
        (g_raw,) = ctx.saved_tensors
        return (gout[:, None] * g_raw).T.contiguous(), None


class SamplingLogLike(_function_base()):
    """logL(ln amp, f0, fdot, phi0, cos iota, psi, lam, sin beta), fddot tied."""

    @staticmethod
    def forward(ctx, theta, like):
        raw = sampling_to_raw(theta.detach())
        ll, g = like.ll_and_grad(raw)
        J = basis_jacobian(raw, "sampling", tie_fddot=True)
        g8 = (J.transpose(1, 2) @ g.unsqueeze(-1)).squeeze(-1)  # (B, 8)
        ctx.save_for_backward(g8.to(theta.device))
        return ll.to(theta.device)

    @staticmethod
    def backward(ctx, gout):
        (g8,) = ctx.saved_tensors
        return (gout[:, None] * g8).T.contiguous(), None
