"""GBGPU user API on the gfx950 backend.

Mirrors the reference class (src/gbgpu/gbgpu.py:29-1327): same class, method
and attribute names, argument order, defaults, layouts and error behaviour.
The waveform array passes that the reference runs as numpy/cupy code
(src/gbgpu/gbgpu.py:108-540) are hand-written HIP kernels here, reached
through the C ABI (include/gbgpu_hip.h); torch supplies device memory and the
stream.  Arrays returned on the device are torch tensors (the reference
returns cupy arrays on its GPU backend).
"""

from __future__ import annotations

import abc
import warnings

import numpy as np

from . import _capi
from .cutils import get_backend
from .orbits import EqualArmlengthOrbits, Orbits, orbit_table
from .utils.constants import C_SI, YRSID_SI
from .utils.utility import get_N

_PRECISIONS = {"fp64": _capi.GBGPU_FP64, "mixed": _capi.GBGPU_MIXED}


def _torch():
    import torch

    return torch


N_MAX = 8192  # the reference's get_N caps N here (src/gbgpu/utils/utility.py:268)


def _check_N(N):
    """The waveform kernels take powers of two in [32, 8192]: the band table and
    the SNR rule of get_N only produce those (gbgpu_amd.utils.get_N).  Other
    lengths, which the reference's numpy FFT would take, raise ValueError
    before any device work."""
    if N & (N - 1) or not (32 <= N <= N_MAX):
        raise ValueError(f"N={N}: the HIP kernels support powers of two in [32, {N_MAX}]")


class GBGPUBase(abc.ABC):
    """Generate Galactic-binary waveforms and likelihoods in batches.

    Args:
        orbits: object with ``armlength``, ``t_base`` and ``get_pos(t, sc)``
            (default: analytic ``EqualArmlengthOrbits``).
        force_backend: ``None``, ``"hip"`` or ``"gpu"`` (all select the
            gfx950 backend).
        precision: ``"fp64"`` (plain) or ``"mixed"`` (FP64 phases, start
            bins and likelihood accumulation; FP32 slow part and transform).
    """

    # kernels implement only these hook sets; subclasses with other hooks
    # cannot run on the GPU path
    _kernel_variant = "circular"

    def __init__(self, orbits: Orbits = None, force_backend=None, precision: str = "fp64", transform: str = "lds"):
        self.backend = get_backend(force_backend)
        self.d_d = None
        self.orbits = orbits
        self.precision = precision
        self.transform = transform
        self._tables = {}
        self._wave_state = None  # (params tensor, N, T, tdi2, precision) for lazy templates
        self._XYZf = None
        self._AETf = None

    # ----------------------------------------------------------- backend ---
    @property
    def xp(self):
        return self.backend.xp

    @property
    def get_ll_func(self):
        """get_ll native func (src/gbgpu/gbgpu.py:76-79)."""
        return self.backend.get_ll

    @property
    def fill_global_func(self):
        """fill_global native func (src/gbgpu/gbgpu.py:81-84)."""
        return self.backend.fill_global

    @property
    def orbits(self):
        return self._orbits

    @orbits.setter
    def orbits(self, orbits):
        if orbits is None:
            self._orbits = EqualArmlengthOrbits()
        elif not (hasattr(orbits, "get_pos") and hasattr(orbits, "armlength")):
            raise ValueError("Input orbits must provide armlength, t_base and get_pos(t, sc)")
        else:
            self._orbits = orbits
        if hasattr(self._orbits, "configure"):
            self._orbits.configure(linear_interp_setup=True)
        self._tables = {}

    @property
    def transform(self):
        """Transform of the slow part for N = 256 calls of run_wave,
        run_wave_ae-style templates and get_ll without a third body: "lds"
        (radix-4 transform in LDS, default), or two 16 x 16 complex matrix
        products on the matrix cores: "mfma_f64" for FP64 calls (FP64
        products and sums), "mfma_f32", "mfma_bf16x3" or "mfma_bf16x6" for
        mixed-precision calls (bf16x3 carries about 16 significant bits, the
        others FP32 accuracy).  A call of the other precision, or any other
        call, uses "lds"."""
        return self._transform

    @transform.setter
    def transform(self, t):
        if t not in _capi.TRANSFORMS:
            raise ValueError(f"transform must be one of {list(_capi.TRANSFORMS)}")
        self._transform = t

    def _transform_code(self, N, precision, num_params):
        p = precision or self.precision
        if int(N) == 256 and num_params == 9 and _capi.TRANSFORM_PRECISION.get(self._transform) == p:
            return _capi.TRANSFORMS[self._transform]
        return 0

    @property
    def precision(self):
        return self._precision

    @precision.setter
    def precision(self, p):
        if p not in _PRECISIONS:
            raise ValueError(f"precision must be one of {list(_PRECISIONS)}")
        self._precision = p

    @property
    def citation(self):
        return "Cornish & Littenberg 2007 (arXiv:0704.1808); Katz et al. 2022 (arXiv:2205.03461)"

    def _device(self):
        return self.backend.device()

    def _table(self, T, N):
        key = (float(T), int(N))
        t = self._tables.get(key)
        if t is None:
            torch = _torch()
            t = torch.from_numpy(orbit_table(self.orbits, T, N)).to(self._device())
            self._tables[key] = t
        return t

    def _cfg(self, N, T, tdi2, precision, transform=0):
        _check_N(N)
        tab = self._table(T, N)
        cfg = _capi.WaveCfg(
            N=int(N),
            T=float(T),
            armlength=float(self.orbits.armlength),
            orbit_table=tab.data_ptr(),
            tdi2=int(bool(tdi2)),
            precision=_PRECISIONS[precision or self.precision],
            transform=int(transform),
        )
        return cfg, tab

    # ---------------------------------------------------------- waveform ---
    def _prepare(self, amp, f0, fdot, fddot, phi0, iota, psi, lam, beta, args, N, T, dt, oversample):
        """Shared front half of run_wave (src/gbgpu/gbgpu.py:177-229)."""
        N_obs = int(T / dt)
        T = N_obs * dt
        if T > np.max(self.orbits.t_base):
            raise ValueError(
                f"Observation time ({T}) is larger than length of time in orbital information "
                f"({np.max(self.orbits.t_base)})"
            )
        base = [np.atleast_1d(np.asarray(x, dtype=np.float64)) for x in (amp, f0, fdot, fddot, phi0, iota, psi, lam, beta)]
        B = len(base[0])
        if any(len(x) != B for x in base):
            raise ValueError("all parameter arrays must have the same length")
        if len(args) == 0:
            add_args = ()
        else:
            add_args = self.prepare_additional_args(*args)
        if N is None:
            N_temp = self.special_get_N(base[0], base[1], T, *args, oversample=oversample)
            N = int(np.max(N_temp))
        N = int(N)
        _check_N(N)
        rows = base + [np.broadcast_to(np.asarray(a, dtype=np.float64), (B,)) for a in add_args]
        torch = _torch()
        params = torch.from_numpy(np.ascontiguousarray(np.stack(rows, axis=0))).to(self._device())
        return params, N, T

    def run_wave(self, amp, f0, fdot, fddot, phi0, iota, psi, lam, beta, *args, N=None, T=4 * YRSID_SI, dt=10.0,
                 oversample=1, tdi2=False, precision=None):
        """Create A/E/T and X/Y/Z templates for a batch (src/gbgpu/gbgpu.py:108-316).

        Outputs (torch, on the GPU): ``XYZf``/``AETf`` (num_bin, 3, N)
        complex128, ``start_inds`` (num_bin,) int64, plus ``df``, ``N``,
        ``num_bin`` -- as the reference sets them.  With
        ``precision="mixed"`` the T channel (``T``, ``T_out``, ``AETf[:, 2]``)
        is NaN: it cancels to O((f/f*)^2) of X, below the FP32 spectra's
        error.
        """
        if self._user_hooks:
            params, N, T, add_args = self._prepare_hooked(amp, f0, fdot, fddot, phi0, iota, psi, lam, beta, args, N,
                                                          T, dt, oversample)
            self._run_wave_params(params, N, T, tdi2, precision, add_args)
            return
        params, N, T = self._prepare(amp, f0, fdot, fddot, phi0, iota, psi, lam, beta, args, N, T, dt, oversample)
        self._run_wave_params(params, N, T, tdi2, precision)

    def _run_wave_params(self, params, N, T, tdi2, precision, add_args=None):
        """add_args: the extra arguments of a derived waveform with its own
        hooks (None otherwise)."""
        torch = _torch()
        B = params.shape[1]
        self.num_bin = B
        self.N = N
        self.df = 1 / T
        hooked = add_args is not None
        if hooked and (precision or self.precision) != "fp64":
            raise ValueError("derived waveforms run in FP64")
        cfg, _tab = self._cfg(N, T, tdi2, "fp64" if hooked else precision,
                              0 if hooked else self._transform_code(N, precision, params.shape[0]))
        dev = params.device
        XYZf = torch.empty((B, 3, N), dtype=torch.complex128, device=dev)
        AETf = torch.empty((B, 3, N), dtype=torch.complex128, device=dev)
        start = torch.empty((B,), dtype=torch.int64, device=dev)
        lib = _capi.load()
        if hooked:
            ext = self._hook_outputs(params, N, T, add_args)
            rc = lib.gbgpu_run_wave_ext(XYZf.data_ptr(), AETf.data_ptr(), start.data_ptr(), params.data_ptr(), 9, B,
                                        cfg, ext.data_ptr(), _capi.stream_handle())
            _capi.check(rc, "gbgpu_run_wave_ext")
        else:
            rc = lib.gbgpu_run_wave(XYZf.data_ptr(), AETf.data_ptr(), start.data_ptr(), params.data_ptr(),
                                    params.shape[0], B, cfg, _capi.stream_handle())
            _capi.check(rc, "gbgpu_run_wave")
        self._XYZf, self._AETf = XYZf, AETf
        self.start_inds = self.kmin = start
        self._wave_state = None

    def _materialize(self):
        if self._wave_state is not None:
            self._run_wave_params(*self._wave_state)

    @property
    def XYZf(self):
        self._materialize()
        return self._XYZf

    @property
    def AETf(self):
        self._materialize()
        return self._AETf

    # output views (src/gbgpu/gbgpu.py:318-346, :542-578)
    @property
    def X_out(self):
        return self.XYZf[:, 0].T.flatten()

    @property
    def Y_out(self):
        return self.XYZf[:, 1].T.flatten()

    @property
    def Z_out(self):
        return self.XYZf[:, 2].T.flatten()

    @property
    def A_out(self):
        return self.AETf[:, 0].T.flatten()

    @property
    def E_out(self):
        return self.AETf[:, 1].T.flatten()

    @property
    def T_out(self):
        return self.AETf[:, 2].T.flatten()

    @property
    def A(self):
        return self.AETf[:, 0]

    @property
    def E(self):
        return self.AETf[:, 1]

    @property
    def T(self):
        return self.AETf[:, 2]

    @property
    def X(self):
        return self.XYZf[:, 0]

    @property
    def Y(self):
        return self.XYZf[:, 1]

    @property
    def Z(self):
        return self.XYZf[:, 2]

    @property
    def freqs(self):
        torch = _torch()
        return (torch.arange(self.N, device=self.start_inds.device)[None, :] + self.start_inds[:, None]) * self.df

    # -------------------------------------------------------- likelihood ---
    def _to_dev(self, x, dtype):
        torch = _torch()
        if isinstance(x, np.ndarray):
            return torch.from_numpy(np.ascontiguousarray(x.astype(dtype, copy=False))).to(self._device())
        if torch.is_tensor(x):
            tdt = {np.complex128: torch.complex128, np.float64: torch.float64, np.int32: torch.int32}[dtype]
            if x.dtype != tdt:
                raise TypeError(f"expected {tdt}, got {x.dtype}")
            return x.to(self._device()).contiguous()
        raise TypeError("Make sure the data arrays are numpy arrays or torch tensors.")

    @staticmethod
    def _check_data(data, data_index, noise_index):
        """Host-only checks of get_ll's data (src/gbgpu/gbgpu.py:650-665)."""
        torch = _torch()
        if not isinstance(data[0], np.ndarray) and not torch.is_tensor(data[0]):
            raise TypeError("Make sure the data arrays are the same type as template arrays (cupy vs numpy).")
        if data[0].ndim == 1:
            if data_index is not None:
                raise ValueError("If inputing 1D data, cannot use data_index kwarg.")
            if noise_index is not None:
                raise ValueError("If inputing 1D data, cannot use noise_index kwarg.")
            return data[0].shape[0]
        elif data[0].ndim == 2:
            return data[0].shape[1]
        raise ValueError("data must be 1D or 2D")

    def _prep_data(self, data, psd, data_index, noise_index, num_bin):
        """Validation and flattening of get_ll's data/psd (src/gbgpu/gbgpu.py:650-686)."""
        torch = _torch()
        data_length = self._check_data(data, data_index, noise_index)
        dA = self._to_dev(data[0], np.complex128).reshape(-1)
        dE = self._to_dev(data[1], np.complex128).reshape(-1)
        SA = self._to_dev(psd[0], np.float64).reshape(-1)
        SE = self._to_dev(psd[1], np.float64).reshape(-1)
        if data_index is None:
            data_index = torch.zeros(num_bin, dtype=torch.int32, device=dA.device)
        if noise_index is None:
            noise_index = torch.zeros(num_bin, dtype=torch.int32, device=dA.device)
        if isinstance(data_index, np.ndarray) and data_index.dtype != np.int32:
            raise AssertionError("data_index must be int32")
        if isinstance(noise_index, np.ndarray) and noise_index.dtype != np.int32:
            raise AssertionError("noise_index must be int32")
        data_index = self._to_dev(data_index, np.int32)
        noise_index = self._to_dev(noise_index, np.int32)
        assert len(data_index) == num_bin
        assert len(data_index) == len(noise_index)
        if num_bin and (int(data_index.min()) < 0 or int(noise_index.min()) < 0):
            raise ValueError("data_index/noise_index must be non-negative")
        assert int(data_index.max()) * data_length <= len(dA)
        assert int(noise_index.max()) * data_length <= len(dA)
        if int(data_index.max()) * data_length >= len(dA) or int(noise_index.max()) * data_length >= len(SA):
            raise ValueError("data_index/noise_index point past the provided streams")
        return dA, dE, SA, SE, data_index, noise_index, data_length

    def get_ll(self, params, data, psd, phase_marginalize=False, start_freq_ind=0, data_index=None,
               noise_index=None, **kwargs):
        """Batched log-likelihood (src/gbgpu/gbgpu.py:580-735).

        Runs the fused waveform + likelihood kernel: templates stay in LDS.
        ``gb.A``/``gb.E``/... remain available afterwards (regenerated on
        first access).  Returns a numpy (num_bin,) float64 array.
        """
        if self.d_d is None:
            raise ValueError(
                "self.d_d attribute must be set before computing log-Likelihood. This attribute is the data "
                "with data inner product (<d|d>)."
            )
        torch = _torch()
        tdi2 = kwargs.pop("tdi2", False)
        precision = kwargs.pop("precision", None)
        N = kwargs.pop("N", None)
        T = kwargs.pop("T", 4 * YRSID_SI)
        dt = kwargs.pop("dt", 10.0)
        oversample = kwargs.pop("oversample", 1)
        if kwargs:
            raise TypeError(f"unexpected keyword arguments {list(kwargs)}")
        params = np.asarray(params) if not torch.is_tensor(params) else params.cpu().numpy()
        self._check_data(data, data_index, noise_index)
        hooked = self._user_hooks
        if N is not None and np.ndim(N) == 1:
            self._check_hooks("get_ll with one N per binary")
            return self._get_ll_uneven(params, data, psd, np.asarray(N), T, dt, tdi2, precision, phase_marginalize,
                                       start_freq_ind, data_index, noise_index)
        if hooked:
            if (precision or self.precision) != "fp64":
                raise ValueError("derived waveforms run in FP64")
            p, N, T, add_args = self._prepare_hooked(*params[:9], params[9:], N, T, dt, oversample)
        else:
            p, N, T = self._prepare(*params[:9], params[9:], N, T, dt, oversample)
            add_args = None
        B = p.shape[1]
        self.num_bin, self.N, self.df = B, N, 1 / T
        dA, dE, SA, SE, di, ni, L = self._prep_data(data, psd, data_index, noise_index, B)
        cfg, _tab = self._cfg(N, T, tdi2, "fp64" if hooked else precision,
                              0 if hooked else self._transform_code(N, precision, p.shape[0]))
        dev = p.device
        ll = torch.empty(B, dtype=torch.float64, device=dev)
        d_h = torch.empty(B, dtype=torch.complex128, device=dev)
        h_h = torch.empty(B, dtype=torch.complex128, device=dev)
        dat = _capi.Data(A=dA.data_ptr(), E=dE.data_ptr(), psdA=SA.data_ptr(), psdE=SE.data_ptr(),
                         data_index=di.data_ptr(), noise_index=ni.data_ptr(), data_length=int(L),
                         start_freq_ind=int(start_freq_ind), df=float(1 / T))
        lib = _capi.load()
        if hooked:
            ext = self._hook_outputs(p, N, T, add_args)
            rc = lib.gbgpu_wave_ll_ext(ll.data_ptr(), d_h.data_ptr(), h_h.data_ptr(), p.data_ptr(), 9, B, cfg, dat,
                                       float(np.real(self.d_d)), int(bool(phase_marginalize)), ext.data_ptr(),
                                       _capi.stream_handle())
            _capi.check(rc, "gbgpu_wave_ll_ext")
        else:
            rc = lib.gbgpu_wave_ll(ll.data_ptr(), d_h.data_ptr(), h_h.data_ptr(), p.data_ptr(), p.shape[0], B, cfg,
                                   dat, float(np.real(self.d_d)), int(bool(phase_marginalize)), _capi.stream_handle())
            _capi.check(rc, "gbgpu_wave_ll")
        # start bins without materialising templates: rint(f0 T) - N/2
        f0 = p[1]
        self.start_inds = self.kmin = (torch.round(f0 * T).to(torch.int64) - N // 2)
        self._wave_state = (p, N, T, tdi2, precision, add_args)
        if phase_marginalize:
            self.non_marg_d_h = d_h.cpu().numpy()
            d_h = torch.abs(d_h).to(torch.complex128)
        self.h_h = h_h
        self.d_h = d_h
        return ll.cpu().numpy()

    def swap_likelihood_difference(self, params_remove, params_add, data, psd, phase_marginalize=False,
                                   start_freq_ind=0, data_index=None, noise_index=None, return_terms=False,
                                   **kwargs):
        """Log-likelihood change of replacing binaries by proposals.

        For each proposal p, ``data`` holds the residual r (data minus the
        current global template, which includes binary ``params_remove[:, p]``)
        and the result is ``logL(r + h_remove - h_add) - logL(r)``, i.e.
        Re<r|h_a> - Re<r|h_r> - (<h_a|h_a> + <h_r|h_r> - 2 Re<h_r|h_a>)/2 with
        the reference's inner product (src/gbgpu/cutils/gbgpu_utils.cu:156-183).
        The reference has no such method; arguments mirror ``get_ll``
        (src/gbgpu/gbgpu.py:580-735): params are (num_params, P), ``data`` and
        ``psd`` are [A, E] (1-D, or 2-D with per-proposal data_index /
        noise_index), ``start_freq_ind`` as in get_ll.  Remaining kwargs go to
        the waveform (N, T, dt, tdi2, precision).  Returns (P,) float64, plus
        (P, 5) terms when ``return_terms``.
        """
        if phase_marginalize:
            raise NotImplementedError("phase marginalisation is not defined for a swap difference")
        torch = _torch()
        self._check_hooks("swap_likelihood_difference")
        tdi2 = kwargs.pop("tdi2", False)
        precision = kwargs.pop("precision", None)
        N = kwargs.pop("N", None)
        T = kwargs.pop("T", 4 * YRSID_SI)
        dt = kwargs.pop("dt", 10.0)
        oversample = kwargs.pop("oversample", 1)
        if kwargs:
            raise TypeError(f"unexpected keyword arguments {list(kwargs)}")
        pr = np.atleast_2d(np.asarray(params_remove, dtype=np.float64))
        pa = np.atleast_2d(np.asarray(params_add, dtype=np.float64))
        if pr.shape != pa.shape:
            raise ValueError("params_remove and params_add must have the same shape")
        self._check_data(data, data_index, noise_index)
        P = pr.shape[1]
        pairs = np.empty((pr.shape[0], 2 * P))
        pairs[:, 0::2] = pr
        pairs[:, 1::2] = pa
        if N is None:
            N_r = self.special_get_N(pr[0], pr[1], int(T / dt) * dt, *pr[9:], oversample=oversample)
            N_a = self.special_get_N(pa[0], pa[1], int(T / dt) * dt, *pa[9:], oversample=oversample)
            N = int(max(np.max(N_r), np.max(N_a)))
        p, N, T = self._prepare(*pairs[:9], pairs[9:], N, T, dt, oversample)
        dA, dE, SA, SE, di, ni, L = self._prep_data(data, psd, data_index, noise_index, P)
        cfg, _tab = self._cfg(N, T, tdi2, precision)
        dev = p.device
        dll = torch.empty(P, dtype=torch.float64, device=dev)
        terms = torch.empty((P, 5), dtype=torch.float64, device=dev) if return_terms else None
        dat = _capi.Data(A=dA.data_ptr(), E=dE.data_ptr(), psdA=SA.data_ptr(), psdE=SE.data_ptr(),
                         data_index=di.data_ptr(), noise_index=ni.data_ptr(), data_length=int(L),
                         start_freq_ind=int(start_freq_ind), df=float(1 / T))
        rc = _capi.load().gbgpu_swap_ll_diff(dll.data_ptr(), _capi.ptr(terms), p.data_ptr(), p.shape[0], P, cfg, dat,
                                             _capi.stream_handle())
        _capi.check(rc, "gbgpu_swap_ll_diff")
        if return_terms:
            return dll.cpu().numpy(), terms.cpu().numpy()
        return dll.cpu().numpy()

    def _get_ll_uneven(self, params, data, psd, N, T, dt, tdi2, precision, phase_marginalize, start_freq_ind,
                       data_index, noise_index):
        """get_ll with one N per binary (powers of two in [32, 8192]): one
        gbgpu_wave_ll_batched call buckets the binaries by N on the device and
        runs each bucket's fused kernel; same arithmetic as the uniform get_ll
        (no FP64 re-evaluation of loud binaries).  Third-body rows, tdi2 and
        data_index / noise_index as in get_ll."""
        from .batch import BatchedLogLike

        if len(N) != params.shape[1]:
            raise ValueError("N must have one entry per binary")
        for n in np.unique(N):
            _check_N(int(n))
        if T > np.max(self.orbits.t_base):
            raise ValueError(
                f"Observation time ({T}) is larger than length of time in orbital information "
                f"({np.max(self.orbits.t_base)})"
            )
        if params.shape[0] > 9:
            # the extra rows through the class's own hook, as run_wave does
            params = np.vstack([params[:9], np.array(self.prepare_additional_args(*params[9:]))])
        # 2-D streams without indices: every binary uses stream 0 (data_index
        # and noise_index default to zero, src/gbgpu/gbgpu.py:675-678), and a
        # window reaching past data_length is clipped there
        if data[0].ndim == 2 and data_index is None:
            data = [data[0][0], data[1][0]]
        if psd[0].ndim == 2 and noise_index is None:
            psd = [psd[0][0], psd[1][0]]
        bl = BatchedLogLike(self, data, psd, T, dt, precision=precision or self.precision,
                            start_freq_ind=start_freq_ind, N_levels=(int(N.min()),), f_edges=(),
                            phase_marginalize=phase_marginalize, guard_tol=None, tdi2=tdi2,
                            data_index=data_index, noise_index=noise_index)
        ll, d_h, h_h = bl(params, N=N.astype(np.int32), d_h=True)
        self.num_bin, self.N, self.df = params.shape[1], N, 1 / bl.T
        if phase_marginalize:
            self.non_marg_d_h = d_h.cpu().numpy()
            d_h = _torch().abs(d_h).to(_torch().complex128)
        self.h_h, self.d_h = h_h, d_h
        self._wave_state = None
        return ll.cpu().numpy()

    # -------------------------------------------------- global templates ---
    def fill_global_template(self, group_index, templates, A, E, start_inds, N=None, start_freq_ind=0):
        """Sum binaries into per-group A/E arrays (src/gbgpu/gbgpu.py:737-837).

        Deterministic (bit-reproducible) summation.  As in the reference,
        ``templates`` is overwritten with the sum of the given binaries.
        """
        torch = _torch()
        total_groups, nchannels, data_length = templates.shape
        if nchannels < 2:
            raise ValueError("Calculates for A and E channels.")
        elif nchannels > 2:
            warnings.warn("Only calculating A and E channels here currently.")
        if not torch.is_tensor(templates) or templates.device.type != "cuda":
            raise TypeError("Make sure the data arrays are the same type as template arrays (cupy vs numpy).")
        dev = templates.device
        group_index = torch.as_tensor(np.asarray(group_index) if not torch.is_tensor(group_index) else group_index,
                                      dtype=torch.int32).to(dev).contiguous()
        num_bin = len(group_index)
        tA = torch.zeros((total_groups, data_length), dtype=torch.complex128, device=dev)
        tE = torch.zeros((total_groups, data_length), dtype=torch.complex128, device=dev)
        # host arrays and host tensors are copied to the templates' device
        A = (A if torch.is_tensor(A) else torch.from_numpy(np.asarray(A))).to(dev)
        E = (E if torch.is_tensor(E) else torch.from_numpy(np.asarray(E))).to(dev)
        if A.ndim > 2 or E.ndim > 2:
            raise ValueError("A_in, E_in have maximum allowable dimension of 2.")
        lib = _capi.load()
        if A.ndim == 2:
            assert E.ndim == 2
            N = A.shape[1]
            if A.stride(1) != 1:
                A = A.contiguous()
            if E.stride(1) != 1 or E.stride(0) != A.stride(0):
                E = E.contiguous()
                A = A.contiguous()
            st = torch.as_tensor(start_inds if torch.is_tensor(start_inds) else np.asarray(start_inds),
                                 dtype=torch.int64).to(dev).contiguous()
            rc = lib.gbgpu_fill_global_rows(tA.data_ptr(), tE.data_ptr(), A.data_ptr(), E.data_ptr(), A.stride(0),
                                            st.data_ptr(), int(start_freq_ind), int(N), num_bin,
                                            group_index.data_ptr(), int(data_length), int(total_groups),
                                            _capi.stream_handle())
            _capi.check(rc, "gbgpu_fill_global_rows")
        else:
            if N is None:
                raise ValueError(
                    "If providing a 1D flattened array for A and E, the N kwarg also needs to be provided."
                )
            st = (torch.as_tensor(start_inds if torch.is_tensor(start_inds) else np.asarray(start_inds)).to(dev)
                  - start_freq_ind).to(torch.int32).contiguous()
            self.fill_global_func(tA.reshape(-1), tE.reshape(-1), A.contiguous(), E.contiguous(), st, N, num_bin,
                                  group_index, data_length, total_groups)
        templates[:, 0] = tA
        templates[:, 1] = tE

    def generate_global_template(self, params, group_index, templates, start_freq_ind=0, **kwargs):
        """run_wave then fill_global_template (src/gbgpu/gbgpu.py:839-878).

        Only the A and E spectra are computed for the fill
        (``gbgpu_run_wave_ae``); the other run_wave outputs (``XYZf``,
        ``AETf``, ``A``, ...) are computed on first access, as after get_ll.
        """
        torch = _torch()
        params = np.asarray(params).T
        N = kwargs.pop("N", None)
        T = kwargs.pop("T", 4 * YRSID_SI)
        dt = kwargs.pop("dt", 10.0)
        oversample = kwargs.pop("oversample", 1)
        tdi2 = kwargs.pop("tdi2", False)
        precision = kwargs.pop("precision", None)
        if kwargs:
            raise TypeError(f"unexpected keyword arguments {list(kwargs)}")
        if self._user_hooks:
            # derived waveform: its hooked run_wave, then the fill
            self.run_wave(*params[:9], *params[9:], N=N, T=T, dt=dt, oversample=oversample, tdi2=tdi2,
                          precision=precision)
            self.fill_global_template(group_index, templates, self.AETf[:, 0], self.AETf[:, 1], self.start_inds,
                                      self.N, start_freq_ind=start_freq_ind)
            return
        p, N, T = self._prepare(*params[:9], params[9:], N, T, dt, oversample)
        B = p.shape[1]
        self.num_bin, self.N, self.df = B, N, 1 / T
        cfg, _tab = self._cfg(N, T, tdi2, precision, self._transform_code(N, precision, p.shape[0]))
        AE = torch.empty((B, 2, N), dtype=torch.complex128, device=p.device)
        start = torch.empty((B,), dtype=torch.int64, device=p.device)
        rc = _capi.load().gbgpu_run_wave_ae(AE.data_ptr(), start.data_ptr(), p.data_ptr(), p.shape[0], B, cfg,
                                            _capi.stream_handle())
        _capi.check(rc, "gbgpu_run_wave_ae")
        self.start_inds = self.kmin = start
        self._wave_state = (p, N, T, tdi2, precision, None)
        self.fill_global_template(group_index, templates, AE[:, 0], AE[:, 1], start, N,
                                  start_freq_ind=start_freq_ind)

    def inject_signal(self, *args, fmax=None, T=4.0 * YRSID_SI, dt=10.0, **kwargs):
        """Place binary 0 into zero A/E arrays on [0, fmax] (src/gbgpu/gbgpu.py:880-946)."""
        if fmax is None:
            fmax = 1 / (2 * dt)
        if T > np.max(self.orbits.t_base):
            raise ValueError(
                f"Observation time ({T}) is larger than length of time in orbital information "
                f"({np.max(self.orbits.t_base)})"
            )
        N_obs = int(T / dt)
        T = N_obs * dt
        kwargs["T"] = T
        kwargs["dt"] = dt
        self.df = df = 1 / T
        f = np.arange(0.0, fmax + df, df)
        num = len(f)
        A_out = np.zeros(num, dtype=np.complex128)
        E_out = np.zeros(num, dtype=np.complex128)
        self.run_wave(*args, **kwargs)
        start = int(self.start_inds[0].item())
        A_temp = self.A[0].cpu().numpy()
        E_temp = self.E[0].cpu().numpy()
        A_out[start : start + self.N] = A_temp
        E_out[start : start + self.N] = E_temp
        return A_out, E_out

    # ---------------------------------------------------- Fisher (FD) ---
    def _apply_parameter_transforms(self, params, parameter_transforms):
        """src/gbgpu/gbgpu.py:948-955."""
        for ind_trans, trans in parameter_transforms.items():
            if isinstance(ind_trans, int):
                params[ind_trans] = trans(params[ind_trans])
            else:
                params[np.asarray(ind_trans)] = trans(*params[np.asarray(ind_trans)])
        return params

    # ------------------------------------------------------- derivatives ---
    def _deriv_setup(self, params, N, T, dt, oversample, kwargs_left):
        if kwargs_left:
            raise TypeError(f"unexpected keyword arguments {list(kwargs_left)}")
        self._check_hooks("the exact derivative kernels")
        if type(self)._kernel_variant != "circular":
            raise NotImplementedError("derivatives are implemented for circular (9-parameter) binaries")
        torch = _torch()
        params = np.asarray(params) if not torch.is_tensor(params) else params.detach().cpu().numpy()
        params = np.atleast_2d(np.asarray(params, dtype=np.float64))
        if params.shape[0] != 9:
            raise ValueError("derivatives need (9, num_bin) parameters")
        return self._prepare(*params[:9], (), N, T, dt, oversample)

    def get_ll_grad(self, params, data, psd, phase_marginalize=False, start_freq_ind=0, basis="sampling",
                    tie_fddot=False, return_gpu=False, **kwargs):
        """Log-likelihood and its gradient from the adjoint kernel (FP64).

        Same likelihood and data conventions as ``get_ll`` (one A/E stream).
        Returns ``(loglike (B,), grad (B, M))``: with basis "sampling" the
        gradient is with respect to (ln amp, f0, fdot, fddot, phi0, cos iota,
        psi, lam, sin beta); with "raw" with respect to the run_wave
        parameters; ``tie_fddot`` folds fddot = 11/3 fdot^2/f0 into f0 and
        fdot and drops its entry (M = 8).
        """
        from .derivatives import basis_jacobian

        if self.d_d is None:
            raise ValueError("self.d_d attribute must be set before computing log-Likelihood.")
        torch = _torch()
        N = kwargs.pop("N", None)
        T = kwargs.pop("T", 4 * YRSID_SI)
        dt = kwargs.pop("dt", 10.0)
        tdi2 = kwargs.pop("tdi2", False)
        oversample = kwargs.pop("oversample", 1)
        p, N, T = self._deriv_setup(params, N, T, dt, oversample, kwargs)
        B = p.shape[1]
        self._check_data(data, None, None)
        dA, dE, SA, SE, di, ni, L = self._prep_data(data, psd, None, None, B)
        cfg, _tab = self._cfg(N, T, tdi2, "fp64")
        ll = torch.empty(B, dtype=torch.float64, device=p.device)
        g = torch.empty((B, 9), dtype=torch.float64, device=p.device)
        dat = _capi.Data(A=dA.data_ptr(), E=dE.data_ptr(), psdA=SA.data_ptr(), psdE=SE.data_ptr(), data_index=None,
                         noise_index=None, data_length=int(L), start_freq_ind=int(start_freq_ind), df=float(1 / T))
        rc = _capi.load().gbgpu_wave_ll_grad(ll.data_ptr(), g.data_ptr(), None, None, p.data_ptr(), 9, B, cfg, dat,
                                             float(np.real(self.d_d)), int(bool(phase_marginalize)),
                                             _capi.stream_handle())
        _capi.check(rc, "gbgpu_wave_ll_grad")
        J = basis_jacobian(p, basis, tie_fddot)
        g = (J.transpose(1, 2) @ g.unsqueeze(-1)).squeeze(-1)
        if return_gpu:
            return ll, g
        return ll.cpu().numpy(), g.cpu().numpy()

    def fisher_matrix(self, params, psd, data=None, start_freq_ind=0, basis="sampling", tie_fddot=False,
                      return_gpu=False, **kwargs):
        """Fisher matrices from exact forward tangents (FP64, N <= 1024).

        F_kl = 4 df sum_{A,E} sum_f Re(conj(dh_k) dh_l) / S over each
        template's bins within the PSD streams ``psd`` = [S_A, S_E] (indexed
        like get_ll's data, from ``start_freq_ind``).  Returns (B, M, M) in
        the basis chosen as for ``get_ll_grad``.
        """
        from .derivatives import basis_jacobian

        torch = _torch()
        N = kwargs.pop("N", None)
        T = kwargs.pop("T", 4 * YRSID_SI)
        dt = kwargs.pop("dt", 10.0)
        tdi2 = kwargs.pop("tdi2", False)
        oversample = kwargs.pop("oversample", 1)
        p, N, T = self._deriv_setup(params, N, T, dt, oversample, kwargs)
        B = p.shape[1]
        SA = self._to_dev(psd[0], np.float64).reshape(-1)
        SE = self._to_dev(psd[1], np.float64).reshape(-1)
        if data is not None:
            dA = self._to_dev(data[0], np.complex128).reshape(-1)
            dE = self._to_dev(data[1], np.complex128).reshape(-1)
            if dA.numel() != SA.numel():
                raise ValueError("data and psd streams must have the same length")
        cfg, _tab = self._cfg(N, T, tdi2, "fp64")
        F = torch.empty((B, 9, 9), dtype=torch.float64, device=p.device)
        dat = _capi.Data(A=None if data is None else dA.data_ptr(), E=None if data is None else dE.data_ptr(),
                         psdA=SA.data_ptr(), psdE=SE.data_ptr(), data_index=None, noise_index=None,
                         data_length=int(SA.numel()), start_freq_ind=int(start_freq_ind), df=float(1 / T))
        rc = _capi.load().gbgpu_fisher(F.data_ptr(), None, None, p.data_ptr(), 9, B, cfg, dat,
                                       float(np.real(self.d_d or 0.0)), _capi.stream_handle())
        _capi.check(rc, "gbgpu_fisher")
        J = basis_jacobian(p, basis, tie_fddot)
        F = J.transpose(1, 2) @ F @ J
        return F if return_gpu else F.cpu().numpy()

    def information_matrix(self, params, eps=1e-9, parameter_transforms={}, inds=None, N=1024, psd_func=None,
                           psd_kwargs={}, easy_central_difference=False, return_gpu=False, method="exact",
                           **kwargs):
        """Information matrix (src/gbgpu/gbgpu.py:957-1138).

        As the reference, the derivative with respect to parameter i is taken
        with a multiplicative step, dh_i = x_i dh/dx_i (x = ``params``, mapped
        to waveform parameters by ``parameter_transforms`` after the step),
        and F_ij = 4 df sum_{A,E} Re(conj(dh_i) dh_j) / psd_func(f) over the
        template bins.  ``method="exact"`` takes dh/dx from the forward-tangent
        kernel (the transforms' Jacobian by the reference's stencil on the
        transforms alone); ``method="fd"`` differences HIP run_wave outputs with
        the reference's 5-point (or central) stencil.  ``psd_func`` defaults to
        the analytic A-channel PSD.
        """
        torch = _torch()
        if psd_func is None:
            from .noise import A1TDISens

            psd_func = A1TDISens.get_Sn
        if N is None:
            raise ValueError("N must be provided up front for Infromation Matrix.")
        params = np.atleast_2d(np.asarray(params, dtype=np.float64))
        num_params, num_bins = params.shape
        if inds is None:
            inds = np.arange(num_params)
        inds = np.asarray(inds)
        exact_ok = (method == "exact" and num_params == 9 and N <= 1024 and type(self)._kernel_variant == "circular"
                    and not self._user_hooks)
        if method not in ("exact", "fd"):
            raise ValueError("method must be 'exact' or 'fd'")
        if not exact_ok:
            return self._information_matrix_fd(params, eps, parameter_transforms, inds, N, psd_func, psd_kwargs,
                                               easy_central_difference, return_gpu, **kwargs)
        from .derivatives import basis_jacobian

        T = kwargs.pop("T", 4 * YRSID_SI)
        dt = kwargs.pop("dt", 10.0)
        tdi2 = kwargs.pop("tdi2", False)
        oversample = kwargs.pop("oversample", 1)
        if kwargs:
            raise TypeError(f"unexpected keyword arguments {list(kwargs)}")
        raw = self._apply_parameter_transforms(params.copy(), parameter_transforms)
        # d raw / d ln x: identity scaled by x without transforms, else the
        # reference's stencil applied to the transforms
        if parameter_transforms:
            Jx = np.zeros((num_bins, 9, num_params))
            for j in range(num_params):
                def tr(m):
                    q = np.ones_like(params)
                    q[j] += m * eps
                    return self._apply_parameter_transforms(q * params, parameter_transforms)

                if easy_central_difference:
                    d = (tr(1) - tr(-1)) / (2 * eps)
                else:
                    d = (-tr(2) + tr(-2) + 8 * (tr(1) - tr(-1))) / (12 * eps)
                Jx[:, :, j] = d.T
        else:
            Jx = np.broadcast_to(np.eye(9), (num_bins, 9, 9)) * params.T[:, None, :]
        p, N, T = self._prepare(*raw[:9], (), N, T, dt, oversample)
        df = 1.0 / T
        start = np.rint(raw[1] * T).astype(np.int64) - N // 2
        k0, k1 = int(start.min()), int(start.max()) + N
        fgrid = (k0 + np.arange(k1 - k0)) * df
        S = np.asarray(psd_func(fgrid, **psd_kwargs), dtype=np.float64)
        St = torch.from_numpy(np.ascontiguousarray(S)).to(p.device)
        cfg, _tab = self._cfg(N, T, tdi2, "fp64")
        F = torch.empty((num_bins, 9, 9), dtype=torch.float64, device=p.device)
        dat = _capi.Data(A=None, E=None, psdA=St.data_ptr(), psdE=St.data_ptr(), data_index=None, noise_index=None,
                         data_length=int(St.numel()), start_freq_ind=k0, df=float(df))
        rc = _capi.load().gbgpu_fisher(F.data_ptr(), None, None, p.data_ptr(), 9, num_bins, cfg, dat, 0.0,
                                       _capi.stream_handle())
        _capi.check(rc, "gbgpu_fisher")
        J = basis_jacobian(p, "raw") @ torch.from_numpy(np.ascontiguousarray(Jx)).to(p.device)
        info = J.transpose(1, 2) @ F @ J
        sel = torch.from_numpy(inds).to(p.device)
        info = info[:, sel][:, :, sel]
        self.N, self.df, self.num_bin = N, df, num_bins
        if return_gpu:
            return info
        return info.cpu().numpy()

    def _information_matrix_fd(self, params, eps, parameter_transforms, inds, N, psd_func, psd_kwargs,
                               easy_central_difference, return_gpu, **kwargs):
        """The reference's finite-difference stencil over HIP run_wave outputs."""
        torch = _torch()
        kwargs["N"] = N
        num_params, num_bins = params.shape
        nd = len(inds)
        dev = self._device()
        dh = torch.zeros((num_bins, nd, 2, N), dtype=torch.complex128, device=dev)

        starts = []

        def wave(p):
            self.run_wave(*p, **kwargs)
            starts.append(self.start_inds.clone())
            return torch.stack([self.A, self.E], dim=1).clone()

        for i, ind in enumerate(inds):
            def step(m):
                p = np.ones_like(params)
                p[ind] += m * eps
                p *= params
                return self._apply_parameter_transforms(p, parameter_transforms)

            starts.clear()
            up1 = wave(step(1))
            dn1 = wave(step(-1))
            if easy_central_difference:
                dh[:, i] = (up1 - dn1) / (2 * eps)
            else:
                up2 = wave(step(2))
                dn2 = wave(step(-2))
                dh[:, i] = (-up2 + dn2 + 8 * (up1 - dn1)) / (12 * eps)
            # every stencil point must put each binary on the same bins (the
            # reference compares start bins across binaries instead,
            # src/gbgpu/gbgpu.py:1092-1095, which only holds for identical ones)
            if not all(bool(torch.equal(s, starts[0])) for s in starts[1:]):
                raise ValueError(
                    "The user should decrease steps size (eps) because the frequency bins are changing during "
                    "derivative calculation, which is not allowed."
                )
        fr = self.freqs.cpu().numpy()
        psd = torch.from_numpy(np.asarray(psd_func(fr, **psd_kwargs), dtype=np.float64)).to(dev)
        info = torch.zeros((num_bins, nd, nd), dtype=torch.float64, device=dev)
        for i in range(nd):
            for j in range(i, nd):
                ip = 4 * self.df * torch.sum((dh[:, i].conj() * dh[:, j]).real / psd[:, None, :], dim=(1, 2))
                info[:, i, j] = ip
                info[:, j, i] = ip
        if return_gpu:
            return info
        return info.cpu().numpy()

    # ------------------------------------------------------------- hooks ---
    @property
    def _user_hooks(self):
        """True when a class outside this package defines the slow-part hooks
        (shift_frequency / add_to_argS, src/gbgpu/gbgpu.py:1140-1229):
        run_wave, get_ll and generate_global_template then evaluate them on
        device arrays between gbgpu_hook_inputs and gbgpu_run_wave_ext /
        gbgpu_wave_ll_ext.  The package's own classes (circular, third body)
        evaluate their hooks inside the kernels."""
        pkg = __name__.split(".")[0]
        for cls in type(self).__mro__:
            if cls is object or cls.__module__.split(".")[0] == pkg:
                continue
            if {"add_to_argS", "shift_frequency"} & set(vars(cls)):
                return True
        return False

    def _check_hooks(self, what="this method"):
        if self._user_hooks:
            raise NotImplementedError(
                f"{type(self).__name__} defines its own slow-part hooks: {what} runs the kernels' own circular "
                "and third-body hook sets only (run_wave, get_ll and generate_global_template take derived waveforms)"
            )

    def _prepare_hooked(self, amp, f0, fdot, fddot, phi0, iota, psi, lam, beta, args, N, T, dt, oversample):
        """_prepare for a derived waveform: the 9 base rows go to the kernels;
        the extra arguments pass through prepare_additional_args to the hooks
        (src/gbgpu/gbgpu.py:197-207)."""
        T_adj = int(T / dt) * dt
        add_args = self.prepare_additional_args(*args) if len(args) else ()
        if N is None:
            N = int(np.max(self.special_get_N(np.atleast_1d(amp), np.atleast_1d(f0), T_adj, *args,
                                              oversample=oversample)))
        params, N, T = self._prepare(amp, f0, fdot, fddot, phi0, iota, psi, lam, beta, (), N, T, dt, oversample)
        return params, N, T, add_args

    def _hook_outputs(self, params, N, T, add_args):
        """Runs the class's hooks on device arrays and returns (B, 2, 3, N):
        the hooked f_i, then argS, for every binary, spacecraft and sample
        (src/gbgpu/gbgpu.py:441-453, :497-505)."""
        torch = _torch()
        if N > 2048:
            raise ValueError(f"N={N}: derived waveforms take N <= 2048")
        B = params.shape[1]
        cfg, _tab = self._cfg(N, T, False, "fp64")
        xi, fi, argS = (torch.empty((B, 3, N), dtype=torch.float64, device=params.device) for _ in range(3))
        rc = _capi.load().gbgpu_hook_inputs(xi.data_ptr(), fi.data_ptr(), argS.data_ptr(), params.data_ptr(), 9, B,
                                            cfg, _capi.stream_handle())
        _capi.check(rc, "gbgpu_hook_inputs")
        fi = self.shift_frequency(fi, xi, *add_args)
        argS = self.add_to_argS(argS, params[1], params[2], params[3], xi, *add_args)
        out = []
        for name, x in (("shift_frequency", fi), ("add_to_argS", argS)):
            x = torch.as_tensor(x).to(device=params.device, dtype=torch.float64)
            if tuple(x.shape) != (B, 3, N):
                raise ValueError(f"{name} must return a ({B}, 3, {N}) array, got {tuple(x.shape)}")
            out.append(x)
        return torch.stack(out, dim=1).contiguous()

    @abc.abstractmethod
    def add_to_argS(self, argS, f0, fdot, fddot, xi, *args):
        """Phase hook (src/gbgpu/gbgpu.py:1140-1167); evaluated inside the kernel."""

    @abc.abstractmethod
    def prepare_additional_args(self, *args):
        """src/gbgpu/gbgpu.py:1169-1184."""

    @abc.abstractmethod
    def special_get_N(self, amp, f0, T, *args, oversample=1):
        """src/gbgpu/gbgpu.py:1186-1210."""

    @abc.abstractmethod
    def shift_frequency(self, fi, xi, *args):
        """Frequency hook (src/gbgpu/gbgpu.py:1212-1229); evaluated inside the kernel."""


class GBGPU(GBGPUBase):
    """Circular Galactic binaries (src/gbgpu/gbgpu.py:1233-1327)."""

    def prepare_additional_args(self, *args):
        # the reference returns (does not raise) this ValueError (gbgpu.py:1255-1257)
        raise ValueError(
            "If providing more args than the base args, must be a class derived from GBGPUBase with an adjusted "
            "'prepare_additional_args' method."
        )

    def special_get_N(self, amp, f0, T, *args, oversample=1):
        return get_N(amp, f0, T, oversample=oversample)

    def shift_frequency(self, fi, xi, *args):
        return fi

    def add_to_argS(self, argS, f0, fdot, fddot, xi, *args):
        return argS
