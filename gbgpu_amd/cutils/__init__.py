"""Backend registry: the plugin slot of the reference
(src/gbgpu/cutils/__init__.py:13-166, src/gbgpu/__init__.py:38-44).

The reference declares ``GBGPUBackendMethods(get_ll, fill_global, xp)`` and
backend classes that import ``gbgpu_backend_<name>.utils``.  ``gpubackendtools``
(the reference's backend manager) is not available, so this module provides
a registry of the same shape with one backend, ``"hip"``, whose methods call
the gfx950 library through the C ABI.  Its array namespace is torch on the
GPU (device memory and streams only; all compute is in the HIP library).
"""

from __future__ import annotations

import dataclasses
import typing

from .. import _capi


class BackendUnavailableException(RuntimeError):
    """Raised when a backend cannot be loaded (src/gbgpu/cutils/__init__.py:47-50)."""


@dataclasses.dataclass
class GBGPUBackendMethods:
    """Same fields as the reference's GBGPUBackendMethods plus the fused
    entry points this backend adds."""

    get_ll: typing.Callable[..., None]
    fill_global: typing.Callable[..., None]
    xp: typing.Any
    run_wave: typing.Callable[..., None] = None
    wave_ll: typing.Callable[..., None] = None
    fill_global_rows: typing.Callable[..., None] = None


class GBGPUBackend:
    name: str = "base"

    def __init__(self, methods: GBGPUBackendMethods):
        assert isinstance(methods, GBGPUBackendMethods)
        self.get_ll = methods.get_ll
        self.fill_global = methods.fill_global
        self.xp = methods.xp
        self.run_wave = methods.run_wave
        self.wave_ll = methods.wave_ll
        self.fill_global_rows = methods.fill_global_rows


def _hip_get_ll(d_h, h_h, A_template, E_template, A_data, E_data, A_psd, E_psd, df, start_ind, M, num_bin,
                data_index, noise_index, data_length):
    """Positional twin of the reference's ``backend.get_ll`` (src/gbgpu/gbgpu.py:696-712)."""
    lib = _capi.load()
    rc = lib.gbgpu_get_ll(
        _capi.ptr(d_h), _capi.ptr(h_h), _capi.ptr(A_template), _capi.ptr(E_template), _capi.ptr(A_data),
        _capi.ptr(E_data), _capi.ptr(A_psd), _capi.ptr(E_psd), float(df), _capi.ptr(start_ind), int(M),
        int(num_bin), _capi.ptr(data_index), _capi.ptr(noise_index), int(data_length), _capi.stream_handle(),
    )
    _capi.check(rc, "gbgpu_get_ll")


def _hip_fill_global(A_glob, E_glob, A_template, E_template, start_ind, M, num_bin, group_index, data_length,
                     num_groups=None):
    """Positional twin of the reference's ``backend.fill_global`` (src/gbgpu/gbgpu.py:823-833).

    ``num_groups`` defaults to len(A_glob) // data_length (A_glob is the
    flattened (num_groups, data_length) buffer the reference passes)."""
    if num_groups is None:
        num_groups = A_glob.numel() // int(data_length)
    lib = _capi.load()
    rc = lib.gbgpu_fill_global(
        _capi.ptr(A_glob), _capi.ptr(E_glob), _capi.ptr(A_template), _capi.ptr(E_template), _capi.ptr(start_ind),
        int(M), int(num_bin), _capi.ptr(group_index), int(data_length), int(num_groups), _capi.stream_handle(),
    )
    _capi.check(rc, "gbgpu_fill_global")


class GBGPUHipBackend(GBGPUBackend):
    """gfx950 backend: hand-written HIP kernels behind the C ABI."""

    name = "hip"
    _backend_name = "gbgpu_backend_hip"
    _name = "gbgpu_hip"

    def __init__(self):
        try:
            import torch
        except ImportError as e:  # pragma: no cover
            raise BackendUnavailableException("'hip' backend requires torch (ROCm) for device memory") from e
        super().__init__(
            GBGPUBackendMethods(
                get_ll=_hip_get_ll,
                fill_global=_hip_fill_global,
                xp=torch,
            )
        )

    @staticmethod
    def lib():
        return _capi.load()

    @staticmethod
    def device():
        import torch

        if not torch.cuda.is_available():
            raise BackendUnavailableException("'hip' backend needs a ROCm GPU; none is visible")
        return torch.device("cuda", torch.cuda.current_device())


KNOWN_BACKENDS = {
    "hip": GBGPUHipBackend,
}

# aliases accepted by force_backend, as the reference accepts "gpu"/"cuda12x"
BACKEND_ALIASES = {"hip": "hip", "gpu": "hip", "gbgpu_hip": "hip", None: "hip"}


def get_backend(name=None) -> GBGPUBackend:
    key = BACKEND_ALIASES.get(name, name)
    if isinstance(name, tuple):  # ('gbgpu', name) as parallelbase.py:8-11 wraps it
        key = BACKEND_ALIASES.get(name[1], name[1])
    if key not in KNOWN_BACKENDS:
        raise BackendUnavailableException(f"unknown backend {name!r}; available: {list(KNOWN_BACKENDS)}")
    return KNOWN_BACKENDS[key]()
