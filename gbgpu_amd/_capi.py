"""ctypes binding of the C ABI in include/gbgpu_hip.h.

The library ``gbgpu_amd/libgbgpu_hip.so`` is built in-tree by
``gbgpu_amd/build.py``.  Loading fails loudly (``RuntimeError``) when it is
missing: there is no CPU fallback behind the HIP backend.  Every wrapper
raises ``RuntimeError`` carrying ``gbgpu_last_error()`` when the library
returns a non-zero status.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libgbgpu_hip.so"

GBGPU_FP64 = 0
GBGPU_MIXED = 1
# transform of the slow part (gbgpu_wave_cfg.transform, gbgpu_batch_cfg.transform)
TRANSFORMS = {"lds": 0, "mfma_f32": 1, "mfma_bf16x3": 2, "mfma_bf16x6": 3, "mfma_f64": 4}
# the precision policy each matrix-core transform belongs to
TRANSFORM_PRECISION = {"mfma_f32": "mixed", "mfma_bf16x3": "mixed", "mfma_bf16x6": "mixed", "mfma_f64": "fp64"}
GBGPU_BATCH_BUCKETS = 0
GBGPU_BATCH_PERSISTENT = 1
ORBIT_ROWS = 22
NOALLOC_MAX_BIN = 65536
N_MAX = 8192
NBUCKET = 9  # N buckets of the batched path: N = 32 << k, k < 9

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_double = ctypes.c_double
c_longlong = ctypes.c_longlong


class WaveCfg(ctypes.Structure):
    """``gbgpu_wave_cfg``."""

    _fields_ = [
        ("N", c_int),
        ("T", c_double),
        ("armlength", c_double),
        ("orbit_table", c_void_p),
        ("tdi2", c_int),
        ("precision", c_int),
        ("transform", c_int),
    ]


class Data(ctypes.Structure):
    """``gbgpu_data``."""

    _fields_ = [
        ("A", c_void_p),
        ("E", c_void_p),
        ("psdA", c_void_p),
        ("psdE", c_void_p),
        ("data_index", c_void_p),
        ("noise_index", c_void_p),
        ("data_length", c_int),
        ("start_freq_ind", c_int),
        ("df", c_double),
    ]


class BatchCfg(ctypes.Structure):
    """``gbgpu_batch_cfg``."""

    _fields_ = [
        ("T", c_double),
        ("armlength", c_double),
        ("tdi2", c_int),
        ("precision", c_int),
        ("orbit_tables", c_void_p * NBUCKET),
        ("strategy", c_int),
        ("guard_tol", c_double),
        ("transform", c_int),
        ("bucket_capacity", c_int * NBUCKET),
    ]


# name -> (restype, argtypes) for every symbol declared in include/gbgpu_hip.h
SIGNATURES = {
    "gbgpu_get_ll": (
        c_int,
        [c_void_p] * 8 + [c_double, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p],
    ),
    "gbgpu_fill_global": (
        c_int,
        [c_void_p] * 5 + [c_int, c_int, c_void_p, c_int, c_int, c_void_p],
    ),
    "gbgpu_fill_global_rows": (
        c_int,
        [c_void_p] * 4 + [c_longlong, c_void_p, c_longlong, c_int, c_int, c_void_p, c_int, c_int, c_void_p],
    ),
    "gbgpu_run_wave": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(WaveCfg), c_void_p],
    ),
    "gbgpu_fma_probe": (c_int, [c_int, ctypes.POINTER(c_double), c_void_p]),
    "gbgpu_bw_probe": (c_int, [c_int, ctypes.POINTER(c_double), c_void_p]),
    "gbgpu_hook_inputs": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(WaveCfg), c_void_p],
    ),
    "gbgpu_run_wave_ext": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(WaveCfg), c_void_p, c_void_p],
    ),
    "gbgpu_wave_ll_ext": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(WaveCfg), ctypes.POINTER(Data),
         c_double, c_int, c_void_p, c_void_p],
    ),
    "gbgpu_run_wave_ae": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(WaveCfg), c_void_p],
    ),
    "gbgpu_wave_ll": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(WaveCfg), ctypes.POINTER(Data),
         c_double, c_int, c_void_p],
    ),
    "gbgpu_wave_ll_noalloc": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(WaveCfg), ctypes.POINTER(Data),
         c_double, c_int, c_void_p],
    ),
    "gbgpu_swap_ll_diff": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(WaveCfg), ctypes.POINTER(Data), c_void_p],
    ),
    "gbgpu_wave_ll_grad": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(WaveCfg),
         ctypes.POINTER(Data), c_double, c_int, c_void_p],
    ),
    "gbgpu_fisher": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(WaveCfg), ctypes.POINTER(Data),
         c_double, c_void_p],
    ),
    "gbgpu_batch_workspace_bytes": (c_longlong, [c_int, c_int]),
    "gbgpu_wave_ll_batched": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, ctypes.POINTER(BatchCfg),
         ctypes.POINTER(Data), c_double, c_int, c_void_p, c_void_p],
    ),
    "gbgpu_batch_guard_counts": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "gbgpu_batch_set_variants": (c_int, [c_int, c_void_p]),
    "gbgpu_batch_autotune": (
        c_int,
        [c_void_p, c_void_p, c_int, c_int, c_void_p, ctypes.POINTER(BatchCfg), ctypes.POINTER(Data), c_double,
         c_void_p, c_void_p, c_void_p],
    ),
    "gbgpu_perturb_params": (
        c_int,
        [c_void_p, c_void_p, c_int, c_int, c_double, ctypes.c_ulonglong, ctypes.c_ulonglong, c_void_p, c_int,
         c_void_p, c_void_p, c_void_p],
    ),
    "gbgpu_last_error": (ctypes.c_char_p, []),
    "gbgpu_abi_version": (c_int, []),
    "gbgpu_kernel_timing": (c_int, [c_int]),
    "gbgpu_kernel_timing_read": (c_int, [ctypes.POINTER(c_double), ctypes.POINTER(c_int)]),
}

_lib = None


def load(path: os.PathLike | None = None):
    """Load (once) and return the ctypes library handle."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    # GBGPU_HIP_LIB selects another build of the same ABI (e.g. for A/B timing)
    p = Path(path) if path is not None else Path(os.environ.get("GBGPU_HIP_LIB", LIB_PATH))
    if not p.exists():
        raise RuntimeError(
            f"HIP library {p} is missing: build it with `python -m gbgpu_amd.build` "
            "(the hip backend has no CPU fallback)"
        )
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().gbgpu_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def ptr(t) -> int | None:
    """Device pointer of a torch GPU tensor (None passes NULL).

    Host tensors are refused: the kernels would dereference a host address
    (a GPU page fault with XNACK off)."""
    if t is None:
        return None
    if getattr(getattr(t, "device", None), "type", None) != "cuda":
        raise TypeError(f"expected a GPU (cuda) tensor, got {type(t).__name__} on "
                        f"{getattr(t, 'device', 'host')}")
    return t.data_ptr()


def stream_handle(device=None) -> int:
    import torch

    return torch.cuda.current_stream(device).cuda_stream


class KernelTimer:
    """Context manager summing the library's dominant-kernel durations
    (HIP events on the launch stream) over the calls made inside it."""

    def __enter__(self):
        check(load().gbgpu_kernel_timing(1), "gbgpu_kernel_timing")
        self.total_ms = 0.0
        self.count = 0
        return self

    def __exit__(self, *exc):
        t = c_double(0.0)
        n = c_int(0)
        rc = load().gbgpu_kernel_timing_read(ctypes.byref(t), ctypes.byref(n))
        load().gbgpu_kernel_timing(0)
        check(rc, "gbgpu_kernel_timing_read")
        self.total_ms, self.count = t.value, n.value
        return False

    @property
    def mean_ms(self):
        return self.total_ms / self.count if self.count else float("nan")

