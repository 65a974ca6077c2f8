"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"\b(gbgpu_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("gbgpu_get_ll", "gbgpu_fill_global", "gbgpu_run_wave", "gbgpu_wave_ll", "gbgpu_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from gbgpu_amd import _capi

    lib = _capi.load()
    for n in declared_functions():
        assert hasattr(lib, n), n
    # the ctypes table covers exactly the declared functions
    assert set(_capi.SIGNATURES) == set(declared_functions())
    assert lib.gbgpu_abi_version() == 9


def test_invalid_arguments_return_errors_without_gpu_work():
    from gbgpu_amd import _capi

    lib = _capi.load()
    cfg = _capi.WaveCfg(N=100, T=1e7, armlength=2.5e9, orbit_table=1, tdi2=0, precision=0)
    rc = lib.gbgpu_run_wave(None, None, 1, 1, 9, 4, cfg, None)
    assert rc == -1 and b"power of two" in lib.gbgpu_last_error()
    cfg.N = 256
    rc = lib.gbgpu_run_wave(None, None, 1, 1, 7, 4, cfg, None)
    assert rc == -1 and b"num_params" in lib.gbgpu_last_error()
    rc = lib.gbgpu_get_ll(None, None, None, None, None, None, None, None, 1.0, None, 4, -1, None, None, 10, None)
    assert rc == -1
    rc = lib.gbgpu_fill_global(None, None, None, None, None, 4, 3, None, 10, -2, None)
    assert rc == -1
    # zero-size batches are no-ops
    assert lib.gbgpu_get_ll(None, None, None, None, None, None, None, None, 1.0, None, 4, 0, None, None, 10, None) == 0


@pytest.mark.parametrize("precision,N,num_params,transform,ok", [
    (0, 256, 9, 1, False),   # matrix-core transforms are mixed-precision only
    (1, 128, 9, 2, False),   # ... at N = 256 only
    (1, 256, 14, 3, False),  # ... without a third body
    (1, 256, 9, 4, False),   # unknown transform
    (1, 256, 9, -1, False),
    (1, 256, 9, 3, True),
])
def test_transform_combinations_checked_before_gpu_work(precision, N, num_params, transform, ok):
    """Unsupported transform combinations return GBGPU_E_ARG before any HIP
    call; a supported one passes the checks (num_bin = 0: nothing runs)."""
    from gbgpu_amd import _capi

    lib = _capi.load()
    cfg = _capi.WaveCfg(N=N, T=1e7, armlength=2.5e9, orbit_table=1, tdi2=0, precision=precision,
                        transform=transform)
    rc = lib.gbgpu_run_wave(1, 1, 1, 1, num_params, 0 if ok else 4, cfg, None)
    if ok:
        assert rc == 0
    else:
        assert rc == -1 and b"transform" in lib.gbgpu_last_error()
    # swap and derivatives take the LDS transform only
    if ok:
        rc = lib.gbgpu_swap_ll_diff(1, None, 1, 9, 4, cfg, None, None)
        assert rc == -1 and b"transform" in lib.gbgpu_last_error()


def test_empty_batches_are_noops_on_every_entry_point():
    """num_bin = 0 returns OK after argument checks and launches nothing
    (dummy non-null pointers are never dereferenced)."""
    import ctypes

    from gbgpu_amd import _capi

    lib = _capi.load()
    P = 8  # any non-null address
    cfg = _capi.WaveCfg(N=256, T=1e7, armlength=2.5e9, orbit_table=P, tdi2=0, precision=0)
    dat = _capi.Data(A=P, E=P, psdA=P, psdE=P, data_index=None, noise_index=None, data_length=10,
                     start_freq_ind=0, df=1e-7)
    assert lib.gbgpu_run_wave(None, P, P, P, 9, 0, cfg, None) == 0
    assert lib.gbgpu_run_wave_ae(P, P, P, 9, 0, cfg, None) == 0
    # the A/E-only entry needs both outputs
    assert lib.gbgpu_run_wave_ae(None, P, P, 9, 4, cfg, None) == -1
    assert b"AEf" in lib.gbgpu_last_error()
    assert lib.gbgpu_wave_ll(P, None, None, P, 9, 0, cfg, dat, 0.0, 0, None) == 0
    assert lib.gbgpu_wave_ll_noalloc(P, None, None, P, 9, 0, cfg, dat, 0.0, 0, None) == 0
    # the allocation-free entry refuses batches beyond its fixed scratch
    assert lib.gbgpu_wave_ll_noalloc(P, None, None, P, 9, _capi.NOALLOC_MAX_BIN + 1, cfg, dat, 0.0, 0, None) == -1
    assert b"GBGPU_NOALLOC_MAX_BIN" in lib.gbgpu_last_error()
    assert lib.gbgpu_swap_ll_diff(P, None, P, 9, 0, cfg, dat, None) == 0
    assert lib.gbgpu_wave_ll_grad(P, P, None, None, P, 9, 0, cfg, dat, 0.0, 0, None) == 0
    assert lib.gbgpu_fisher(P, None, None, P, 9, 0, cfg, dat, 0.0, None) == 0
    assert lib.gbgpu_fill_global(P, P, P, P, P, 4, 0, P, 10, 1, None) == 0
    bcfg = _capi.BatchCfg(T=1e7, armlength=2.5e9, tdi2=0, precision=1)
    assert lib.gbgpu_wave_ll_batched(P, None, None, P, 9, 0, P, bcfg, dat, 0.0, 0, None, None) == 0
    assert lib.gbgpu_perturb_params(P, P, 9, 0, 1e-3, 1, 0, None, 0, None, None, None) == 0
    # derivatives reject what they do not support, before any GPU work
    cfg.precision = 1
    assert lib.gbgpu_wave_ll_grad(P, P, None, None, P, 9, 4, cfg, dat, 0.0, 0, None) == -1
    assert b"FP64" in lib.gbgpu_last_error()
    cfg.precision = 0
    assert lib.gbgpu_wave_ll_grad(P, P, None, None, P, 14, 4, cfg, dat, 0.0, 0, None) == -1
    cfg.N = 2048
    assert lib.gbgpu_fisher(P, None, None, P, 9, 4, cfg, dat, 0.0, None) == -1
    assert b"1024" in lib.gbgpu_last_error()
    del ctypes


def test_N_limits_checked_before_gpu_work():
    """Waveform and likelihood entries take powers of two up to 8192; swap and
    gradients stop at 2048 and Fisher matrices at 1024, with GBGPU_E_ARG
    before any HIP call (dummy pointers are never dereferenced)."""
    from gbgpu_amd import _capi

    lib = _capi.load()
    P = 8
    dat = _capi.Data(A=P, E=P, psdA=P, psdE=P, data_index=None, noise_index=None, data_length=10,
                     start_freq_ind=0, df=1e-7)
    cfg = _capi.WaveCfg(N=16384, T=1e7, armlength=2.5e9, orbit_table=P, tdi2=0, precision=0)
    assert lib.gbgpu_run_wave(P, P, P, P, 9, 4, cfg, None) == -1
    assert b"[32, 8192]" in lib.gbgpu_last_error()
    assert lib.gbgpu_wave_ll(P, None, None, P, 9, 4, cfg, dat, 0.0, 0, None) == -1
    for N in (4096, 8192):
        cfg.N = N
        # accepted by the waveform entries (nothing runs for an empty batch)
        assert lib.gbgpu_run_wave(P, P, P, P, 9, 0, cfg, None) == 0
        assert lib.gbgpu_wave_ll(P, None, None, P, 9, 0, cfg, dat, 0.0, 0, None) == 0
        assert lib.gbgpu_swap_ll_diff(P, None, P, 9, 4, cfg, dat, None) == -1
        assert b"2048" in lib.gbgpu_last_error()
        assert lib.gbgpu_wave_ll_grad(P, P, None, None, P, 9, 4, cfg, dat, 0.0, 0, None) == -1
        assert b"2048" in lib.gbgpu_last_error()
        assert lib.gbgpu_fisher(P, None, None, P, 9, 4, cfg, dat, 0.0, None) == -1


def test_matrix_core_transforms_checked_against_precision():
    """GBGPU_TRANSFORM_MFMA_F64 goes with FP64 only and the FP32/bf16
    variants with mixed only, all at N = 256 without a third body: other
    pairs return GBGPU_E_ARG before any HIP call (dummy pointers)."""
    from gbgpu_amd import _capi

    lib = _capi.load()
    P = 8
    dat = _capi.Data(A=P, E=P, psdA=P, psdE=P, data_index=None, noise_index=None, data_length=10,
                     start_freq_ind=0, df=1e-7)
    T = _capi.TRANSFORMS
    for prec, tr, ok in ((0, T["mfma_f64"], True), (1, T["mfma_f64"], False), (0, T["mfma_f32"], False),
                         (1, T["mfma_bf16x6"], True), (0, 5, False)):
        cfg = _capi.WaveCfg(N=256, T=1e7, armlength=2.5e9, orbit_table=P, tdi2=0, precision=prec, transform=tr)
        # an empty batch runs nothing, so accepted pairs return 0
        rc = lib.gbgpu_wave_ll(P, None, None, P, 9, 0, cfg, dat, 0.0, 0, None)
        assert (rc == 0) == ok, (prec, tr, rc, lib.gbgpu_last_error())
        if not ok:
            assert b"transform" in lib.gbgpu_last_error()
    cfg = _capi.WaveCfg(N=512, T=1e7, armlength=2.5e9, orbit_table=P, tdi2=0, precision=0,
                        transform=T["mfma_f64"])
    assert lib.gbgpu_run_wave(P, P, P, P, 9, 0, cfg, None) == -1
    assert b"N = 256" in lib.gbgpu_last_error()
    assert lib.gbgpu_run_wave(P, P, P, P, 14, 0, cfg, None) == -1


def test_missing_library_fails_loudly(tmp_path):
    from gbgpu_amd import _capi

    with pytest.raises(RuntimeError, match="missing"):
        _capi.load(tmp_path / "nope.so")
