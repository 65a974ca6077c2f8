"""Host-side behaviour of the GBGPU mirror that needs no GPU: validation
errors raised before any device work, as the reference raises them."""

import numpy as np
import pytest

from gbgpu_amd import GBGPU, GBGPUThirdBody
from gbgpu_amd.cutils import BackendUnavailableException, get_backend
from tests.helpers import YR


def _params(B=3):
    return np.array([np.full(B, v) for v in (1e-22, 2e-3, 1e-18, 0.0, 0.1, 0.2, 0.3, 0.4, 0.5)])


def test_get_ll_requires_d_d():
    gb = GBGPU()
    with pytest.raises(ValueError, match="d_d"):
        gb.get_ll(_params(), [np.zeros(10, complex)] * 2, [np.ones(10)] * 2)


def test_observation_time_beyond_orbits():
    gb = GBGPU()
    with pytest.raises(ValueError, match="Observation time"):
        gb.run_wave(*_params(), N=128, T=20 * YR, dt=15.0)
    with pytest.raises(ValueError, match="Observation time"):
        gb.inject_signal(*_params(1)[:, 0], N=128, T=20 * YR, dt=15.0)


def test_data_index_with_1d_data_rejected():
    gb = GBGPU()
    gb.d_d = 0.0
    with pytest.raises(ValueError, match="data_index"):
        gb.get_ll(_params(), [np.zeros(10, complex)] * 2, [np.ones(10)] * 2, data_index=np.zeros(3, np.int32))
    with pytest.raises(ValueError, match="noise_index"):
        gb.get_ll(_params(), [np.zeros(10, complex)] * 2, [np.ones(10)] * 2, noise_index=np.zeros(3, np.int32))


def test_wrong_data_type_rejected():
    gb = GBGPU()
    gb.d_d = 0.0
    with pytest.raises(TypeError):
        gb.get_ll(_params(), [[0j] * 10] * 2, [np.ones(10)] * 2)


def test_extra_args_on_circular_class_rejected():
    gb = GBGPU()
    with pytest.raises(ValueError, match="prepare_additional_args"):
        gb.run_wave(*_params(), *np.ones((5, 3)), N=128, T=YR, dt=15.0)


def test_unknown_backend_and_precision():
    with pytest.raises(BackendUnavailableException):
        get_backend("cuda12x")
    with pytest.raises(ValueError):
        GBGPU(precision="fp16")
    assert get_backend("gpu").name == "hip"
    assert get_backend(("gbgpu", "hip")).name == "hip"


def test_third_body_prepare_and_get_N():
    gb = GBGPUThirdBody()
    A2, varpi, e2, n2, T2 = gb.prepare_additional_args(19.5, 0.0, 0.3, 0.6, 0.5)
    assert np.isclose(n2[0], 2 * np.pi / (0.6 * 31558149.763545603))
    assert np.isclose(T2[0], 0.5 * 31558149.763545603)
    # a 0.1-yr outer orbit over 4 yr needs >= 2/P2 samples per yr -> N doubles from the base
    N = gb.special_get_N(np.array([1e-30]), np.array([2e-3]), 4 * YR, 0.0, 0.0, 0.0, np.array([0.1]), 0.0)
    assert N[0] >= 4 * 2 / 0.1


def test_fill_global_template_validation():
    gb = GBGPU()
    with pytest.raises(ValueError, match="A and E"):
        gb.fill_global_template(np.zeros(2, np.int32), np.zeros((1, 1, 10), complex), np.zeros((2, 4)),
                                np.zeros((2, 4)), np.zeros(2, int))
    with pytest.raises(TypeError):
        gb.fill_global_template(np.zeros(2, np.int32), np.zeros((1, 2, 10), complex), np.zeros((2, 4)),
                                np.zeros((2, 4)), np.zeros(2, int))


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message, before touching a device."""
    import json
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parent.parent
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_host_tensors_are_refused_as_device_pointers():
    """A host (CPU) tensor never reaches a kernel as a pointer: the ctypes
    layer refuses it before any call."""
    import torch

    from gbgpu_amd import _capi

    with pytest.raises(TypeError, match="GPU"):
        _capi.ptr(torch.zeros(4, dtype=torch.complex128))
    with pytest.raises(TypeError, match="GPU"):
        _capi.ptr(np.zeros(4))
    assert _capi.ptr(None) is None


def test_persistent_strategy_refused():
    """The persistent batch strategy is switched off (its kernel faulted with
    a cause not yet shown); asking for it is an error, with or without the
    mixed-precision guard (which only the per-bucket launches ever had)."""
    from gbgpu_amd.batch import BatchedLogLike

    gb = GBGPU()
    gb.d_d = 0.0
    z = [np.zeros(10, complex)] * 2
    for guard in (1e-3, None):
        with pytest.raises(ValueError, match="buckets"):
            BatchedLogLike(gb, z, [np.ones(10)] * 2, YR, 15.0, precision="mixed", strategy="persistent",
                           guard_tol=guard)


def test_persistent_strategy_refused_by_the_library():
    """The C ABI refuses it too, before any GPU work."""
    from gbgpu_amd import _capi

    lib = _capi.load()
    P = 8
    bcfg = _capi.BatchCfg(T=1e7, armlength=2.5e9, tdi2=0, precision=0, strategy=_capi.GBGPU_BATCH_PERSISTENT)
    dat = _capi.Data(A=P, E=P, psdA=P, psdE=P, data_index=None, noise_index=None, data_length=10, start_freq_ind=0,
                     df=1e-7)
    assert lib.gbgpu_wave_ll_batched(P, None, None, P, 9, 4, P, bcfg, dat, 0.0, 0, None, None) == -1
    assert b"persistent" in lib.gbgpu_last_error()


def test_derived_waveform_hooks_detected_and_unsupported_paths_refused():
    """A subclass defined outside the package with its own slow-part hooks
    takes the derived-waveform path (gbgpu_hook_inputs -> hooks -> *_ext);
    the package's own classes do not.  Paths the kernels cannot hook refuse
    before any device work."""
    import numpy as np
    import pytest

    from gbgpu_amd import GBGPU, GBGPUBase, GBGPUThirdBody

    class Shifted(GBGPU):
        def shift_frequency(self, fi, xi, *args):
            return fi * 1.0001

    class OnlyN(GBGPU):
        def special_get_N(self, amp, f0, T, *args, oversample=1):
            return super().special_get_N(amp, f0, T, *args, oversample=oversample)

    assert Shifted()._user_hooks
    assert not OnlyN()._user_hooks
    assert not GBGPU()._user_hooks and not GBGPUThirdBody()._user_hooks
    assert issubclass(Shifted, GBGPUBase)
    gb = Shifted()
    p = np.ones((9, 2))
    A = np.zeros(100, dtype=np.complex128)
    with pytest.raises(NotImplementedError, match="swap_likelihood_difference"):
        gb.swap_likelihood_difference(p, p, [A, A], [np.ones(100)] * 2, N=64)
    gb.d_d = 0.0
    with pytest.raises(NotImplementedError, match="one N per binary"):
        gb.get_ll(p, [A, A], [np.ones(100)] * 2, N=np.array([64, 64]))
