"""The allocation-free fused likelihood (gbgpu_wave_ll_noalloc) against
gbgpu_wave_ll, through the C ABI.

Both entry points run the same kernels on the same per-binary constants and
twiddle values, so the bar is bit-identical loglike, d_h and h_h (which is
within every tolerance the fused likelihood is held to).  Two calls in a row
share the library's fixed scratch: the second, smaller batch must not see
anything the first one left behind.
"""

import numpy as np
import pytest

from tests.helpers import YR, draw_params, make_data

pytestmark = pytest.mark.gpu

DT = 15.0


def _call(lib_fn, gb, p, N, dat, d_d, precision, torch):
    from gbgpu_amd import _capi

    B = p.shape[1]
    cfg, _tab = gb._cfg(N, YR_ADJ, False, precision)
    ll = torch.full((B,), np.nan, dtype=torch.float64, device="cuda")
    d_h = torch.full((B,), np.nan, dtype=torch.complex128, device="cuda")
    h_h = torch.full((B,), np.nan, dtype=torch.complex128, device="cuda")
    rc = lib_fn(ll.data_ptr(), d_h.data_ptr(), h_h.data_ptr(), p.data_ptr(), p.shape[0], B, cfg, dat, d_d, 0,
                _capi.stream_handle())
    _capi.check(rc, lib_fn.__name__)
    torch.cuda.synchronize()
    return ll.cpu().numpy(), d_h.cpu().numpy(), h_h.cpu().numpy()


YR_ADJ = int(YR / DT) * DT


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_wave_ll_on_consecutive_calls(precision):
    import torch

    from gbgpu_amd import GBGPU, _capi

    lib = _capi.load()
    rng = np.random.default_rng(900 + (precision == "mixed"))
    inj = draw_params(64, rng)
    A, E, SA, SE, df, L = make_data(inj, 256, YR, DT, rng)
    gb = GBGPU()
    d_d = float(4 * df * (np.sum(np.abs(A) ** 2 / SA) + np.sum(np.abs(E) ** 2 / SE)))
    # first call: a sorted batch (>= 8192 binaries) at N = 256; second call:
    # a smaller, unsorted batch at N = 128 reusing the same scratch
    for B, N in ((20000, 256), (3000, 128)):
        print(f"noalloc {precision} B={B} N={N}", flush=True)
        p = torch.from_numpy(draw_params(B, rng)).to("cuda")
        dA, dE, S_A, S_E, di, ni, Ld = gb._prep_data([A, E], [SA, SE], None, None, B)
        dat = _capi.Data(A=dA.data_ptr(), E=dE.data_ptr(), psdA=S_A.data_ptr(), psdE=S_E.data_ptr(),
                         data_index=di.data_ptr(), noise_index=ni.data_ptr(), data_length=int(Ld),
                         start_freq_ind=0, df=float(1 / YR_ADJ))
        ref = _call(lib.gbgpu_wave_ll, gb, p, N, dat, d_d, precision, torch)
        got = _call(lib.gbgpu_wave_ll_noalloc, gb, p, N, dat, d_d, precision, torch)
        for r, g in zip(ref, got):
            assert np.all(np.isfinite(g))
            assert np.array_equal(r, g)


def test_noalloc_rejects_oversized_batch():
    from gbgpu_amd import GBGPU, _capi

    lib = _capi.load()
    gb = GBGPU()
    cfg, _tab = gb._cfg(256, YR_ADJ, False, "fp64")
    P = 8  # never dereferenced: the size check comes first
    dat = _capi.Data(A=P, E=P, psdA=P, psdE=P, data_index=None, noise_index=None, data_length=10, start_freq_ind=0,
                     df=1e-7)
    rc = lib.gbgpu_wave_ll_noalloc(P, None, None, P, 9, _capi.NOALLOC_MAX_BIN + 1, cfg, dat, 0.0, 0, None)
    assert rc == -1 and b"GBGPU_NOALLOC_MAX_BIN" in lib.gbgpu_last_error()
