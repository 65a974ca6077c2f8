"""Large batches: above 2^17 binaries the allocating entry points order by
start bin with the library's LSD radix sort, above 2^21 with rocPRIM's
onesweep (smaller batches merge-sort; below 8192 binaries nothing is
sorted).

* fused likelihood: each binary's arithmetic does not depend on the order
  the binaries run in, so one 300,000-binary call must equal, bit for bit,
  the same binaries scored in unsorted chunks of 4000;
* fill: 300,000 (LSD) and 2,200,000 (onesweep) synthetic rows summed into
  one 4-yr stream, against the serial C restatement of the reference's
  fill_global (gbgpu_utils.cu:104-122) to 1e-12 of the largest bin, and
  bit-identical on a repeat.
"""

import ctypes

import numpy as np
import pytest

from tests.helpers import YR, draw_params, make_data, oracle_c

pytestmark = pytest.mark.gpu

DT = 15.0
YR_ADJ = int(YR / DT) * DT
BIG = 300_000


def _wave_ll(lib, gb, p, N, dat, d_d, precision, torch):
    from gbgpu_amd import _capi

    B = p.shape[1]
    cfg, _tab = gb._cfg(N, YR_ADJ, False, precision)
    ll = torch.full((B,), np.nan, dtype=torch.float64, device="cuda")
    d_h = torch.full((B,), np.nan, dtype=torch.complex128, device="cuda")
    h_h = torch.full((B,), np.nan, dtype=torch.complex128, device="cuda")
    rc = lib.gbgpu_wave_ll(ll.data_ptr(), d_h.data_ptr(), h_h.data_ptr(), p.data_ptr(), p.shape[0], B, cfg, dat,
                           d_d, 0, _capi.stream_handle())
    _capi.check(rc, "gbgpu_wave_ll")
    return ll, d_h, h_h


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_large_sorted_batch_matches_unsorted_chunks(precision):
    import torch

    from gbgpu_amd import GBGPU, _capi

    lib = _capi.load()
    rng = np.random.default_rng(950 + (precision == "mixed"))
    N = 256
    inj = draw_params(64, rng, f_lo=1e-3, f_hi=6e-3)
    A, E, SA, SE, df, L = make_data(inj, N, YR, DT, rng)
    gb = GBGPU()
    d_d = float(4 * df * (np.sum(np.abs(A) ** 2 / SA) + np.sum(np.abs(E) ** 2 / SE)))
    p = torch.from_numpy(draw_params(BIG, rng, f_lo=1e-3, f_hi=6e-3)).to("cuda")
    p[:, :64] = torch.from_numpy(inj).to("cuda")  # a few loud templates among the quiet ones
    dA, dE, S_A, S_E, di, ni, Ld = gb._prep_data([A, E], [SA, SE], None, None, BIG)
    dat = _capi.Data(A=dA.data_ptr(), E=dE.data_ptr(), psdA=S_A.data_ptr(), psdE=S_E.data_ptr(),
                     data_index=di.data_ptr(), noise_index=ni.data_ptr(), data_length=int(Ld), start_freq_ind=0,
                     df=float(1 / YR_ADJ))
    print(f"large batch {precision} B={BIG} N={N}: one call", flush=True)
    whole = _wave_ll(lib, gb, p, N, dat, d_d, precision, torch)
    print(f"large batch {precision}: chunks of 4000", flush=True)
    parts = [_wave_ll(lib, gb, p[:, i:i + 4000].contiguous(), N, dat, d_d, precision, torch)
             for i in range(0, BIG, 4000)]
    torch.cuda.synchronize()
    for k, name in enumerate(("loglike", "d_h", "h_h")):
        a = whole[k].cpu().numpy()
        b = torch.cat([q[k] for q in parts]).cpu().numpy()
        assert np.isfinite(a).all(), name
        assert np.array_equal(a, b), name
    # the injected templates are loud: their logL is far above the quiet ones'
    ll = whole[0].cpu().numpy()
    assert np.median(ll[:64]) > np.median(ll[64:])


@pytest.mark.parametrize("B,M", [(BIG, 128), (2_200_000, 32)])
def test_fill_large_batch_matches_oracle(B, M):
    import torch

    from gbgpu_amd import _capi

    lib = _capi.load()
    c = oracle_c()
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    rng = np.random.default_rng(952 + B)
    L = 4_207_681
    start = rng.integers(-M // 2, L - M // 2, B).astype(np.int64)  # some rows cross either end
    rows_A = (rng.standard_normal((B, M)) + 1j * rng.standard_normal((B, M))).astype(np.complex128)
    rows_E = (rng.standard_normal((B, M)) + 1j * rng.standard_normal((B, M))).astype(np.complex128)
    tA_rows = torch.from_numpy(rows_A).to("cuda")
    tE_rows = torch.from_numpy(rows_E).to("cuda")
    st = torch.from_numpy(start).to("cuda")
    g0 = torch.zeros(B, dtype=torch.int32, device="cuda")
    runs = []
    for rep in range(2):
        print(f"fill large batch B={B} M={M} rep={rep}", flush=True)
        gA = torch.zeros(L, dtype=torch.complex128, device="cuda")
        gE = torch.zeros(L, dtype=torch.complex128, device="cuda")
        rc = lib.gbgpu_fill_global_rows(gA.data_ptr(), gE.data_ptr(), tA_rows.data_ptr(), tE_rows.data_ptr(), M,
                                        st.data_ptr(), 0, M, B, g0.data_ptr(), L, 1, _capi.stream_handle())
        _capi.check(rc, "gbgpu_fill_global_rows")
        torch.cuda.synchronize()
        runs.append((gA.cpu().numpy(), gE.cpu().numpy()))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    oA = np.zeros(L, np.complex128)
    oE = np.zeros(L, np.complex128)
    At = np.ascontiguousarray(rows_A.T)
    Et = np.ascontiguousarray(rows_E.T)
    c.oracle_fill_global(P(oA), P(oE), P(At), P(Et), P(start.astype(np.int32)), M, B,
                         P(np.zeros(B, np.int32)), L)
    hA, hE = runs[0]
    for h, o in ((hA, oA), (hE, oE)):
        assert np.abs(h - o).max() <= 1e-12 * np.abs(o).max()
