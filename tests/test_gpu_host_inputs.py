"""Host-side input handling that needs a GPU: host tensors copied before a
launch, index validation, and the per-binary-N path on 2-D streams."""

import numpy as np
import pytest

from oracle import fastgb as F
from tests.helpers import YR, draw_params, make_data

pytestmark = pytest.mark.gpu

DT = 15.0


def test_fill_global_template_accepts_host_tensors():
    """CPU torch tensors for A/E are moved to the templates' device (never
    passed to a kernel as host pointers); the result equals the numpy input's."""
    import torch

    from gbgpu_amd import GBGPU

    rng = np.random.default_rng(41)
    B, N, L = 20, 128, 30_000
    p = draw_params(B, rng)
    o = F.run_wave(*p, N=N, T=YR, dt=DT)
    k0 = int(o["start_inds"].min()) - 100
    st = o["start_inds"] - k0
    gb = GBGPU()
    out = []
    for conv in (lambda x: x, torch.from_numpy):
        tm = torch.zeros((1, 2, L), dtype=torch.complex128, device="cuda")
        gb.fill_global_template(np.zeros(B, np.int32), tm, conv(o["AETf"][:, 0].copy()),
                                conv(o["AETf"][:, 1].copy()), st)
        out.append(tm.cpu().numpy())
    assert np.array_equal(out[0], out[1])
    ref = F.fill_global(1, L, o["AETf"][:, 0], o["AETf"][:, 1], st, np.zeros(B, int))
    assert np.array_equal(out[0], ref)


def test_negative_stream_index_rejected():
    from gbgpu_amd import GBGPU

    rng = np.random.default_rng(42)
    p = draw_params(4, rng)
    gb = GBGPU()
    gb.d_d = 0.0
    d = [np.zeros((2, 5000), complex)] * 2
    S = [np.ones((2, 5000))] * 2
    with pytest.raises(ValueError, match="non-negative"):
        gb.get_ll(p, d, S, N=128, T=YR, dt=DT, data_index=np.array([0, 1, -1, 0], np.int32))


def test_per_binary_N_on_2d_streams_uses_stream_0_and_clips():
    """Per-binary N with 2-D data and no indices: every binary scores against
    stream 0 (the reference's default indices) and a window reaching past
    data_length is clipped there, exactly as the uniform path does."""
    from gbgpu_amd import GBGPU

    rng = np.random.default_rng(43)
    T = YR
    Tadj = F.adjust_T(T, DT)
    B = 16
    p = draw_params(B, rng)
    N = np.where(np.arange(B) % 2 == 0, 128, 256).astype(np.int32)
    A, E, SA, SE, df, Lf = make_data(p[:, :4], 256, T, DT, rng)
    k0 = 60_000
    L = int(np.rint(p[1].max() * Tadj)) - k0 + 40  # the highest binary straddles the top of stream 0
    sl = slice(k0, k0 + L)
    dA = np.stack([A[sl], rng.standard_normal(L) * 1e-20 + 0j])
    dE = np.stack([E[sl], rng.standard_normal(L) * 1e-20 + 0j])
    S2 = [np.stack([SA[sl], SA[sl] * 2]), np.stack([SE[sl], SE[sl] * 2])]
    gb = GBGPU()
    gb.d_d = 1.0
    ll = gb.get_ll(p, [dA, dE], S2, N=N, T=T, dt=DT, start_freq_ind=k0)
    ref = np.empty(B)
    for n in np.unique(N):
        sel = N == n
        ref[sel] = gb.get_ll(p[:, sel], [dA, dE], S2, N=int(n), T=T, dt=DT, start_freq_ind=k0)
    assert np.array_equal(ll, ref)
