"""Writes tests/golden/fastgb_small.npz: small input/output vectors of the CPU
FP64 oracle (oracle/fastgb.py), frozen so that

  * tests/test_golden.py can check that the oracle still reproduces them
    (the oracle itself is pinned by tests/test_oracle.py and
    tests/test_oracle_physics.py), and
  * the GPU parity tests can compare the HIP path with them directly.

The reference's own tests hold no numeric vectors (they assert finiteness
only), and its Python cannot be imported here (lisatools and
gpubackendtools are absent), so the vectors come from the oracle.

Run from the repository root:  python tests/golden/make_fixtures.py
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import fastgb as F  # noqa: E402
from tests.helpers import YR, draw_params  # noqa: E402

OUT = Path(__file__).resolve().parent / "fastgb_small.npz"
DT = 15.0


def main():
    rng = np.random.default_rng(20261204)
    out = {"dt": np.float64(DT), "T_in": np.float64(YR)}

    # circular binaries over 0.3-30 mHz (every transfer-series tier), N = 128
    p = draw_params(6, rng, f_lo=3e-4, f_hi=3e-2, log_f=True)
    o = F.run_wave(*p, N=128, T=YR, dt=DT)
    out.update(circ_params=p, circ_N=np.int64(128), circ_XYZf=o["XYZf"], circ_AETf=o["AETf"],
               circ_start=o["start_inds"])

    # third body (A2, varpi, e2, P2, T2 as GBGPUThirdBody takes them), N = 256
    p = draw_params(4, rng)
    third = np.array([[40.0, 80.0, 15.0, 60.0], [0.3, 1.9, 4.0, 5.5], [0.0, 0.3, 0.0, 0.6],
                      [1.5, 3.0, 7.0, 2.2], [0.1, 0.7, -0.4, 1.3]])
    o = F.run_wave(*p, *third, N=256, T=YR, dt=DT)
    out.update(third_params=np.concatenate([p, third]), third_N=np.int64(256), third_XYZf=o["XYZf"],
               third_start=o["start_inds"])

    # second-generation TDI factor, N = 128
    p = draw_params(3, rng)
    o = F.run_wave(*p, N=128, T=YR, dt=DT, tdi2=True)
    out.update(tdi2_params=p, tdi2_N=np.int64(128), tdi2_XYZf=o["XYZf"], tdi2_start=o["start_inds"])

    # likelihood against a data window: bins [s_lo, s_lo + Lw) of a 1-yr A/E
    # stream holding noise plus three injected binaries, start_freq_ind = s_lo
    T = F.adjust_T(YR, DT)
    df = 1.0 / T
    s_lo, Lw, N = 63_000, 8192, 256
    inj = draw_params(3, rng, f_lo=2.02e-3, f_hi=2.18e-3)
    o = F.run_wave(*inj, N=N, T=YR, dt=DT)
    glob = F.fill_global(1, Lw, o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"] - s_lo, np.zeros(3, int))
    from gbgpu_amd.noise import A1TDISens, noise_realisation

    S = A1TDISens.get_Sn((np.arange(Lw) + s_lo) * df)
    A = glob[0, 0] + noise_realisation(S, df, rng)
    E = glob[0, 1] + noise_realisation(S, df, rng)
    d_d = 4 * df * (np.sum(np.abs(A) ** 2 / S) + np.sum(np.abs(E) ** 2 / S))
    p = draw_params(6, rng, f_lo=2.0e-3, f_hi=2.2e-3)
    p[:, :3] = inj  # the injected sources themselves, then three others
    o = F.run_wave(*p, N=N, T=YR, dt=DT)
    d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"] - s_lo, A, E, S, S, df)
    out.update(ll_params=p, ll_N=np.int64(N), ll_start_freq_ind=np.int64(s_lo), ll_A=A, ll_E=E, ll_S=S,
               ll_d_d=np.float64(d_d), ll_d_h=d_h, ll_h_h=h_h, ll_loglike=F.loglike(d_d, d_h, h_h),
               ll_loglike_marg=F.loglike(d_d, d_h, h_h, phase_marginalize=True))
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
