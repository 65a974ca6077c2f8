"""Benchmark: FastGB waveform + log-likelihood throughput on one MI355X.

Workload (one "step" = one fused wave+logL call over the whole batch):
10^6 binaries, N = 256, Tobs = 4 yr, dt = 15 s, f0 ~ U[1, 6] mHz, amp
log-uniform [1e-23, 1e-21], chirp mass U[0.2, 1] Msun (fdot, fddot from it),
isotropic sky/orientation, scored against one 4,207,681-bin complex128 A/E
stream of analytic-PSD Gaussian noise plus 1000 injected binaries.
Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  ``value`` = binaries/s of the plain FP64 HIP path;
the mixed-precision path on the same batch is reported under ``mixed``.
``roofline`` is for the dominant kernel (the fused FP64 wave+logL kernel):
its duration comes from HIP events around the timed launches; its HBM
traffic and executed FLOPs come from two short rocprofv3 counter runs of this
script in a child process (``--counters-child``), outside the timed region,
scaled by what the same counters read for a 1 GiB device copy (null when
rocprofv3 is unavailable).  ``cpu_baseline`` times the numpy FP64 oracle on
a bounded sample of the same workload, on one host thread and on every host
core this process may use; it runs before anything touches the GPU.
Parity of the HIP path with the oracle is established by the test suite
(tests/test_gpu_parity.py, tests/test_gpu_golden.py); this script reports
the mixed path's differences from the FP64 HIP path.
"""

from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

YRSID_SI = 31558149.763545603
JULIAN_YEAR_S = 31557600.0  # the workload's "1 yr" (4 yr -> 4,207,681 bins at dt = 15 s)
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 30  # the calibration copy: 1 GiB read, 1 GiB written


def flop_per_binary(N: int) -> float:
    """Algorithmic FLOP per binary of wave+logL: N*(343 + 15 log2 N) + 200
    (slow part + TDI + A/E 313/sample, likelihood 30/sample, transform
    5 N log2 N per channel; FMA = 2 FLOP; transcendentals not counted)."""
    return N * (343 + 15 * np.log2(N)) + 200


def bytes_per_binary(N: int) -> float:
    """Fused wave+logL HBM bytes per binary: data A/E (32 B) + PSD A/E (16 B)
    per sample, plus 9 params, 1 output, indices (112 B)."""
    return 48 * N + 112


G_SI = 6.67408e-11
C_SI = 299792458.0
MSUN_SI = 1.98848e30


def draw_params(B, rng, f_lo=1e-3, f_hi=6e-3, log_f=False):
    """(9, B) seeded source parameters of the workloads: amp log-uniform in
    [1e-23, 1e-21]; f0 uniform (or log-uniform) in [f_lo, f_hi]; chirp mass
    U[0.2, 1] Msun with fdot = 96/5 pi^(8/3) (G Mc/c^3)^(5/3) f0^(11/3) and
    fddot = 11/3 fdot^2 / f0; phi0 U[0, 2 pi); cos iota U[-1, 1]; psi U[0, pi);
    lambda U[0, 2 pi); sin beta U[-1, 1]."""
    amp = 10 ** rng.uniform(-23, -21, B)
    f0 = 10 ** rng.uniform(np.log10(f_lo), np.log10(f_hi), B) if log_f else rng.uniform(f_lo, f_hi, B)
    Mc = rng.uniform(0.2, 1.0, B) * MSUN_SI
    fdot = (96.0 / 5.0) * np.pi ** (8 / 3) * (G_SI * Mc / C_SI**3) ** (5 / 3) * f0 ** (11 / 3)
    fddot = 11.0 / 3.0 * fdot**2 / f0
    phi0 = rng.uniform(0, 2 * np.pi, B)
    iota = np.arccos(rng.uniform(-1, 1, B))
    psi = rng.uniform(0, np.pi, B)
    lam = rng.uniform(0, 2 * np.pi, B)
    beta = np.arcsin(rng.uniform(-1, 1, B))
    return np.array([amp, f0, fdot, fddot, phi0, iota, psi, lam, beta])


def make_workload(B, N, T, dt, seed):
    rng = np.random.default_rng(seed)
    params = draw_params(B, rng, f_lo=1e-3, f_hi=6e-3)
    return params, rng


def build_data(gb, params, N, T, dt, rng, n_inject=1000):
    import torch

    from gbgpu_amd.noise import A1TDISens, noise_realisation

    Tadj = int(T / dt) * dt
    df = 1.0 / Tadj
    L = int(Tadj / dt) // 2 + 1
    f = np.arange(L) * df
    f[0] = df
    S = A1TDISens.get_Sn(f)
    dev = torch.device("cuda", 0)
    tmpl = torch.zeros((1, 2, L), dtype=torch.complex128, device=dev)
    gb.generate_global_template(params[:, :n_inject].T, np.zeros(n_inject, np.int32), tmpl, N=N, T=T, dt=dt)
    A = tmpl[0, 0] + torch.from_numpy(noise_realisation(S, df, rng)).to(dev)
    E = tmpl[0, 1] + torch.from_numpy(noise_realisation(S, df, rng)).to(dev)
    St = torch.from_numpy(S).to(dev)
    d_d = float((4 * df * (torch.sum(A.abs() ** 2 / St) + torch.sum(E.abs() ** 2 / St))).item())
    return A.contiguous(), E.contiguous(), St, St.clone(), d_d, df, L


def time_calls(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(steps):
        fn()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return ev0.elapsed_time(ev1) / 1e3 / steps, wall / steps


def _cpu_chunks(args):
    """Worker of the CPU baseline: the numpy FP64 oracle (run_wave + the
    brute-force get_ll) over chunks of the workload for budget_s seconds on
    one thread.  Returns the number of binaries done and the elapsed time."""
    N, T, dt, budget_s, chunk, seed = args
    try:
        from threadpoolctl import threadpool_limits

        threadpool_limits(1)
    except Exception:  # pragma: no cover
        pass
    from gbgpu_amd.noise import A1TDISens
    from oracle import fastgb as F

    rng = np.random.default_rng(seed)
    Tadj = int(T / dt) * dt
    df = 1 / Tadj
    L = int(Tadj / dt) // 2 + 1
    f = np.arange(L) * df
    f[0] = df
    S = A1TDISens.get_Sn(f)
    dA = (rng.standard_normal(L) + 1j * rng.standard_normal(L)) * np.sqrt(S / (8 * df))
    dE = (rng.standard_normal(L) + 1j * rng.standard_normal(L)) * np.sqrt(S / (8 * df))
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        p = draw_params(chunk, rng, f_lo=1e-3, f_hi=6e-3)
        o = F.run_wave(*p, N=N, T=T, dt=dt)
        d_h, h_h = F.inner_terms(o["AETf"][:, 0], o["AETf"][:, 1], o["start_inds"], dA, dE, S, S, df)
        F.loglike(1.0, d_h, h_h)
        done += chunk
    return done, time.perf_counter() - t0


def cpu_baseline(N, T, dt, budget_s=10.0, chunk=500):
    """The CPU path timed on this host: the numpy FP64 oracle on one thread,
    then one single-threaded worker per usable core (os.sched_getaffinity,
    at most 16) for the same budget.  Runs before the GPU is touched, so the
    worker processes are plain forks of a CPU-only interpreter."""
    import multiprocessing as mp

    os.environ.setdefault("OMP_NUM_THREADS", "1")
    one_done, one_el = _cpu_chunks((N, T, dt, budget_s, chunk, 1234))
    cores = min(16, len(os.sched_getaffinity(0)))
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_chunks, [(N, T, dt, budget_s, chunk, 1000 + i) for i in range(cores)])
    # the workers' timed intervals run concurrently (each after its own set-up)
    all_done = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    return dict(
        value=all_done / wall,
        unit="binaries/s",
        cores=cores,
        kind="port",
        sample=f"{all_done} binaries of the same workload (N={N}, 4 yr, f0 U[1,6] mHz) through the numpy FP64 "
        f"oracle run_wave + get_ll in chunks of {chunk}: {cores} single-threaded worker processes for "
        f"{wall:.1f} s",
        single_core={"value": one_done / one_el, "unit": "binaries/s", "cores": 1,
                     "sample": f"{one_done} binaries in {one_el:.1f} s on 1 host thread"},
    )


def mixed_vs_fp64(gb, params, A, E, SA, SE, N, T, dt, n=4096):
    """Differences of the mixed-precision path from the FP64 HIP path (which
    the test suite pins to the oracle) on a sample of the batch: log-L
    absolute difference and the largest per-binary relative L2 difference of
    the A and E spectra."""
    p = params[:, :n]
    out = {}
    ll = {}
    aet = {}
    for prec in ("fp64", "mixed"):
        ll[prec] = gb.get_ll(p, [A, E], [SA, SE], N=N, T=T, dt=dt, precision=prec)
        gb.run_wave(*p, N=N, T=T, dt=dt, precision=prec)
        aet[prec] = gb.AETf.cpu().numpy()
    werr = {}
    for c, name in enumerate("AE"):
        num = np.sqrt(np.sum(np.abs(aet["mixed"][:, c] - aet["fp64"][:, c]) ** 2, axis=1))
        den = np.sqrt(np.sum(np.abs(aet["fp64"][:, c]) ** 2, axis=1))
        werr[name] = float(np.max(num / den))
    out.update(n=n, max_abs_logL_diff=float(np.max(np.abs(ll["mixed"] - ll["fp64"]))),
               max_rel_wave_diff_per_channel=werr)
    return out


def uneven_batch(gb, A, E, SA, SE, d_d, B, T, dt, n_calls, seed=77):
    """Uneven-N batch: f0 log-uniform in [0.1, 20] mHz, N by the frequency
    rule (128/256/512/1024/2048 split at 1/3/6/10 mHz), a captured graph of
    n_calls likelihood calls at seeded 1e-3 relative jumps; FP64, mixed with
    the FP64 accuracy guard at 1e-3 (the default), and mixed without it."""
    import torch

    from gbgpu_amd.batch import BatchedLogLike
    from gbgpu_amd.utils.utility import baseline_N
    rng = np.random.default_rng(seed)
    p = draw_params(B, rng, f_lo=1e-4, f_hi=20e-3, log_f=True)
    N = baseline_N(p[1])
    res = {"num_bin": B, "calls": n_calls,
           "N_fractions": {int(n): float(np.mean(N == n)) for n in np.unique(N)},
           "mean_N": float(N.mean())}
    lls = {}
    # the persistent strategy is switched off in this build (its kernel took
    # a GPU memory fault whose cause has not been shown)
    res["persistent_strategy"] = "disabled"
    for prec, strat, guard in (("fp64", "buckets", None), ("mixed", "buckets", 1e-3),
                               ("mixed_unguarded", "buckets", None)):
        if True:
            print(f"[bench] uneven batch {prec} {strat}", file=sys.stderr, flush=True)
            bl = BatchedLogLike(gb, [A, E], [SA, SE], T, dt, precision=prec.split("_")[0], strategy=strat,
                                guard_tol=guard)
            shapes = bl.autotune(p) if (strat == "buckets" and prec in ("fp64", "mixed")) else None
            g, out = bl.capture_loop(p, n_calls, sigma=1e-3, seed=seed)
            g.replay()
            torch.cuda.synchronize()
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            reps = 2
            ev0.record()
            for _ in range(reps):
                g.replay()
            ev1.record()
            torch.cuda.synchronize()
            sec = ev0.elapsed_time(ev1) / 1e3 / reps
            r = {"binaries_per_s": n_calls * B / sec, "ms_per_call": sec / n_calls * 1e3}
            if shapes is not None:
                r["autotuned_tile_shape_per_N"] = shapes
            if guard:
                r["guard_tol"] = guard
                r["fp64_reevaluated_per_call"] = sum(bl.guard_counts().values())
            if strat == "buckets":
                res[prec] = r
                lls[prec] = out.clone()
            else:
                res[prec]["persistent_strategy"] = r
            del g, out, bl
            torch.cuda.empty_cache()
    res["speedup_mixed_vs_fp64"] = res["mixed"]["binaries_per_s"] / res["fp64"]["binaries_per_s"]
    res["speedup_mixed_unguarded_vs_fp64"] = (res["mixed_unguarded"]["binaries_per_s"] /
                                              res["fp64"]["binaries_per_s"])
    for key in ("mixed", "mixed_unguarded"):
        d = (lls[key] - lls["fp64"]).abs()
        res[f"logL_abs_err_{key}_vs_fp64"] = {"max": float(d.max()), "median": float(d.median()),
                                             "frac_le_1e-3": float((d <= 1e-3).double().mean())}
    # per-channel waveform difference of the mixed path from the FP64 HIP
    # path (pinned to the oracle by the tests), on a sample of every N bucket
    # (unperturbed parameters); the mixed path has no T channel
    errs = {"A": 0.0, "E": 0.0}
    for n in np.unique(N):
        sel = np.flatnonzero(N == n)[:256]
        aet = {}
        for prec in ("fp64", "mixed"):
            gb.run_wave(*p[:, sel], N=int(n), T=T, dt=dt, precision=prec)
            aet[prec] = gb.AETf.cpu().numpy()
        for c, name in enumerate("AE"):
            num = np.sqrt(np.sum(np.abs(aet["mixed"][:, c] - aet["fp64"][:, c]) ** 2, axis=1))
            den = np.sqrt(np.sum(np.abs(aet["fp64"][:, c]) ** 2, axis=1))
            errs[name] = max(errs[name], float(np.max(num / den)))
    res["mixed_max_rel_wave_diff_vs_fp64_per_channel"] = errs
    return res


def differentiation(gb, A, E, SA, SE, d_d, T, dt, steps, warmup, B=100_000, seed=3):
    """Differentiation workload: 1e5 binaries, N = 256, 4 yr; adjoint
    gradient (9 sampling-basis parameters) and per-binary 9x9 Fisher
    matrices, timed separately; plus a Richardson central-difference check
    of the gradient on 8 binaries."""
    import torch

    from gbgpu_amd import _capi
    from gbgpu_amd.derivatives import FastGBLikelihood

    N = 256
    rng = np.random.default_rng(seed)
    p = draw_params(B, rng, f_lo=1e-3, f_hi=6e-3)
    like = FastGBLikelihood(gb, [A, E], [SA, SE], N=N, T=T, dt=dt)
    pd = torch.from_numpy(np.ascontiguousarray(p)).cuda()
    print("[bench] differentiation", file=sys.stderr, flush=True)
    fwd_cfg, _t = gb._cfg(N, like.T, False, "fp64")
    ll = torch.empty(B, dtype=torch.float64, device="cuda")
    lib = _capi.load()

    def fwd():
        _capi.check(lib.gbgpu_wave_ll(ll.data_ptr(), None, None, pd.data_ptr(), 9, B, fwd_cfg, like.data, like.d_d, 0,
                                      _capi.stream_handle()), "wave_ll")

    res = {"workload": "1e5 binaries, N=256, Tobs=4 yr, f0 U[1,6] mHz; gradient w.r.t. (ln amp, f0, fdot, fddot, "
                       "phi0, cos iota, psi, lam, sin beta); 9x9 Fisher per binary", "num_bin": B}
    for name, fn in (("forward", fwd), ("gradient", lambda: like.ll_and_grad(pd)), ("fisher", lambda: like.fisher(pd))):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        with _capi.KernelTimer() as kt:
            sec, _wall = time_calls(fn, steps, 0)
        res[name] = {"per_s": B / sec, "ms_per_call": sec * 1e3, "kernel_ms": kt.mean_ms}
    res["gradient_cost_vs_forward"] = res["forward"]["per_s"] / res["gradient"]["per_s"]
    # central-difference check on 8 binaries (steps moving the phase ~1e-2 rad)
    q = p[:, :8].copy()
    _, g = gb.get_ll_grad(q, [A, E], [SA, SE], N=N, T=T, dt=dt)
    F = gb.fisher_matrix(q, [SA, SE], N=N, T=T, dt=dt)
    K = np.array([np.log(q[0]), q[1], q[2], q[3], q[4], np.cos(q[5]), q[6], q[7], np.sin(q[8])])
    Ta = like.T
    h = np.array([1e-3, 1e-2 / (2 * np.pi * Ta), 1e-2 / (np.pi * Ta**2), 1e-2 / (np.pi / 3 * Ta**3), 1e-2, 1e-3,
                  1e-2, 1e-3, 1e-3])

    def from_K(Kp):
        return np.array([np.exp(Kp[0]), Kp[1], Kp[2], Kp[3], Kp[4], np.arccos(Kp[5]), Kp[6], Kp[7],
                         np.arcsin(Kp[8])])

    gfd = np.zeros_like(g)
    for k in range(9):
        D = []
        for m in (1.0, 0.5, 0.25):
            v = []
            for sg in (1, -1):
                Kp = K.copy()
                Kp[k] = Kp[k] + sg * m * h[k]
                v.append(gb.get_ll(from_K(Kp), [A, E], [SA, SE], N=N, T=T, dt=dt))
            D.append((v[0] - v[1]) / (2 * m * h[k]))
        gfd[:, k] = (16 * (4 * D[2] - D[1]) / 3 - (4 * D[1] - D[0]) / 3) / 15
    scale = np.maximum(np.abs(gfd), np.sqrt(np.abs(np.diagonal(F, axis1=1, axis2=2))))
    res["fd_check"] = {"n": 8, "max_rel_err": float(np.max(np.abs(g - gfd) / scale)),
                       "stencil": "Richardson central differences, steps moving the phase ~1e-2 rad"}
    return res


def dropin_get_ll(params, A, E, SA, SE, d_d, df, L, N, T, dt, steps, warmup, ll_fused):
    """The reference's two-step path on the GPU: run_wave materialises the A/E
    templates in its (N, B)-major layout, then the drop-in native get_ll
    (gbgpu_get_ll; src/gbgpu/cutils/gbgpu_utils.cu:125-210) reduces them
    against the data and PSD windows.  HBM-bound: 80 N + 44 bytes per binary
    (templates 32 N, data 32 N, PSD 16 N; start index, outputs)."""
    import torch

    from gbgpu_amd import GBGPU, _capi

    print("[bench] drop-in get_ll", file=sys.stderr, flush=True)
    B = params.shape[1]
    gbw = GBGPU()
    gbw.run_wave(*params, N=N, T=T, dt=dt)
    A_t, E_t = gbw.A_out, gbw.E_out
    start = torch.as_tensor(gbw.start_inds, device=A.device).to(torch.int32)
    d_h = torch.zeros(B, dtype=torch.complex128, device=A.device)
    h_h = torch.zeros_like(d_h)
    lib = _capi.load()
    st = _capi.stream_handle()

    def call():
        rc = lib.gbgpu_get_ll(d_h.data_ptr(), h_h.data_ptr(), A_t.data_ptr(), E_t.data_ptr(), A.data_ptr(),
                              E.data_ptr(), SA.data_ptr(), SE.data_ptr(), df, start.data_ptr(), N, B, None, None,
                              L, st)
        _capi.check(rc, "gbgpu_get_ll")

    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    with _capi.KernelTimer() as kt:
        sec, _ = time_calls(call, steps, 0)
    ll = -0.5 * (d_d + h_h.real - 2.0 * d_h.real)
    err = float((ll - ll_fused).abs().max())
    kernel_s = kt.mean_ms / 1e3
    nbytes = (80.0 * N + 44.0) * B
    out = dict(
        workload=f"gbgpu_get_ll on run_wave templates: {B} binaries, N={N}, (N, B)-major A/E",
        binaries_per_s=B / sec,
        ms_per_call=sec * 1e3,
        kernel="k_get_ll",
        kernel_ms=kernel_s * 1e3,
        bytes_per_binary=80.0 * N + 44.0,
        algorithmic_GB_per_s=nbytes / kernel_s / 1e9,
        hbm_peak_GB_per_s=HBM_PEAK_GBS,
        frac=nbytes / kernel_s / 1e9 / HBM_PEAK_GBS,
        max_abs_logL_diff_vs_fused=err,
    )
    del gbw, A_t, E_t, d_h, h_h
    torch.cuda.empty_cache()
    return out


def global_and_swap(gb, SA, SE, T, dt, B=10_000_000, P=100_000, chunk_samples=1 << 28, seed=11):
    """Global template of 1e7 binaries (f0 log-uniform 0.1-10 mHz, N by the
    frequency rule) summed deterministically into one 4-yr A/E pair, twice
    (bit-identical check); then swap_likelihood_difference for 1e5 proposals
    (remove binary, add a 1e-3 relative perturbed copy) against a noise
    residual, circular and third-body (circular orbit: A2 U[0, 100] m/s,
    P2 log-uniform 1-10 yr, phase U[0, 2 pi))."""
    import torch

    from gbgpu_amd import GBGPUThirdBody, _capi
    from gbgpu_amd.noise import noise_realisation
    from gbgpu_amd.utils.utility import baseline_N

    lib = _capi.load()
    st = _capi.stream_handle()
    rng = np.random.default_rng(seed)
    p = draw_params(B, rng, f_lo=1e-4, f_hi=1e-2, log_f=True)
    Nb = baseline_N(p[1])
    Tadj = int(T / dt) * dt
    L = SA.numel()
    dev = SA.device
    res = {"workload": "fill_global_template of 1e7 binaries (f0 log-uniform 0.1-10 mHz, N 128-1024 by frequency) "
                       "into one 4-yr A/E pair; swap of 1e5 proposals, circular and third-body",
           "num_bin": B, "N_fractions": {int(n): float(np.mean(Nb == n)) for n in np.unique(Nb)}}
    buckets = [(int(n), np.flatnonzero(Nb == n)) for n in np.unique(Nb)]
    # chunks of chunk_samples template samples per N bucket (C_n = chunk_samples / n binaries)
    minN = min(n for n, _ in buckets)
    Cmax = max(1, chunk_samples // minN)
    aet = torch.empty(2 * chunk_samples, dtype=torch.complex128, device=dev)
    starts = torch.empty(Cmax, dtype=torch.int64, device=dev)
    gidx = torch.zeros(Cmax, dtype=torch.int32, device=dev)
    pd_all = torch.from_numpy(np.ascontiguousarray(p)).to(dev)

    def generate():
        total = torch.zeros((2, L), dtype=torch.complex128, device=dev)
        fill_ms, fill_n, tmpl_bins = 0.0, 0, 0
        for n, idx in buckets:
            cfg, _tab = gb._cfg(n, Tadj, False, "fp64")
            C = max(1, chunk_samples // n)
            for c0 in range(0, len(idx), C):
                sel = torch.from_numpy(idx[c0:c0 + C]).to(dev)
                pc = pd_all[:, sel].contiguous()
                m = pc.shape[1]
                a = aet[: m * 2 * n].view(m, 2, n)
                _capi.check(lib.gbgpu_run_wave_ae(a.data_ptr(), starts.data_ptr(), pc.data_ptr(), 9, m, cfg, st),
                            "run_wave_ae")
                # the gather-sum adds into the (zeroed) global arrays; chunks and
                # buckets always come in the same order, so the sum is reproducible
                with _capi.KernelTimer() as kt:
                    _capi.check(lib.gbgpu_fill_global_rows(total[0].data_ptr(), total[1].data_ptr(),
                                                           a[:, 0].data_ptr(), a[:, 1].data_ptr(), 2 * n,
                                                           starts.data_ptr(), 0, n, m, gidx.data_ptr(), L, 1, st),
                                "fill_global_rows")
                fill_ms += kt.total_ms
                fill_n += kt.count
                tmpl_bins += m * n
        return total, fill_ms, tmpl_bins

    print("[bench] global template 1e7", file=sys.stderr, flush=True)
    generate()  # warm (twiddles, tables, allocator)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tot1, fill_ms, tmpl_bins = generate()
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    tot2, _, _ = generate()
    torch.cuda.synchronize()
    # fill traffic per chunk: template A/E rows read (32 B/bin), start bins
    # (8 B/row), and the touched part of the A/E output read and written
    # (64 B/bin, counted once per bucket over the whole stream: an upper bound)
    fill_bytes = 32.0 * tmpl_bins + 8.0 * B + 64.0 * L * len(buckets)
    res["generate"] = {"seconds": t_gen, "binaries_per_s": B / t_gen}
    res["fill"] = {"kernel_ms_total": fill_ms, "algorithmic_GB_per_s": fill_bytes / (fill_ms / 1e3) / 1e9,
                   "hbm_peak_GB_per_s": HBM_PEAK_GBS, "template_bins": tmpl_bins}
    res["bit_identical_repeat"] = bool(torch.equal(tot1, tot2))
    del aet, tot2
    torch.cuda.empty_cache()

    # swap against a noise residual
    df = 1.0 / Tadj
    r_A = torch.from_numpy(noise_realisation(SA.cpu().numpy(), df, rng)).to(dev)
    r_E = torch.from_numpy(noise_realisation(SE.cpu().numpy(), df, rng)).to(dev)
    pr = p[:, :P].copy()
    pa = pr * (1.0 + 1e-3 * rng.standard_normal(pr.shape))
    NP = Nb[:P]
    dat = _capi.Data(A=r_A.data_ptr(), E=r_E.data_ptr(), psdA=SA.data_ptr(), psdE=SE.data_ptr(), data_index=None,
                     noise_index=None, data_length=L, start_freq_ind=0, df=df)
    gb3 = GBGPUThirdBody()
    A2 = rng.uniform(0, 100, P)
    P2 = 10 ** rng.uniform(0, 1, P)
    varpi = rng.uniform(0, 2 * np.pi, P)
    extra = np.array(gb3.prepare_additional_args(A2, varpi, np.zeros(P), P2, np.zeros(P)))
    for label, nparams in (("circular", 9), ("third_body", 14)):
        print(f"[bench] swap {label}", file=sys.stderr, flush=True)
        calls = []
        for n in np.unique(NP):
            sel = np.flatnonzero(NP == n)
            q_r, q_a = pr[:, sel], pa[:, sel]
            if nparams == 14:
                q_r = np.concatenate([q_r, extra[:, sel]])
                q_a = np.concatenate([q_a, extra[:, sel]])
            pairs = np.empty((nparams, 2 * len(sel)))
            pairs[:, 0::2] = q_r
            pairs[:, 1::2] = q_a
            cfg, tab = gb._cfg(int(n), Tadj, False, "fp64")
            calls.append((torch.from_numpy(pairs).to(dev), len(sel), cfg, tab,
                          torch.empty(len(sel), dtype=torch.float64, device=dev)))

        def swap_all():
            for pp, m, cfg, _tab, dll in calls:
                _capi.check(lib.gbgpu_swap_ll_diff(dll.data_ptr(), None, pp.data_ptr(), nparams, m, cfg, dat, st),
                            "swap")

        swap_all()
        torch.cuda.synchronize()
        with _capi.KernelTimer() as kt:
            sec, _ = time_calls(swap_all, 5, 0)
        # algorithmic HBM bytes per proposal: both binaries' data and PSD
        # windows (A, E: 48 B per bin each), both parameter columns and the result
        swap_bytes = float(np.sum(2 * 48.0 * NP + 2 * 8.0 * nparams + 8.0))
        kms = kt.total_ms / 5
        res[f"swap_{label}"] = {"proposals": P, "proposals_per_s": P / sec, "ms_per_pass": sec * 1e3,
                                "kernel_ms_per_pass": kms,
                                "algorithmic_GB_per_s": swap_bytes / (kms / 1e3) / 1e9,
                                "hbm_peak_GB_per_s": HBM_PEAK_GBS}
    return res


# kernel-name prefixes of the fused likelihood at N = 256 with the LDS
# transform; the template's trailing arguments (transform variant) follow
K_FP64 = "k_wave<256, 1, false, gbgpu::F64"
K_MIXED = "k_wave<256, 1, false, gbgpu::Mixed"


def _is_kernel(full_name, prefix):
    """full_name is the launch of `prefix` with the LDS transform (its last
    template argument, if any, is 0)."""
    i = full_name.find(prefix)
    if i < 0:
        return False
    rest = full_name[i + len(prefix):]
    return rest.startswith(">") or rest.startswith(", 0>")


def counters_child(args):
    """--counters-child: the launches the counter runs read.  A 1 GiB
    device-to-device elementwise copy (the calibration), then two launches
    each of the fused FP64 and mixed kernels on the default workload."""
    import torch

    from gbgpu_amd import GBGPU, _capi

    src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.mul(src, 1.0, out=dst)  # reads 1 GiB, writes 1 GiB
    torch.cuda.synchronize()
    del src, dst
    B, N, dt = args.num_bin, 256, 15.0
    T = 4 * JULIAN_YEAR_S
    params, rng = make_workload(B, N, T, dt, seed=2024)
    gb = GBGPU()
    A, E, SA, SE, d_d, df, L = build_data(gb, params, N, T, dt, rng)
    p_dev = torch.from_numpy(np.ascontiguousarray(params)).cuda()
    ll = torch.empty(B, dtype=torch.float64, device="cuda")
    dat = _capi.Data(A=A.data_ptr(), E=E.data_ptr(), psdA=SA.data_ptr(), psdE=SE.data_ptr(), data_index=None,
                     noise_index=None, data_length=L, start_freq_ind=0, df=df)
    lib = _capi.load()
    for prec in ("fp64", "mixed"):
        cfg, _tab = gb._cfg(N, int(T / dt) * dt, False, prec)
        for _ in range(2):
            _capi.check(lib.gbgpu_wave_ll(ll.data_ptr(), None, None, p_dev.data_ptr(), 9, B, cfg, dat, d_d, 0,
                                          _capi.stream_handle()), "gbgpu_wave_ll")
        torch.cuda.synchronize()


def _read_counter_csv(d):
    """{(dispatch, kernel name): {counter: value summed over instances}}."""
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    out = {}
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                key = (int(r["Dispatch_Id"]), r["Kernel_Name"])
                c = out.setdefault(key, {})
                c[r["Counter_Name"]] = c.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def _run_counters(counters, num_bin, timeout=240):
    exe = shutil.which("rocprofv3")
    if exe is None:
        raise RuntimeError("rocprofv3 not found")
    d = tempfile.mkdtemp(prefix="gbgpu_pmc_", dir="/tmp")
    cmd = [exe, "--pmc", *counters, "-d", d, "-o", "run", "--output-format", "csv", "--", sys.executable,
           os.path.abspath(__file__), "--counters-child", "--num-bin", str(num_bin)]
    proc = subprocess.Popen(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), stdout=subprocess.PIPE,
                            stderr=subprocess.STDOUT, start_new_session=True)
    try:
        log, _ = proc.communicate(timeout=timeout)
    except subprocess.TimeoutExpired:
        os.killpg(proc.pid, signal.SIGKILL)
        proc.communicate()
        raise RuntimeError(f"counter run {counters} timed out")
    rows = _read_counter_csv(d)
    shutil.rmtree(d, ignore_errors=True)
    if proc.returncode != 0 and not rows:
        raise RuntimeError(f"counter run {counters} failed: {log[-400:]!r}")
    return rows


def measure_counters(num_bin):
    """HBM traffic and executed FLOPs of the fused kernels from two counter
    runs (FETCH_SIZE with the VALU FLOP counters, then WRITE_SIZE), each
    scaled by what it reads for the 1 GiB calibration copy."""
    fetch = _run_counters(["FETCH_SIZE", "SQ_INSTS_VALU_FLOPS_FP64", "SQ_INSTS_VALU_FLOPS_FP32", "SQ_WAVES"],
                          num_bin)
    write = _run_counters(["WRITE_SIZE"], num_bin)

    def copy_value(rows, counter):
        cand = [c[counter] for (_d, k), c in rows.items() if "elementwise" in k and counter in c]
        return max(cand)

    def kernel_mean(rows, name, counter):
        v = [c[counter] for (_d, k), c in rows.items() if _is_kernel(k, name) and counter in c]
        if not v:
            raise RuntimeError(f"no {counter} rows for {name}")
        return sum(v) / len(v)

    f_copy = copy_value(fetch, "FETCH_SIZE") * 1024.0
    w_copy = copy_value(write, "WRITE_SIZE") * 1024.0
    ff, wf = COPY_BYTES / f_copy, COPY_BYTES / w_copy
    out = {"copy_fetch_factor": ff, "copy_write_factor": wf}
    for name, key in ((K_FP64, "fp64"), (K_MIXED, "mixed")):
        waves = kernel_mean(fetch, name, "SQ_WAVES")
        out[key] = {
            "traffic_bytes": kernel_mean(fetch, name, "FETCH_SIZE") * 1024.0 * ff
            + kernel_mean(write, name, "WRITE_SIZE") * 1024.0 * wf,
            # the FLOP counters count per wavefront instruction (FMA = 2):
            # times 64 lanes, per wavefront (one binary per wavefront at N = 256)
            "executed_fp64_flop_per_binary": kernel_mean(fetch, name, "SQ_INSTS_VALU_FLOPS_FP64") / waves * 64,
            "executed_fp32_flop_per_binary": kernel_mean(fetch, name, "SQ_INSTS_VALU_FLOPS_FP32") / waves * 64,
        }
    return out


def fma_probe(precision):
    """Attainable FMA rate (TFLOP/s): the library's 8-chain FMA kernel, FP64
    (precision 0) or FP32 without packed forms (1)."""
    import ctypes

    from gbgpu_amd import _capi

    out = ctypes.c_double(0.0)
    _capi.check(_capi.load().gbgpu_fma_probe(int(precision), ctypes.byref(out), _capi.stream_handle()),
                "gbgpu_fma_probe")
    return out.value


def bw_probe(mode):
    """Attainable HBM streaming rate (GB/s): the library's probe over a 1 GiB
    buffer, one 16-byte non-temporal access per thread; mode 0 reads it, mode
    1 copies it (read + written bytes)."""
    import ctypes

    from gbgpu_amd import _capi

    out = ctypes.c_double(0.0)
    _capi.check(_capi.load().gbgpu_bw_probe(int(mode), ctypes.byref(out), _capi.stream_handle()), "gbgpu_bw_probe")
    return out.value


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # the GPU clock settles over the first ~8 launches of a heavy kernel
    # (5.9 -> 4.95 ms per launch at the default workload): warm up past that
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--num-bin", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=8.0, help="seconds per CPU-baseline leg")
    ap.add_argument("--no-counters", action="store_true", help="skip the rocprofv3 traffic/FLOP counter runs")
    ap.add_argument("--counters-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None, help="time one precision only")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-N batch")
    ap.add_argument("--uneven-calls", type=int, default=100, help="calls in the captured uneven-N loop")
    ap.add_argument("--no-deriv", action="store_true", help="skip the differentiation workload")
    ap.add_argument("--no-global", action="store_true", help="skip the global-template and swap workload")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in run_wave + native get_ll path")
    ap.add_argument("--no-transforms", action="store_true", help="skip the matrix-core transform variants")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.counters_child:
        counters_child(args)
        return
    B, N, dt = args.num_bin, 256, 15.0
    T = 4 * JULIAN_YEAR_S

    # the CPU path first, before anything touches the GPU
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(N, T, dt, budget_s=args.cpu_budget)

    import torch

    from gbgpu_amd import GBGPU, _capi

    print(f"[bench] workload B={B} N={N} T=4yr", file=sys.stderr, flush=True)
    params, rng = make_workload(B, N, T, dt, seed=2024)
    gb = GBGPU()
    A, E, SA, SE, d_d, df, L = build_data(gb, params, N, T, dt, rng)
    gb.d_d = d_d
    dev = A.device
    p_dev = torch.from_numpy(np.ascontiguousarray(params)).to(dev)
    ll = torch.empty(B, dtype=torch.float64, device=dev)
    lib = _capi.load()
    Tadj = int(T / dt) * dt

    def runner(precision, transform="lds"):
        cfg, _tab = gb._cfg(N, Tadj, False, precision, _capi.TRANSFORMS[transform])
        dat = _capi.Data(A=A.data_ptr(), E=E.data_ptr(), psdA=SA.data_ptr(), psdE=SE.data_ptr(), data_index=None,
                         noise_index=None, data_length=L, start_freq_ind=0, df=df)
        st = _capi.stream_handle()

        def call():
            rc = lib.gbgpu_wave_ll(ll.data_ptr(), None, None, p_dev.data_ptr(), 9, B, cfg, dat, d_d, 0, st)
            _capi.check(rc, "gbgpu_wave_ll")

        return call, _tab

    call64, _t64 = runner("fp64")
    print("[bench] timing fp64", file=sys.stderr, flush=True)
    if args.only == "mixed":
        sec64, wall64, kernel_s = float("nan"), float("nan"), float("nan")
    else:
        for _ in range(args.warmup):
            call64()
        torch.cuda.synchronize()
        with _capi.KernelTimer() as kt64:
            sec64, wall64 = time_calls(call64, args.steps, 0)
        kernel_s = kt64.mean_ms / 1e3
    value = B / sec64
    flops = flop_per_binary(N) * B
    achieved_tflops = flops / kernel_s / 1e12

    # the FP64 policy's matrix-core transform (same workload, same arithmetic
    # type), against the LDS transform timed above
    fp64_transforms = None
    if args.only != "mixed" and not args.no_transforms:
        print("[bench] timing fp64 transform=mfma_f64", file=sys.stderr, flush=True)
        call64()
        torch.cuda.synchronize()
        ll_lds = ll.cpu().numpy()
        callf, _tf = runner("fp64", "mfma_f64")
        for _ in range(max(1, args.warmup)):
            callf()
        torch.cuda.synchronize()
        with _capi.KernelTimer() as ktf:
            secf, _ = time_calls(callf, args.steps, 0)
        fp64_transforms = {"mfma_f64": dict(
            value=B / secf, ms_per_step=secf * 1e3, kernel_ms=ktf.mean_ms, speedup_vs_lds_transform=sec64 / secf,
            max_abs_logL_diff_vs_lds=float(np.max(np.abs(ll.cpu().numpy() - ll_lds))),
            kernel="k_wave<256, MODE_LL, circular, F64, GBGPU_TRANSFORM_MFMA_F64>")}

    mixed = None
    if args.only != "fp64":
        callm, _tm = runner("mixed")
        for _ in range(max(1, args.warmup)):
            callm()
        torch.cuda.synchronize()
        print("[bench] timing mixed", file=sys.stderr, flush=True)
        with _capi.KernelTimer() as ktm:
            secm, _ = time_calls(callm, args.steps, 0)
        mixed = dict(value=B / secm, ms_per_step=secm * 1e3, kernel_ms=ktm.mean_ms, speedup_vs_fp64=sec64 / secm,
                     kernel="k_wave<256, MODE_LL, circular, Mixed>")
        if args.only is None:
            mixed["vs_fp64_hip"] = mixed_vs_fp64(gb, params, A, E, SA, SE, N, T, dt)
        if not args.no_transforms:
            ll64 = None
            if args.only is None:
                call64()
                torch.cuda.synchronize()
                ll64 = ll.cpu().numpy()
            mixed["transforms"] = {}
            for tr in ("mfma_f32", "mfma_bf16x3", "mfma_bf16x6"):
                print(f"[bench] timing mixed transform={tr}", file=sys.stderr, flush=True)
                callt, _tt = runner("mixed", tr)
                for _ in range(max(1, args.warmup)):
                    callt()
                torch.cuda.synchronize()
                with _capi.KernelTimer() as ktt:
                    sect, _ = time_calls(callt, args.steps, 0)
                r = dict(value=B / sect, kernel_ms=ktt.mean_ms, speedup_vs_lds_transform=secm / sect)
                if ll64 is not None:
                    r["max_abs_logL_diff_vs_fp64_hip"] = float(np.max(np.abs(ll.cpu().numpy() - ll64)))
                    gb.transform = tr
                    r["vs_fp64_hip"] = mixed_vs_fp64(gb, params, A, E, SA, SE, N, T, dt)
                    gb.transform = "lds"
                mixed["transforms"][tr] = r

    print("[bench] bandwidth and FMA probes", file=sys.stderr, flush=True)
    read_gbs = bw_probe(0)
    copy_gbs = bw_probe(1)
    fma64 = fma_probe(0)
    fma32 = fma_probe(1)

    dropin = None
    if not args.no_dropin and args.only is None:
        call64()
        torch.cuda.synchronize()
        dropin = dropin_get_ll(params, A, E, SA, SE, d_d, df, L, N, T, dt, args.steps, args.warmup, ll.clone())
        # read-dominated (templates, data and PSD in; two numbers per binary out)
        dropin["attainable_GB_per_s"] = read_gbs
        dropin["frac_of_attainable"] = dropin["algorithmic_GB_per_s"] / read_gbs

    uneven = None
    if not args.no_uneven:
        uneven = uneven_batch(gb, A, E, SA, SE, d_d, args.num_bin, T, dt, args.uneven_calls)

    deriv = None
    if not args.no_deriv:
        deriv = differentiation(gb, A, E, SA, SE, d_d, T, dt, args.steps, args.warmup)

    glob_res = None
    if not args.no_global:
        glob_res = global_and_swap(gb, SA, SE, T, dt)
        # the gather-sum reads the template rows and writes the (small) output
        glob_res["fill"]["attainable_read_GB_per_s"] = read_gbs
        glob_res["fill"]["frac_of_attainable"] = glob_res["fill"]["algorithmic_GB_per_s"] / read_gbs

    del A, E, SA, SE, p_dev, ll, gb
    torch.cuda.empty_cache()
    counters, counters_note = None, None
    if not args.no_counters:
        print("[bench] counter runs (rocprofv3, two passes)", file=sys.stderr, flush=True)
        try:
            counters = measure_counters(B)
        except Exception as e:  # no profiler on this host, or it failed: traffic stays null
            counters_note = f"counter runs unavailable: {str(e)[:200]}"

    roof = {
        "bound": "valu",
        "compute": "FP64 vector ALU; MI355X's FP64 vector and matrix peaks are both 78.6 TFLOP/s",
        "achieved": achieved_tflops,
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved_tflops / FP64_PEAK_TFLOPS,
        "traffic": counters["fp64"]["traffic_bytes"] if counters else None,
        "traffic_note": ("HBM bytes per launch: rocprofv3 FETCH_SIZE x %.3f + WRITE_SIZE x %.3f, the factors "
                         "being 1 GiB over what each counter reads for a 1 GiB device copy; the algorithmic bytes "
                         "are 48 N + 112 per binary = %.3g, served mostly from L2 and the Infinity Cache"
                         % (counters["copy_fetch_factor"], counters["copy_write_factor"], bytes_per_binary(N) * B)
                         if counters else counters_note),
        "kernel": "k_wave<256, MODE_LL, circular, F64>",
        "kernel_ms": kernel_s * 1e3,
        "flop_per_binary": flop_per_binary(N),
        "algorithmic_hbm_GBs": bytes_per_binary(N) * B / kernel_s / 1e9,
    }
    roof["attainable_fma_TFLOPs"] = fma64
    roof["attainable_note"] = ("FP64 FMA rate of an 8-independent-chain kernel on this GPU (gbgpu_fma_probe); "
                               "FP32 without packed forms: %.1f TFLOP/s" % fma32)
    if counters:
        ex = counters["fp64"]["executed_fp64_flop_per_binary"]
        roof["executed_flop_per_binary"] = ex
        roof["executed_TFLOPs"] = ex * B / kernel_s / 1e12
        roof["executed_frac"] = roof["executed_TFLOPs"] / FP64_PEAK_TFLOPS
        roof["executed_frac_of_attainable"] = roof["executed_TFLOPs"] / fma64
    if mixed is not None and counters:
        km = mixed["kernel_ms"] / 1e3
        c = counters["mixed"]
        mixed["roofline"] = {
            "bound": "valu",
            "peak_fp32_TFLOPs": FP32_PEAK_TFLOPS,
            "peak_fp64_TFLOPs": FP64_PEAK_TFLOPS,
            "executed_fp32_flop_per_binary": c["executed_fp32_flop_per_binary"],
            "executed_fp64_flop_per_binary": c["executed_fp64_flop_per_binary"],
            "executed_fp32_TFLOPs": c["executed_fp32_flop_per_binary"] * B / km / 1e12,
            "executed_fp64_TFLOPs": c["executed_fp64_flop_per_binary"] * B / km / 1e12,
            "attainable_fp32_fma_TFLOPs": fma32,
            "traffic": c["traffic_bytes"],
        }

    out = {
        "metric": "binaries/sec (wave+logL), plain FP64 vs batched mixed, at <=1e-6 rel err",
        "value": value,
        "unit": "binaries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded parameters; analytic-PSD Gaussian noise plus 1000 injected binaries)",
        "config": {
            "workload": "plain FP64 fused wave+logL, uniform batch: 1e6 binaries, N=256, Tobs=4 yr, dt=15 s, "
            "f0 U[1,6] mHz, one 4,207,681-bin A/E stream",
            "num_bin": B,
            "N": N,
            "Tobs_s": Tadj,
            "data_length": L,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": "HIP vs oracle: tests/test_gpu_parity.py, tests/test_gpu_golden.py, tests/test_gpu_configs_4yr.py",
        "attainable_hbm_GB_per_s": {"read": read_gbs, "copy": copy_gbs,
                                    "note": "gbgpu_bw_probe: 1 GiB, one 16-byte non-temporal access per thread"},
        "fp64_transforms": fp64_transforms,
        "mixed": mixed,
        "uneven_batch": uneven,
        "dropin_get_ll": dropin,
        "differentiation": deriv,
        "global_template_and_swap": glob_res,
        "wall_ms_per_step": wall64 * 1e3,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
