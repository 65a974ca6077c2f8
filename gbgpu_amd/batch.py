"""Uneven-N likelihood batches and their graph-captured sampler loop.

``BatchedLogLike`` keeps the data, PSD, per-N orbit tables and the library
workspace resident on the GPU and evaluates the fused waveform + likelihood
for a batch in which every binary has its own N (``gbgpu_wave_ll_batched``:
device-side N bucketing, then one launch per bucket).  ``run_loop``
evaluates a sequence of calls whose parameters are seeded Gaussian relative
jumps of a base point
(``gbgpu_perturb_params``; N is re-derived from the jumped f0 by the step
rule), either eagerly or as one captured hipGraph (``capture_loop``).
"""

from __future__ import annotations

import numpy as np

from . import _capi

DEFAULT_F_EDGES = (1e-3, 3e-3, 6e-3, 10e-3)
DEFAULT_N_LEVELS = (128, 256, 512, 1024, 2048)


def _torch():
    import torch

    return torch


class BatchedLogLike:
    """Resident uneven-N log-likelihood evaluator.

    Args:
        gb: a ``GBGPU`` instance (supplies orbits, d_d and the device).
        data, psd: [A, E] arrays, numpy or torch: 1-D, or 2-D (streams,
            data_length) selected per binary by data_index / noise_index.
        T, dt: observation time and cadence (T is reduced to int(T/dt)*dt).
        precision: "fp64" or "mixed".
        N_levels, f_edges: the N step rule used when N is derived from f0:
            N_levels[i] for f0 < f_edges[i], the last level above.
        strategy: "buckets" (one launch of the uniform kernel per N bucket).
            "persistent" (one persistent launch over all buckets) is
            switched off in this build and raises ValueError.
        guard_tol: mixed precision with the "buckets" strategy: binaries
            whose log-likelihood error bound 2.5e-6 (<h|h> + |<d|h>|) exceeds
            guard_tol are evaluated in FP64 within the same call -- routed
            there at set-up when an estimate of <h|h> already exceeds the
            bound, else re-evaluated after the mixed pass (None or 0
            disables; loud templates otherwise carry an absolute error that
            grows with SNR^2).
        transform: transform of the mixed N = 256 bucket ("lds",
            "mfma_f32", "mfma_bf16x3", "mfma_bf16x6"; see GBGPU.transform);
            None takes gb.transform.  Every other bucket, every FP64 pass and
            every third-body batch uses the LDS transform.
        tdi2: second-generation TDI factor, as get_ll's ``tdi2``.
        data_index, noise_index: (B,) int32 stream of each binary for 2-D
            data / psd (None: stream 0), as get_ll's; fixed for the
            evaluator, so every call takes B binaries in the same order.

    Parameters are (9, B) as get_ll's, or (14, B) for third-body binaries
    with rows 9-13 in kernel form (A2, varpi, e2, n2 = 2 pi / P2 [1/s],
    T2 [s]): ``kernel_params`` converts the user rows (P2, T2 in years).
    N per binary is a power of two in [32, 8192].
    """

    STRATEGIES = {"buckets": _capi.GBGPU_BATCH_BUCKETS, "persistent": _capi.GBGPU_BATCH_PERSISTENT}

    def __init__(self, gb, data, psd, T, dt, precision="mixed", start_freq_ind=0, N_levels=DEFAULT_N_LEVELS,
                 f_edges=DEFAULT_F_EDGES, phase_marginalize=False, strategy="buckets", guard_tol=1e-3,
                 transform=None, tdi2=False, data_index=None, noise_index=None):
        torch = _torch()
        if gb.d_d is None:
            raise ValueError("gb.d_d must be set before building a likelihood")
        self.gb = gb
        self.T = int(T / dt) * dt
        self.precision = precision
        self.N_levels = np.asarray(N_levels, dtype=np.int32)
        self.f_edges = np.asarray(f_edges, dtype=np.float64)
        if len(self.N_levels) != len(self.f_edges) + 1:
            raise ValueError("need len(N_levels) == len(f_edges) + 1")
        self.phase_marginalize = bool(phase_marginalize)
        if strategy not in self.STRATEGIES:
            raise ValueError(f"strategy must be one of {list(self.STRATEGIES)}")
        if strategy == "persistent":
            raise ValueError("the persistent batch strategy is disabled in this build; use strategy='buckets'")
        self.guard_tol = float(guard_tol or 0.0) if precision == "mixed" else 0.0
        if self.guard_tol > 0 and strategy == "persistent":
            raise ValueError("the accuracy guard runs with the 'buckets' strategy only; pass guard_tol=None "
                             "for an unguarded persistent launch")
        dev = gb._device()
        if data[0].ndim not in (1, 2) or psd[0].ndim not in (1, 2):
            raise ValueError("data and psd must be 1-D or 2-D")
        self.L = int(data[0].shape[-1])
        if int(psd[0].shape[-1]) != self.L:
            raise ValueError("data and psd must have the same length")
        if data[0].ndim == 1 and data_index is not None or psd[0].ndim == 1 and noise_index is not None:
            raise ValueError("data_index / noise_index need 2-D data / psd")
        self.A = gb._to_dev(data[0], np.complex128).reshape(-1)
        self.E = gb._to_dev(data[1], np.complex128).reshape(-1)
        self.SA = gb._to_dev(psd[0], np.float64).reshape(-1)
        self.SE = gb._to_dev(psd[1], np.float64).reshape(-1)
        self.data_index = self._index(gb, data_index, self.A.numel())
        self.noise_index = self._index(gb, noise_index, self.SA.numel())
        transform = gb.transform if transform is None else transform
        if transform not in _capi.TRANSFORMS:
            raise ValueError(f"transform must be one of {list(_capi.TRANSFORMS)}")
        self.cfg = _capi.BatchCfg(T=float(self.T), armlength=float(gb.orbits.armlength), tdi2=int(bool(tdi2)),
                                  precision=1 if precision == "mixed" else 0, strategy=self.STRATEGIES[strategy],
                                  guard_tol=self.guard_tol,
                                  transform=_capi.TRANSFORMS[transform]
                                  if _capi.TRANSFORM_PRECISION.get(transform) == "mixed" and precision == "mixed"
                                  else 0)
        self._tabs = {}
        for N in self.N_levels:
            self._add_table(int(N))
        self.data = _capi.Data(A=self.A.data_ptr(), E=self.E.data_ptr(), psdA=self.SA.data_ptr(),
                               psdE=self.SE.data_ptr(), data_index=_capi.ptr(self.data_index),
                               noise_index=_capi.ptr(self.noise_index), data_length=self.L,
                               start_freq_ind=int(start_freq_ind), df=float(1.0 / self.T))
        self.dev = dev
        self._ws = None
        self._last_ws = None  # workspace of the last launch (eager or captured)
        self._last_B = 0
        self._graph = None

    def _index(self, gb, idx, total):
        """Per-binary stream index on the device (None: NULL, stream 0)."""
        if idx is None:
            return None
        t = gb._to_dev(np.asarray(idx, dtype=np.int32) if not _torch().is_tensor(idx) else idx, np.int32)
        if t.numel() and (int(t.min()) < 0 or (int(t.max()) + 1) * self.L > total):
            raise ValueError("data_index/noise_index point outside the provided streams")
        return t

    def kernel_params(self, params):
        """(14, B) third-body parameters as get_ll takes them (P2, T2 in
        years) -> the kernel rows (n2 = 2 pi / P2 [1/s], T2 [s]); (9, B) are
        returned unchanged."""
        p = np.asarray(params, dtype=np.float64)
        if p.shape[0] == 9:
            return p
        return np.vstack([p[:9], np.array(self.gb.prepare_additional_args(*p[9:]))])

    def _check_batch(self, p):
        if p.shape[0] not in (9, 14):
            raise ValueError("params must have 9 rows, or 14 for third-body binaries")
        B = p.shape[1]
        for idx in (self.data_index, self.noise_index):
            if idx is not None and idx.numel() != B:
                raise ValueError("data_index / noise_index must have one entry per binary")

    def _add_table(self, N):
        k = int(np.log2(N)) - 5
        if N not in self._tabs:
            if N & (N - 1) or not 32 <= N <= _capi.N_MAX:
                raise ValueError(f"N={N}: powers of two in [32, {_capi.N_MAX}] only")
            self._tabs[N] = self.gb._table(self.T, N)
            self.cfg.orbit_tables[k] = self._tabs[N].data_ptr()

    def _set_capacity(self, Nn, slack=False):
        """Per-bucket size bounds for the launches (``bucket_capacity``):
        the exact bucket sizes of the host-side N array Nn, or with ``slack``
        (N recomputed on the device between calls, as in the sampler loop)
        those plus 1% and 1024 for binaries that cross a bucket edge.  A
        bound that a call exceeds only costs speed: the binaries beyond it are
        evaluated by the strided kernel."""
        counts = np.bincount(np.log2(np.asarray(Nn)).astype(np.int64) - 5, minlength=_capi.NBUCKET)
        B = len(Nn)
        for k in range(_capi.NBUCKET):
            c = int(counts[k])
            if slack:
                c = min(B, c + c // 100 + 1024)
            self.cfg.bucket_capacity[k] = c

    def _workspace(self, B):
        """Workspace for B binaries.  A buffer whose address a captured graph
        holds is never replaced: a larger batch after capture gets a new
        buffer, and the graph's stays referenced by self._graph."""
        torch = _torch()
        need = _capi.load().gbgpu_batch_workspace_bytes(B, self.cfg.precision)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(int(need), dtype=torch.uint8, device=self.dev)
        return self._ws

    def autotune(self, params, N=None):
        """Time the kernel variants per N bucket on this batch and keep the
        fastest (process-wide, per precision); returns {N: variant}: 0 =
        min(N/4, 256) threads per binary, 1 = half of that, and for the mixed
        N = 256 bucket also 2 / 4 = the "mfma_f32" / "mfma_bf16x6" transforms
        (FP32-accurate; the two-piece "mfma_bf16x3", variant 3, only on
        request).  Synchronises; not capturable."""
        torch = _torch()
        p = params if torch.is_tensor(params) else torch.from_numpy(np.ascontiguousarray(params, np.float64))
        p = p.to(self.dev).contiguous()
        self._check_batch(p)
        B = p.shape[1]
        if N is None:
            N = self.N_of(p[1].cpu().numpy())
        Nn = np.asarray(N.cpu().numpy() if torch.is_tensor(N) else N, dtype=np.int32)
        for n in np.unique(Nn):
            self._add_table(int(n))
        Nt = torch.from_numpy(Nn).to(self.dev)
        self._set_capacity(Nn)
        ll = torch.empty(B, dtype=torch.float64, device=self.dev)
        ws = self._workspace(B)
        chosen = np.zeros(_capi.NBUCKET, dtype=np.int32)
        rc = _capi.load().gbgpu_batch_autotune(ll.data_ptr(), p.data_ptr(), p.shape[0], B, Nt.data_ptr(), self.cfg,
                                               self.data,
                                               float(np.real(self.gb.d_d)), ws.data_ptr(), chosen.ctypes.data,
                                               _capi.stream_handle())
        _capi.check(rc, "gbgpu_batch_autotune")
        return {32 << k: int(c) for k, c in enumerate(chosen) if c >= 0}

    def set_shapes(self, shapes):
        """Pin the per-bucket variants ({N: 0 or 1}, and 2-4 for the mixed
        N = 256 bucket as in autotune; others 0) for this precision."""
        v = np.zeros(_capi.NBUCKET, dtype=np.int32)
        for n, c in shapes.items():
            v[int(np.log2(n)) - 5] = c
        _capi.check(_capi.load().gbgpu_batch_set_variants(self.cfg.precision, v.ctypes.data),
                    "gbgpu_batch_set_variants")

    def guard_counts(self):
        """Binaries per N (dict) the accuracy guard evaluated in FP64 in the
        last call -- routed there at set-up by the <h|h> estimate, or
        re-evaluated after the mixed pass (synchronises the current stream)."""
        counts = np.zeros(_capi.NBUCKET, dtype=np.int32)
        if self._last_ws is None:
            return {}
        _capi.check(_capi.load().gbgpu_batch_guard_counts(self._last_ws.data_ptr(), self._last_B, self.cfg.precision,
                                                          counts.ctypes.data, _capi.stream_handle()),
                    "gbgpu_batch_guard_counts")
        return {32 << k: int(c) for k, c in enumerate(counts) if c}

    def N_of(self, f0):
        """N from f0 by the step rule."""
        return self.N_levels[np.searchsorted(self.f_edges, np.abs(np.asarray(f0)), side="right")]

    def __call__(self, params, N=None, d_h=False):
        """log-likelihood of (9, B) or (14, B) kernel-form params with
        per-binary N (default: from f0)."""
        torch = _torch()
        p = params if torch.is_tensor(params) else torch.from_numpy(np.ascontiguousarray(params, np.float64))
        p = p.to(self.dev).contiguous()
        self._check_batch(p)
        B = p.shape[1]
        if N is None:
            N = self.N_of(p[1].cpu().numpy())
        Nn = np.asarray(N.cpu().numpy() if torch.is_tensor(N) else N, dtype=np.int32)
        for n in np.unique(Nn):
            self._add_table(int(n))
        Nt = torch.from_numpy(Nn).to(self.dev) if not torch.is_tensor(N) else N.to(self.dev, torch.int32)
        self._set_capacity(Nn)
        ll = torch.empty(B, dtype=torch.float64, device=self.dev)
        dh = torch.empty(B, dtype=torch.complex128, device=self.dev) if d_h else None
        hh = torch.empty(B, dtype=torch.complex128, device=self.dev) if d_h else None
        self._launch(ll, dh, hh, p, Nt, B)
        return (ll, dh, hh) if d_h else ll

    def _launch(self, ll, dh, hh, p, Nt, B):
        ws = self._workspace(B)
        self._last_B = B
        self._last_ws = ws
        rc = _capi.load().gbgpu_wave_ll_batched(ll.data_ptr(), _capi.ptr(dh), _capi.ptr(hh), p.data_ptr(), p.shape[0], B,
                                                Nt.data_ptr(), self.cfg, self.data, float(np.real(self.gb.d_d)),
                                                int(self.phase_marginalize), ws.data_ptr(), _capi.stream_handle())
        _capi.check(rc, "gbgpu_wave_ll_batched")

    # ------------------------------------------------------ sampler loop ---
    def _loop_buffers(self, params, n_calls):
        torch = _torch()
        base = params if torch.is_tensor(params) else torch.from_numpy(np.ascontiguousarray(params, np.float64))
        base = base.to(self.dev).contiguous()
        self._check_batch(base)
        B = base.shape[1]
        out = torch.empty((n_calls, B), dtype=torch.float64, device=self.dev)
        pj = torch.empty_like(base)
        Nj = torch.empty(B, dtype=torch.int32, device=self.dev)
        edges = np.ascontiguousarray(self.f_edges)
        levels = np.ascontiguousarray(self.N_levels)
        # the jumps move a few binaries across bucket edges: bounds with slack
        self._set_capacity(self.N_of(base[1].cpu().numpy()), slack=True)
        return base, out, pj, Nj, edges, levels

    def _loop_body(self, base, out, pj, Nj, edges, levels, n_calls, sigma, seed):
        lib = _capi.load()
        B = base.shape[1]
        st = _capi.stream_handle()
        for i in range(n_calls):
            rc = lib.gbgpu_perturb_params(pj.data_ptr(), base.data_ptr(), base.shape[0], B, float(sigma), int(seed),
                                          int(i), Nj.data_ptr(), len(edges), edges.ctypes.data, levels.ctypes.data,
                                          st)
            _capi.check(rc, "gbgpu_perturb_params")
            self._launch(out[i], None, None, pj, Nj, B)

    def run_loop(self, params, n_calls, sigma=1e-3, seed=0):
        """n_calls evaluations at seeded jumps of ``params``; (n_calls, B) tensor."""
        bufs = self._loop_buffers(params, n_calls)
        self._workspace(bufs[0].shape[1])
        self._loop_body(*bufs, n_calls, sigma, seed)
        return bufs[1]

    def capture_loop(self, params, n_calls, sigma=1e-3, seed=0):
        """Capture the n_calls loop in one graph; returns (graph, out).

        ``graph.replay()`` re-runs every call and rewrites ``out``.  One eager
        call is made first so that the per-N tables exist before capture.
        """
        torch = _torch()
        bufs = self._loop_buffers(params, n_calls)
        base, out = bufs[0], bufs[1]
        self._workspace(base.shape[1])
        self._loop_body(*bufs, 1, sigma, seed)  # warm: twiddles, occupancy query
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self._loop_body(*bufs, n_calls, sigma, seed)
        torch.cuda.current_stream().wait_stream(s)
        # the graph holds this workspace's address: keep it alive with the
        # graph, and give later eager calls a buffer of their own
        self._graph = (g, bufs, self._ws)
        self._ws = None
        return g, out
