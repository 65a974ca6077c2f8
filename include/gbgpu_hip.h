/* gbgpu_hip.h -- C ABI of the MI355X (gfx950) FastGB waveform + likelihood library.
 *
 * All pointers are DEVICE pointers unless stated otherwise; complex arrays are
 * interleaved {re, im} float64 pairs (numpy complex128 layout, the same as the
 * reference's gcmplx::complex<double>, src/gbgpu/cutils/global.h:66).  Every
 * entry point is asynchronous on the caller's HIP stream (``stream`` is a
 * hipStream_t passed as void*, NULL = default stream) and never synchronises
 * the device.  Per-call scratch (per-binary constants, sort keys) is
 * stream-ordered: hipMallocAsync / hipFreeAsync on the call's stream, except
 * for gbgpu_wave_ll_batched given a workspace, which allocates nothing.  The
 * first call for a given N on a device also creates that N's twiddle table
 * (a blocking upload, cached until the process ends).  Every call returns 0
 * on success or a negative GBGPU_E* code; gbgpu_last_error() then holds a
 * message for the calling thread.
 */
#ifndef GBGPU_HIP_H
#define GBGPU_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gbgpu_cmplx {
    double re, im;
} gbgpu_cmplx;

enum {
    GBGPU_OK = 0,
    GBGPU_E_ARG = -1,     /* invalid argument (size, pointer, unsupported N) */
    GBGPU_E_HIP = -2,     /* HIP runtime error */
    GBGPU_E_ALLOC = -3,   /* cache allocation failed */
};

enum { GBGPU_FP64 = 0, GBGPU_MIXED = 1 };

/* Length-N transform of the slow part (the reference's xp.fft.fft,
 * src/gbgpu/gbgpu.py:378).  GBGPU_TRANSFORM_LDS: radix-4 Stockham transform
 * in LDS, every N and precision.  The MFMA variants evaluate the transform as
 * two 16 x 16 complex matrix products on the matrix cores, in registers, at
 * N = 256 without a third body only (other combinations return
 * GBGPU_E_ARG).  For GBGPU_MIXED: _F32 with FP32 matrix instructions,
 * _BF16X3 / _BF16X6 with bf16 matrix instructions on a two-piece (three
 * products, about 16 significant bits) or three-piece (six products, FP32
 * level) split of each operand.  For GBGPU_FP64: _F64 with FP64 matrix
 * instructions (FP64 products and sums; results agree with the LDS
 * transform to rounding).  MI355X's FP64 matrix and vector peaks are equal
 * and the 16 x 16 products take about four times the arithmetic of the
 * radix-4 transform, so _F64 is slower than the LDS transform there; it is
 * kept selectable for GPUs whose FP64 matrix rate exceeds their vector
 * rate. */
enum {
    GBGPU_TRANSFORM_LDS = 0,
    GBGPU_TRANSFORM_MFMA_F32 = 1,
    GBGPU_TRANSFORM_MFMA_BF16X3 = 2,
    GBGPU_TRANSFORM_MFMA_BF16X6 = 3,
    GBGPU_TRANSFORM_MFMA_F64 = 4
};

/* Parameter rows of the (num_params, num_bin) float64 parameter array:
 * amp, f0, fdot, fddot, phi0, iota, psi, lam, beta   (src/gbgpu/gbgpu.py:108-120)
 * and, when num_params == 14, the third-body rows A2, varpi, e2, n2, T2 with
 * n2 = 2*pi/(P2*YRSID_SI) and T2 in seconds, i.e. already passed through
 * prepare_additional_args (src/gbgpu/thirdbody.py:39-80). */
#define GBGPU_NPARAMS 9
#define GBGPU_NPARAMS_THIRD 14
#define GBGPU_ORBIT_ROWS 22

/* Largest N of the waveform entry points (the reference's get_N caps N at
 * 8192, src/gbgpu/utils/utility.py:268).  gbgpu_run_wave, gbgpu_run_wave_ae,
 * gbgpu_wave_ll and gbgpu_wave_ll_noalloc take every power of two in
 * [32, GBGPU_N_MAX]; gbgpu_swap_ll_diff, gbgpu_wave_ll_grad and the batched
 * path take N <= 2048, gbgpu_fisher N <= 1024 (GBGPU_E_ARG otherwise). */
#define GBGPU_N_MAX 8192

/* Waveform configuration shared by every waveform entry point. */
typedef struct gbgpu_wave_cfg {
    int N;                     /* samples per binary: power of two, 32..GBGPU_N_MAX */
    double T;                  /* observation time, already int(T/dt)*dt (gbgpu.py:177-179) */
    double armlength;          /* metres; f* = c / (2 pi armlength) (gbgpu.py:243) */
    const double *orbit_table; /* (22, N) float64: t_n, guiding centre Pc (xyz), offsets P_s - Pc
                                  for s = 1..3 (xyz), arms r12, r13, r23 (xyz) */
    int tdi2;                  /* 1: multiply by 2i sin(2wL) e^{-2iwL} (gbgpu.py:301-304) */
    int precision;             /* GBGPU_FP64 or GBGPU_MIXED */
    int transform;             /* GBGPU_TRANSFORM_*; 0 (LDS) when zero-initialised.  Only
                                  gbgpu_run_wave, gbgpu_run_wave_ae, gbgpu_wave_ll and
                                  gbgpu_wave_ll_noalloc accept a non-zero value */
} gbgpu_wave_cfg;

/* Data against which templates are scored (get_ll's data/psd arguments,
 * src/gbgpu/gbgpu.py:580-712). */
typedef struct gbgpu_data {
    const gbgpu_cmplx *A, *E;  /* (num_streams, data_length) */
    const double *psdA, *psdE; /* (num_psds, data_length) */
    const int *data_index;     /* (num_bin,) or NULL = all zero */
    const int *noise_index;    /* (num_bin,) or NULL = all zero */
    int data_length;
    int start_freq_ind;        /* global bin of element 0 of every stream */
    double df;                 /* bin width, 1/T */
} gbgpu_data;

/* ---- drop-in replacements of the reference's native entry points ------- */

/* Replaces get_ll_wrap (src/gbgpu/cutils/gbgpu_utils.hh:9-13,
 * gbgpu_utils.cu:188-210).  Same arguments and meaning; templates are
 * (M, num_bin) major.  Bins outside [0, data_length) contribute nothing
 * (the reference reads out of bounds there).  64-bit flat indexing. */
int gbgpu_get_ll(gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const gbgpu_cmplx *A_template,
                 const gbgpu_cmplx *E_template, const gbgpu_cmplx *A_data, const gbgpu_cmplx *E_data,
                 const double *A_psd, const double *E_psd, double df, const int *start_ind, int M,
                 int num_bin, const int *data_index, const int *noise_index, int data_length,
                 void *stream);

/* Replaces fill_global_wrap (src/gbgpu/cutils/gbgpu_utils.hh:6-7,
 * gbgpu_utils.cu:104-122): A_glob[g*L + start_b + i] += template[i*num_bin + b]
 * for 0 <= start_b + i < L.  Deterministic: binaries are sorted by (group,
 * start) and each output bin adds its contributions in a fixed order (four
 * contiguous runs of the sorted list, combined in order), so results are
 * bit-identical across runs
 * (the reference's GPU build uses unordered atomics).  num_groups (the
 * leading dimension of the caller's templates array, templates.shape[0] at
 * src/gbgpu/gbgpu.py:778) is the one argument the reference does not pass;
 * binaries whose group index lies outside [0, num_groups) are skipped.
 * Uses stream-ordered scratch (hipMallocAsync) of 2*M*num_bin complex plus
 * sort buffers. */
int gbgpu_fill_global(gbgpu_cmplx *A_glob, gbgpu_cmplx *E_glob, const gbgpu_cmplx *A_template,
                      const gbgpu_cmplx *E_template, const int *start_ind, int M, int num_bin,
                      const int *group_index, int data_length, int num_groups, void *stream);

/* fill_global over binary-major rows (no transposition): sample i of binary
 * b is A_rows[b*row_stride + i]; start_inds are int64 as run_wave produces
 * them and are shifted by -start_freq_ind (gbgpu.py:804).  group_index may be
 * NULL (all zero).  Same deterministic summation as gbgpu_fill_global.  Backs
 * fill_global_template for 2-D inputs and generate_global_template
 * (src/gbgpu/gbgpu.py:737-878). */
int gbgpu_fill_global_rows(gbgpu_cmplx *A_glob, gbgpu_cmplx *E_glob, const gbgpu_cmplx *A_rows,
                           const gbgpu_cmplx *E_rows, long long row_stride, const long long *start_inds,
                           long long start_freq_ind, int M, int num_bin, const int *group_index,
                           int data_length, int num_groups, void *stream);

/* ---- waveform-level entry points (replace the xp array code) ----------- */

/* Replaces GBGPUBase.run_wave's array passes (src/gbgpu/gbgpu.py:108-316,
 * :348-540): writes XYZf and AETf as (num_bin, 3, N) complex128 (either may be
 * NULL), fftshifted and scaled exactly as the reference, and start_inds
 * (num_bin,) int64 = rint(f0*T) - N/2.  With GBGPU_MIXED the T channel
 * AETf[:, 2] is NaN: (X + Y + Z)/sqrt3 cancels to O((f/f*)^2) of X, below
 * the FP32 spectra's error. */
int gbgpu_run_wave(gbgpu_cmplx *XYZf, gbgpu_cmplx *AETf, long long *start_inds, const double *params,
                   int num_params, int num_bin, const gbgpu_wave_cfg *cfg, void *stream);

/* The A and E outputs of run_wave only, for callers that consume nothing
 * else -- generate_global_template (src/gbgpu/gbgpu.py:839-878), which runs
 * run_wave (:108-316) and fills A and E: AEf (num_bin, 2, N) complex128,
 * fftshifted and scaled as gbgpu_run_wave's AETf[:, 0:2], and start_inds.
 * Two channels go through the transform instead of three, and a third less
 * is written. */
int gbgpu_run_wave_ae(gbgpu_cmplx *AEf, long long *start_inds, const double *params, int num_params,
                      int num_bin, const gbgpu_wave_cfg *cfg, void *stream);

/* Fused run_wave + get_ll (src/gbgpu/gbgpu.py:580-735): per binary computes
 * d_h, h_h (complex, as get_ll_wrap) and loglike = -1/2 Re(d_d + h_h - 2 d_h),
 * with |d_h| in place of d_h when phase_marginalize (gbgpu.py:714-728).
 * Templates never leave LDS.  d_h / h_h may be NULL. */
int gbgpu_wave_ll(double *loglike, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const double *params,
                  int num_params, int num_bin, const gbgpu_wave_cfg *cfg, const gbgpu_data *data,
                  double d_d, int phase_marginalize, void *stream);

/* Largest num_bin gbgpu_wave_ll_noalloc accepts. */
#define GBGPU_NOALLOC_MAX_BIN 65536

/* gbgpu_wave_ll without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight sampler loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and kernels as gbgpu_wave_ll; the per-binary
 * constants, sort keys and twiddle tables live in a fixed 24 MiB device
 * array of the library (one per GPU, held until the process ends; the first
 * call on a device looks up its address and copies the twiddle tables in,
 * blocking).  num_bin <= GBGPU_NOALLOC_MAX_BIN, else GBGPU_E_ARG.  Calls
 * never allocate, free or memset; every call fully rewrites the part of the
 * array it reads, so nothing carries over from the previous call.  All
 * calls on one GPU share the array: calls that overlap in time on one GPU
 * (concurrent streams or host threads) race and must be serialised by the
 * caller. */
int gbgpu_wave_ll_noalloc(double *loglike, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const double *params,
                          int num_params, int num_bin, const gbgpu_wave_cfg *cfg, const gbgpu_data *data,
                          double d_d, int phase_marginalize, void *stream);

/* swap_likelihood_difference (no native counterpart in the reference; its
 * likelihood algebra, src/gbgpu/gbgpu.py:728 with the inner product of
 * src/gbgpu/cutils/gbgpu_utils.cu:156-183): for proposal p, remove binary
 * h_r and add h_a against the residual r held in `data`:
 *   dll[p] = Re<r|h_a> - Re<r|h_r> - (<h_a|h_a> + <h_r|h_r> - 2 Re<h_r|h_a>)/2.
 * params_pairs is (num_params, 2*num_prop) with column 2p the removed and
 * column 2p+1 the added binary (same N).  terms (num_prop, 5) may be NULL;
 * else it receives Re<r|h_r>, Re<r|h_a>, <h_r|h_r>, <h_a|h_a>, Re<h_r|h_a>.
 * data->data_index / noise_index are per proposal. */
int gbgpu_swap_ll_diff(double *dll, double *terms, const double *params_pairs, int num_params, int num_prop,
                       const gbgpu_wave_cfg *cfg, const gbgpu_data *data, void *stream);

/* ---- derivatives (FP64, 9-parameter sources) ------------------------------ */

/* Parameter basis of every gradient and Fisher output ("sampling basis"):
 *   0 ln amp, 1 f0, 2 fdot, 3 fddot, 4 phi0, 5 cos iota, 6 psi, 7 lambda,
 *   8 sin beta,
 * with f0, fdot and fddot independent (a tied fddot = 11/3 fdot^2/f0 follows
 * by the chain rule on the host). */
#define GBGPU_NGRAD 9

/* Adjoint gradient of the fused likelihood (no counterpart in the reference,
 * which has no derivatives beyond finite differences): loglike, d_h, h_h as
 * gbgpu_wave_ll, and grad (num_bin, 9) row-major = d loglike / d theta in the
 * sampling basis (for the phase-marginalised likelihood when
 * phase_marginalize).  The backward pass pushes the residual back through the
 * transform and recomputes the slow part; cfg->precision must be GBGPU_FP64. */
int gbgpu_wave_ll_grad(double *loglike, double *grad, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const double *params,
                       int num_params, int num_bin, const gbgpu_wave_cfg *cfg, const gbgpu_data *data,
                       double d_d, int phase_marginalize, void *stream);

/* Replaces the finite-difference information_matrix
 * (src/gbgpu/gbgpu.py:957-1138, F_ij = 4 df sum Re(conj(dh_i) dh_j)/S over
 * A and E, :1119-1132) with exact forward tangents: fisher (num_bin, 9, 9)
 * row-major in the sampling basis, against the PSD streams of `data`
 * (data->A / data->E may be NULL).  With data, grad (num_bin, 9) and loglike
 * (num_bin) are also available (either may be NULL; phase not marginalised).
 * N <= 1024; cfg->precision must be GBGPU_FP64. */
int gbgpu_fisher(double *fisher, double *grad, double *loglike, const double *params, int num_params, int num_bin,
                 const gbgpu_wave_cfg *cfg, const gbgpu_data *data, double d_d, void *stream);

/* ---- uneven-N batches -------------------------------------------------- */

/* Configuration of a batch whose binaries carry their own N. */
enum { GBGPU_BATCH_BUCKETS = 0, GBGPU_BATCH_PERSISTENT = 1 };

typedef struct gbgpu_batch_cfg {
    double T;                      /* observation time, already int(T/dt)*dt */
    double armlength;              /* metres */
    int tdi2;
    int precision;                 /* GBGPU_FP64 or GBGPU_MIXED */
    const double *orbit_tables[9]; /* table for N = 32 << k (k < 9), NULL when no binary uses it */
    int strategy;                  /* GBGPU_BATCH_BUCKETS: one launch per N bucket (default);
                                      GBGPU_BATCH_PERSISTENT: one persistent launch */
    double guard_tol;              /* mixed precision with GBGPU_BATCH_BUCKETS: binaries whose
                                      log-likelihood error bound 2.5e-6 (<h|h> + |<d|h>|) exceeds
                                      guard_tol are evaluated in FP64: routed there at set-up when
                                      an estimate of <h|h> from amplitude, frequency and the PSD
                                      already exceeds the bound, else re-evaluated after the mixed
                                      pass when its own result does; <= 0 disables */
    int transform;                 /* GBGPU_TRANSFORM_* for the mixed N = 256 bucket (every
                                      other bucket and every FP64 pass uses the LDS transform) */
    int bucket_capacity[9];        /* optional bound on the number of binaries with N = 32 << k
                                      (0: none).  A bucket with a bound runs as one pass of
                                      workgroups sized for it, with the registers of the uniform
                                      kernel; binaries beyond the bound are still evaluated, by
                                      the strided kernel that every bucket without a bound uses.
                                      Results do not depend on it. */
} gbgpu_batch_cfg;

/* Workspace bytes gbgpu_wave_ll_batched needs for num_bin binaries. */
long long gbgpu_batch_workspace_bytes(int num_bin, int precision);

/* Fused wave + log-likelihood for a batch with per-binary N (N_bin, device
 * int32, powers of two in [32, 8192]): binaries are bucketed on the device
 * by N and start bin, then either each bucket runs the uniform fused kernel
 * over its device-side segment (GBGPU_BATCH_BUCKETS) or one persistent
 * launch pulls N-bucketed work items from an atomic queue, largest N first
 * (GBGPU_BATCH_PERSISTENT, switched off in this build: GBGPU_E_ARG).
 * num_params is 9 or 14 (third-body rows); cfg->tdi2 and the data's
 * data_index / noise_index apply as in gbgpu_wave_ll.  The reference needs
 * one run_wave + get_ll call per distinct N (src/gbgpu/gbgpu.py:205-207,
 * gbgpu_utils.cu:126).
 * Outputs as gbgpu_wave_ll; binaries with an unsupported N or a NULL orbit
 * table get NaN.  workspace: gbgpu_batch_workspace_bytes() of device memory,
 * or NULL to allocate stream-ordered scratch.  With a workspace and after one
 * earlier call on the device (which creates the per-N twiddle tables), the
 * call performs no allocation, memset or host synchronisation, so it can be
 * captured in a hipGraph. */
int gbgpu_wave_ll_batched(double *loglike, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const double *params,
                          int num_params, int num_bin, const int *N_bin, const gbgpu_batch_cfg *cfg,
                          const gbgpu_data *data, double d_d, int phase_marginalize, void *workspace,
                          void *stream);

/* Autotuner of the per-bucket launches (GBGPU_BATCH_BUCKETS): runs the batch
 * with each kernel variant (0: N/4 threads per binary, in workgroups of
 * max(N/4, 256) threads; 1: half as many threads per binary, twice the
 * binaries per workgroup; for the GBGPU_MIXED N = 256 bucket, when
 * cfg->transform is GBGPU_TRANSFORM_LDS, also 2 and 4: the
 * GBGPU_TRANSFORM_MFMA_F32 and _BF16X6 transforms) three times, times
 * every bucket with HIP events, and keeps the fastest variant per (precision,
 * bucket) for all later calls in the process.  chosen (9 ints, N = 32 << k,
 * may be NULL) receives the pick, -1 for buckets without an orbit table.
 * loglike receives the last run's results.  Synchronises `stream`; call it
 * outside graph capture.  No counterpart in the reference. */
int gbgpu_batch_autotune(double *loglike, const double *params, int num_params, int num_bin, const int *N_bin,
                         const gbgpu_batch_cfg *cfg, const gbgpu_data *data, double d_d, void *workspace,
                         int *chosen, void *stream);

/* Sets the per-bucket variants (9 ints: 0 or 1 as for the autotuner; for the
 * GBGPU_MIXED N = 256 bucket also 1 + t, the matrix-core transform t =
 * GBGPU_TRANSFORM_MFMA_*) used by GBGPU_BATCH_BUCKETS calls of the given
 * precision. */
int gbgpu_batch_set_variants(int precision, const int *variants);

/* Per-bucket counts (9 ints, N = 32 << k) of binaries the accuracy guard
 * evaluated in FP64 (routed at set-up plus re-evaluated after the mixed
 * pass) in the last gbgpu_wave_ll_batched call that used this workspace.
 * Diagnostic: copies to host memory and synchronises `stream`. */
int gbgpu_batch_guard_counts(const void *workspace, int num_bin, int precision, int *counts, void *stream);

/* Sampler utility with no counterpart in the reference: seeded Gaussian
 * relative jumps out = in * (1 + sigma_rel * z) for every parameter row
 * (counter-based: the same seed and call_index give the same draw), and,
 * when N_out is non-NULL, N per binary from f0 (row 1) by a step rule:
 * N_levels[i] for f0 < f_edges[i], N_levels[n_edges] above (n_edges <= 8). */
int gbgpu_perturb_params(double *out, const double *in, int num_params, int num_bin, double sigma_rel,
                         unsigned long long seed, unsigned long long call_index, int *N_out, int n_edges,
                         const double *f_edges, const int *N_levels, void *stream);

/* ---- derived waveforms ------------------------------------------------ */

/* The reference's extension API (src/gbgpu/gbgpu.py:1140-1229) lets a
 * subclass change the slow part through two hooks on (num_bin, 3, N) arrays
 * (called at src/gbgpu/gbgpu.py:453 and :505):
 *   fi = shift_frequency(fi, xi, *args), argS = add_to_argS(argS, f0, fdot, fddot, xi, *args).
 * gbgpu_hook_inputs writes their inputs for a batch: xi[b, s, n] =
 * t_n - k.P_s(t_n)/c, the unshifted fi = f0 + fdot xi + fddot xi^2/2 and the
 * base argS = phi0 + (om - 2 pi rint(f0 T)/T) t_n + pi fdot xi^2
 * + (pi/3) fddot xi^3, each (num_bin, 3, N) float64.  The caller runs the
 * hooks and passes fi_argS (num_bin, 2, 3, N) float64 -- per binary the
 * hooked fi, then argS -- to gbgpu_run_wave_ext / gbgpu_wave_ll_ext, which
 * are gbgpu_run_wave / gbgpu_wave_ll with those values in place of the
 * computed ones.  9 parameter rows, GBGPU_FP64, GBGPU_TRANSFORM_LDS,
 * N <= 2048. */
int gbgpu_hook_inputs(double *xi, double *fi, double *argS, const double *params, int num_params, int num_bin,
                      const gbgpu_wave_cfg *cfg, void *stream);
int gbgpu_run_wave_ext(gbgpu_cmplx *XYZf, gbgpu_cmplx *AETf, long long *start_inds, const double *params,
                       int num_params, int num_bin, const gbgpu_wave_cfg *cfg, const double *fi_argS,
                       void *stream);
int gbgpu_wave_ll_ext(double *loglike, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const double *params, int num_params,
                      int num_bin, const gbgpu_wave_cfg *cfg, const gbgpu_data *data, double d_d,
                      int phase_marginalize, const double *fi_argS, void *stream);

/* ---- library -------------------------------------------------------- */

/* Message describing the last failure on the calling thread. */
const char *gbgpu_last_error(void);

/* ABI version of this header (9). */
int gbgpu_abi_version(void);

/* Measurement hook with no counterpart in the reference: while enabled,
 * every call records a pair of HIP events on its stream immediately around
 * the launch of its dominant kernel (the waveform, likelihood, gradient,
 * Fisher, persistent batch, gather-sum or swap kernel).  Enabling or
 * disabling drops the pairs recorded so far.  Not for use under graph
 * capture. */
int gbgpu_kernel_timing(int enable);

/* Waits for the recorded pairs, returns their summed duration (ms) and
 * count, and drops them. */
int gbgpu_kernel_timing_read(double *total_ms, int *count);

/* Measurement hook with no counterpart in the reference: the FMA rate this
 * GPU sustains (TFLOP/s, FMA = 2 FLOP) in a kernel of 8 independent FMA
 * chains per thread, FP64 (GBGPU_FP64) or FP32 without packed forms
 * (GBGPU_MIXED).  Synchronises `stream` (about 30 ms); for rooflines. */
int gbgpu_fma_probe(int precision, double *tflops, void *stream);

/* Measurement hook with no counterpart in the reference: the HBM streaming
 * rate (GB/s) of a 1 GiB buffer, one 16-byte non-temporal access per thread
 * over contiguous 4 KiB chunks per workgroup; mode 0 reads it, mode 1 copies
 * it (read + write bytes counted).
 * Allocates 1-2 GiB stream-ordered and synchronises `stream`; for
 * rooflines. */
int gbgpu_bw_probe(int mode, double *gbps, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GBGPU_HIP_H */
