"""CPU FP64 oracle for the FastGB waveform + likelihood path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``gbgpu_amd``) imports,
links or executes anything under ``oracle/``; only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use it, and
only as the checker or the timed CPU baseline.

Contents:
  fastgb.py   numpy restatement of the reference's Python waveform
              (src/gbgpu/gbgpu.py:108-540), likelihood and global-template
              algebra, third-body hooks (src/gbgpu/thirdbody.py) and the
              finite-difference information matrix.
  cref.c      plain-C restatement of the reference's native get_ll and
              fill_global (src/gbgpu/cutils/gbgpu_utils.cu:55-210), serial as
              in the reference's CPU build.
  ref_shim.cpp + Makefile  build the reference's own gbgpu_utils.cu (host
              path) into oracle/_ref/libgbgpu_ref.so when /root/reference is
              present, to check cref.c against it.

How each function is pinned (the checks live in tests/):
  * start bins: closed form rint(f0*T) - N/2 and exact integer vectors,
    including half-integer f0*T (round-half-even);
  * AET: linear closed form;
  * FFT / fftshift / scaling: numpy.fft plus the 0.5*T/N normalisation;
  * slow part: invariants that a dropped term, wrong sign or swapped index
    breaks -- phase covariance h(phi0+d) = e^{i d} h(phi0), linearity in amp,
    psi -> psi+pi/2 negates h, psi -> psi+pi and lam -> lam+2pi leave it
    unchanged, cyclic X->Y->Z under spacecraft relabelling, the
    long-wavelength (f << f*) limit against a direct time-domain TDI
    evaluation, and a direct time-domain single-link response check;
  * orbits: closed forms (arm length ~ L to O(ec), 1-yr period, 60-degree
    plane tilt);
  * get_ll / fill_global: brute-force numpy sums, cref.c, and the
    reference's own compiled code in oracle/_ref;
  * swap: brute force logL(r + h_r - h_a) - logL(r);
  * third body: circular closed-form v_LOS and its analytic phase integral;
    Kepler solver against scipy root finding;
  * constants: notebook known answers (get_amplitude, get_fdot).

Parity unpinned (nothing available here can fix them: they would need
lisatools, which is absent, and the reference's tests hold no golden
vectors):
  * absolute waveform values with the reference's default orbits
    (lisatools EqualArmlengthOrbits, interpolated from an orbit file); this
    oracle and the product both use the analytic Cornish & Rubbo orbits with
    the LISA.h constants, and every check above holds for any orbit model;
  * get_N's SNR adjustment and the default PSD of information_matrix, which
    use this build's own analytic A/X-channel sensitivity in place of
    lisatools' A1TDISens / X1TDISens.
"""
