/* Plain-C restatement of the reference's native likelihood and global-template
 * loops.  TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline); the
 * product never links it.
 *
 * Complex numbers are interleaved {re, im} doubles, layout-compatible with
 * numpy complex128 and the reference's gcmplx::complex<double>
 * (src/gbgpu/cutils/global.h:66).
 */
#include <stddef.h>
#include <stdint.h>

/* Restates get_ll (src/gbgpu/cutils/gbgpu_utils.cu:125-185) on the CPU path
 * (serial over binaries, serial over samples).  Templates are (M, num_bin)
 * major: sample i of binary b is at i*num_bin + b.  Flat indices are 64-bit.
 * Unlike the reference, bins outside [0, data_length) are skipped instead of
 * read out of bounds. */
void oracle_get_ll(double *d_h, double *h_h, const double *A_template, const double *E_template,
                   const double *A_data, const double *E_data, const double *A_psd, const double *E_psd,
                   double df, const int *start_ind, int M, int num_bin, const int *data_index,
                   const int *noise_index, int data_length)
{
    for (int b = 0; b < num_bin; b++) {
        const int64_t s = start_ind[b];
        const int64_t doff = (int64_t)data_index[b] * data_length;
        const int64_t noff = (int64_t)noise_index[b] * data_length;
        double dhr = 0.0, dhi = 0.0, hhr = 0.0, hhi = 0.0;
        for (int i = 0; i < M; i++) {
            const int64_t j = s + i;
            if (j < 0 || j >= data_length) continue;
            const double SA = A_psd[noff + j], SE = E_psd[noff + j];
            const int64_t t = ((int64_t)i * num_bin + b) * 2;
            const double har = A_template[t], hai = A_template[t + 1];
            const double her = E_template[t], hei = E_template[t + 1];
            const double dar = A_data[(doff + j) * 2], dai = A_data[(doff + j) * 2 + 1];
            const double der = E_data[(doff + j) * 2], dei = E_data[(doff + j) * 2 + 1];
            /* conj(d) * h / S */
            dhr += (dar * har + dai * hai) / SA;
            dhi += (dar * hai - dai * har) / SA;
            dhr += (der * her + dei * hei) / SE;
            dhi += (der * hei - dei * her) / SE;
            /* conj(h) * h / S */
            hhr += (har * har + hai * hai) / SA;
            hhi += (har * hai - hai * har) / SA;
            hhr += (her * her + hei * hei) / SE;
            hhi += (her * hei - hei * her) / SE;
        }
        d_h[2 * b] = 4.0 * df * dhr;
        d_h[2 * b + 1] = 4.0 * df * dhi;
        h_h[2 * b] = 4.0 * df * hhr;
        h_h[2 * b + 1] = 4.0 * df * hhi;
    }
}

/* Restates fill_global (src/gbgpu/cutils/gbgpu_utils.cu:55-101): serial in
 * binary order, += into the caller's buffers, out-of-range bins skipped. */
void oracle_fill_global(double *A_glob, double *E_glob, const double *A_template, const double *E_template,
                        const int *start_ind, int M, int num_bin, const int *group_index, int data_length)
{
    for (int b = 0; b < num_bin; b++) {
        const int64_t s = start_ind[b];
        const int64_t goff = (int64_t)group_index[b] * data_length;
        for (int i = 0; i < M; i++) {
            const int64_t j = s + i;
            if (j < 0 || j >= data_length) continue;
            const int64_t t = ((int64_t)i * num_bin + b) * 2;
            A_glob[(goff + j) * 2] += A_template[t];
            A_glob[(goff + j) * 2 + 1] += A_template[t + 1];
            E_glob[(goff + j) * 2] += E_template[t];
            E_glob[(goff + j) * 2 + 1] += E_template[t + 1];
        }
    }
}
