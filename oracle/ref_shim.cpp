// extern "C" forwarders onto the reference's own host build of
// src/gbgpu/cutils/gbgpu_utils.cu (get_ll_wrap / fill_global_wrap declared at
// src/gbgpu/cutils/gbgpu_utils.hh:6-13), so ctypes can call them by plain
// names.  TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ from the sources
// where they lie under /root/reference and used to check oracle/cref.c.
#include "gbgpu_utils.hh"

extern "C" {

void ref_get_ll(void *d_h, void *h_h, void *A_template, void *E_template, void *A_data, void *E_data,
                double *A_psd, double *E_psd, double df, int *start_ind, int M, int num_bin,
                int *data_index, int *noise_index, int data_length)
{
    get_ll_wrap((cmplx *)d_h, (cmplx *)h_h, (cmplx *)A_template, (cmplx *)E_template, (cmplx *)A_data,
                (cmplx *)E_data, A_psd, E_psd, df, start_ind, M, num_bin, data_index, noise_index,
                data_length);
}

void ref_fill_global(void *A_glob, void *E_glob, void *A_template, void *E_template, int *start_ind, int M,
                     int num_bin, int *group_index, int data_length)
{
    fill_global_wrap((cmplx *)A_glob, (cmplx *)E_glob, (cmplx *)A_template, (cmplx *)E_template, start_ind, M,
                     num_bin, group_index, data_length);
}
}
