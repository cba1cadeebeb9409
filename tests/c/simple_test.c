/* A C caller of the NFFT3 plan API, written the way the reference's examples
 * use it (examples/nfft/simple_test.c.in, examples/solver/simple_test.c.in):
 * nothing here knows about HIP.  It is compiled against include/nfft3.h and
 * linked against libnfft3_hip.so; exit status 0 means every fast transform
 * matched the direct NDFT within the reference's error bound. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "nfft3.h"

static int check_2d_double(void)
{
  nfft_plan p;
  int N[2] = {32, 14}, n[2] = {64, 32};
  int M = N[0] * N[1];
  nfft_init_guru(&p, 2, N, M, n, 7,
                 PRE_PHI_HUT | PRE_FULL_PSI | MALLOC_F_HAT | MALLOC_X | MALLOC_F |
                     FFTW_INIT | FFT_OUT_OF_PLACE,
                 FFTW_ESTIMATE | FFTW_DESTROY_INPUT);
  nfft_vrand_shifted_unit_double(p.x, p.d * p.M_total);
  if (p.flags & PRE_ONE_PSI)
    nfft_precompute_one_psi(&p);
  nfft_vrand_unit_complex(p.f_hat, p.N_total);
  if (nfft_check(&p) != 0)
    return 1;
  fftw_complex *ref = (fftw_complex *)nfft_malloc(sizeof(fftw_complex) * (size_t)M);
  nfft_trafo_direct(&p);
  for (int j = 0; j < M; j++)
  {
    ref[j][0] = p.f[j][0];
    ref[j][1] = p.f[j][1];
  }
  nfft_trafo(&p);
  double e1 = nfft_error_l_infty_complex(ref, p.f, M);
  fftw_complex *refh = (fftw_complex *)nfft_malloc(sizeof(fftw_complex) * (size_t)p.N_total);
  nfft_adjoint_direct(&p);
  for (int k = 0; k < p.N_total; k++)
  {
    refh[k][0] = p.f_hat[k][0];
    refh[k][1] = p.f_hat[k][1];
  }
  nfft_adjoint(&p);
  double e2 = nfft_error_l_infty_complex(refh, p.f_hat, p.N_total);
  printf("2d double m=7: trafo err %.3e adjoint err %.3e\n", e1, e2);
  nfft_free(ref);
  nfft_free(refh);
  nfft_finalize(&p);
  return !(e1 < 1e-9 && e2 < 1e-9);
}

static int check_1d_float(void)
{
  nfftf_plan p;
  nfftf_init_1d(&p, 14, 19);
  nfftf_vrand_shifted_unit_double(p.x, p.M_total);
  if (p.flags & PRE_ONE_PSI)
    nfftf_precompute_one_psi(&p);
  nfftf_vrand_unit_complex(p.f_hat, p.N_total);
  if (nfftf_check(&p) != 0)
    return 1;
  fftwf_complex ref[19];
  nfftf_trafo_direct(&p);
  for (int j = 0; j < 19; j++)
  {
    ref[j][0] = p.f[j][0];
    ref[j][1] = p.f[j][1];
  }
  nfftf_trafo(&p);
  float e = nfftf_error_l_infty_complex(ref, p.f, 19);
  printf("1d float default m: trafo err %.3e\n", (double)e);
  nfftf_finalize(&p);
  return !(e < 1e-3f);
}

int main(void)
{
  int bad = check_2d_double();
  bad |= check_1d_float();
  printf(bad ? "FAIL\n" : "OK\n");
  return bad;
}
