/* A consumer of the plan API that wraps an nfft_plan in a matrix-vector plan
 * of its own and hands that to the solver, the way the reference's
 * kernel/mri/mri.c builds its inhomogeneity plans (mri.c:159-262: an inner
 * nfft_plan, the wrapper's own mv_trafo / mv_adjoint behind the MV prefix,
 * the inner plan's f / f_hat re-pointed at the wrapper's vectors on every
 * call because the solver swaps them).  Written in C99 complex arithmetic
 * against <complex.h>, as mri.c and the Julia shim are.
 *
 * The wrapper is a coil-weighted NFFT: (A f_hat)_j = c_j (F f_hat)_j with a
 * smooth complex sensitivity c_j per node.  The program runs the same CGNR
 * twice through the wrapper, once over the fast transforms and once over
 * the direct NDFT, and exits 0 when the iterates agree within the window
 * error of m = 6 and the normal-equation residual falls. */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nfft3.h"

typedef struct
{
  /* matrix-vector prefix: the solver reads only these fields */
  NFFT_INT N_total;
  NFFT_INT M_total;
  double _Complex *f_hat;
  double _Complex *f;
  void (*mv_trafo)(void *);
  void (*mv_adjoint)(void *);
  /* the wrapper's own state */
  nfft_plan plan;
  double _Complex *c;
  double _Complex *tmp;
  int direct;
} coil_plan;

static void coil_trafo(void *v)
{
  coil_plan *ths = (coil_plan *)v;
  /* the solver swapped ths->f / ths->f_hat since the last call */
  ths->plan.f_hat = ths->f_hat;
  ths->plan.f = ths->f;
  if (ths->direct)
    nfft_trafo_direct(&ths->plan);
  else
    nfft_trafo(&ths->plan);
  for (NFFT_INT j = 0; j < ths->M_total; j++)
    ths->f[j] *= ths->c[j];
}

static void coil_adjoint(void *v)
{
  coil_plan *ths = (coil_plan *)v;
  for (NFFT_INT j = 0; j < ths->M_total; j++)
    ths->tmp[j] = conj(ths->c[j]) * ths->f[j];
  ths->plan.f = ths->tmp;
  ths->plan.f_hat = ths->f_hat;
  if (ths->direct)
    nfft_adjoint_direct(&ths->plan);
  else
    nfft_adjoint(&ths->plan);
  ths->plan.f = ths->f;
}

static void coil_init(coil_plan *ths, int N0, int N1, int M, int m)
{
  int N[2] = {N0, N1}, n[2] = {2 * N0, 2 * N1};
  nfft_init_guru(&ths->plan, 2, N, M, n, m,
                 PRE_PHI_HUT | PRE_PSI | MALLOC_X | MALLOC_F_HAT | MALLOC_F | FFTW_INIT |
                     FFT_OUT_OF_PLACE,
                 FFTW_ESTIMATE | FFTW_DESTROY_INPUT);
  ths->N_total = ths->plan.N_total;
  ths->M_total = ths->plan.M_total;
  ths->f_hat = ths->plan.f_hat;
  ths->f = ths->plan.f;
  ths->mv_trafo = coil_trafo;
  ths->mv_adjoint = coil_adjoint;
  ths->c = (double _Complex *)nfft_malloc(sizeof(double _Complex) * (size_t)M);
  ths->tmp = (double _Complex *)nfft_malloc(sizeof(double _Complex) * (size_t)M);
  ths->direct = 0;
}

static void coil_finalize(coil_plan *ths)
{
  nfft_free(ths->c);
  nfft_free(ths->tmp);
  /* hand the plan back the vectors it allocated, whichever the solver left */
  ths->plan.f = ths->f;
  ths->plan.f_hat = ths->f_hat;
  nfft_finalize(&ths->plan);
}

/* CGNR from zero for `steps` steps; iterates into out (steps x N_total) and
 * dot_z_hat_iter per step into dz (steps + 1). */
static void run_cgnr(coil_plan *cp, const double _Complex *y, int steps, double _Complex *out,
                     double *dz)
{
  solver_plan_complex ip;
  solver_init_advanced_complex(&ip, (nfft_mv_plan_complex *)cp, CGNR);
  memcpy(ip.y, y, sizeof(double _Complex) * (size_t)cp->M_total);
  memset(ip.f_hat_iter, 0, sizeof(double _Complex) * (size_t)cp->N_total);
  solver_before_loop_complex(&ip);
  dz[0] = ip.dot_z_hat_iter;
  for (int it = 0; it < steps; it++)
  {
    solver_loop_one_step_complex(&ip);
    memcpy(out + (size_t)it * cp->N_total, ip.f_hat_iter,
           sizeof(double _Complex) * (size_t)cp->N_total);
    dz[it + 1] = ip.dot_z_hat_iter;
  }
  solver_finalize_complex(&ip);
}

int main(void)
{
  const int N0 = 24, N1 = 20, M = 1500, m = 6, steps = 12;
  coil_plan cp;
  coil_init(&cp, N0, N1, M, m);
  if (nfft_check(&cp.plan) != 0)
  {
    printf("check: %s\n", nfft_check(&cp.plan));
    return 1;
  }
  nfft_srand48(1234);
  nfft_vrand_shifted_unit_double(cp.plan.x, 2 * M);
  nfft_precompute_one_psi(&cp.plan);
  for (int j = 0; j < M; j++)
  {
    const double x0 = cp.plan.x[2 * j], x1 = cp.plan.x[2 * j + 1];
    cp.c[j] = (1.0 + 0.5 * x1) * cexp(2.0 * M_PI * I * 0.3 * x0);
  }

  /* measurements: the direct operator applied to a random image */
  double _Complex *f_true = (double _Complex *)nfft_malloc(sizeof(double _Complex) * (size_t)cp.N_total);
  double _Complex *y = (double _Complex *)nfft_malloc(sizeof(double _Complex) * (size_t)M);
  for (NFFT_INT k = 0; k < cp.N_total; k++)
    f_true[k] = (nfft_drand48() - 0.5) + I * (nfft_drand48() - 0.5);
  memcpy(cp.f_hat, f_true, sizeof(double _Complex) * (size_t)cp.N_total);
  cp.direct = 1;
  cp.mv_trafo(&cp);
  memcpy(y, cp.f, sizeof(double _Complex) * (size_t)M);

  double _Complex *it_fast = (double _Complex *)nfft_malloc(sizeof(double _Complex) * (size_t)steps * cp.N_total);
  double _Complex *it_direct = (double _Complex *)nfft_malloc(sizeof(double _Complex) * (size_t)steps * cp.N_total);
  double dz_fast[64], dz_direct[64];
  cp.direct = 0;
  run_cgnr(&cp, y, steps, it_fast, dz_fast);
  cp.direct = 1;
  run_cgnr(&cp, y, steps, it_direct, dz_direct);

  int ok = 1;
  for (int it = 0; it < steps; it++)
  {
    double num = 0.0, den = 0.0;
    for (NFFT_INT k = 0; k < cp.N_total; k++)
    {
      const double _Complex a = it_fast[(size_t)it * cp.N_total + k];
      const double _Complex b = it_direct[(size_t)it * cp.N_total + k];
      num += creal((a - b) * conj(a - b));
      den += creal(b * conj(b));
    }
    const double rel = sqrt(num / den);
    printf("step %2d: fast vs direct iterate rel l2 %.3e, residual %.3e\n", it + 1, rel,
           sqrt(dz_fast[it + 1] / dz_fast[0]));
    if (!(rel < 1e-8))
      ok = 0;
  }
  if (!(dz_fast[steps] < 1e-4 * dz_fast[0]))
    ok = 0;
  /* the image the solver reached against the one that made the data */
  double num = 0.0, den = 0.0;
  for (NFFT_INT k = 0; k < cp.N_total; k++)
  {
    const double _Complex e = it_fast[(size_t)(steps - 1) * cp.N_total + k] - f_true[k];
    num += creal(e * conj(e));
    den += creal(f_true[k] * conj(f_true[k]));
  }
  printf("image rel error after %d steps: %.3e\n", steps, sqrt(num / den));

  nfft_free(f_true);
  nfft_free(y);
  nfft_free(it_fast);
  nfft_free(it_direct);
  coil_finalize(&cp);
  printf(ok ? "mv_caller: OK\n" : "mv_caller: FAILED\n");
  return ok ? 0 : 1;
}
