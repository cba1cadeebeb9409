/*
 * nfft3.h -- C ABI of the MI355X NFFT library (libnfft3_hip.so).
 *
 * A drop-in for the hot path of NFFT3's public header (include/nfft3.h of the
 * reference): the plan structs are laid out field for field as there, so
 * existing C callers that read and write p.x, p.f, p.f_hat, p.M_total, ...
 * and the solver's ip.y, ip.w, ip.f_hat_iter, ip.dot_r_iter compile and run
 * unchanged.  Every entry point below cites the reference declaration it
 * replaces.  Two precision families are exported: nfftf_ / solverf_ (float,
 * FP32 bulk arithmetic on the GPU) and nfft_ / solver_ (double, FP64 on the
 * GPU).  The long-double family nfftl_ has no GPU counterpart and is not
 * exported.
 *
 * FFTW is not used.  The reference's header includes <fftw3.h> because the
 * plan embeds fftw_complex and fftw_plan; this header defines ABI-identical
 * types itself unless the real <fftw3.h> was included first.  The
 * my_fftw_plan1 field holds the library's own device context instead of an
 * FFTW plan.
 *
 * Device-resident extensions (no reference counterpart) are in nfft3_hip.h.
 */
#ifndef NFFT3_HIP_COMPAT_H
#define NFFT3_HIP_COMPAT_H

#include <stddef.h>

#ifndef FFTW3_H
/* Layout of FFTW's own definitions: two reals, re then im; plans are opaque
 * pointers.  As in FFTW's header, a C caller that included <complex.h> first
 * gets the C99 complex types (same size and layout), so code such as
 * p->f_hat[k] = c * cexp(...) compiles as it does against the reference. */
#if !defined(FFTW_NO_Complex) && defined(_Complex_I) && defined(complex) && defined(I)
typedef double _Complex fftw_complex;
typedef float _Complex fftwf_complex;
#else
typedef double fftw_complex[2];
typedef float fftwf_complex[2];
#endif
typedef struct fftw_plan_s *fftw_plan;
typedef struct fftwf_plan_s *fftwf_plan;
#endif

#ifdef __cplusplus
extern "C"
{
#endif

/* ptrdiff_t-sized integer (reference include/nfft3.h:49-51) */
typedef ptrdiff_t NFFT_INT;

/* ----------------------------------------------------------------------- */
/* Plan layouts (reference include/nfft3.h:54-60 MACRO_MV_PLAN and          */
/* include/nfft3.h:109-161 X(plan)).  Order and types must not change.      */
/* ----------------------------------------------------------------------- */
#define NFFT_HIP_MV_FIELDS(RC)                                               \
  NFFT_INT N_total;           /* number of Fourier coefficients */           \
  NFFT_INT M_total;           /* number of nodes */                          \
  RC *f_hat;                  /* Fourier coefficients, N_total */            \
  RC *f;                      /* samples, M_total */                         \
  void (*mv_trafo)(void *);   /* set to nfft[f]_trafo by init */             \
  void (*mv_adjoint)(void *); /* set to nfft[f]_adjoint by init */

#define NFFT_HIP_PLAN_FIELDS(R, C, PLAN)                                     \
  NFFT_HIP_MV_FIELDS(C)                                                      \
  NFFT_INT d;           /* dimension */                                      \
  NFFT_INT *N;          /* bandwidths, even */                               \
  R *sigma;             /* oversampling n[t]/N[t] */                         \
  NFFT_INT *n;          /* oversampled FFT lengths */                        \
  NFFT_INT n_total;     /* product of n */                                   \
  NFFT_INT m;           /* window cut-off */                                 \
  R *b;                 /* Kaiser-Bessel shape per dimension */              \
  NFFT_INT K;           /* PRE_LIN_PSI table size (kept for layout) */       \
  unsigned flags;       /* PRE_*, MALLOC_*, ... */                           \
  unsigned fftw_flags;  /* accepted and ignored */                           \
  R *x;                 /* nodes, x[j*d+t] in [-0.5,0.5) */                  \
  R MEASURE_TIME_t[3];  /* seconds of D, FFT, B of the last call */          \
  PLAN my_fftw_plan1;   /* library device context (opaque) */                \
  PLAN my_fftw_plan2;   /* unused, NULL */                                   \
  R **c_phi_inv;        /* deconvolution factors, with PRE_PHI_HUT */        \
  R *psi;               /* host window table: not allocated (NULL) */        \
  NFFT_INT *psi_index_g;/* not allocated (NULL) */                           \
  NFFT_INT *psi_index_f;/* not allocated (NULL) */                           \
  C *g;                 /* oversampled grid lives on the device: NULL */     \
  C *g_hat;             /* NULL */                                           \
  C *g1;                /* NULL */                                           \
  C *g2;                /* NULL */                                           \
  R *spline_coeffs;     /* NULL */                                           \
  NFFT_INT *index_x;    /* node order lives on the device: NULL */

typedef struct { NFFT_HIP_MV_FIELDS(fftwf_complex) } nfftf_mv_plan_complex;
typedef struct { NFFT_HIP_MV_FIELDS(float) } nfftf_mv_plan_double;
typedef struct { NFFT_HIP_PLAN_FIELDS(float, fftwf_complex, fftwf_plan) } nfftf_plan;

typedef struct { NFFT_HIP_MV_FIELDS(fftw_complex) } nfft_mv_plan_complex;
typedef struct { NFFT_HIP_MV_FIELDS(double) } nfft_mv_plan_double;
typedef struct { NFFT_HIP_PLAN_FIELDS(double, fftw_complex, fftw_plan) } nfft_plan;

/* ----------------------------------------------------------------------- */
/* Plan flags (reference include/nfft3.h:195-208), same bit values.         */
/* ----------------------------------------------------------------------- */
#define PRE_PHI_HUT                (1U << 0)
#define FG_PSI                     (1U << 1)
#define PRE_LIN_PSI                (1U << 2)
#define PRE_FG_PSI                 (1U << 3)
#define PRE_PSI                    (1U << 4)
#define PRE_FULL_PSI               (1U << 5)
#define MALLOC_X                   (1U << 6)
#define MALLOC_F_HAT               (1U << 7)
#define MALLOC_F                   (1U << 8)
#define FFT_OUT_OF_PLACE           (1U << 9)
#define FFTW_INIT                  (1U << 10)
#define NFFT_SORT_NODES            (1U << 11)
#define NFFT_OMP_BLOCKWISE_ADJOINT (1U << 12)
#define PRE_ONE_PSI (PRE_LIN_PSI | PRE_FG_PSI | PRE_PSI | PRE_FULL_PSI)

/* FFTW planner flags are accepted for source compatibility and ignored. */
#ifndef FFTW3_H
#define FFTW_MEASURE (0U)
#define FFTW_DESTROY_INPUT (1U << 0)
#define FFTW_ESTIMATE (1U << 6)
#endif

/* ----------------------------------------------------------------------- */
/* Plan API (reference include/nfft3.h:163-187).                            */
/* ----------------------------------------------------------------------- */
#define NFFT_HIP_DECLARE_API(P, PLAN)                                        \
  void P##trafo_direct(const PLAN *ths);                                     \
  void P##adjoint_direct(const PLAN *ths);                                   \
  void P##trafo(PLAN *ths);                                                  \
  void P##trafo_1d(PLAN *ths);                                               \
  void P##trafo_2d(PLAN *ths);                                               \
  void P##trafo_3d(PLAN *ths);                                               \
  void P##adjoint(PLAN *ths);                                                \
  void P##adjoint_1d(PLAN *ths);                                             \
  void P##adjoint_2d(PLAN *ths);                                             \
  void P##adjoint_3d(PLAN *ths);                                             \
  void P##init_1d(PLAN *ths, int N1, int M);                                 \
  void P##init_2d(PLAN *ths, int N1, int N2, int M);                         \
  void P##init_3d(PLAN *ths, int N1, int N2, int N3, int M);                 \
  void P##init(PLAN *ths, int d, int *N, int M);                             \
  void P##init_guru(PLAN *ths, int d, int *N, int M, int *n, int m,          \
                    unsigned flags, unsigned fftw_flags);                    \
  void P##init_lin(PLAN *ths, int d, int *N, int M, int *n, int m, int K,    \
                   unsigned flags, unsigned fftw_flags);                     \
  void P##precompute_one_psi(PLAN *ths);                                     \
  void P##precompute_psi(PLAN *ths);                                         \
  void P##precompute_full_psi(PLAN *ths);                                    \
  void P##precompute_fg_psi(PLAN *ths);                                      \
  void P##precompute_lin_psi(PLAN *ths);                                     \
  const char *P##check(PLAN *ths);                                           \
  void P##finalize(PLAN *ths);

NFFT_HIP_DECLARE_API(nfftf_, nfftf_plan)
NFFT_HIP_DECLARE_API(nfft_, nfft_plan)

/* ----------------------------------------------------------------------- */
/* Memory and exit hooks (reference include/nfft3.h:69-87).                 */
/* ----------------------------------------------------------------------- */
#define NFFT_HIP_DECLARE_MALLOC(P)                                           \
  void *P##malloc(size_t n);                                                 \
  void P##free(void *p);                                                     \
  void P##die(const char *s);                                                \
  typedef void *(*P##malloc_type_function)(size_t n);                        \
  typedef void (*P##free_type_function)(void *p);                            \
  typedef void (*P##die_type_function)(const char *errString);              \
  extern P##malloc_type_function P##malloc_hook;                             \
  extern P##free_type_function P##free_hook;                                 \
  extern P##die_type_function P##die_hook;

NFFT_HIP_DECLARE_MALLOC(nfftf_)
NFFT_HIP_DECLARE_MALLOC(nfft_)

/* ----------------------------------------------------------------------- */
/* Solver (reference include/nfft3.h:758-829).                              */
/* ----------------------------------------------------------------------- */
#define NFFT_HIP_SOLVER_FIELDS(R, C, MV)                                     \
  MV *mv;                 /* transform (prefix of an nfft plan) */           \
  unsigned flags;         /* CGNR, PRECOMPUTE_WEIGHT, ... */                 \
  R *w;                   /* weights, M_total, with PRECOMPUTE_WEIGHT */     \
  R *w_hat;               /* damping, N_total, with PRECOMPUTE_DAMP */       \
  C *y;                   /* right-hand side, M_total */                     \
  C *f_hat_iter;          /* iterate, N_total */                             \
  C *r_iter;              /* residual y - A f_hat_iter, M_total */           \
  C *z_hat_iter;          /* A^H W r, N_total */                             \
  C *p_hat_iter;          /* search direction, N_total */                    \
  C *v_iter;              /* A W_hat p, M_total */                           \
  R alpha_iter;                                                              \
  R beta_iter;                                                               \
  R dot_r_iter;                                                              \
  R dot_r_iter_old;                                                          \
  R dot_z_hat_iter;                                                          \
  R dot_z_hat_iter_old;                                                      \
  R dot_p_hat_iter;                                                          \
  R dot_v_iter;

typedef struct { NFFT_HIP_SOLVER_FIELDS(float, fftwf_complex, nfftf_mv_plan_complex) } solverf_plan_complex;
typedef struct { NFFT_HIP_SOLVER_FIELDS(double, fftw_complex, nfft_mv_plan_complex) } solver_plan_complex;

/* solver flags (reference include/nfft3.h:823-829) */
#define LANDWEBER             (1U << 0)
#define STEEPEST_DESCENT      (1U << 1)
#define CGNR                  (1U << 2)
#define CGNE                  (1U << 3)
#define NORMS_FOR_LANDWEBER   (1U << 4)
#define PRECOMPUTE_WEIGHT     (1U << 5)
#define PRECOMPUTE_DAMP       (1U << 6)

#define NFFT_HIP_DECLARE_SOLVER(S, PLAN, MV)                                 \
  void S##init_advanced_complex(PLAN *ths, MV *mv, unsigned flags);          \
  void S##init_complex(PLAN *ths, MV *mv);                                   \
  void S##before_loop_complex(PLAN *ths);                                    \
  void S##loop_one_step_complex(PLAN *ths);                                  \
  void S##finalize_complex(PLAN *ths);

NFFT_HIP_DECLARE_SOLVER(solverf_, solverf_plan_complex, nfftf_mv_plan_complex)
NFFT_HIP_DECLARE_SOLVER(solver_, solver_plan_complex, nfft_mv_plan_complex)

/* Real-valued solver on an mv plan with real f_hat and f (reference
 * include/nfft3.h:786-819, kernel/solver/solver.c:392-714): the same four
 * methods, flags and weights; y, the iterate and every work vector hold one
 * real per entry.  Runs on the host around the caller's mv_trafo /
 * mv_adjoint, with the reference's pointer swaps. */
typedef struct { NFFT_HIP_SOLVER_FIELDS(float, float, nfftf_mv_plan_double) } solverf_plan_double;
typedef struct { NFFT_HIP_SOLVER_FIELDS(double, double, nfft_mv_plan_double) } solver_plan_double;

#define NFFT_HIP_DECLARE_SOLVER_DOUBLE(S, PLAN, MV)                          \
  void S##init_advanced_double(PLAN *ths, MV *mv, unsigned flags);           \
  void S##init_double(PLAN *ths, MV *mv);                                    \
  void S##before_loop_double(PLAN *ths);                                     \
  void S##loop_one_step_double(PLAN *ths);                                   \
  void S##finalize_double(PLAN *ths);

NFFT_HIP_DECLARE_SOLVER_DOUBLE(solverf_, solverf_plan_double, nfftf_mv_plan_double)
NFFT_HIP_DECLARE_SOLVER_DOUBLE(solver_, solver_plan_double, nfft_mv_plan_double)

/* ----------------------------------------------------------------------- */
/* Utilities used by callers of the path (reference include/nfft3.h:839-894) */
/* ----------------------------------------------------------------------- */
#define NFFT_HIP_DECLARE_UTIL(P, R, C)                                       \
  R P##drand48(void);                                                        \
  void P##srand48(long int seed);                                            \
  void P##vrand_unit_complex(C *x, const NFFT_INT n);                        \
  void P##vrand_shifted_unit_double(R *x, const NFFT_INT n);                 \
  void P##vrand_real(R *x, const NFFT_INT n, const R a, const R b);          \
  void P##vpr_double(R *x, const NFFT_INT n, const char *text);              \
  void P##vpr_complex(C *x, const NFFT_INT n, const char *text);             \
  NFFT_INT P##get_num_threads(void);                                         \
  void P##set_num_threads(NFFT_INT nthreads);                                \
  NFFT_INT P##has_threads_enabled(void);                                     \
  R P##clock_gettime_seconds(void);                                          \
  R P##error_l_infty_complex(const C *x, const C *y, const NFFT_INT n);      \
  R P##error_l_infty_1_complex(const C *x, const C *y, const NFFT_INT n,     \
                               const C *z, const NFFT_INT m);                \
  NFFT_INT P##exp2i(const NFFT_INT a);                                       \
  NFFT_INT P##next_power_of_2(const NFFT_INT N);                             \
  R P##dot_complex(C *x, NFFT_INT n);                                        \
  void P##upd_axpy_complex(C *x, R a, C *y, NFFT_INT n);                     \
  void P##fftshift_complex(C *x, NFFT_INT d, NFFT_INT *N);                   \
  void P##fftshift_complex_int(C *x, int d, int *N);                         \
  void P##get_version(unsigned *major, unsigned *minor, unsigned *patch);    \
  const char *P##get_window_name(void);                                      \
  NFFT_INT P##get_default_window_cut_off(void);

NFFT_HIP_DECLARE_UTIL(nfftf_, float, fftwf_complex)
NFFT_HIP_DECLARE_UTIL(nfft_, double, fftw_complex)

/* ----------------------------------------------------------------------- */
/* Damping factors for the solver's w_hat (PRECOMPUTE_DAMP): the library's  */
/* exported helpers of reference include/infft.h:1506-1520 (bodies          */
/* kernel/util/damp.c:25-142).  Factor kk is the sum of a kernel's Fourier  */
/* coefficients at kk and kk + 1: Fejer 2/N^2 (1 - |2kk+1|/N), Jackson      */
/* (power 4, n = (N/2+1)/2), generalised Jackson (power 8, n = (N/2+3)/4),  */
/* Sobolev |k|^(-2 mu), multiquadric (k^2 + c^2)^(-mu).                     */
/* ----------------------------------------------------------------------- */
#define NFFT_HIP_DECLARE_DAMP(P, R)                                          \
  R P##modified_fejer(const NFFT_INT N, const NFFT_INT kk);                  \
  R P##modified_jackson2(const NFFT_INT N, const NFFT_INT kk);               \
  R P##modified_jackson4(const NFFT_INT N, const NFFT_INT kk);               \
  R P##modified_sobolev(const R mu, const NFFT_INT kk);                      \
  R P##modified_multiquadric(const R mu, const R c, const NFFT_INT kk);

NFFT_HIP_DECLARE_DAMP(nfftf_, float)
NFFT_HIP_DECLARE_DAMP(nfft_, double)

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* NFFT3_HIP_COMPAT_H */
