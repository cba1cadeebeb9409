// Iterative inverse NFFT (reference kernel/solver/solver.c): Landweber,
// steepest descent, CGNR and CGNE with optional weights W (PRECOMPUTE_WEIGHT)
// and damping W_hat (PRECOMPUTE_DAMP), for the solverf_ and solver_ families.
//
// Two paths behind the same entry points:
//  * device CGNR -- when the method is CGNR and the mv plan is one of this
//    library's NFFT plans, all vectors live on the GPU, the transforms run
//    device to device, inner products are reduced in FP64 and alpha / beta
//    are formed on the device, so one step ends with a single wait.  After
//    every step the host copies of f_hat_iter and all scalars (alpha_iter,
//    beta_iter, dot_*) are current; r_iter, z_hat_iter, p_hat_iter and v_iter
//    are refreshed on the host at before_loop and at every step as well,
//    unless the caller lowers the sync level (solver[f]_set_sync in
//    nfft3_hip.h).  y, w and w_hat are read from the host at before_loop.
//  * host path -- any other mv plan or method: the reference recurrences
//    (solver.c:81-125, 128-292, 295-344) on host vectors, transforms through
//    mv_trafo / mv_adjoint with the CSWAP pointer swaps of the reference.
//
// CGNR step (solver.c:232-292):
//   v = A W_hat p; dot_v = v^H W v; alpha = dot_z / dot_v
//   f_hat += alpha W_hat p; r -= alpha v; dot_r = r^H W r
//   z = A^H W r; beta = dot_z_new / dot_z_old; p = beta p + z
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "../../include/nfft3.h"
#include "../../include/nfft3_hip.h"
#include "engine.hpp"

namespace nfft_hip {

// ------------------------------------------------------------------ kernels
constexpr int kRedBlocks = 512;
constexpr int kRedThreads = 256;

// scalar slots (FP64) on the device
enum Scal { S_DOT_R, S_DOT_R_OLD, S_DOT_Z, S_DOT_Z_OLD, S_DOT_V, S_DOT_P, S_ALPHA, S_BETA, S_COUNT };

// partial sums of sum_k w_k |x_k|^2 (w == nullptr: w = 1)
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_dot_w(const cplx<R> *__restrict__ x,
                                                      const R *__restrict__ w, long n,
                                                      double *__restrict__ part) {
  __shared__ double sh[kRedThreads];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const cplx<R> v = x[i];
    const double a = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    s += w ? (double)w[i] * a : a;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Finish a dot product into slot `dst`; then, by `op`:
//   1: old[dst_old] = previous value of dst before overwriting
//   alpha_from: alpha = S_DOT_Z / S_DOT_V          (op & 2)
//   beta_from:  beta  = S_DOT_Z / S_DOT_Z_OLD      (op & 4)
__global__ __launch_bounds__(kRedThreads) void k_dot_finish(const double *__restrict__ part,
                                                            int nparts, double *__restrict__ sc,
                                                            int dst, int dst_old, int op) {
  __shared__ double sh[kRedThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (dst_old >= 0) sc[dst_old] = sc[dst];
    sc[dst] = sh[0];
    if (op & 2) sc[S_ALPHA] = sc[S_DOT_Z] / sc[S_DOT_V];
    if (op & 4) sc[S_BETA] = sc[S_DOT_Z] / sc[S_DOT_Z_OLD];
  }
}

// x = (w ? w . y : y)
template <typename R>
__global__ void k_cp_w(cplx<R> *__restrict__ x, const R *__restrict__ w,
                       const cplx<R> *__restrict__ y, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx<R> v = y[i];
  const R s = w ? w[i] : R(1);
  x[i] = cplx<R>{s * v.x, s * v.y};
}

// x = a x + y with host constant a (before_loop: r = y - A f)
template <typename R>
__global__ void k_axpy_const(cplx<R> *__restrict__ x, R a, const cplx<R> *__restrict__ y,
                             long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx<R> v = x[i], u = y[i];
  x[i] = cplx<R>{a * v.x + u.x, a * v.y + u.y};
}

// x += sgn * sc[slot] * (w ? w . y : y)
template <typename R>
__global__ void k_xpay_dev(cplx<R> *__restrict__ x, const double *__restrict__ sc, int slot,
                           R sgn, const R *__restrict__ w, const cplx<R> *__restrict__ y,
                           long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R a = sgn * (R)sc[slot];
  if (w) a *= w[i];
  const cplx<R> v = x[i], u = y[i];
  x[i] = cplx<R>{v.x + a * u.x, v.y + a * u.y};
}

// p = sc[S_BETA] p + z
template <typename R>
__global__ void k_axpy_beta(cplx<R> *__restrict__ p, const double *__restrict__ sc,
                            const cplx<R> *__restrict__ z, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const R b = (R)sc[S_BETA];
  const cplx<R> v = p[i], u = z[i];
  p[i] = cplx<R>{b * v.x + u.x, b * v.y + u.y};
}

static unsigned blocks_for(long n) { return (unsigned)std::max<long>(1, ceil_div(n, 256)); }

// ------------------------------------------------------------------ device CGNR state
template <typename R> struct DevCgnr {
  Engine<R> *e = nullptr;
  long M = 0, N = 0;
  cplx<R> *y = nullptr, *fh = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *v = nullptr;
  cplx<R> *tmpM = nullptr, *tmpN = nullptr;
  R *w = nullptr, *w_hat = nullptr;
  double *part = nullptr, *sc = nullptr;
  double *sc_host = nullptr;  // pinned
  int sync_level = 2;         // 0 scalars, 1 + f_hat_iter, 2 + all vectors
  // The kernels of one step (everything but the copies back to the host)
  // as a hipGraph, captured on the first step and replayed with one launch
  // afterwards (NFFT_HIP_CGNR_GRAPH=0 at solver creation: eager launches).
  // Re-captured when the plan's stream changes; not used while stage timing
  // is on (its events are created on demand).
  bool use_graph = true;
  hipGraphExec_t step_exec = nullptr;
  hipStream_t step_stream = nullptr;

  void alloc(Engine<R> *eng, long M_, long N_, bool weighted, bool damped) {
    e = eng;
    M = M_;
    N = N_;
    auto a = [&](auto **ptr, size_t count, size_t elem) {
      NFFT_HIP_CHECK(R, hipMalloc((void **)ptr, elem * (count ? count : 1)));
    };
    a(&y, M, sizeof(cplx<R>));
    a(&r, M, sizeof(cplx<R>));
    a(&v, M, sizeof(cplx<R>));
    a(&tmpM, M, sizeof(cplx<R>));
    a(&fh, N, sizeof(cplx<R>));
    a(&z, N, sizeof(cplx<R>));
    a(&p, N, sizeof(cplx<R>));
    a(&tmpN, N, sizeof(cplx<R>));
    if (weighted) a(&w, M, sizeof(R));
    if (damped) a(&w_hat, N, sizeof(R));
    a(&part, kRedBlocks, sizeof(double));
    a(&sc, S_COUNT, sizeof(double));
    NFFT_HIP_CHECK(R, hipHostMalloc((void **)&sc_host, S_COUNT * sizeof(double), hipHostMallocDefault));
    NFFT_HIP_CHECK(R, hipMemset(sc, 0, S_COUNT * sizeof(double)));
  }
  void release() {
    if (step_exec) hipGraphExecDestroy(step_exec);
    for (void *q : {(void *)y, (void *)r, (void *)v, (void *)tmpM, (void *)fh, (void *)z,
                    (void *)p, (void *)tmpN, (void *)w, (void *)w_hat, (void *)part, (void *)sc})
      if (q) hipFree(q);
    if (sc_host) hipHostFree(sc_host);
  }

  void dot(const cplx<R> *x, const R *wt, long n, int dst, int dst_old, int op) {
    hipStream_t s = e->stream;
    const int nb = (int)std::min<long>(kRedBlocks, std::max<long>(1, ceil_div(n, kRedThreads)));
    hipLaunchKernelGGL(k_dot_w<R>, dim3(nb), dim3(kRedThreads), 0, s, x, wt, n, part);
    hipLaunchKernelGGL(k_dot_finish, dim3(1), dim3(kRedThreads), 0, s, part, nb, sc, dst, dst_old,
                       op);
  }
};

template <typename R> struct Registry {
  std::mutex mu;
  std::unordered_map<void *, DevCgnr<R> *> map;
};
template <typename R>
Registry<R> &registry() {
  static Registry<R> r;
  return r;
}
template <typename R>
DevCgnr<R> *find_dev(void *ip) {
  auto &rg = registry<R>();
  std::lock_guard<std::mutex> g(rg.mu);
  auto it = rg.map.find(ip);
  return it == rg.map.end() ? nullptr : it->second;
}

// ------------------------------------------------------------------ host BLAS-1
// K reals per vector entry: 2 for complex vectors, 1 for real ones
template <typename R, int K>
static double hdot(const R *x, const R *w, long n) {  // sum_k w_k |x_k|^2
  double s = 0;
  for (long k = 0; k < n; k++) {
    double a = 0;
    for (int c = 0; c < K; c++) a += (double)x[K * k + c] * x[K * k + c];
    s += w ? (double)w[k] * a : a;
  }
  return s;
}
template <typename R, int K>
static void hcp_w(R *x, const R *w, const R *y, long n) {  // x = w y
  for (long k = 0; k < n; k++) {
    const R s = w ? w[k] : R(1);
    for (int c = 0; c < K; c++) x[K * k + c] = s * y[K * k + c];
  }
}
template <typename R, int K>
static void haxpy(R *x, R a, const R *y, long n) {  // x = a x + y
  for (long k = 0; k < K * n; k++) x[k] = a * x[k] + y[k];
}
template <typename R, int K>
static void hxpawy(R *x, R a, const R *w, const R *y, long n) {  // x += a w y
  for (long k = 0; k < n; k++) {
    const R s = w ? a * w[k] : a;
    for (int c = 0; c < K; c++) x[K * k + c] += s * y[K * k + c];
  }
}

// ------------------------------------------------------------------ solver API
template <typename R, typename SPLAN, typename MV, typename C>
struct SolverApi {
  // reals per vector entry (complex solver 2, real solver 1)
  static constexpr int K = (int)(sizeof(C) / sizeof(R));

  // Vectors of the complex solver are page-locked (the device CGNR copies
  // them every step); the real solver's are plain nfft[f]_malloc memory, as
  // in the reference (solver.c:392-428), so it needs no GPU at all.
  static void *valloc(size_t bytes) {
    if constexpr (K == 2) return host_array_alloc<R>(bytes);
    else if constexpr (sizeof(R) == 4) return nfftf_malloc(bytes);
    else return nfft_malloc(bytes);
  }
  static void vfree(void *p) {
    if constexpr (K == 2) host_array_free<R>(p);
    else if constexpr (sizeof(R) == 4) nfftf_free(p);
    else nfft_free(p);
  }

  static void init_advanced(SPLAN *ths, MV *mv, unsigned flags) {
    ths->mv = mv;
    ths->flags = flags;
    const long M = (long)mv->M_total, N = (long)mv->N_total;
    auto cvec = [&](long n) { return (C *)valloc(sizeof(C) * (size_t)n); };
    ths->y = cvec(M);
    ths->r_iter = cvec(M);
    ths->f_hat_iter = cvec(N);
    ths->p_hat_iter = cvec(N);
    ths->z_hat_iter = nullptr;
    ths->v_iter = nullptr;
    ths->w = nullptr;
    ths->w_hat = nullptr;
    if (flags & LANDWEBER) ths->z_hat_iter = ths->p_hat_iter;
    if (flags & STEEPEST_DESCENT) {
      ths->z_hat_iter = ths->p_hat_iter;
      ths->v_iter = cvec(M);
    }
    if (flags & CGNR) {
      ths->z_hat_iter = cvec(N);
      ths->v_iter = cvec(M);
    }
    if (flags & CGNE) ths->z_hat_iter = ths->p_hat_iter;
    if (flags & PRECOMPUTE_WEIGHT) ths->w = (R *)valloc(sizeof(R) * (size_t)M);
    if (flags & PRECOMPUTE_DAMP) ths->w_hat = (R *)valloc(sizeof(R) * (size_t)N);
    ths->alpha_iter = ths->beta_iter = 0;
    ths->dot_r_iter = ths->dot_r_iter_old = 0;
    ths->dot_z_hat_iter = ths->dot_z_hat_iter_old = 0;
    ths->dot_p_hat_iter = ths->dot_v_iter = 0;

    // the real-valued solver always runs on the host: its mv plan is the
    // caller's own (this library's plans are complex)
    Engine<R> *e = K == 2 ? engine_of_mv<R>(mv) : nullptr;
    if ((flags & CGNR) && e && !e->direct_only) {
      DevCgnr<R> *dv = new DevCgnr<R>();
      dv->alloc(e, M, N, flags & PRECOMPUTE_WEIGHT, flags & PRECOMPUTE_DAMP);
      if (const char *gv = std::getenv("NFFT_HIP_CGNR_GRAPH")) dv->use_graph = std::strcmp(gv, "0") != 0;
      auto &rg = registry<R>();
      std::lock_guard<std::mutex> g(rg.mu);
      rg.map[(void *)ths] = dv;
    }
  }

  // ---------------------------------------------------------------- device CGNR
  static void pull_scalars(SPLAN *ths, DevCgnr<R> *dv) {
    ths->dot_r_iter = (R)dv->sc_host[S_DOT_R];
    ths->dot_r_iter_old = (R)dv->sc_host[S_DOT_R_OLD];
    ths->dot_z_hat_iter = (R)dv->sc_host[S_DOT_Z];
    ths->dot_z_hat_iter_old = (R)dv->sc_host[S_DOT_Z_OLD];
    ths->dot_v_iter = (R)dv->sc_host[S_DOT_V];
    ths->alpha_iter = (R)dv->sc_host[S_ALPHA];
    ths->beta_iter = (R)dv->sc_host[S_BETA];
  }

  static void sync_out(SPLAN *ths, DevCgnr<R> *dv) {
    hipStream_t s = dv->e->stream;
    const size_t cM = sizeof(C) * (size_t)dv->M, cN = sizeof(C) * (size_t)dv->N;
    NFFT_HIP_CHECK(R, hipMemcpyAsync(dv->sc_host, dv->sc, S_COUNT * sizeof(double),
                                     hipMemcpyDeviceToHost, s));
    if (dv->sync_level >= 1 && cN)
      NFFT_HIP_CHECK(R, hipMemcpyAsync(ths->f_hat_iter, dv->fh, cN, hipMemcpyDeviceToHost, s));
    if (dv->sync_level >= 2) {
      if (cM) {
        NFFT_HIP_CHECK(R, hipMemcpyAsync(ths->r_iter, dv->r, cM, hipMemcpyDeviceToHost, s));
        NFFT_HIP_CHECK(R, hipMemcpyAsync(ths->v_iter, dv->v, cM, hipMemcpyDeviceToHost, s));
      }
      if (cN) {
        NFFT_HIP_CHECK(R, hipMemcpyAsync(ths->z_hat_iter, dv->z, cN, hipMemcpyDeviceToHost, s));
        NFFT_HIP_CHECK(R, hipMemcpyAsync(ths->p_hat_iter, dv->p, cN, hipMemcpyDeviceToHost, s));
      }
    }
    NFFT_HIP_CHECK(R, hipStreamSynchronize(s));
    pull_scalars(ths, dv);
  }

  static void dev_before_loop(SPLAN *ths, DevCgnr<R> *dv) {
    Engine<R> *e = dv->e;
    hipStream_t s = e->stream;
    MV *mv = ths->mv;
    const long M = dv->M, N = dv->N;
    const size_t cM = sizeof(C) * (size_t)M, cN = sizeof(C) * (size_t)N;
    // node data as the plan's trafo would use them
    ensure_nodes_of_mv<R>((void *)mv);
    if (cM) NFFT_HIP_CHECK(R, hipMemcpyAsync(dv->y, ths->y, cM, hipMemcpyHostToDevice, s));
    if (cN) NFFT_HIP_CHECK(R, hipMemcpyAsync(dv->fh, ths->f_hat_iter, cN, hipMemcpyHostToDevice, s));
    if (dv->w) NFFT_HIP_CHECK(R, hipMemcpyAsync(dv->w, ths->w, sizeof(R) * M, hipMemcpyHostToDevice, s));
    if (dv->w_hat)
      NFFT_HIP_CHECK(R, hipMemcpyAsync(dv->w_hat, ths->w_hat, sizeof(R) * N, hipMemcpyHostToDevice, s));
    // the reference copies f_hat_iter into mv->f_hat first (solver.c:83)
    if (cN) std::memcpy(mv->f_hat, ths->f_hat_iter, cN);
    // r = y - A f_hat_iter
    e->trafo(dv->fh, dv->r);
    if (M) hipLaunchKernelGGL(k_axpy_const<R>, dim3(blocks_for(M)), dim3(256), 0, s, dv->r, R(-1), dv->y, M);
    dv->dot(dv->r, dv->w, M, S_DOT_R, -1, 0);
    // z = A^H W r
    if (M) hipLaunchKernelGGL(k_cp_w<R>, dim3(blocks_for(M)), dim3(256), 0, s, dv->tmpM, dv->w, dv->r, M);
    e->adjoint(dv->tmpM, dv->z);
    dv->dot(dv->z, dv->w_hat, N, S_DOT_Z, -1, 0);
    // p = z
    if (cN) NFFT_HIP_CHECK(R, hipMemcpyAsync(dv->p, dv->z, cN, hipMemcpyDeviceToDevice, s));
    NFFT_HIP_CHECK(R, hipGetLastError());
    sync_out(ths, dv);
  }

  static void dev_step(SPLAN *ths, DevCgnr<R> *dv) {
    Engine<R> *e = dv->e;
    hipStream_t s = e->stream;
    if (dv->use_graph && !e->timing && !e->deferred) {
      if (!dv->step_exec || dv->step_stream != s) {
        if (dv->step_exec) NFFT_HIP_CHECK(R, hipGraphExecDestroy(dv->step_exec));
        dv->step_exec = nullptr;
        hipGraph_t graph;
        NFFT_HIP_CHECK(R, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        dev_step_kernels(dv);
        NFFT_HIP_CHECK(R, hipStreamEndCapture(s, &graph));
        NFFT_HIP_CHECK(R, hipGraphInstantiate(&dv->step_exec, graph, nullptr, nullptr, 0));
        NFFT_HIP_CHECK(R, hipGraphDestroy(graph));
        dv->step_stream = s;
      }
      NFFT_HIP_CHECK(R, hipGraphLaunch(dv->step_exec, s));
    } else {
      dev_step_kernels(dv);
    }
    sync_out(ths, dv);
  }

  // the kernels of one CGNR step, enqueued on the plan's stream: no
  // allocation, host copy or memset, so the sequence can be captured
  static void dev_step_kernels(DevCgnr<R> *dv) {
    Engine<R> *e = dv->e;
    hipStream_t s = e->stream;
    const long M = dv->M, N = dv->N;
    const unsigned bM = blocks_for(M), bN = blocks_for(N);
    // v = A W_hat p
    const cplx<R> *pin = dv->p;
    if (dv->w_hat) {
      hipLaunchKernelGGL(k_cp_w<R>, dim3(bN), dim3(256), 0, s, dv->tmpN, dv->w_hat, dv->p, N);
      pin = dv->tmpN;
    }
    e->trafo(pin, dv->v);
    // dot_v, alpha = dot_z / dot_v
    dv->dot(dv->v, dv->w, M, S_DOT_V, -1, 2);
    // f_hat += alpha W_hat p ; r -= alpha v
    if (N) hipLaunchKernelGGL(k_xpay_dev<R>, dim3(bN), dim3(256), 0, s, dv->fh, dv->sc, (int)S_ALPHA, R(1), dv->w_hat, dv->p, N);
    if (M) hipLaunchKernelGGL(k_xpay_dev<R>, dim3(bM), dim3(256), 0, s, dv->r, dv->sc, (int)S_ALPHA, R(-1), (const R *)nullptr, dv->v, M);
    dv->dot(dv->r, dv->w, M, S_DOT_R, -1, 0);
    // z = A^H W r
    if (M) hipLaunchKernelGGL(k_cp_w<R>, dim3(bM), dim3(256), 0, s, dv->tmpM, dv->w, dv->r, M);
    e->adjoint(dv->tmpM, dv->z);
    // dot_z (keeping the old one), beta = new / old
    dv->dot(dv->z, dv->w_hat, N, S_DOT_Z, S_DOT_Z_OLD, 4);
    // p = beta p + z
    if (N) hipLaunchKernelGGL(k_axpy_beta<R>, dim3(bN), dim3(256), 0, s, dv->p, dv->sc, dv->z, N);
    NFFT_HIP_CHECK(R, hipGetLastError());
  }

  // ---------------------------------------------------------------- host path
#define NFFT_HIP_CSWAP(a, b) \
  do {                       \
    C *t_ = (a);             \
    (a) = (b);               \
    (b) = t_;                \
  } while (0)

  static void host_before_loop(SPLAN *ths) {
    MV *mv = ths->mv;
    const long M = (long)mv->M_total, N = (long)mv->N_total;
    std::memcpy(mv->f_hat, ths->f_hat_iter, sizeof(C) * (size_t)N);
    NFFT_HIP_CSWAP(ths->r_iter, mv->f);
    mv->mv_trafo(mv);
    NFFT_HIP_CSWAP(ths->r_iter, mv->f);
    haxpy<R, K>((R *)ths->r_iter, R(-1), (const R *)ths->y, M);
    if (!(ths->flags & LANDWEBER) || (ths->flags & NORMS_FOR_LANDWEBER))
      ths->dot_r_iter = (R)hdot<R, K>((const R *)ths->r_iter, ths->flags & PRECOMPUTE_WEIGHT ? ths->w : nullptr, M);
    hcp_w<R, K>((R *)mv->f, ths->flags & PRECOMPUTE_WEIGHT ? ths->w : nullptr, (const R *)ths->r_iter, M);
    NFFT_HIP_CSWAP(ths->z_hat_iter, mv->f_hat);
    mv->mv_adjoint(mv);
    NFFT_HIP_CSWAP(ths->z_hat_iter, mv->f_hat);
    if (!(ths->flags & LANDWEBER) || (ths->flags & NORMS_FOR_LANDWEBER))
      ths->dot_z_hat_iter = (R)hdot<R, K>((const R *)ths->z_hat_iter, ths->flags & PRECOMPUTE_DAMP ? ths->w_hat : nullptr, N);
    if (ths->flags & CGNE) ths->dot_p_hat_iter = ths->dot_z_hat_iter;
    if (ths->flags & CGNR) std::memcpy(ths->p_hat_iter, ths->z_hat_iter, sizeof(C) * (size_t)N);
  }

  static void host_landweber(SPLAN *ths) {
    MV *mv = ths->mv;
    const long M = (long)mv->M_total, N = (long)mv->N_total;
    const R *wh = ths->flags & PRECOMPUTE_DAMP ? ths->w_hat : nullptr;
    const R *w = ths->flags & PRECOMPUTE_WEIGHT ? ths->w : nullptr;
    hxpawy<R, K>((R *)ths->f_hat_iter, ths->alpha_iter, wh, (const R *)ths->z_hat_iter, N);
    std::memcpy(mv->f_hat, ths->f_hat_iter, sizeof(C) * (size_t)N);
    NFFT_HIP_CSWAP(ths->r_iter, mv->f);
    mv->mv_trafo(mv);
    NFFT_HIP_CSWAP(ths->r_iter, mv->f);
    haxpy<R, K>((R *)ths->r_iter, R(-1), (const R *)ths->y, M);
    if (ths->flags & NORMS_FOR_LANDWEBER) ths->dot_r_iter = (R)hdot<R, K>((const R *)ths->r_iter, w, M);
    hcp_w<R, K>((R *)mv->f, w, (const R *)ths->r_iter, M);
    NFFT_HIP_CSWAP(ths->z_hat_iter, mv->f_hat);
    mv->mv_adjoint(mv);
    NFFT_HIP_CSWAP(ths->z_hat_iter, mv->f_hat);
    if (ths->flags & NORMS_FOR_LANDWEBER)
      ths->dot_z_hat_iter = (R)hdot<R, K>((const R *)ths->z_hat_iter, wh, N);
  }

  static void host_steepest(SPLAN *ths) {
    MV *mv = ths->mv;
    const long M = (long)mv->M_total, N = (long)mv->N_total;
    const R *wh = ths->flags & PRECOMPUTE_DAMP ? ths->w_hat : nullptr;
    const R *w = ths->flags & PRECOMPUTE_WEIGHT ? ths->w : nullptr;
    hcp_w<R, K>((R *)mv->f_hat, wh, (const R *)ths->z_hat_iter, N);
    NFFT_HIP_CSWAP(ths->v_iter, mv->f);
    mv->mv_trafo(mv);
    NFFT_HIP_CSWAP(ths->v_iter, mv->f);
    ths->dot_v_iter = (R)hdot<R, K>((const R *)ths->v_iter, w, M);
    ths->alpha_iter = ths->dot_z_hat_iter / ths->dot_v_iter;
    hxpawy<R, K>((R *)ths->f_hat_iter, ths->alpha_iter, wh, (const R *)ths->z_hat_iter, N);
    hxpawy<R, K>((R *)ths->r_iter, -ths->alpha_iter, nullptr, (const R *)ths->v_iter, M);
    ths->dot_r_iter = (R)hdot<R, K>((const R *)ths->r_iter, w, M);
    hcp_w<R, K>((R *)mv->f, w, (const R *)ths->r_iter, M);
    NFFT_HIP_CSWAP(ths->z_hat_iter, mv->f_hat);
    mv->mv_adjoint(mv);
    NFFT_HIP_CSWAP(ths->z_hat_iter, mv->f_hat);
    ths->dot_z_hat_iter = (R)hdot<R, K>((const R *)ths->z_hat_iter, wh, N);
  }

  static void host_cgnr(SPLAN *ths) {
    MV *mv = ths->mv;
    const long M = (long)mv->M_total, N = (long)mv->N_total;
    const R *wh = ths->flags & PRECOMPUTE_DAMP ? ths->w_hat : nullptr;
    const R *w = ths->flags & PRECOMPUTE_WEIGHT ? ths->w : nullptr;
    hcp_w<R, K>((R *)mv->f_hat, wh, (const R *)ths->p_hat_iter, N);
    NFFT_HIP_CSWAP(ths->v_iter, mv->f);
    mv->mv_trafo(mv);
    NFFT_HIP_CSWAP(ths->v_iter, mv->f);
    ths->dot_v_iter = (R)hdot<R, K>((const R *)ths->v_iter, w, M);
    ths->alpha_iter = ths->dot_z_hat_iter / ths->dot_v_iter;
    hxpawy<R, K>((R *)ths->f_hat_iter, ths->alpha_iter, wh, (const R *)ths->p_hat_iter, N);
    hxpawy<R, K>((R *)ths->r_iter, -ths->alpha_iter, nullptr, (const R *)ths->v_iter, M);
    ths->dot_r_iter = (R)hdot<R, K>((const R *)ths->r_iter, w, M);
    hcp_w<R, K>((R *)mv->f, w, (const R *)ths->r_iter, M);
    NFFT_HIP_CSWAP(ths->z_hat_iter, mv->f_hat);
    mv->mv_adjoint(mv);
    NFFT_HIP_CSWAP(ths->z_hat_iter, mv->f_hat);
    ths->dot_z_hat_iter_old = ths->dot_z_hat_iter;
    ths->dot_z_hat_iter = (R)hdot<R, K>((const R *)ths->z_hat_iter, wh, N);
    ths->beta_iter = ths->dot_z_hat_iter / ths->dot_z_hat_iter_old;
    haxpy<R, K>((R *)ths->p_hat_iter, ths->beta_iter, (const R *)ths->z_hat_iter, N);
  }

  static void host_cgne(SPLAN *ths) {
    MV *mv = ths->mv;
    const long M = (long)mv->M_total, N = (long)mv->N_total;
    const R *wh = ths->flags & PRECOMPUTE_DAMP ? ths->w_hat : nullptr;
    const R *w = ths->flags & PRECOMPUTE_WEIGHT ? ths->w : nullptr;
    ths->alpha_iter = ths->dot_r_iter / ths->dot_p_hat_iter;
    hxpawy<R, K>((R *)ths->f_hat_iter, ths->alpha_iter, wh, (const R *)ths->p_hat_iter, N);
    hcp_w<R, K>((R *)mv->f_hat, wh, (const R *)ths->p_hat_iter, N);
    mv->mv_trafo(mv);
    hxpawy<R, K>((R *)ths->r_iter, -ths->alpha_iter, nullptr, (const R *)mv->f, M);
    ths->dot_r_iter_old = ths->dot_r_iter;
    ths->dot_r_iter = (R)hdot<R, K>((const R *)ths->r_iter, w, M);
    ths->beta_iter = ths->dot_r_iter / ths->dot_r_iter_old;
    hcp_w<R, K>((R *)mv->f, w, (const R *)ths->r_iter, M);
    mv->mv_adjoint(mv);
    haxpy<R, K>((R *)ths->p_hat_iter, ths->beta_iter, (const R *)mv->f_hat, N);
    ths->dot_p_hat_iter = (R)hdot<R, K>((const R *)ths->p_hat_iter, wh, N);
  }
#undef NFFT_HIP_CSWAP

  static void before_loop(SPLAN *ths) {
    clear_launch_error();
    if (DevCgnr<R> *dv = K == 2 ? find_dev<R>(ths) : nullptr) dev_before_loop(ths, dv);
    else host_before_loop(ths);
  }

  static void loop_one_step(SPLAN *ths) {
    clear_launch_error();
    if (ths->flags & LANDWEBER) host_landweber(ths);
    if (ths->flags & STEEPEST_DESCENT) host_steepest(ths);
    if (ths->flags & CGNR) {
      if (DevCgnr<R> *dv = K == 2 ? find_dev<R>(ths) : nullptr) dev_step(ths, dv);
      else host_cgnr(ths);
    }
    if (ths->flags & CGNE) host_cgne(ths);
  }

  static void finalize(SPLAN *ths) {
    {
      auto &rg = registry<R>();
      std::lock_guard<std::mutex> g(rg.mu);
      auto it = rg.map.find((void *)ths);
      if (it != rg.map.end()) {
        hipStreamSynchronize(it->second->e->stream);
        it->second->release();
        delete it->second;
        rg.map.erase(it);
      }
    }
    if (ths->flags & PRECOMPUTE_WEIGHT) vfree(ths->w);
    if (ths->flags & PRECOMPUTE_DAMP) vfree(ths->w_hat);
    if (ths->flags & CGNR) {
      vfree(ths->v_iter);
      vfree(ths->z_hat_iter);
    }
    if (ths->flags & STEEPEST_DESCENT) vfree(ths->v_iter);
    vfree(ths->p_hat_iter);
    vfree(ths->f_hat_iter);
    vfree(ths->r_iter);
    vfree(ths->y);
  }

  static void set_sync(SPLAN *ths, int level) {
    if (DevCgnr<R> *dv = K == 2 ? find_dev<R>(ths) : nullptr) dv->sync_level = level < 0 ? 0 : level > 2 ? 2 : level;
  }

  static void sync_all(SPLAN *ths) {
    if (DevCgnr<R> *dv = K == 2 ? find_dev<R>(ths) : nullptr) {
      int keep = dv->sync_level;
      dv->sync_level = 2;
      sync_out(ths, dv);
      dv->sync_level = keep;
    }
  }
};

}  // namespace nfft_hip

using nfft_hip::SolverApi;

#define NFFT_HIP_EXPORT_SOLVER(S, R, SPLAN, MV, C)                                              \
  extern "C" {                                                                                  \
  void S##init_advanced_complex(SPLAN *ths, MV *mv, unsigned flags) {                           \
    SolverApi<R, SPLAN, MV, C>::init_advanced(ths, mv, flags);                                  \
  }                                                                                             \
  void S##init_complex(SPLAN *ths, MV *mv) { SolverApi<R, SPLAN, MV, C>::init_advanced(ths, mv, CGNR); } \
  void S##before_loop_complex(SPLAN *ths) { SolverApi<R, SPLAN, MV, C>::before_loop(ths); }     \
  void S##loop_one_step_complex(SPLAN *ths) { SolverApi<R, SPLAN, MV, C>::loop_one_step(ths); } \
  void S##finalize_complex(SPLAN *ths) { SolverApi<R, SPLAN, MV, C>::finalize(ths); }           \
  void S##set_sync_complex(SPLAN *ths, int level) { SolverApi<R, SPLAN, MV, C>::set_sync(ths, level); } \
  void S##sync_complex(SPLAN *ths) { SolverApi<R, SPLAN, MV, C>::sync_all(ths); }               \
  }

NFFT_HIP_EXPORT_SOLVER(solverf_, float, solverf_plan_complex, nfftf_mv_plan_complex, fftwf_complex)
NFFT_HIP_EXPORT_SOLVER(solver_, double, solver_plan_complex, nfft_mv_plan_complex, fftw_complex)

#define NFFT_HIP_EXPORT_SOLVER_DOUBLE(S, R, SPLAN, MV)                                          \
  extern "C" {                                                                                  \
  void S##init_advanced_double(SPLAN *ths, MV *mv, unsigned flags) {                            \
    SolverApi<R, SPLAN, MV, R>::init_advanced(ths, mv, flags);                                  \
  }                                                                                             \
  void S##init_double(SPLAN *ths, MV *mv) { SolverApi<R, SPLAN, MV, R>::init_advanced(ths, mv, CGNR); } \
  void S##before_loop_double(SPLAN *ths) { SolverApi<R, SPLAN, MV, R>::before_loop(ths); }      \
  void S##loop_one_step_double(SPLAN *ths) { SolverApi<R, SPLAN, MV, R>::loop_one_step(ths); }  \
  void S##finalize_double(SPLAN *ths) { SolverApi<R, SPLAN, MV, R>::finalize(ths); }            \
  }

NFFT_HIP_EXPORT_SOLVER_DOUBLE(solverf_, float, solverf_plan_double, nfftf_mv_plan_double)
NFFT_HIP_EXPORT_SOLVER_DOUBLE(solver_, double, solver_plan_double, nfft_mv_plan_double)
