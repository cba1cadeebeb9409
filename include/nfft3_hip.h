/*
 * nfft3_hip.h -- device-resident extensions of libnfft3_hip.so.
 *
 * None of these has a counterpart in the reference: NFFT3's trafo/adjoint
 * read and write host arrays only.  They serve callers that keep f_hat and f
 * in GPU memory (the benchmark, an iterative solver, a pipeline of GPU
 * stages) and want no PCIe copy per call.  Pointers named *_dev are HIP
 * device pointers holding interleaved complex values in the plan's precision
 * (fftwf_complex layout for nfftf_, fftw_complex for nfft_).  All work is
 * enqueued on the plan's stream; the calls return without waiting unless
 * noted.
 */
#ifndef NFFT3_HIP_EXT_H
#define NFFT3_HIP_EXT_H

#include "nfft3.h"

#ifdef __cplusplus
extern "C"
{
#endif

/* An instantiated hipGraph of batched transforms (see P##graph_trafo). */
typedef struct nfft_hip_graph_s nfft_hip_graph;

#define NFFT_HIP_DECLARE_EXT(P, PLAN, REAL)                                  \
  /* f_dev <- A f_hat_dev (same math as P##trafo) */                          \
  void P##trafo_device(PLAN *ths, const void *f_hat_dev, void *f_dev);       \
  /* f_hat_dev <- A^H f_dev (same math as P##adjoint) */                      \
  void P##adjoint_device(PLAN *ths, const void *f_dev, void *f_hat_dev);     \
  /* Run the plan's work on this hipStream_t from now on (NULL: the plan's   \
   * own stream). */                                                         \
  void P##set_stream(PLAN *ths, void *hip_stream);                           \
  void *P##get_stream(PLAN *ths);                                            \
  /* Wait for all work of the plan. */                                       \
  void P##synchronize(PLAN *ths);                                            \
  /* With on != 0, every trafo/adjoint records HIP events around its D, FFT  \
   * and B stages and, once the call's work has finished, stores their       \
   * durations in seconds in MEASURE_TIME_t[0..2] (the reference's           \
   * --enable-measure-time fields).  The *_device calls then wait for their  \
   * work before returning. */                                               \
  void P##set_timing(PLAN *ths, int on);                                     \
  /* Deferred stage timing for throughput runs: between timing_begin and     \
   * timing_end every trafo / adjoint records HIP events around its stages   \
   * on the plan's stream without waiting.  timing_end waits and writes     \
   * out[0..8] = number of trafo calls, number of adjoint calls, then the    \
   * summed seconds of trafo D, trafo FFT, trafo B (one interpolation        \
   * kernel), adjoint B^T spreading (one kernel), adjoint tile reduction     \
   * (one kernel, 0 on the atomic path), adjoint FFT, adjoint D^T. */        \
  void P##timing_begin(PLAN *ths);                                           \
  void P##timing_end(PLAN *ths, double *out);                                \
  /* Number of B-stage work items (sub-problems) built by the last           \
   * precompute, for diagnostics. */                                         \
  long P##num_subproblems(PLAN *ths);                                       \
  /* Gridding reconstruction f_hat = A^H (w .* f), w real (NULL: plain     \
   * adjoint); the weighted adjoint of the reference's MRI applications     \
   * (applications/mri/mri2d/reconstruct_data_gridding.c:54-68, which      \
   * scales f by the weights of weights.dat and calls nfft_adjoint).  f is  \
   * not modified.  The host form reads ths->f and w, writes ths->f_hat. */ \
  void P##gridding_device(PLAN *ths, const void *f_dev, const REAL *w_dev,  \
                          void *f_hat_dev);                                  \
  void P##gridding(PLAN *ths, const REAL *w);                                \
  /* nrhs right-hand sides sharing the plan's nodes (e.g. MRI coils):       \
   * vector r of f_hat_dev starts at r N_total, of f_dev at r M_total. */    \
  void P##trafo_batch_device(PLAN *ths, const void *f_hat_dev, void *f_dev, \
                             int nrhs);                                      \
  void P##adjoint_batch_device(PLAN *ths, const void *f_dev,                \
                               void *f_hat_dev, int nrhs);                   \
  /* adjoint_batch_device that allocates nothing.  No counterpart in the    \
   * reference; for callers that cannot allocate device memory (code under  \
   * HIP graph capture, tight serving loops).  The batch grids and spreading \
   * tiles live in one static 64 MiB device array per GPU, held until the    \
   * process ends, which each chunk of right-hand sides reuses; a right-hand \
   * side needing more than 64 MiB (grid plus tiles) is refused.  The plan   \
   * needs a PRE_* psi flag with precompute_one_psi done and stage timing    \
   * off.  Returns 0 when the work is enqueued, else a negative code (-1     \
   * nrhs < 0, -2 plan not ready, -3 scratch address unknown, -4 too large)  \
   * with a message on stderr, and enqueues nothing.  Calls on one GPU share \
   * the array: calls that overlap in time on one GPU (other streams, other  \
   * threads) race and must be serialised by the caller. */                 \
  int P##adjoint_batch_device_noalloc(PLAN *ths, const void *f_dev,         \
                                      void *f_hat_dev, int nrhs);            \
  /* Capture the batched transforms of count plans (nrhs each, device       \
   * arrays f_hat_dev[i] / f_dev[i] of plan i) into one hipGraph.  Plans on  \
   * different streams run concurrently on replay; the graph replays on     \
   * plans[0]'s stream.  Node data are built before the capture; stage      \
   * timing must be off.  Replays read the same device arrays. */           \
  nfft_hip_graph *P##graph_trafo(PLAN **plans, int count,                   \
                                 const void **f_hat_dev, void **f_dev,       \
                                 int nrhs);                                  \
  nfft_hip_graph *P##graph_adjoint(PLAN **plans, int count,                 \
                                   const void **f_dev, void **f_hat_dev,     \
                                   int nrhs);

NFFT_HIP_DECLARE_EXT(nfftf_, nfftf_plan, float)
NFFT_HIP_DECLARE_EXT(nfft_, nfft_plan, double)

/* Diagnostics: the split of one bin of cnt nodes into B-stage work items,
 * as the node binning computes it (at most max_per_sub nodes per item, at
 * most max_items items, 0: no limit).  Returns the item count k and writes
 * first[0..k-1] (offset in the bin) and count[0..k-1]; host only. */
int nfft_hip_work_split(int cnt, int max_per_sub, int max_items, int *first, int *count);

/* Replay a captured graph (enqueued, no wait), wait for it, release it. */
void nfft_hip_graph_launch(nfft_hip_graph *graph);
void nfft_hip_graph_synchronize(nfft_hip_graph *graph);
void nfft_hip_graph_destroy(nfft_hip_graph *graph);

/* CGNR on one of this library's plans keeps all solver vectors on the GPU.
 * The sync level says what is copied back to the host plan fields after
 * before_loop and after every step: 0 the scalars (alpha_iter, beta_iter,
 * dot_*), 1 also f_hat_iter, 2 (default) also r_iter, z_hat_iter,
 * p_hat_iter and v_iter -- the reference's full host state.  *_sync_complex
 * copies everything once.  Both are no-ops for solvers on the host path. */
#define NFFT_HIP_DECLARE_SOLVER_EXT(S, SPLAN)                                \
  void S##set_sync_complex(SPLAN *ths, int level);                           \
  void S##sync_complex(SPLAN *ths);

NFFT_HIP_DECLARE_SOLVER_EXT(solverf_, solverf_plan_complex)
NFFT_HIP_DECLARE_SOLVER_EXT(solver_, solver_plan_complex)

/* Density-compensation weights w (M_total floats) for the weighted CGNR
 * (solverf_init_advanced_complex with PRECOMPUTE_WEIGHT).  No counterpart in
 * the reference, which reads such weights from files.  Float family only.
 *
 * Pipe-Menon fixed point w <- w / (B B^T w) with the plan's own Kaiser-Bessel
 * interpolation B and spreading B^T, `iters` times from the w given (the
 * usual start is w = 1).  The _device form takes a device array and is
 * enqueued on the plan's stream; both return after the work is done. */
void nfftf_density_pipe_menon(nfftf_plan *ths, float *w, int iters);
void nfftf_density_pipe_menon_device(nfftf_plan *ths, float *w_dev, int iters);

/* Learned weights: a 3-layer MLP (80 -> 64 -> 64 -> 1, ReLU) on per-node
 * density features (d = 2 plans): the radius 2|x| and, on coarse grids of
 * 32, 64 and 128 cells per dimension, log(1 + count / mean count) of the node
 * counts in the 5 x 5 cells around the node (76 values, zero-padded to 80,
 * row-major [M_total][80] floats).  The model predicts log w; its dense
 * layers run on BF16 MFMA with FP32 accumulation.  Parameters are a flat
 * float array of nfft_dcf_num_params() values: W1 [80][64], b1 [64],
 * W2 [64][64], b2 [64], w3 [64], b3 (weights in [input][output] order). */
typedef struct nfft_dcf_model_s nfft_dcf_model;
int nfft_dcf_num_features(void);
int nfft_dcf_num_params(void);
void nfftf_dcf_features_device(nfftf_plan *ths, float *feat_dev);
nfft_dcf_model *nfft_dcf_model_create(unsigned seed);
void nfft_dcf_model_destroy(nfft_dcf_model *model);
void nfft_dcf_model_get_params(const nfft_dcf_model *model, float *params);
void nfft_dcf_model_set_params(nfft_dcf_model *model, const float *params);
/* Mean squared error of the predicted log w against target_dev[rows] and its
 * gradient with respect to every parameter (grad: nfft_dcf_num_params()
 * floats on the host, may be NULL); the parameters are not changed. */
float nfft_dcf_model_loss_grad(nfft_dcf_model *model, const float *feat_dev,
                               const float *target_dev, int rows, float *grad);
/* One Adam step (beta 0.9 / 0.999) on that loss; returns the loss before
 * the step. */
float nfft_dcf_model_train_step(nfft_dcf_model *model, const float *feat_dev,
                                const float *target_dev, int rows, float lr);
/* out_dev[rows] = predicted log w */
void nfft_dcf_model_predict(nfft_dcf_model *model, const float *feat_dev, int rows,
                            float *out_dev);
/* w_dev = exp(prediction) on the plan's nodes */
void nfftf_density_learned_device(nfftf_plan *ths, nfft_dcf_model *model, float *w_dev);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* NFFT3_HIP_EXT_H */
