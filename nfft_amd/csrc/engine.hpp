// Device-side state of one NFFT plan and the launchers of its kernels.
//
// One Engine<R> lives behind each nfft_plan / nfftf_plan (in the plan's
// my_fftw_plan1 field).  Data layout in HBM:
//   g        oversampled grid, n_total complex, row-major (last dim fastest),
//            transformed in place (D -> FFT -> B and B^T -> FFT -> D^T)
//   fhat, f  device images of the plan's f_hat (N_total) and f (M_total)
//   cell     d x M ints, SoA: wrapped cell floor(n_t x_t) mod n_t of every node,
//            in bin-sorted order
//   frac     d x M reals, SoA: n_t x_t - floor(n_t x_t) in [0,1), same order
//   perm     M ints: bin-sorted position -> node index of the caller's order
//   subs     work items of the B kernels: (bin, first sorted node, count)
#pragma once

#include <array>
#include <vector>

#include "common.hpp"

namespace nfft_hip {

constexpr int kMaxDim = 6;

// Geometry shared by the node-binning and B kernels.
struct BinGeom {
  int d;
  int m;
  int n[kMaxDim];   // oversampled grid lengths
  int bs[kMaxDim];  // bin edge in cells
  int nb[kMaxDim];  // bins per dimension
  int E[kMaxDim];   // LDS tile edge: bs + 2m + 1
  long M;
  // spreading tiles go to scratch in zone order (see spread_impl.hpp):
  // needs bs_t >= m + 1 and bs_t dividing n_t
  int zoned;
  // zoned only, built once per plan: zmap[o] = tile cell written to scratch
  // position o; zdesc[8 k + (0, 1..3, 4..6)] = offset in the tile, core
  // offset and size per dimension of the zone that neighbour k (digits
  // 0, 1, 2 = left, self, right per dimension) contributes to a bin's core
  const int *zmap;
  const int *zdesc;
  // batched launches (blockIdx.y = right-hand side): per right-hand side
  // offsets of the grid, the node vector and the scratch tiles; 0 otherwise
  long bs_g, bs_f, bs_t;
  // right-hand sides of a batched launch whose workgroups each take several
  // (the 2-D multi-right-hand-side spreading kernel); set at its launch
  int nrhs;
  // interpolation slots (see nodes.hip): slots[k] = (first node, count) of
  // up to slot_nodes consecutive nodes of one cell; islots[item] = (first
  // slot, slot count); null when the plan's interpolation kernel does not
  // use them
  const int2 *slots;
  const int2 *islots;
  // LDS row pitch (cells) of the 3-D lane-per-node interpolation tile:
  // E[2], or padded so that lanes reading windows of different cells of one
  // plane spread over more LDS banks; ipp its plane pitch (E[1] ip2, or
  // more)
  int ip2;
  int ipp;
  // 3-D run-merging spreading (spread_runs.hpp): the LDS tile's plane pitch
  // (cells; E[1] E[2] or padded) and the lane table, sp_lane[lane + 64 slot]
  // = the (dimension-0 tap, dimension-2 pair) index the lane owns (255:
  // none); sp_gc = cells per run along dimension 2
  int sp_pp;
  int sp_gc;
  unsigned char sp_lane[192];
};

// Kaiser-Bessel window of one dimension, scaled by 1/phi(0) so that the B
// weights stay O(1) (the scale is folded into the deconvolution factors).
template <typename R> struct WinParams {
  // Per dimension t and tap w, the scaled window psi(frac + m - w) as a
  // polynomial in (frac - 1/2) of degree poly_deg_v<R>, coefficients lowest
  // first at poly[(t * (2m+2) + w) * (poly_deg_v<R> + 1) + k]; fitted on the
  // host in long double (Chebyshev interpolation).  FP32, degree 8: the
  // largest error is 1.7e-8 of the peak at m = 4, sigma = 2, 1.9e-7 at m = 2
  // and below 5e-8 for m >= 3, against FP32's 6e-8 rounding of values near
  // the peak.  FP64, degree 20: at most 1.9e-16 of the peak for m = 1..16 and
  // sigma = 1.25, 2, 4 with the FP64 Horner evaluation, FP64 rounding.
  // The table is mirror-symmetric: tap 2m+1-w at t is tap w at -t, so its
  // coefficients are those of tap w with odd powers negated.
  const R *poly;
  // every dimension has the same sigma_t, hence the same tap polynomials
  int poly_shared;
};
constexpr int kPolyDeg = 8;
constexpr int kPolyDeg64 = 20;
template <typename R> constexpr int poly_deg_v = sizeof(R) == 4 ? kPolyDeg : kPolyDeg64;
// widest window (taps) with interpolation slots of up to four nodes and
// spreading runs of kRunCells cells; wider windows use pairs and one-cell runs
constexpr int kSlotQuadMaxW = 10;
constexpr int kRunCells = 3;
// right-hand sides per workgroup of the batched 2-D run-merging spreading
// (32 coils on 256^2: two 0.92 ms per batched trafo + adjoint, four 0.97,
// eight 1.30, one per workgroup 1.16)
constexpr int kRhsPerGroup = 2;
// wavefronts per workgroup of the 2-D run-merging spreading, each on its own
// share of the item's nodes and its own copy of the LDS tile
constexpr int kRunWaves2d = 2;
constexpr int kPolyCoef = kPolyDeg + 1;

// Twiddles and radix schedule of the FFT along one grid dimension.
template <typename R> struct FftAxis {
  int len;
  int nstages;
  int radix[20];
  cplx<R> *twiddle;  // exp(-2 pi i k / len), k < len
};

// Where an FFT pass reads its lines from and writes them to.  The band of
// dimension t is the N_t grid positions (k - N_t/2) mod n_t, k < N_t, that
// D fills and D^T reads.
enum FftIn { kFromGrid = 0, kFromGridBand = 1, kFromCoeffs = 2 };
enum FftOut { kToGrid = 0, kToGridBand = 1, kToCoeffs = 2 };

// LDS line stride (complex elements) of the FFT passes: odd, with one pad
// element per 16 (fft.hip)
__host__ __device__ inline int lds_line_stride(int n) { return n + (n >> 4) + 1; }

// One pruned pass of the oversampled FFT along dimension t (fft.hip): the
// lines are all combinations of the other dimensions' coordinates, each
// either over the full grid or over the band only (dimensions whose band is
// all that is nonzero, or all that is needed).
template <typename R> struct FftPass {
  int len;        // n_t
  long stride;    // element stride of dimension t in the grid
  int Nband;      // N_t
  int in_mode, out_mode;
  int nod;        // other dimensions (d - 1), outer first
  int cnt[kMaxDim - 1];     // lines along each other dimension
  int band[kMaxDim - 1];    // 1: that coordinate runs over the band
  int N_od[kMaxDim - 1], n_od[kMaxDim - 1];
  long gstr[kMaxDim - 1];   // grid strides of the other dimensions
  long fstr[kMaxDim - 1];   // f_hat strides of the other dimensions
  long fstr_ax;   // f_hat stride of dimension t
  const R *phi_ax;
  const R *phi_od[kMaxDim - 1];
  long nlines;
  int lines_per_block;
  int nstages;
  int radix[20];
  const cplx<R> *twiddle;
  int twiddle_in_lds;  // copy the twiddle table to LDS (when it fits 64 KB)
  // batched launches (blockIdx.y = right-hand side): offsets of the grid and
  // of the coefficient vectors per right-hand side
  long bs_g, bs_c;
  // Four-step split of an axis longer than one workgroup's LDS lines (or
  // whose length has a factor no single stage handles), n = n1 n2 with the
  // axis index j1 + n1 j2:
  //   four = 1: lines (other coordinates, j1), elements j2, length n2; the
  //             result k2 is multiplied by w_n^(j1 k2) and written back in
  //             place (position j1 + n1 k2); the input's band test uses the
  //             axis index j1 + n1 j2
  //   four = 2: lines (other coordinates, k2), elements j1 at j1 + n1 k2,
  //             length n1; the result k1 is the axis index k2 + n2 k1, where
  //             it is written (to the other grid buffer) with the band rule
  //             of out_mode
  // four = 0: an ordinary pass (nsub = 1).
  int four;
  int nsub;     // lines per line of the other coordinates: n1 or n2
  int nfull;    // n
  int fn1, fn2;
  const cplx<R> *twiddle_full;  // exp(-2 pi i k / n), k < n
  // grid buffer the pass reads and the one it writes: 0 the plan's grid, 1
  // its second grid (four-step plans only)
  int src, dst;
  // lines per workgroup relative to the launch's default choice: times
  // 2^lpb_shift (FFTW_MEASURE plans time -1, 0, +1)
  int lpb_shift;
};

template <typename R> struct Engine {
  // geometry
  int d = 0;
  static_assert(kMaxDim == 6, "unit extents below");
  long N[kMaxDim] = {1, 1, 1, 1, 1, 1};
  long n[kMaxDim] = {1, 1, 1, 1, 1, 1};
  long N_total = 0, n_total = 0, M = 0;
  int m = 0;
  bool direct_only = false;  // some N_t <= m or n_t <= 2m+2: NDFT path

  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  bool timing = false;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  bool deferred = false;  // timing_begin .. timing_end
  std::vector<std::array<hipEvent_t, 5>> log_ev;
  std::vector<int> log_kind;  // 0 trafo, 1 adjoint
  size_t log_n = 0;

  // device buffers
  cplx<R> *g = nullptr;
  cplx<R> *fhat = nullptr;
  cplx<R> *f = nullptr;
  R *xdev = nullptr;
  R *phi_inv = nullptr;  // sum_t N_t scaled deconvolution factors
  R *poly = nullptr;     // window polynomials (WinParams::poly)
  int *cell = nullptr;
  R *frac = nullptr;
  int *perm = nullptr;
  int4 *subs = nullptr;  // interpolation work items
  long nsub = 0;
  long subs_cap = 0;
  // spreading work items; per bin its first item and item count
  int4 *sp_subs = nullptr;
  long sp_nsub = 0, sp_subs_cap = 0;
  int *sp_first = nullptr, *sp_count = nullptr;
  int *bin_order = nullptr;  // zone reduction's bin order (inside sp_first's allocation)
  long bin_arrays_cap = 0;
  int spread_max_per_sub = 2048;
  int spread_max_items = 0;  // spreading items per bin at most (0: no limit)
  // spreading tiles of all work items (no HBM atomics); null: atomic flush
  cplx<R> *tiles = nullptr;
  long tiles_cap = 0;
  bool tile_scratch = false;
  int *scratch_i = nullptr;  // binning scratch
  int *zone_tab = nullptr;   // geom.zmap and geom.zdesc
  int2 *slots = nullptr, *islots = nullptr;  // geom.slots, geom.islots
  int2 *slots_raw = nullptr;                 // slots in node order before sorting
  long slots_cap = 0, islots_cap = 0, nslots = 0;
  // FP32 3-D interpolation reads the taps of two same-cell nodes once
  // (NFFT_HIP_PAIRS=0: off).  In 2-D the extra window work outweighs the
  // saved LDS reads (cfg2-like plans ran 64% slower in B), so 2-D keeps the
  // single-node kernel.
  int pair_interp = 1;
  bool use_slots() const { return sizeof(R) == 4 && d == 3 && pair_interp; }
  // nodes per interpolation slot: 4 (quads where a cell has that many;
  // wavefronts whose slots hold at most two run the pair code), or 2
  // (NFFT_HIP_SLOT=2)
  int slot_nodes = 4;
  // each work item's slots ordered by node count, largest first, and dealt
  // over LDS banks (NFFT_HIP_SLOTSORT=0: node order)
  int slot_sort = 1;
  // device scratch of the density-weight helpers (features, Pipe-Menon),
  // grown on demand and kept with the plan so repeated calls allocate nothing
  char *aux = nullptr;
  size_t aux_cap = 0;
  long scratch_cap = 0;
  bool nodes_ready = false;

  BinGeom geom{};
  WinParams<R> win{};
  int max_per_sub = 1024;
  // B kernel shapes: 0 lane per node (lane), 1 wave per node (wave); for
  // spreading also 2 private tile per wavefront (priv), 3 row-partitioned
  // tile (rows), the FP32 3-D run-merging kernels 10 (runs2sbc: two node
  // streams per wavefront) and 13 (runs2rs: one stream, runs load and store
  // their tile cells), and the 2-D run-merging kernel 14 (runs2d).
  // Set at plan creation; NFFT_HIP_INTERP / NFFT_HIP_SPREAD = name override.
  int interp_variant = 0;
  int spread_variant = 1;
  // lane-per-node FP32 interpolation evaluates tap pairs from the shared
  // mirror-symmetric window table when win.poly_shared (NFFT_HIP_SYMWIN=0: off)
  int sym_window = 1;
  std::vector<FftAxis<R>> axes;
  std::vector<FftPass<R>> fwd_passes;  // trafo: D fused into the first pass
  std::vector<FftPass<R>> bwd_passes;  // adjoint: D^T fused into the last pass
  // four-step axes: their twiddle tables (sub-lengths and full length) and
  // the second grid their passes write; trafo_grid = the buffer holding the
  // trafo's transformed grid (0 g, 1 g2)
  std::vector<cplx<R> *> extra_twiddles;
  cplx<R> *g2 = nullptr;
  int trafo_grid = 0;
  cplx<R> *grid_after_trafo() const { return trafo_grid ? g2 : g; }

  void create(int d_, const long *N_, const long *n_, long M_, int m_);
  void set_bins(const int *bs);
  void destroy();
  void upload_nodes(const R *x_host);
  void build_nodes();  // from xdev
  // FFTW_MEASURE plans: time, on the plan's own nodes, the bin edges (B and
  // B^T together), the B and B^T kernel shapes, the spreading work-item size
  // and the FFT lines per workgroup once, and keep the fastest of each
  // (overrides in the environment win)
  bool tuned = false;
  void autotune();
  void trafo(const cplx<R> *fhat_dev, cplx<R> *f_dev);
  // B^T alone: g <- B^T f (spreading, then the tile reduction or a zeroed
  // grid with atomics); B alone: f <- B g.  Node data must be built.
  void spread_to_grid(const cplx<R> *f_dev);
  void interp_from_grid(cplx<R> *f_dev);
  void adjoint(const cplx<R> *f_dev, cplx<R> *fhat_dev);
  // nrhs right-hand sides with the same nodes in batched launches (grid y =
  // right-hand side) when the plan's kernels support it, else one by one
  bool batchable() const;
  void reserve_batch(int nrhs);
  void trafo_batch(const cplx<R> *fhat_dev, cplx<R> *f_dev, int nrhs);
  void adjoint_batch(const cplx<R> *f_dev, cplx<R> *fhat_dev, int nrhs);
  void adjoint_batch_chunks(const cplx<R> *f_dev, cplx<R> *fhat_dev, int nrhs, int k,
                            cplx<R> *grids, cplx<R> *tls);
  // adjoint_batch with its batch grids and tiles in the library's static
  // per-GPU scratch array (plan.hip) instead of plan allocations: 0 when the
  // work is enqueued, -3 when the scratch address is unknown, -4 when one
  // right-hand side needs more than the array holds
  int device = 0;
  void *noalloc_scratch = nullptr;  // device address of the scratch array
  long noalloc_bytes_per_rhs() const;
  int adjoint_batch_noalloc(const cplx<R> *f_dev, cplx<R> *fhat_dev, int nrhs);
  cplx<R> *bgrid = nullptr;   // batched grids
  cplx<R> *btiles = nullptr;  // batched spreading tiles
  long bgrid_cap = 0, btiles_cap = 0;  // in complex values
  int batch_chunk(int nrhs) const;
  void trafo_direct(const cplx<R> *fhat_dev, cplx<R> *f_dev);
  void adjoint_direct(const cplx<R> *f_dev, cplx<R> *fhat_dev);
  void stage_times(double out[4]);
  hipEvent_t *call_events(int kind);
  void timing_begin();
  void timing_end(double *out);
};

// Host arrays owned by the library (plan MALLOC_* arrays, solver vectors):
// page-locked unless the family's malloc hook is set.
template <typename R> void *host_array_alloc(size_t bytes);
template <typename R> void host_array_free(void *p);
// Engine behind a plan whose mv_trafo is this library's trafo, else nullptr.
template <typename R> Engine<R> *engine_of_mv(void *mv);
// Node data of such a plan, built as its trafo would build them.
template <typename R> void ensure_nodes_of_mv(void *mv);

// ---- launchers (explicitly instantiated for float and double) ----
template <typename R>
void launch_fill_zero(cplx<R> *p, long count, hipStream_t s);
template <typename R>
void launch_deconv_to_grid(const cplx<R> *fhat, cplx<R> *g, const R *phi_inv, int d,
                           const long *N, const long *n, hipStream_t s);
template <typename R>
void launch_deconv_from_grid(const cplx<R> *g, cplx<R> *fhat, const R *phi_inv, int d,
                             const long *N, const long *n, hipStream_t s);
// nb > 1: nb right-hand sides in one launch (grid y), offsets from the
// pass's / geometry's batch strides; gb / tiles override e.geom / e.tiles
template <typename R>
void launch_fft_pass(cplx<R> *g, const cplx<R> *fhat_in, cplx<R> *fhat_out, const FftPass<R> &p,
                     int sign, hipStream_t s, int nb = 1, cplx<R> *g2 = nullptr);
template <typename R>
void launch_interp(const cplx<R> *g, cplx<R> *f, const Engine<R> &e, hipStream_t s, int nb = 1,
                   const BinGeom *gb = nullptr);
template <typename R>
void launch_spread(const cplx<R> *f, cplx<R> *g, const Engine<R> &e, hipStream_t s, int nb = 1,
                   const BinGeom *gb = nullptr, cplx<R> *tiles = nullptr);
template <typename R>
void launch_gather_tiles(const Engine<R> &e, cplx<R> *g, hipStream_t s, int nb = 1,
                         const BinGeom *gb = nullptr, cplx<R> *tiles = nullptr);
template <typename R>
bool batch_kernels(const Engine<R> &e);
template <typename R>
void launch_ndft(const R *x, const cplx<R> *in, cplx<R> *out, int d, const long *N, long M,
                 int adjoint, hipStream_t s);
template <typename R>
void bin_nodes(Engine<R> &e);

}  // namespace nfft_hip
