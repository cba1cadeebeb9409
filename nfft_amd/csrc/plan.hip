// The NFFT plan API of the C ABI (include/nfft3.h) for the nfftf_ (float)
// and nfft_ (double) families, and the Engine that runs a plan on the GPU.
//
// Reference behaviour followed (kernel/nfft/nfft.c):
//   init_help 5950-6046   geometry, sigma, b, MALLOC_* ownership, phi_hut
//   init / init_guru / init_lin / init_1d,2d,3d 6048-6167   defaults: n_t =
//       2 next_power_of_2(N_t), m = 4 (float) / 8 (double), flags
//   precompute_one_psi 5938-5948, trafo 5655-5701, adjoint 5703-5749 (direct
//       NDFT whenever some N_t <= m or n_t <= 2m+2), check 6169-6207,
//       finalize 6209-6270
// Host arrays stay authoritative: every trafo/adjoint copies f_hat / f at the
// plan's current pointers (the solver swaps them between calls).  Node data
// are taken from x at precompute time when a PRE_* psi flag is set, and from
// the current x on every call otherwise, as the reference does.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>

#include "../../include/nfft3.h"
#include "../../include/nfft3_hip.h"
#include "engine.hpp"

// instantiated hipGraph of batched transforms (include/nfft3_hip.h)
struct nfft_hip_graph_s {
  hipGraphExec_t exec;
  hipStream_t stream;
};

namespace nfft_hip {

// ------------------------------------------------------------------ scratch
// Device memory of the allocation-free batched adjoint
// (nfft[f]_adjoint_batch_device_noalloc): the batch grids and spreading
// tiles of one chunk of right-hand sides.  One copy per GPU the library runs
// on, held until the process ends.  On the host the array's name is only a
// stand-in for the symbol; its device address on each GPU comes from
// hipGetSymbolAddress, looked up once per device when the first plan is
// created there, and only that address is handed to kernels.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_mu;
static void *g_scratch_dev[kNoAllocMaxDevices];

// device address of `scratch` on device dev (the current device), looked up
// on the first call; nullptr when HIP does not provide it
static void *scratch_setup(int dev) {
  if (dev < 0 || dev >= kNoAllocMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch_dev[dev] = ptr;
  }
  return g_scratch_dev[dev];
}

// ------------------------------------------------------------------ helpers
// I0 by its power series sum_k ((x/2)^2)^k / (k!)^2, in long double.
static long double bessel_i0(long double x) {
  long double q = x * x / 4.0L, term = 1.0L, sum = 1.0L;
  for (int k = 1; k < 4000; k++) {
    term *= q / ((long double)k * (long double)k);
    sum += term;
    if (term < sum * 1e-21L) break;
  }
  return sum;
}

static const long double kPiL = 3.141592653589793238462643383279502884L;

// PHI_HUT(n, k) = I0(m sqrt(b^2 - (2 pi k / n)^2))  (include/infft.h:208)
static long double phi_hut(long n, long k, int m, long double b) {
  long double w = 2.0L * kPiL * (long double)k / (long double)n;
  return bessel_i0((long double)m * sqrtl(b * b - w * w));
}

// The window of include/infft.h:209-215 divided by its peak, in long double:
// sinh(b s)/(pi s) for s = sqrt(m^2 - y^2) > 0, sin(b s)/(pi s) beyond |y| = m.
static long double psi_scaled(long double y, int m, long double b) {
  const long double q = (long double)m * m - y * y;
  const long double peak = sinhl(b * m) / (kPiL * m);
  long double v;
  if (q > 0) v = sinhl(b * sqrtl(q)) / (kPiL * sqrtl(q));
  else if (q < 0) v = sinl(b * sqrtl(-q)) / (kPiL * sqrtl(-q));
  else v = b / kPiL;
  return v / peak;
}

// Per tap w = 0..2m+1 of one dimension, psi(frac + m - w) for frac in [0,1)
// as a polynomial in t = frac - 1/2 of degree poly_deg_v<R> (8 for FP32, 20
// for FP64): Chebyshev interpolation at 32 (FP32) or 48 (FP64) points,
// converted to monomials in t.  The window is an entire function of y^2, so
// the fit converges geometrically.
template <typename R>
static void fit_window_polys(int m, long double b, std::vector<R> &out) {
  constexpr int DEG = poly_deg_v<R>;
  constexpr int K = DEG <= 8 ? 32 : 48;
  const size_t first = out.size();
  for (int w = 0; w < 2 * m + 2; w++) {
    long double f[K], u[K];
    for (int j = 0; j < K; j++) {
      u[j] = cosl(kPiL * (j + 0.5L) / K);  // u = 2t in [-1, 1]
      f[j] = psi_scaled(0.5L * u[j] + 0.5L + m - w, m, b);
    }
    long double c[DEG + 1];
    for (int k = 0; k <= DEG; k++) {
      long double s = 0;
      for (int j = 0; j < K; j++) s += f[j] * cosl(k * kPiL * (j + 0.5L) / K);
      c[k] = s * 2.0L / K;
    }
    c[0] *= 0.5L;
    // Chebyshev series in u -> monomials in u (T_{k+1} = 2u T_k - T_{k-1})
    long double mono[DEG + 1] = {0}, tkm1[DEG + 1] = {0}, tk[DEG + 1] = {0};
    tkm1[0] = 1;  // T0
    tk[1] = 1;    // T1
    for (int i = 0; i <= DEG; i++) mono[i] += c[0] * tkm1[i];
    if (DEG >= 1)
      for (int i = 0; i <= DEG; i++) mono[i] += c[1] * tk[i];
    for (int k = 2; k <= DEG; k++) {
      long double tn[DEG + 1] = {0};
      for (int i = 0; i <= DEG; i++) {
        if (i > 0) tn[i] += 2 * tk[i - 1];
        tn[i] -= tkm1[i];
      }
      for (int i = 0; i <= DEG; i++) {
        mono[i] += c[k] * tn[i];
        tkm1[i] = tk[i];
        tk[i] = tn[i];
      }
    }
    // u = 2t: coefficient of t^i is mono[i] 2^i
    for (int i = 0; i <= DEG; i++) out.push_back((R)(mono[i] * ldexpl(1.0L, i)));
  }
  // psi is even, so psi(frac + m - (2m+1-w)) = psi((1 - frac) + m - w): the
  // upper taps are the lower ones at -t.  Written as exact mirrors so that
  // kernels may evaluate a tap pair from one polynomial's even and odd parts.
  for (int w = m + 1; w < 2 * m + 2; w++)
    for (int i = 0; i <= DEG; i++)
      out[first + w * (DEG + 1) + i] = (i % 2 ? -1 : 1) * out[first + (2 * m + 1 - w) * (DEG + 1) + i];
}

// factor n into FFT radices: 8, 4, 2, the small primes with unrolled
// butterflies, then any other prime factor (a generic stage, O(n p) per
// line); at most 20 stages
static bool fft_radices(long n, std::vector<int> &out) {
  out.clear();
  while (n % 8 == 0 && n > 1) { out.push_back(8); n /= 8; }
  while (n % 4 == 0 && n > 1) { out.push_back(4); n /= 4; }
  while (n % 2 == 0 && n > 1) { out.push_back(2); n /= 2; }
  for (int p : {3, 5, 7, 11, 13})
    while (n % p == 0) { out.push_back(p); n /= p; }
  for (long p = 17; p * p <= n; p += 2)
    while (n % p == 0) { out.push_back((int)p); n /= p; }
  if (n > 1) out.push_back((int)n);
  return out.size() <= 20;
}

// Lane table and plane pitch of the 3-D run-merging spreading kernel for W
// taps, runs of gc cells along dimension 2 and an E0 x E1 x E2 tile.  Its
// lanes own (dimension-0 tap, dimension-2 pair) cells and one flush
// instruction updates one dimension-1 row of every lane's cells, at tile
// offsets pa pp + pb from a common base.  Lanes whose 8-byte cells share an
// LDS bank serialise, so the table and the padding of the plane pitch (up
// to what keeps the same number of tiles per CU in 160 KB of LDS) minimise
// the bank collisions of those offsets: per 16-lane group the largest count
// of one residue mod 16, plus per 32-lane group of one residue mod 32
// (8-byte units; either grouping may be the one the LDS serves 64-bit
// accesses in).  Greedy placement, then pairwise swaps from a fixed seed;
// deterministic, cached per geometry.
static void spread_lane_table(int W, int gc, int E0, int E1, int E2, int &pp_out,
                              unsigned char *table) {
  static std::mutex mu;
  static std::map<std::array<int, 5>, std::pair<int, std::vector<unsigned char>>> cache;
  const std::array<int, 5> key{W, gc, E0, E1, E2};
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it == cache.end()) {
    const int hw = (W + gc) / 2, pairs = W * hw, slots = (pairs + 63) / 64, nl = 64 * slots;
    const int plane = E1 * E2;
    const long lds = 160 * 1024 / 8;  // cells
    const long per_cu = lds / ((long)E0 * plane);
    const int pmax = per_cu > 0 ? (int)std::min<long>(32, lds / per_cu / E0 - plane) : 0;
    auto cost = [&](const std::vector<int> &t, int pp) {
      int c = 0;
      for (int g = 0; g < nl; g += 16) {
        int h[16] = {0}, mx = 0;
        for (int l = g; l < g + 16; l++)
          if (t[l] < pairs) mx = std::max(mx, ++h[((t[l] / hw) * pp + 2 * (t[l] % hw)) & 15]);
        c += mx;
      }
      for (int g = 0; g < nl; g += 32) {
        int h[32] = {0}, mx = 0;
        for (int l = g; l < g + 32; l++)
          if (t[l] < pairs) mx = std::max(mx, ++h[((t[l] / hw) * pp + 2 * (t[l] % hw)) & 31]);
        c += mx;
      }
      return c;
    };
    int best_pp = plane, best_c = 1 << 30;
    std::vector<int> best;
    for (int pad = 0; pad <= std::max(0, pmax); pad++) {
      const int pp = plane + pad;
      std::vector<int> t(nl, 255);
      for (int q = 0; q < pairs; q++) t[q] = q;
      int c = cost(t, pp);
      unsigned s = 12345u;
      for (int it2 = 0; it2 < 3000 && c > 0; it2++) {
        s = s * 1664525u + 1013904223u;
        const int i = (int)((s >> 8) % (unsigned)nl);
        s = s * 1664525u + 1013904223u;
        const int j = (int)((s >> 8) % (unsigned)nl);
        if (t[i] == t[j]) continue;
        std::swap(t[i], t[j]);
        const int c2 = cost(t, pp);
        if (c2 <= c) c = c2;
        else std::swap(t[i], t[j]);
      }
      if (c < best_c) {
        best_c = c;
        best_pp = pp;
        best = t;
      }
    }
    std::vector<unsigned char> tab(192, 255);
    for (int l = 0; l < nl && l < 192; l++) tab[l] = (unsigned char)best[l];
    it = cache.emplace(key, std::make_pair(best_pp, tab)).first;
  }
  pp_out = it->second.first;
  std::memcpy(table, it->second.second.data(), 192);
}

// Bin edges bs[t] (and with them the LDS tile edge bs + 2m + 1): the
// geometry shared by the node binning and the B / B^T kernels, whether B^T
// reduces its tiles through scratch, and the zone tables of that reduction.
// Node data must be rebuilt after a change (build_nodes).
template <typename R>
void Engine<R>::set_bins(const int *bs) {
  geom.d = d;
  geom.m = m;
  geom.M = M;
  for (int t = 0; t < kMaxDim; t++) {
    geom.n[t] = t < d ? (int)n[t] : 1;
    geom.bs[t] = t < d ? bs[t] : 1;
    geom.nb[t] = t < d ? (int)ceil_div(n[t], bs[t]) : 1;
    geom.E[t] = t < d ? bs[t] + 2 * m + 1 : 1;
  }
  geom.ip2 = geom.E[2];
  geom.ipp = geom.E[1] * geom.E[2];
  if (d == 3) {
    // interpolation tile rows padded to 4 (mod 8) cells (an 8 x 8 plane of
    // distinct cells then maps to 32 banks of 8 bytes 2-way instead of 4-way
    // at 17 cells), when the tile still fits; NFFT_HIP_IPITCH=0: unpadded,
    // NFFT_HIP_IPITCH=p (p >= E[2]): rows p cells apart; NFFT_HIP_IPLANE=q:
    // q more cells between planes
    const char *ip = std::getenv("NFFT_HIP_IPITCH");
    const int want = ip ? std::atoi(ip) : -1;
    int p2 = geom.E[2];
    if (want >= geom.E[2]) {
      p2 = want;
    } else if (want != 0) {
      // (at most 3 cells of padding per row: measured at m = 4, 17 -> 20)
      while (p2 % 8 != 4) p2++;
      if (p2 - geom.E[2] > 3) p2 = geom.E[2];
    }
    int pp = geom.E[1] * p2;
    if (const char *v = std::getenv("NFFT_HIP_IPLANE")) pp += std::max(0, std::atoi(v));
    if ((size_t)geom.E[0] * pp * sizeof(cplx<R>) <= 64 * 1024) {
      geom.ip2 = p2;
      geom.ipp = pp;
    }
  }
  // the run-merging spreading kernel's lane table (3-D, 2m + 2 taps; runs
  // of kRunCells cells along dimension 2 where the bin edge allows)
  geom.sp_gc = 1;
  geom.sp_pp = geom.E[1] * geom.E[2];
  std::memset(geom.sp_lane, 255, sizeof(geom.sp_lane));
  if (d == 3 && 2 * m + 2 <= 20) {
    geom.sp_gc = geom.bs[2] >= kRunCells && 2 * m + 2 <= kSlotQuadMaxW ? kRunCells : 1;
    spread_lane_table(2 * m + 2, geom.sp_gc, geom.E[0], geom.E[1], geom.E[2], geom.sp_pp,
                      geom.sp_lane);
  }
  // B^T writes its LDS tiles to scratch and reduces them without atomics
  // when the tile fits LDS and never wraps onto itself (E_t <= n_t)
  {
    size_t tb = sizeof(cplx<R>);
    bool fits = true;
    for (int t = 0; t < d; t++) {
      tb *= (size_t)geom.E[t];
      fits = fits && geom.E[t] <= n[t];
    }
    // (the tile reductions cover d <= 3; d = 4 adds its tiles with atomics)
    tile_scratch = fits && tb <= 64 * 1024 && d <= 3;
    bool zoned = tile_scratch;
    long core = 1;
    for (int t = 0; t < d; t++) {
      zoned = zoned && geom.bs[t] >= m + 1 && n[t] % geom.bs[t] == 0;
      core *= geom.bs[t];
    }
    zoned = zoned && core <= 4096;  // the zone reduction's cells per workgroup
    if (const char *z = std::getenv("NFFT_HIP_ZONES")) zoned = zoned && std::strcmp(z, "0") != 0;
    geom.zoned = zoned ? 1 : 0;
    geom.zmap = geom.zdesc = nullptr;
    if (zone_tab) NFFT_HIP_CHECK(R, hipFree(zone_tab));
    zone_tab = nullptr;
    if (zoned) {
      // per dimension the tile splits into the part in the left neighbour's
      // core (m cells), the own core (bs), the part in the right
      // neighbour's core (m + 1); zones are written in base-3 order of
      // their (z_0, ..., z_{d-1}) digits, each row-major
      auto zstart = [&](int t, int z) { return z == 0 ? 0 : (z == 1 ? m : m + geom.bs[t]); };
      auto zsize = [&](int t, int z) { return z == 0 ? m : (z == 1 ? geom.bs[t] : m + 1); };
      int zc = 1;
      for (int t = 0; t < d; t++) zc *= 3;
      std::vector<int> tab;
      std::vector<long> zbase(zc + 1, 0);
      for (int zz = 0; zz < zc; zz++) {
        int dg[kMaxDim] = {}, r = zz;
        for (int t = d - 1; t >= 0; t--) {
          dg[t] = r % 3;
          r /= 3;
        }
        long cnt = 1;
        for (int t = 0; t < d; t++) cnt *= zsize(t, dg[t]);
        for (long q = 0; q < cnt; q++) {
          long rem = q, src = 0, stride = 1;
          for (int t = d - 1; t >= 0; t--) {
            const int sz = zsize(t, dg[t]);
            src += (zstart(t, dg[t]) + rem % sz) * stride;
            rem /= sz;
            stride *= geom.E[t];
          }
          tab.push_back((int)src);
        }
        zbase[zz + 1] = zbase[zz] + cnt;
      }
      const size_t nmap = tab.size();
      // neighbour k with digits dg (0 left, 1 self, 2 right) contributes
      // its zone 2 - dg_t per dimension, landing at core offset bs - m
      // (right neighbour's left part goes to the top of the core) or 0
      for (int k = 0; k < zc; k++) {
        int dg[kMaxDim] = {}, r = k;
        for (int t = d - 1; t >= 0; t--) {
          dg[t] = r % 3;
          r /= 3;
        }
        int zidx = 0;
        int off[3] = {0, 0, 0}, size[3] = {1, 1, 1};
        for (int t = 0; t < d; t++) {
          const int z = 2 - dg[t];
          zidx = zidx * 3 + z;
          size[t] = zsize(t, z);
          off[t] = dg[t] == 2 ? geom.bs[t] - m : 0;
        }
        tab.push_back((int)zbase[zidx]);
        for (int t = 0; t < 3; t++) tab.push_back(off[t]);
        for (int t = 0; t < 3; t++) tab.push_back(size[t]);
        tab.push_back(0);
      }
      NFFT_HIP_CHECK(R, hipMalloc(&zone_tab, sizeof(int) * tab.size()));
      NFFT_HIP_CHECK(R, hipMemcpy(zone_tab, tab.data(), sizeof(int) * tab.size(),
                                  hipMemcpyHostToDevice));
      geom.zmap = zone_tab;
      geom.zdesc = zone_tab + nmap;
    }
  }
}

template <typename R>
void Engine<R>::create(int d_, const long *N_, const long *n_, long M_, int m_) {
  d = d_;
  M = M_;
  m = m_;
  N_total = 1;
  n_total = 1;
  for (int t = 0; t < d; t++) {
    N[t] = N_[t];
    n[t] = n_[t];
    N_total *= N[t];
    n_total *= n[t];
    if (N[t] <= m || n[t] <= 2 * m + 2) direct_only = true;
  }
  // an oversampled grid shorter than N would wrap frequencies onto each other
  for (int t = 0; t < d; t++)
    if (N[t] < 1 || (!direct_only && n[t] < N[t]))
      family<R>::fatal("nfft: need N[%d] >= 1 and n[%d] >= N[%d] (got N = %ld, n = %ld)", t, t, t,
                       N[t], n[t]);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    family<R>::fatal("nfft: no HIP device available (this library runs on the GPU only)");
  NFFT_HIP_CHECK(R, hipGetDevice(&device));
  noalloc_scratch = scratch_setup(device);
  NFFT_HIP_CHECK(R, hipStreamCreate(&own_stream));
  stream = own_stream;
  for (auto &e : ev) NFFT_HIP_CHECK(R, hipEventCreate(&e));

  NFFT_HIP_CHECK(R, hipMalloc(&fhat, sizeof(cplx<R>) * (size_t)(N_total > 0 ? N_total : 1)));
  NFFT_HIP_CHECK(R, hipMalloc(&f, sizeof(cplx<R>) * (size_t)(M > 0 ? M : 1)));
  NFFT_HIP_CHECK(R, hipMalloc(&xdev, sizeof(R) * (size_t)(M > 0 ? M * d : 1)));
  if (direct_only) return;

  if (m > 16) family<R>::fatal("nfft: window cut-off m = %d exceeds the supported 16", m);

  NFFT_HIP_CHECK(R, hipMalloc(&g, sizeof(cplx<R>) * (size_t)n_total));
  NFFT_HIP_CHECK(R, hipMalloc(&cell, sizeof(int) * (size_t)(M > 0 ? M * d : 1)));
  NFFT_HIP_CHECK(R, hipMalloc(&frac, sizeof(R) * (size_t)(M > 0 ? M * d : 1)));
  NFFT_HIP_CHECK(R, hipMalloc(&perm, sizeof(int) * (size_t)(M > 0 ? M : 1)));

  // window parameters and scaled deconvolution factors
  std::vector<R> pinv;
  for (int t = 0; t < d; t++) {
    long double sigma = (long double)n[t] / (long double)N[t];
    long double b = kPiL * (2.0L - 1.0L / sigma);
    long double peak = sinhl(b * m) / (kPiL * m);  // phi(0)
    for (long ks = 0; ks < N[t]; ks++) pinv.push_back((R)(peak / phi_hut(n[t], ks - N[t] / 2, m, b)));
  }
  NFFT_HIP_CHECK(R, hipMalloc(&phi_inv, sizeof(R) * pinv.size()));
  NFFT_HIP_CHECK(R, hipMemcpy(phi_inv, pinv.data(), sizeof(R) * pinv.size(), hipMemcpyHostToDevice));
  {
    std::vector<R> pc;
    for (int t = 0; t < d; t++) {
      long double sigma = (long double)n[t] / (long double)N[t];
      long double b = kPiL * (2.0L - 1.0L / sigma);
      fit_window_polys<R>(m, b, pc);
    }
    NFFT_HIP_CHECK(R, hipMalloc(&poly, sizeof(R) * pc.size()));
    NFFT_HIP_CHECK(R, hipMemcpy(poly, pc.data(), sizeof(R) * pc.size(), hipMemcpyHostToDevice));
    win.poly = poly;
    const size_t per_dim = pc.size() / d;
    win.poly_shared = 1;
    for (int t = 1; t < d; t++)
      if (!std::equal(pc.begin(), pc.begin() + per_dim, pc.begin() + t * per_dim)) win.poly_shared = 0;
  }

  // bins: start from an edge that gives the grid enough bins to fill the
  // GPU (1-D 256 cells, 2-D n/32 clamped to [16, 64], 3-D 16), then shrink
  // the largest edge until the LDS tile fits the budget
  const size_t budget = 48 * 1024;
  int bs[kMaxDim];
  for (int t = 0; t < kMaxDim; t++) bs[t] = 1;
  for (int t = 0; t < d; t++) {
    const long init = d == 1 ? 256 : d == 2 ? std::min<long>(64, std::max<long>(16, n[t] / 32)) : 16;
    bs[t] = (int)std::min<long>(n[t], init);
  }
  auto bytes = [&]() {
    size_t s = sizeof(cplx<R>);
    for (int t = 0; t < d; t++) s *= (size_t)(bs[t] + 2 * m + 1);
    return s;
  };
  while (bytes() > budget) {
    int tmax = 0;
    for (int t = 1; t < d; t++)
      if (bs[t] > bs[tmax]) tmax = t;
    if (bs[tmax] == 1) break;
    bs[tmax] = (bs[tmax] + 1) / 2;
  }
  // NFFT_HIP_BINS=e0,e1,...: bin edges instead of the default
  if (const char *v = std::getenv("NFFT_HIP_BINS")) {
    int k = 0;
    for (const char *q = v; *q && k < d; k++) {
      const int e = std::atoi(q);
      if (e >= 1) bs[k] = (int)std::min<long>(n[k], e);
      while (*q && *q != ',') q++;
      if (*q == ',') q++;
    }
  }
  set_bins(bs);
  // nodes per work item; larger 3-D spreading items write fewer tiles for
  // the reduction (measured on the 3-D kooshball: 2048 -> 4096 halves the
  // tile reduction at equal spreading time, larger items end the launch late)
  max_per_sub = d == 3 ? 2048 : 512;
  spread_max_per_sub = d == 3 ? 4096 : 1024;
  if (d == 2) {
    // 2-D spreading items: about 1024 of them for a plan's nodes, 256..1024
    // nodes each (few large items leave dense regions to a handful of
    // wavefronts; many small ones multiply the tiles to reduce)
    long per = 256;
    while (per < 1024 && per * 1024 < M) per *= 2;
    spread_max_per_sub = (int)per;
  }
  if (const char *v = std::getenv("NFFT_HIP_ITEM")) max_per_sub = std::max(64, std::atoi(v));
  if (const char *v = std::getenv("NFFT_HIP_SPREAD_ITEM"))
    spread_max_per_sub = std::max(64, std::atoi(v));
  if (const char *v = std::getenv("NFFT_HIP_SPREAD_CAP")) spread_max_items = std::max(0, std::atoi(v));
  auto variant = [](const char *name, int dflt) {
    const char *v = std::getenv(name);
    if (!v) return dflt;
    return std::strcmp(v, "runs2sbc") == 0  ? 10
           : std::strcmp(v, "runs2rs") == 0 ? 13
           : std::strcmp(v, "runs2d") == 0 ? 14
           : std::strcmp(v, "rows") == 0 ? 3
           : std::strcmp(v, "priv") == 0 ? 2
           : std::strcmp(v, "wave") == 0 ? 1
           : std::strcmp(v, "lane") == 0 ? 0
                                         : dflt;
  };
  interp_variant = variant("NFFT_HIP_INTERP", 0);
  if (const char *v = std::getenv("NFFT_HIP_SYMWIN")) sym_window = std::atoi(v) != 0;
  if (const char *v = std::getenv("NFFT_HIP_PAIRS")) pair_interp = std::atoi(v) != 0;
  if (const char *v = std::getenv("NFFT_HIP_SLOT")) slot_nodes = std::atoi(v) >= 4 ? 4 : 2;
  if (const char *v = std::getenv("NFFT_HIP_SLOTSORT")) slot_sort = std::atoi(v) != 0;
  // float: plain FP32 read-modify-write by wavefronts that own disjoint
  // cells, same-cell runs summed in registers (3-D: two wavefronts per tile
  // split by row parity; 2-D: one wavefront per tile; 1-D: private tiles);
  // double: wavefronts sharing a tile with ds_add_f64
  spread_variant = variant("NFFT_HIP_SPREAD", sizeof(R) == 4 ? (d == 3 ? 10 : (d == 2 ? 14 : 2)) : 1);

  // FFT: twiddles per dimension, then the pruned passes of both directions.
  // An axis whose lines fit one workgroup's LDS (cap elements) and whose
  // length factors into the stage radices is one pass; a longer one is
  // split into two four-step halves n = n1 n2 (split[t] = n1), each fitting.
  const int cap = sizeof(R) == 4 ? 4096 : 2048;  // elements per workgroup
  auto twiddle_table = [&](long len) {
    std::vector<cplx<R>> tw(len);
    for (long k = 0; k < len; k++) {
      long double a = -2.0L * kPiL * (long double)k / (long double)len;
      tw[k] = cplx<R>{(R)cosl(a), (R)sinl(a)};
    }
    cplx<R> *dev = nullptr;
    NFFT_HIP_CHECK(R, hipMalloc(&dev, sizeof(cplx<R>) * tw.size()));
    NFFT_HIP_CHECK(R, hipMemcpy(dev, tw.data(), sizeof(cplx<R>) * tw.size(), hipMemcpyHostToDevice));
    return dev;
  };
  auto make_axis = [&](long len) {
    FftAxis<R> ax{};
    ax.len = (int)len;
    std::vector<int> rad;
    fft_radices(len, rad);
    ax.nstages = (int)rad.size();
    for (size_t i = 0; i < rad.size(); i++) ax.radix[i] = rad[i];
    ax.twiddle = twiddle_table(len);
    return ax;
  };
  long split[kMaxDim] = {};
  std::vector<FftAxis<R>> sub1(d), sub2(d);
  std::vector<cplx<R> *> tw_full(d, nullptr);
  for (int t = 0; t < d; t++) {
    std::vector<int> rad;
    if (n[t] <= cap && fft_radices(n[t], rad)) {
      axes.push_back(make_axis(n[t]));
      continue;
    }
    // the most balanced n1 n2 = n with both halves fitting and factoring
    long best = 0;
    for (long a = 2; a <= cap && a < n[t]; a++) {
      if (n[t] % a) continue;
      const long b = n[t] / a;
      std::vector<int> ra, rb;
      if (b > cap || !fft_radices(a, ra) || !fft_radices(b, rb)) continue;
      if (!best || std::labs(a - b) < std::labs(best - n[t] / best)) best = a;
    }
    if (!best)
      family<R>::fatal("nfft: FFT length n[%d] = %ld is above %d and not a product of two "
                       "lengths up to %d", t, n[t], cap, cap);
    split[t] = best;
    axes.push_back(FftAxis<R>{});  // placeholder, the halves carry the tables
    axes.back().len = (int)n[t];
    sub1[t] = make_axis(n[t] / best);  // first half: length n2
    sub2[t] = make_axis(best);         // second half: length n1
    tw_full[t] = twiddle_table(n[t]);
    extra_twiddles.push_back(sub1[t].twiddle);
    extra_twiddles.push_back(sub2[t].twiddle);
    extra_twiddles.push_back(tw_full[t]);
  }
  bool four_step = false;
  for (int t = 0; t < d; t++) four_step = four_step || split[t] > 0;
  if (four_step) NFFT_HIP_CHECK(R, hipMalloc(&g2, sizeof(cplx<R>) * (size_t)n_total));
  {
    long gstride[kMaxDim], fstride[kMaxDim];
    int phioff[kMaxDim];
    long gs = 1, fs = 1;
    for (int t = d - 1; t >= 0; t--) {
      gstride[t] = gs;
      fstride[t] = fs;
      gs *= n[t];
      fs *= N[t];
    }
    int off = 0;
    for (int t = 0; t < d; t++) {
      phioff[t] = off;
      off += (int)N[t];
    }
    // band of dimension s on the lines of the pass over t: trafo, s < t (not
    // yet transformed, zero outside the band); adjoint, s < t (already
    // cropped to the band)
    auto make_pass = [&](int t, int in_mode, int out_mode) {
      FftPass<R> p{};
      p.bs_g = n_total;
      p.bs_c = N_total;
      p.len = (int)n[t];
      p.stride = gstride[t];
      p.Nband = (int)N[t];
      p.in_mode = in_mode;
      p.out_mode = out_mode;
      p.nod = 0;
      p.nlines = 1;
      for (int s = 0; s < d; s++) {
        if (s == t) continue;
        const int j = p.nod++;
        p.band[j] = s < t ? 1 : 0;
        p.cnt[j] = (int)(p.band[j] ? N[s] : n[s]);
        p.N_od[j] = (int)N[s];
        p.n_od[j] = (int)n[s];
        p.gstr[j] = gstride[s];
        p.fstr[j] = fstride[s];
        p.phi_od[j] = phi_inv + phioff[s];
        p.nlines *= p.cnt[j];
      }
      p.fstr_ax = fstride[t];
      p.phi_ax = phi_inv + phioff[t];
      const FftAxis<R> &ax = axes[t];
      p.nsub = 1;
      p.nfull = (int)n[t];
      p.fn1 = p.fn2 = 1;
      // lines per workgroup that fit the LDS budget (the launch may use
      // fewer, see launch_fft_pass)
      p.lines_per_block = (int)std::max<long>(1, std::min<long>(256, cap / n[t]));
      p.nstages = ax.nstages;
      for (int i = 0; i < p.nstages; i++) p.radix[i] = ax.radix[i];
      p.twiddle = ax.twiddle;
      const size_t line_bytes = sizeof(cplx<R>) * (size_t)p.lines_per_block * lds_line_stride(p.len);
      p.twiddle_in_lds = line_bytes + sizeof(cplx<R>) * (size_t)p.len <= 64 * 1024 ? 1 : 0;
      return p;
    };
    // a four-step axis: its two halves in place of one pass.  The first
    // keeps the pass's input rule and writes every position; the second
    // reads every position and keeps the output rule.  cur = the buffer
    // holding the data; the second half writes the other one.
    auto halves = [&](FftPass<R> p, int t, int &cur, std::vector<FftPass<R>> &out) {
      if (!split[t]) {
        p.src = p.dst = cur;
        out.push_back(p);
        return;
      }
      const long n1 = split[t], n2 = n[t] / split[t];
      FftPass<R> a = p, b = p;
      for (int h = 0; h < 2; h++) {
        FftPass<R> &q = h == 0 ? a : b;
        const FftAxis<R> &ax = h == 0 ? sub1[t] : sub2[t];
        q.four = h + 1;
        q.fn1 = (int)n1;
        q.fn2 = (int)n2;
        q.nfull = (int)n[t];
        q.twiddle_full = tw_full[t];
        q.len = ax.len;
        q.nsub = (int)(h == 0 ? n1 : n2);
        q.nlines = p.nlines * q.nsub;
        q.nstages = ax.nstages;
        for (int i = 0; i < q.nstages; i++) q.radix[i] = ax.radix[i];
        q.twiddle = ax.twiddle;
        q.lines_per_block = (int)std::max<long>(1, std::min<long>(256, cap / q.len));
        const size_t lb = sizeof(cplx<R>) * (size_t)q.lines_per_block * lds_line_stride(q.len);
        q.twiddle_in_lds = lb + sizeof(cplx<R>) * (size_t)q.len <= 64 * 1024 ? 1 : 0;
      }
      a.out_mode = kToGrid;
      b.in_mode = kFromGrid;
      a.src = a.dst = cur;
      b.src = cur;
      b.dst = 1 - cur;
      cur = 1 - cur;
      out.push_back(a);
      out.push_back(b);
    };
    int cur = 0;
    for (int t = d - 1; t >= 0; t--)
      halves(make_pass(t, t == d - 1 ? kFromCoeffs : kFromGridBand, kToGrid), t, cur, fwd_passes);
    trafo_grid = cur;
    cur = 0;
    for (int t = 0; t < d; t++)
      halves(make_pass(t, kFromGrid, t == d - 1 ? kToCoeffs : kToGridBand), t, cur, bwd_passes);
    // NFFT_HIP_FFT_LINES=k: lines per workgroup times 2^k in every pass
    if (const char *v = std::getenv("NFFT_HIP_FFT_LINES")) {
      const int sh = std::max(-4, std::min(4, std::atoi(v)));
      for (auto &ps : fwd_passes) ps.lpb_shift = sh;
      for (auto &ps : bwd_passes) ps.lpb_shift = sh;
    }
  }
}

template <typename R>
void Engine<R>::destroy() {
  if (stream) hipStreamSynchronize(stream);
  for (auto &ax : axes) hipFree(ax.twiddle);
  axes.clear();
  for (auto *tw : extra_twiddles) hipFree(tw);
  extra_twiddles.clear();
  hipFree(g2);
  g2 = nullptr;
  hipFree(g);
  hipFree(fhat);
  hipFree(f);
  hipFree(xdev);
  hipFree(phi_inv);
  hipFree(poly);
  hipFree(cell);
  hipFree(frac);
  hipFree(perm);
  hipFree(subs);
  hipFree(sp_subs);
  hipFree(sp_first);
  hipFree(tiles);
  hipFree(scratch_i);
  hipFree(zone_tab);
  hipFree(aux);
  hipFree(bgrid);
  hipFree(btiles);
  hipFree(slots);
  hipFree(slots_raw);
  hipFree(islots);
  for (auto &e : ev)
    if (e) hipEventDestroy(e);
  for (auto &set : log_ev)
    for (auto &e : set) hipEventDestroy(e);
  if (own_stream) hipStreamDestroy(own_stream);
}

template <typename R>
void Engine<R>::upload_nodes(const R *x_host) {
  if (M == 0) return;
  NFFT_HIP_CHECK(R, hipMemcpyAsync(xdev, x_host, sizeof(R) * (size_t)(M * d),
                                   hipMemcpyHostToDevice, stream));
}

template <typename R>
void Engine<R>::build_nodes() {
  if (!direct_only) {
    bin_nodes<R>(*this);
    if (tile_scratch) {
      long tot = 1;
      for (int t = 0; t < d; t++) tot *= geom.E[t];
      const long need = sp_nsub * tot;
      if (need > tiles_cap) {
        if (tiles) NFFT_HIP_CHECK(R, hipFree(tiles));
        NFFT_HIP_CHECK(R, hipMalloc(&tiles, sizeof(cplx<R>) * (size_t)(need > 0 ? need : 1)));
        tiles_cap = need;
      }
    }
  }
  nodes_ready = true;
}

// events bracketing the stages of one call: the immediate set, a fresh set
// from the deferred log, or none
template <typename R>
hipEvent_t *Engine<R>::call_events(int kind) {
  if (deferred) {
    if (log_n == log_ev.size()) {
      std::array<hipEvent_t, 5> set;
      for (auto &e : set) NFFT_HIP_CHECK(R, hipEventCreate(&e));
      log_ev.push_back(set);
      log_kind.push_back(0);
    }
    log_kind[log_n] = kind;
    return log_ev[log_n++].data();
  }
  return timing ? ev : nullptr;
}

// Event layout of one call.  trafo: E0 | D | E1 | FFT | E2 | B | E3 (E4 = E3).
// adjoint: E0 | B^T spreading | E1 | tile reduction | E2 | FFT | E3 | D^T | E4.
// D and D^T are fused into the FFT passes, so their intervals are empty.
template <typename R>
void Engine<R>::trafo(const cplx<R> *fhat_dev, cplx<R> *f_dev) {
  hipEvent_t *E = call_events(0);
  if (E) NFFT_HIP_CHECK(R, hipEventRecord(E[0], stream));
  // D runs inside the first FFT pass: the D interval stays empty
  if (E) NFFT_HIP_CHECK(R, hipEventRecord(E[1], stream));
  for (auto &ps : fwd_passes) launch_fft_pass<R>(g, fhat_dev, nullptr, ps, -1, stream, 1, g2);
  if (E) NFFT_HIP_CHECK(R, hipEventRecord(E[2], stream));
  launch_interp<R>(grid_after_trafo(), f_dev, *this, stream);
  if (E) NFFT_HIP_CHECK(R, hipEventRecord(E[3], stream));
  if (E) NFFT_HIP_CHECK(R, hipEventRecord(E[4], stream));
  NFFT_HIP_CHECK(R, hipGetLastError());
}

template <typename R>
void Engine<R>::adjoint(const cplx<R> *f_dev, cplx<R> *fhat_dev) {
  hipEvent_t *E = call_events(1);
  // B^T: spreading kernel, then (scratch path) the tile reduction, which
  // writes every grid cell; the atomic path needs a zeroed grid instead
  const bool scratch = tile_scratch && sp_nsub > 0;
  if (!scratch) launch_fill_zero<R>(g, n_total, stream);
  if (E) NFFT_HIP_CHECK(R, hipEventRecord(E[0], stream));
  launch_spread<R>(f_dev, g, *this, stream);
  if (E) NFFT_HIP_CHECK(R, hipEventRecord(E[1], stream));
  if (scratch) launch_gather_tiles<R>(*this, g, stream);
  if (E) NFFT_HIP_CHECK(R, hipEventRecord(E[2], stream));
  // D^T runs inside the last FFT pass: the D^T interval stays empty
  for (auto &ps : bwd_passes) launch_fft_pass<R>(g, nullptr, fhat_dev, ps, +1, stream, 1, g2);
  if (E) NFFT_HIP_CHECK(R, hipEventRecord(E[3], stream));
  if (E) NFFT_HIP_CHECK(R, hipEventRecord(E[4], stream));
  NFFT_HIP_CHECK(R, hipGetLastError());
}

template <typename R>
bool Engine<R>::batchable() const {
  // four-step axes (g2) run their right-hand sides one by one
  return !direct_only && !timing && !deferred && M > 0 && sp_nsub > 0 && !g2 &&
         batch_kernels<R>(*this);
}

// right-hand sides per batched launch: the batch grids and tiles of one
// launch stay below 4 GiB
template <typename R>
int Engine<R>::batch_chunk(int nrhs) const {
  long tot = 1;
  for (int t = 0; t < d; t++) tot *= geom.E[t];
  const long per_rhs = (n_total + sp_nsub * tot) * (long)sizeof(cplx<R>);
  const long cap = std::max(1L, (4L << 30) / std::max(1L, per_rhs));
  return (int)std::min<long>(nrhs, cap);
}

template <typename R>
void Engine<R>::reserve_batch(int nrhs) {
  if (!batchable() || nrhs < 2) return;
  const int k = batch_chunk(nrhs);
  long tot = 1;
  for (int t = 0; t < d; t++) tot *= geom.E[t];
  if ((long)k * n_total > bgrid_cap) {
    NFFT_HIP_CHECK(R, hipStreamSynchronize(stream));
    hipFree(bgrid);
    bgrid = nullptr;
    NFFT_HIP_CHECK(R, hipMalloc(&bgrid, sizeof(cplx<R>) * (size_t)k * n_total));
    bgrid_cap = (long)k * n_total;
  }
  if ((long)k * sp_nsub * tot > btiles_cap) {
    NFFT_HIP_CHECK(R, hipStreamSynchronize(stream));
    hipFree(btiles);
    btiles = nullptr;
    NFFT_HIP_CHECK(R, hipMalloc(&btiles, sizeof(cplx<R>) * (size_t)k * sp_nsub * tot));
    btiles_cap = (long)k * sp_nsub * tot;
  }
}

// nrhs transforms with the plan's nodes: right-hand side r of fhat_dev at
// r N_total, of f_dev at r M; each kernel of the path runs once for a chunk
// of right-hand sides (grid y), on the batch grids
template <typename R>
void Engine<R>::trafo_batch(const cplx<R> *fhat_dev, cplx<R> *f_dev, int nrhs) {
  if (nrhs < 2 || !batchable()) {
    for (int r = 0; r < nrhs; r++) trafo(fhat_dev + (long)r * N_total, f_dev + (long)r * M);
    return;
  }
  reserve_batch(nrhs);
  const int k = batch_chunk(nrhs);
  BinGeom gb = geom;
  gb.bs_g = n_total;
  gb.bs_f = M;
  for (int r0 = 0; r0 < nrhs; r0 += k) {
    const int nb = std::min(k, nrhs - r0);
    for (auto &ps : fwd_passes)
      launch_fft_pass<R>(bgrid, fhat_dev + (long)r0 * N_total, nullptr, ps, -1, stream, nb);
    launch_interp<R>(bgrid, f_dev + (long)r0 * M, *this, stream, nb, &gb);
  }
  NFFT_HIP_CHECK(R, hipGetLastError());
}

template <typename R>
void Engine<R>::adjoint_batch(const cplx<R> *f_dev, cplx<R> *fhat_dev, int nrhs) {
  if (nrhs < 2 || !batchable()) {
    for (int r = 0; r < nrhs; r++) adjoint(f_dev + (long)r * M, fhat_dev + (long)r * N_total);
    return;
  }
  reserve_batch(nrhs);
  adjoint_batch_chunks(f_dev, fhat_dev, nrhs, batch_chunk(nrhs), bgrid, btiles);
}

// batched adjoint in chunks of k right-hand sides on caller-given batch
// grids (k n_total values) and tiles (k sp_nsub E^d values)
template <typename R>
void Engine<R>::adjoint_batch_chunks(const cplx<R> *f_dev, cplx<R> *fhat_dev, int nrhs, int k,
                                     cplx<R> *grids, cplx<R> *tls) {
  long tot = 1;
  for (int t = 0; t < d; t++) tot *= geom.E[t];
  BinGeom gb = geom;
  gb.bs_g = n_total;
  gb.bs_f = M;
  gb.bs_t = sp_nsub * tot;
  for (int r0 = 0; r0 < nrhs; r0 += k) {
    const int nb = std::min(k, nrhs - r0);
    launch_spread<R>(f_dev + (long)r0 * M, grids, *this, stream, nb, &gb, tls);
    launch_gather_tiles<R>(*this, grids, stream, nb, &gb, tls);
    for (auto &ps : bwd_passes)
      launch_fft_pass<R>(grids, nullptr, fhat_dev + (long)r0 * N_total, ps, +1, stream, nb);
  }
  NFFT_HIP_CHECK(R, hipGetLastError());
}

// Bytes of `scratch` that one right-hand side of a batched adjoint needs:
// its grid and its spreading tiles, each rounded up to 256 bytes.
template <typename R>
long Engine<R>::noalloc_bytes_per_rhs() const {
  long tot = 1;
  for (int t = 0; t < d; t++) tot *= geom.E[t];
  auto up = [](long b) { return (b + 255) / 256 * 256; };
  return up(n_total * (long)sizeof(cplx<R>)) + up(sp_nsub * tot * (long)sizeof(cplx<R>));
}

template <typename R>
int Engine<R>::adjoint_batch_noalloc(const cplx<R> *f_dev, cplx<R> *fhat_dev, int nrhs) {
  if (nrhs < 2 || !batchable()) {
    // one by one on the plan's own grid and tiles
    for (int r = 0; r < nrhs; r++) adjoint(f_dev + (long)r * M, fhat_dev + (long)r * N_total);
    return 0;
  }
  if (!noalloc_scratch) return -3;
  const long per = noalloc_bytes_per_rhs();
  if (per > (long)kNoAllocScratchBytes) return -4;
  const int k = (int)std::min<long>(nrhs, (long)kNoAllocScratchBytes / per);
  // k grids, then k tile sets, each block 256-byte aligned
  char *base = static_cast<char *>(noalloc_scratch);
  const long grid_bytes = ((long)k * n_total * (long)sizeof(cplx<R>) + 255) / 256 * 256;
  adjoint_batch_chunks(f_dev, fhat_dev, nrhs, k, reinterpret_cast<cplx<R> *>(base),
                       reinterpret_cast<cplx<R> *>(base + grid_bytes));
  return 0;
}

template <typename R>
void Engine<R>::autotune() {
  tuned = true;
  if (direct_only || M == 0 || timing || deferred) return;
  hipEvent_t a, b;
  NFFT_HIP_CHECK(R, hipEventCreate(&a));
  NFFT_HIP_CHECK(R, hipEventCreate(&b));
  // a candidate's time: the best of three timings of two runs each, after
  // a warm-up run (one timing of small plans swung by up to a third and
  // could pick a slower candidate)
  auto time_of = [&](auto &&run) {
    run();  // warm-up
    float best = 1e30f;
    for (int trial = 0; trial < 3; trial++) {
      NFFT_HIP_CHECK(R, hipEventRecord(a, stream));
      for (int rep = 0; rep < 2; rep++) run();
      NFFT_HIP_CHECK(R, hipEventRecord(b, stream));
      NFFT_HIP_CHECK(R, hipEventSynchronize(b));
      float ms = 0.f;
      NFFT_HIP_CHECK(R, hipEventElapsedTime(&ms, a, b));
      best = std::min(best, ms);
    }
    return best;
  };
  // the plan's own transform pair a few times first, so that the first
  // candidate is not timed on a GPU still raising its clocks (an idle GPU
  // had the default 2-D spreading kernel timed slow and a slower one picked)
  for (int rep = 0; rep < 4; rep++) {
    spread_to_grid(f);
    interp_from_grid(f);
  }
  // inputs: the device image of f (its values do not change the timing)
  // bin edges: the default, and the largest edge halved; in 1-D and 2-D also
  // every edge doubled and every edge halved; all within the LDS budget of
  // the default, nodes rebuilt for each, B and B^T timed together
  if (!std::getenv("NFFT_HIP_BINS") && d <= 3) {
    auto edges = [&]() {
      std::array<int, kMaxDim> e{};
      for (int t = 0; t < kMaxDim; t++) e[t] = geom.bs[t];
      return e;
    };
    const std::array<int, kMaxDim> base = edges();
    auto tile_of = [&](const std::array<int, kMaxDim> &e) {
      size_t s = sizeof(cplx<R>);
      for (int t = 0; t < d; t++) s *= (size_t)(e[t] + 2 * m + 1);
      return s;
    };
    std::vector<std::array<int, kMaxDim>> cand;
    {
      std::array<int, kMaxDim> h = base;
      int tmax = 0;
      for (int t = 1; t < d; t++)
        if (h[t] > h[tmax]) tmax = t;
      h[tmax] = std::max(1, h[tmax] / 2);
      cand.push_back(h);
      if (d <= 2) {
        std::array<int, kMaxDim> up = base, dn = base;
        for (int t = 0; t < d; t++) {
          up[t] = (int)std::min<long>(n[t], 2L * base[t]);
          dn[t] = std::max(1, base[t] / 2);
        }
        cand.push_back(up);
        cand.push_back(dn);
      }
    }
    auto time_b = [&]() {
      return time_of([&]() {
        spread_to_grid(f);
        interp_from_grid(f);
      });
    };
    float best = time_b();
    std::array<int, kMaxDim> pick = base;
    for (const auto &c : cand) {
      if (c == base || tile_of(c) > std::max<size_t>(tile_of(base), 48 * 1024)) continue;
      set_bins(c.data());
      build_nodes();
      const float ms = time_b();
      if (ms < best) {
        best = ms;
        pick = c;
      }
    }
    if (pick != edges()) {
      set_bins(pick.data());
      build_nodes();
    }
  }
  if (!std::getenv("NFFT_HIP_SPREAD")) {
    // float candidates accumulate in plain FP32 (the fixed-point LDS
    // variants are left out); double ones use ds_add_f64
    std::vector<int> cand;
    if (sizeof(R) == 4 && d == 3) cand = {10, 13, 3};
    else if (sizeof(R) == 4 && d == 2) cand = {14, 1, 0, 2};
    else if (sizeof(R) == 8 && d == 2) cand = {14, 1, 0};
    else if (sizeof(R) == 8 && d >= 2) cand = {1, 0};
    float best = 1e30f;
    int pick = spread_variant;
    for (int v : cand) {
      spread_variant = v;
      const float ms = time_of([&]() { spread_to_grid(f); });
      if (ms < best) {
        best = ms;
        pick = v;
      }
    }
    spread_variant = pick;
  }
  // spreading work-item size (nodes per item): half and twice the default,
  // the work lists rebuilt for each (fewer, longer items write fewer tiles
  // to reduce; more, shorter ones balance dense regions better)
  if (!std::getenv("NFFT_HIP_SPREAD_ITEM") && d >= 2 && tile_scratch && sp_nsub > 0) {
    const int base = spread_max_per_sub;
    float best = time_of([&]() { spread_to_grid(f); });
    int pick = base;
    for (int c : {base / 2, base * 2}) {
      if (c < 256 || c > 8192) continue;
      spread_max_per_sub = c;
      build_nodes();
      const float ms = time_of([&]() { spread_to_grid(f); });
      if (ms < best) {
        best = ms;
        pick = c;
      }
    }
    if (spread_max_per_sub != pick) {
      spread_max_per_sub = pick;
      build_nodes();
    }
  }
  if (!std::getenv("NFFT_HIP_INTERP") && d >= 2) {
    float best = 1e30f;
    int pick = interp_variant;
    for (int v : {0, 1}) {
      interp_variant = v;
      const float ms = time_of([&]() { interp_from_grid(f); });
      if (ms < best) {
        best = ms;
        pick = v;
      }
    }
    interp_variant = pick;
  }
  // FFT lines per workgroup: half, the default and twice the default choice
  // of the launch, one setting for all passes of both directions
  if (!std::getenv("NFFT_HIP_FFT_LINES")) {
    auto set_shift = [&](int sh) {
      for (auto &ps : fwd_passes) ps.lpb_shift = sh;
      for (auto &ps : bwd_passes) ps.lpb_shift = sh;
    };
    auto time_fft = [&]() {
      return time_of([&]() {
        for (auto &ps : fwd_passes) launch_fft_pass<R>(g, fhat, nullptr, ps, -1, stream, 1, g2);
        for (auto &ps : bwd_passes) launch_fft_pass<R>(g, nullptr, fhat, ps, +1, stream, 1, g2);
      });
    };
    float best = time_fft();
    int pick = 0;
    for (int sh : {-1, 1}) {
      set_shift(sh);
      const float ms = time_fft();
      if (ms < best) {
        best = ms;
        pick = sh;
      }
    }
    set_shift(pick);
  }
  if (const char *v = std::getenv("NFFT_HIP_TUNE_LOG"); v && std::atoi(v) != 0) {
    std::fprintf(stderr, "[nfft_hip tune] d=%d m=%d M=%ld bins", d, m, M);
    for (int t = 0; t < d; t++) std::fprintf(stderr, "%s%d", t ? "," : " ", geom.bs[t]);
    std::fprintf(stderr, " spread=%d item=%d interp=%d fft_shift=%d\n", spread_variant,
                 spread_max_per_sub, interp_variant,
                 fwd_passes.empty() ? 0 : fwd_passes[0].lpb_shift);
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
}

template <typename R>
void Engine<R>::spread_to_grid(const cplx<R> *f_dev) {
  const bool scratch = tile_scratch && sp_nsub > 0;
  if (!scratch) launch_fill_zero<R>(g, n_total, stream);
  launch_spread<R>(f_dev, g, *this, stream);
  if (scratch) launch_gather_tiles<R>(*this, g, stream);
  NFFT_HIP_CHECK(R, hipGetLastError());
}

template <typename R>
void Engine<R>::interp_from_grid(cplx<R> *f_dev) {
  launch_interp<R>(g, f_dev, *this, stream);
  NFFT_HIP_CHECK(R, hipGetLastError());
}

template <typename R>
void Engine<R>::timing_begin() {
  deferred = true;
  log_n = 0;
}

template <typename R>
void Engine<R>::timing_end(double *out) {
  // out: n_trafo, n_adjoint, trafo D, FFT, B, adjoint spreading, tile
  // reduction, FFT, D^T -- summed seconds
  for (int i = 0; i < 9; i++) out[i] = 0;
  NFFT_HIP_CHECK(R, hipStreamSynchronize(stream));
  for (size_t c = 0; c < log_n; c++) {
    float ms[4];
    for (int i = 0; i < 4; i++)
      NFFT_HIP_CHECK(R, hipEventElapsedTime(&ms[i], log_ev[c][i], log_ev[c][i + 1]));
    if (log_kind[c] == 0) {
      out[0] += 1;
      for (int i = 0; i < 3; i++) out[2 + i] += ms[i] * 1e-3;
    } else {
      out[1] += 1;
      for (int i = 0; i < 4; i++) out[5 + i] += ms[i] * 1e-3;
    }
  }
  deferred = false;
  log_n = 0;
}

template <typename R>
void Engine<R>::trafo_direct(const cplx<R> *fhat_dev, cplx<R> *f_dev) {
  launch_ndft<R>(xdev, fhat_dev, f_dev, d, N, M, 0, stream);
  NFFT_HIP_CHECK(R, hipGetLastError());
}

template <typename R>
void Engine<R>::adjoint_direct(const cplx<R> *f_dev, cplx<R> *fhat_dev) {
  launch_ndft<R>(xdev, f_dev, fhat_dev, d, N, M, 1, stream);
  NFFT_HIP_CHECK(R, hipGetLastError());
}

// stage durations of the last call, in seconds: [0] D, [1] FFT, [2] B
template <typename R>
void Engine<R>::stage_times(double out[4]) {
  float ms[4] = {0, 0, 0, 0};
  NFFT_HIP_CHECK(R, hipEventSynchronize(ev[4]));
  for (int i = 0; i < 4; i++) NFFT_HIP_CHECK(R, hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
  // the four intervals of the event layout above; the caller maps them
  // onto MEASURE_TIME_t
  for (int i = 0; i < 4; i++) out[i] = ms[i] * 1e-3;
}

template struct Engine<float>;
template struct Engine<double>;

// ------------------------------------------------------------------ plan API
template <typename R> struct Fam;
template <> struct Fam<float> {
  static void *alloc(size_t n) { return nfftf_malloc(n); }
  static void release(void *p) { nfftf_free(p); }
  static bool hooked() { return nfftf_malloc_hook != nullptr; }
  static constexpr int default_m = 4;
};
template <> struct Fam<double> {
  static void *alloc(size_t n) { return nfft_malloc(n); }
  static void release(void *p) { nfft_free(p); }
  static bool hooked() { return nfft_malloc_hook != nullptr; }
  static constexpr int default_m = 8;
};

// K of PRE_LIN_PSI for the Kaiser-Bessel window (reference kernel/util/window.c:37-41)
static NFFT_INT m2K(NFFT_INT m) {
  static const int tab[] = {1, 3, 7, 9, 14, 17, 20, 23, 24};
  int j = (int)std::min<NFFT_INT>(m, (NFFT_INT)(sizeof(tab) / sizeof(tab[0]) - 1));
  return (NFFT_INT)(1U << tab[j]) * (m + 2);
}

// Plan-owned x / f / f_hat (MALLOC_*) are page-locked host memory so that the
// per-call copies run at full PCIe rate; a user malloc hook takes precedence.
template <typename R>
void *host_array_alloc(size_t bytes) {
  if (Fam<R>::hooked()) return Fam<R>::alloc(bytes);
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess || !p)
    family<R>::fatal("nfft: out of pinned host memory (%zu bytes)", bytes);
  return p;
}

template <typename R>
void host_array_free(void *p) {
  if (!p) return;
  if (Fam<R>::hooked()) {
    Fam<R>::release(p);
    return;
  }
  hipHostFree(p);
}

// mv_trafo / mv_adjoint of a plan point at the family's own entry points
// (reference kernel/nfft/nfft.c:6044-6045)
static void set_mv(nfftf_plan *p) {
  p->mv_trafo = reinterpret_cast<void (*)(void *)>(nfftf_trafo);
  p->mv_adjoint = reinterpret_cast<void (*)(void *)>(nfftf_adjoint);
}
static void set_mv(nfft_plan *p) {
  p->mv_trafo = reinterpret_cast<void (*)(void *)>(nfft_trafo);
  p->mv_adjoint = reinterpret_cast<void (*)(void *)>(nfft_adjoint);
}

// out = w .* f (complex values times real weights)
template <typename R>
__global__ void k_weight_mul(const cplx<R> *__restrict__ f, const R *__restrict__ w,
                             cplx<R> *__restrict__ out, long M) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < M) out[j] = cplx<R>{w[j] * f[j].x, w[j] * f[j].y};
}

template <typename R, typename PLAN>
struct PlanApi {
  static Engine<R> *eng(const PLAN *p) {
    Engine<R> *e = reinterpret_cast<Engine<R> *>(p->my_fftw_plan1);
    if (!e) family<R>::fatal("nfft: plan used before nfft_init* or after nfft_finalize");
    return e;
  }

  // reference init_help (kernel/nfft/nfft.c:5950-6046)
  static void init_help(PLAN *p) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      family<R>::fatal("nfft: no HIP device available (this library runs on the GPU only)");
    if (p->flags & NFFT_OMP_BLOCKWISE_ADJOINT) p->flags |= NFFT_SORT_NODES;
    if (p->d < 1 || p->d > kMaxDim)
      family<R>::fatal("nfft: dimension d = %td is not supported (1 <= d <= 6)", p->d);
    p->N_total = 1;
    p->n_total = 1;
    for (NFFT_INT t = 0; t < p->d; t++) {
      p->N_total *= p->N[t];
      p->n_total *= p->n[t];
    }
    p->sigma = (R *)Fam<R>::alloc(sizeof(R) * (size_t)p->d);
    p->b = (R *)Fam<R>::alloc(sizeof(R) * (size_t)p->d);
    for (NFFT_INT t = 0; t < p->d; t++) {
      p->sigma[t] = (R)p->n[t] / (R)p->N[t];
      p->b[t] = (R)(kPiL * (2.0L - 1.0L / ((long double)p->n[t] / (long double)p->N[t])));
    }
    if (p->flags & MALLOC_X) p->x = (R *)host_array_alloc<R>(sizeof(R) * (size_t)(p->d * p->M_total));
    if (p->flags & MALLOC_F_HAT)
      p->f_hat = (decltype(p->f_hat))host_array_alloc<R>(2 * sizeof(R) * (size_t)p->N_total);
    if (p->flags & MALLOC_F)
      p->f = (decltype(p->f))host_array_alloc<R>(2 * sizeof(R) * (size_t)p->M_total);
    if (p->flags & PRE_PHI_HUT) {
      p->c_phi_inv = (R **)Fam<R>::alloc(sizeof(R *) * (size_t)p->d);
      for (NFFT_INT t = 0; t < p->d; t++) {
        p->c_phi_inv[t] = (R *)Fam<R>::alloc(sizeof(R) * (size_t)p->N[t]);
        long double b = kPiL * (2.0L - 1.0L / ((long double)p->n[t] / (long double)p->N[t]));
        for (NFFT_INT ks = 0; ks < p->N[t]; ks++)
          p->c_phi_inv[t][ks] = (R)(1.0L / phi_hut(p->n[t], ks - p->N[t] / 2, (int)p->m, b));
      }
    } else {
      p->c_phi_inv = nullptr;
    }
    if ((p->flags & PRE_LIN_PSI) && p->K == 0) p->K = m2K(p->m);
    p->psi = nullptr;
    p->psi_index_g = nullptr;
    p->psi_index_f = nullptr;
    p->g = p->g_hat = p->g1 = p->g2 = nullptr;
    p->spline_coeffs = nullptr;
    p->index_x = nullptr;
    for (int i = 0; i < 3; i++) p->MEASURE_TIME_t[i] = 0;
    p->my_fftw_plan2 = nullptr;

    long N[kMaxDim], n[kMaxDim];
    for (NFFT_INT t = 0; t < p->d; t++) {
      N[t] = (long)p->N[t];
      n[t] = (long)p->n[t];
    }
    Engine<R> *e = new Engine<R>();
    e->create((int)p->d, N, n, (long)p->M_total, (int)p->m);
    p->my_fftw_plan1 = reinterpret_cast<decltype(p->my_fftw_plan1)>(e);
    set_mv(p);
  }

  static void init_guru(PLAN *p, int d, int *N, int M, int *n, int m, unsigned flags,
                        unsigned fftw_flags) {
    p->d = d;
    p->M_total = M;
    p->N = (NFFT_INT *)Fam<R>::alloc(sizeof(NFFT_INT) * (size_t)d);
    p->n = (NFFT_INT *)Fam<R>::alloc(sizeof(NFFT_INT) * (size_t)d);
    for (int t = 0; t < d; t++) {
      p->N[t] = N[t];
      p->n[t] = n[t];
    }
    p->m = m;
    p->flags = flags;
    p->fftw_flags = fftw_flags;
    p->K = 0;
    init_help(p);
  }

  static void init_lin(PLAN *p, int d, int *N, int M, int *n, int m, int K, unsigned flags,
                       unsigned fftw_flags) {
    p->d = d;
    p->M_total = M;
    p->N = (NFFT_INT *)Fam<R>::alloc(sizeof(NFFT_INT) * (size_t)d);
    p->n = (NFFT_INT *)Fam<R>::alloc(sizeof(NFFT_INT) * (size_t)d);
    for (int t = 0; t < d; t++) {
      p->N[t] = N[t];
      p->n[t] = n[t];
    }
    p->m = m;
    p->flags = flags;
    p->fftw_flags = fftw_flags;
    p->K = K;
    init_help(p);
  }

  // defaults of the reference X(init) (kernel/nfft/nfft.c:6048-6087)
  static void init(PLAN *p, int d, int *N, int M) {
    p->d = d;
    p->N = (NFFT_INT *)Fam<R>::alloc(sizeof(NFFT_INT) * (size_t)d);
    p->n = (NFFT_INT *)Fam<R>::alloc(sizeof(NFFT_INT) * (size_t)d);
    for (int t = 0; t < d; t++) {
      p->N[t] = N[t];
      NFFT_INT np2 = 1;
      if (N[t] < 2) np2 = N[t] + 1;
      else
        while (np2 < N[t]) np2 <<= 1;
      p->n[t] = 2 * np2;
    }
    p->M_total = M;
    p->m = Fam<R>::default_m;
    if (d > 1)
      p->flags = PRE_PHI_HUT | PRE_PSI | MALLOC_X | MALLOC_F_HAT | MALLOC_F | FFTW_INIT |
                 NFFT_SORT_NODES;
    else
      p->flags = PRE_PHI_HUT | PRE_PSI | MALLOC_X | MALLOC_F_HAT | MALLOC_F | FFTW_INIT |
                 FFT_OUT_OF_PLACE;
    p->fftw_flags = FFTW_ESTIMATE | FFTW_DESTROY_INPUT;
    p->K = 0;
    init_help(p);
  }

  static void precompute(PLAN *p) {
    clear_launch_error();
    Engine<R> *e = eng(p);
    e->upload_nodes(p->x);
    e->build_nodes();
    // FFTW_MEASURE (no FFTW_ESTIMATE bit) asks for measured plans: here the
    // kernel shapes of B and B^T are timed once on the plan's nodes
    if (!(p->fftw_flags & FFTW_ESTIMATE) && !e->tuned) e->autotune();
  }

  static void ensure_nodes(PLAN *p, Engine<R> *e) {
    if (!(p->flags & PRE_ONE_PSI) || !e->nodes_ready) precompute(p);
  }

  static void store_times(PLAN *p, Engine<R> *e, bool is_adjoint) {
    if (!e->timing) return;
    double t[4];
    e->stage_times(t);
    if (is_adjoint) {  // recorded order: spreading, tile reduction, FFT, D^T
      p->MEASURE_TIME_t[0] = (R)t[3];
      p->MEASURE_TIME_t[1] = (R)t[2];
      p->MEASURE_TIME_t[2] = (R)(t[0] + t[1]);
    } else {
      p->MEASURE_TIME_t[0] = (R)t[0];
      p->MEASURE_TIME_t[1] = (R)t[1];
      p->MEASURE_TIME_t[2] = (R)t[2];
    }
  }

  static void trafo_direct(const PLAN *p) {
    clear_launch_error();
    Engine<R> *e = eng(p);
    e->upload_nodes(p->x);
    if (p->N_total > 0)
      NFFT_HIP_CHECK(R, hipMemcpyAsync(e->fhat, p->f_hat, sizeof(cplx<R>) * (size_t)p->N_total,
                                       hipMemcpyHostToDevice, e->stream));
    e->trafo_direct(e->fhat, e->f);
    if (p->M_total > 0)
      NFFT_HIP_CHECK(R, hipMemcpyAsync(p->f, e->f, sizeof(cplx<R>) * (size_t)p->M_total,
                                       hipMemcpyDeviceToHost, e->stream));
    NFFT_HIP_CHECK(R, hipStreamSynchronize(e->stream));
  }

  static void adjoint_direct(const PLAN *p) {
    clear_launch_error();
    Engine<R> *e = eng(p);
    e->upload_nodes(p->x);
    if (p->M_total > 0)
      NFFT_HIP_CHECK(R, hipMemcpyAsync(e->f, p->f, sizeof(cplx<R>) * (size_t)p->M_total,
                                       hipMemcpyHostToDevice, e->stream));
    e->adjoint_direct(e->f, e->fhat);
    if (p->N_total > 0)
      NFFT_HIP_CHECK(R, hipMemcpyAsync(p->f_hat, e->fhat, sizeof(cplx<R>) * (size_t)p->N_total,
                                       hipMemcpyDeviceToHost, e->stream));
    NFFT_HIP_CHECK(R, hipStreamSynchronize(e->stream));
  }

  static void trafo(PLAN *p) {
    clear_launch_error();
    Engine<R> *e = eng(p);
    if (e->direct_only) {
      trafo_direct(p);
      return;
    }
    ensure_nodes(p, e);
    NFFT_HIP_CHECK(R, hipMemcpyAsync(e->fhat, p->f_hat, sizeof(cplx<R>) * (size_t)p->N_total,
                                     hipMemcpyHostToDevice, e->stream));
    e->trafo(e->fhat, e->f);
    if (p->M_total > 0)
      NFFT_HIP_CHECK(R, hipMemcpyAsync(p->f, e->f, sizeof(cplx<R>) * (size_t)p->M_total,
                                       hipMemcpyDeviceToHost, e->stream));
    NFFT_HIP_CHECK(R, hipStreamSynchronize(e->stream));
    store_times(p, e, false);
  }

  static void adjoint(PLAN *p) {
    clear_launch_error();
    Engine<R> *e = eng(p);
    if (e->direct_only) {
      adjoint_direct(p);
      return;
    }
    ensure_nodes(p, e);
    if (p->M_total > 0)
      NFFT_HIP_CHECK(R, hipMemcpyAsync(e->f, p->f, sizeof(cplx<R>) * (size_t)p->M_total,
                                       hipMemcpyHostToDevice, e->stream));
    e->adjoint(e->f, e->fhat);
    NFFT_HIP_CHECK(R, hipMemcpyAsync(p->f_hat, e->fhat, sizeof(cplx<R>) * (size_t)p->N_total,
                                     hipMemcpyDeviceToHost, e->stream));
    NFFT_HIP_CHECK(R, hipStreamSynchronize(e->stream));
    store_times(p, e, true);
  }

  static void trafo_device(PLAN *p, const void *fh, void *f) {
    clear_launch_error();
    Engine<R> *e = eng(p);
    if (e->direct_only) {
      e->upload_nodes(p->x);  // the NDFT reads the current x (nfft.c:145-297)
      e->trafo_direct((const cplx<R> *)fh, (cplx<R> *)f);
      return;
    }
    ensure_nodes(p, e);
    e->trafo((const cplx<R> *)fh, (cplx<R> *)f);
    if (e->timing) store_times(p, e, false);
  }

  static void adjoint_device(PLAN *p, const void *f, void *fh) {
    clear_launch_error();
    Engine<R> *e = eng(p);
    if (e->direct_only) {
      e->upload_nodes(p->x);  // the NDFT reads the current x (nfft.c:145-297)
      e->adjoint_direct((const cplx<R> *)f, (cplx<R> *)fh);
      return;
    }
    ensure_nodes(p, e);
    e->adjoint((const cplx<R> *)f, (cplx<R> *)fh);
    if (e->timing) store_times(p, e, true);
  }

  // Gridding reconstruction f_hat = A^H (w .* f), the weighted adjoint of the
  // reference's MRI applications (applications/mri/mri2d/
  // reconstruct_data_gridding.c:54-68 scales f by the weights of weights.dat
  // and calls nfft_adjoint).  The weighted values go to the plan's device f
  // buffer; the caller's f is left as it is.  w == NULL: plain adjoint.
  static void gridding_device(PLAN *p, const void *f, const R *w, void *fh) {
    if (!w) {
      adjoint_device(p, f, fh);
      return;
    }
    clear_launch_error();
    Engine<R> *e = eng(p);
    if (e->direct_only) {
      e->upload_nodes(p->x);  // the NDFT reads the current x (nfft.c:145-297)
    } else {
      ensure_nodes(p, e);
    }
    const long M = p->M_total;
    if (M > 0)
      hipLaunchKernelGGL(k_weight_mul<R>, dim3((unsigned)ceil_div(M, 256)), dim3(256), 0,
                         e->stream, (const cplx<R> *)f, w, e->f, M);
    if (e->direct_only) {
      e->adjoint_direct(e->f, (cplx<R> *)fh);
      return;
    }
    e->adjoint(e->f, (cplx<R> *)fh);
    if (e->timing) store_times(p, e, true);
  }

  // host arrays: f_hat = A^H (w .* f) with the plan's f and f_hat
  static void gridding(PLAN *p, const R *w) {
    clear_launch_error();
    Engine<R> *e = eng(p);
    const long M = p->M_total;
    std::vector<cplx<R>> fw((size_t)std::max(0L, M));
    const R *f = reinterpret_cast<const R *>(p->f);
    for (long j = 0; j < M; j++) {
      const R wj = w ? w[j] : R(1);
      fw[j] = cplx<R>{wj * f[2 * j], wj * f[2 * j + 1]};
    }
    if (e->direct_only) {
      e->upload_nodes(p->x);  // the NDFT reads the current x (nfft.c:145-297)
    } else {
      ensure_nodes(p, e);
    }
    if (M > 0)
      NFFT_HIP_CHECK(R, hipMemcpyAsync(e->f, fw.data(), sizeof(cplx<R>) * (size_t)M,
                                       hipMemcpyHostToDevice, e->stream));
    if (e->direct_only) e->adjoint_direct(e->f, e->fhat);
    else e->adjoint(e->f, e->fhat);
    NFFT_HIP_CHECK(R, hipMemcpyAsync(p->f_hat, e->fhat, sizeof(cplx<R>) * (size_t)p->N_total,
                                     hipMemcpyDeviceToHost, e->stream));
    NFFT_HIP_CHECK(R, hipStreamSynchronize(e->stream));
    if (!e->direct_only) store_times(p, e, true);
  }

  // nrhs transforms sharing the plan's nodes: vectors r of f_hat_dev / f_dev
  // at offsets r N_total / r M_total, in batched launches (one launch per
  // kernel for a chunk of right-hand sides) where the plan's kernels allow,
  // else enqueued back to back
  static void trafo_batch_device(PLAN *p, const void *fh, void *f, int nrhs) {
    Engine<R> *e = eng(p);
    if (nrhs < 2 || e->direct_only || e->timing) {
      for (int r = 0; r < nrhs; r++)
        trafo_device(p, (const cplx<R> *)fh + (long)r * p->N_total, (cplx<R> *)f + (long)r * p->M_total);
      return;
    }
    clear_launch_error();
    ensure_nodes(p, e);
    e->trafo_batch((const cplx<R> *)fh, (cplx<R> *)f, nrhs);
  }
  static void adjoint_batch_device(PLAN *p, const void *f, void *fh, int nrhs) {
    Engine<R> *e = eng(p);
    if (nrhs < 2 || e->direct_only || e->timing) {
      for (int r = 0; r < nrhs; r++)
        adjoint_device(p, (const cplx<R> *)f + (long)r * p->M_total, (cplx<R> *)fh + (long)r * p->N_total);
      return;
    }
    clear_launch_error();
    ensure_nodes(p, e);
    e->adjoint_batch((const cplx<R> *)f, (cplx<R> *)fh, nrhs);
  }

  // adjoint_batch_device without any allocation (nfft3_hip.h): the batch
  // grids and tiles of each chunk of right-hand sides live in the library's
  // static scratch array, so the plan must have its node data already (a
  // PRE_* psi flag and precompute_one_psi called) and stage timing off.
  // Returns 0 when the work is enqueued, else a negative code with a message
  // on stderr and nothing enqueued.
  static int adjoint_batch_device_noalloc(PLAN *p, const void *f, void *fh, int nrhs) {
    Engine<R> *e = eng(p);
    const char *why = nullptr;
    int rc = 0;
    if (nrhs < 0) {
      rc = -1;
      why = "nrhs < 0";
    } else if (e->timing || e->deferred) {
      rc = -2;
      why = "stage timing is on (it records into events it creates)";
    } else if (!e->direct_only && (!(p->flags & PRE_ONE_PSI) || !e->nodes_ready)) {
      rc = -2;
      why = "node data not built: use a PRE_* psi flag and call precompute_one_psi first";
    }
    if (!rc && nrhs > 0) {
      clear_launch_error();
      if (e->direct_only) {
        e->upload_nodes(p->x);  // the NDFT reads the current x (nfft.c:145-297)
        for (int r = 0; r < nrhs; r++)
          e->adjoint_direct((const cplx<R> *)f + (long)r * p->M_total,
                            (cplx<R> *)fh + (long)r * p->N_total);
      } else {
        rc = e->adjoint_batch_noalloc((const cplx<R> *)f, (cplx<R> *)fh, nrhs);
        if (rc == -3) why = "the scratch array's device address is unknown on this device";
        if (rc == -4) why = "one right-hand side needs more than the 64 MiB scratch array";
      }
    }
    if (rc)
      std::fprintf(stderr, "%sadjoint_batch_device_noalloc: %s (per right-hand side %ld bytes)\n",
                   sizeof(R) == 4 ? "nfftf_" : "nfft_", why,
                   e->direct_only ? 0L : e->noalloc_bytes_per_rhs());
    return rc;
  }

  // One hipGraph for the batched transforms of `count` plans: captured on
  // plans[0]'s stream, each other plan's stream forks from it and joins it
  // again, so plans on different streams run concurrently on replay.  Node
  // data are built before the capture; nothing in the captured work
  // allocates, copies from the host or clears memory with memset.
  static nfft_hip_graph *capture(PLAN **plans, int count, const void **in, void **out, int nrhs,
                                 bool adjoint) {
    clear_launch_error();
    if (count < 1) family<R>::fatal("nfft: a graph needs at least one plan");
    for (int i = 0; i < count; i++) {
      Engine<R> *e = eng(plans[i]);
      if (e->timing || e->deferred)
        family<R>::fatal("nfft: switch stage timing off before capturing a graph");
      if (e->direct_only) {
        e->upload_nodes(plans[i]->x);
      } else {
        ensure_nodes(plans[i], e);
        e->reserve_batch(nrhs);  // batch buffers exist before the capture
      }
    }
    hipStream_t s0 = eng(plans[0])->stream;
    NFFT_HIP_CHECK(R, hipStreamSynchronize(s0));
    std::vector<hipEvent_t> evs;
    auto new_event = [&]() {
      hipEvent_t ev;
      NFFT_HIP_CHECK(R, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      evs.push_back(ev);
      return ev;
    };
    hipEvent_t fork = new_event();
    NFFT_HIP_CHECK(R, hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    NFFT_HIP_CHECK(R, hipEventRecord(fork, s0));
    for (int i = 0; i < count; i++) {
      PLAN *p = plans[i];
      Engine<R> *e = eng(p);
      hipStream_t si = e->stream;
      if (si != s0) NFFT_HIP_CHECK(R, hipStreamWaitEvent(si, fork, 0));
      if (!e->direct_only && nrhs > 1) {
        if (!adjoint) e->trafo_batch((const cplx<R> *)in[i], (cplx<R> *)out[i], nrhs);
        else e->adjoint_batch((const cplx<R> *)in[i], (cplx<R> *)out[i], nrhs);
      } else {
        for (int r = 0; r < nrhs; r++) {
          if (!adjoint) {
            const cplx<R> *fh = (const cplx<R> *)in[i] + (long)r * p->N_total;
            cplx<R> *f = (cplx<R> *)out[i] + (long)r * p->M_total;
            if (e->direct_only) e->trafo_direct(fh, f);
            else e->trafo(fh, f);
          } else {
            const cplx<R> *f = (const cplx<R> *)in[i] + (long)r * p->M_total;
            cplx<R> *fh = (cplx<R> *)out[i] + (long)r * p->N_total;
            if (e->direct_only) e->adjoint_direct(f, fh);
            else e->adjoint(f, fh);
          }
        }
      }
      if (si != s0) {
        hipEvent_t join = new_event();
        NFFT_HIP_CHECK(R, hipEventRecord(join, si));
        NFFT_HIP_CHECK(R, hipStreamWaitEvent(s0, join, 0));
      }
    }
    hipGraph_t graph;
    NFFT_HIP_CHECK(R, hipStreamEndCapture(s0, &graph));
    auto *h = new nfft_hip_graph();
    h->stream = s0;
    NFFT_HIP_CHECK(R, hipGraphInstantiate(&h->exec, graph, nullptr, nullptr, 0));
    NFFT_HIP_CHECK(R, hipGraphDestroy(graph));
    for (hipEvent_t ev : evs) hipEventDestroy(ev);
    return h;
  }

  // reference X(check) (kernel/nfft/nfft.c:6169-6207)
  static const char *check(PLAN *p) {
    if (!p->f) return "Member f not initialized.";
    if (!p->x) return "Member x not initialized.";
    if (!p->f_hat) return "Member f_hat not initialized.";
    if ((p->flags & PRE_LIN_PSI) && p->K < p->M_total)
      return "Number of nodes too small to use PRE_LIN_PSI.";
    for (NFFT_INT j = 0; j < p->M_total * p->d; j++)
      if (p->x[j] < R(-0.5) || p->x[j] >= R(0.5)) return "ths->x out of range [-0.5,0.5)";
    for (NFFT_INT t = 0; t < p->d; t++) {
      if (p->sigma[t] <= 1) return "Oversampling factor too small";
      if (p->N[t] % 2 == 1) return "polynomial degree N has to be even";
    }
    return nullptr;
  }

  // reference X(finalize) (kernel/nfft/nfft.c:6209-6270)
  static void finalize(PLAN *p) {
    Engine<R> *e = reinterpret_cast<Engine<R> *>(p->my_fftw_plan1);
    if (e) {
      e->destroy();
      delete e;
    }
    p->my_fftw_plan1 = nullptr;
    if ((p->flags & PRE_PHI_HUT) && p->c_phi_inv) {
      for (NFFT_INT t = 0; t < p->d; t++) Fam<R>::release(p->c_phi_inv[t]);
      Fam<R>::release(p->c_phi_inv);
      p->c_phi_inv = nullptr;
    }
    if (p->flags & MALLOC_F) host_array_free<R>(p->f);
    if (p->flags & MALLOC_F_HAT) host_array_free<R>(p->f_hat);
    if (p->flags & MALLOC_X) host_array_free<R>(p->x);
    Fam<R>::release(p->b);
    Fam<R>::release(p->sigma);
    Fam<R>::release(p->n);
    Fam<R>::release(p->N);
  }

  static void set_stream(PLAN *p, void *s) {
    Engine<R> *e = eng(p);
    e->stream = s ? (hipStream_t)s : e->own_stream;
  }
  static void *get_stream(PLAN *p) { return (void *)eng(p)->stream; }
  static void synchronize(PLAN *p) { NFFT_HIP_CHECK(R, hipStreamSynchronize(eng(p)->stream)); }
  static void set_timing(PLAN *p, int on) { eng(p)->timing = on != 0; }
  static long num_subproblems(PLAN *p) { return eng(p)->nsub; }
};

}  // namespace nfft_hip

using nfft_hip::PlanApi;

#define NFFT_HIP_EXPORT_PLAN(P, R, PLAN)                                                       \
  extern "C" {                                                                                \
  void P##init_guru(PLAN *p, int d, int *N, int M, int *n, int m, unsigned fl, unsigned ff) { \
    PlanApi<R, PLAN>::init_guru(p, d, N, M, n, m, fl, ff);                                    \
  }                                                                                           \
  void P##init_lin(PLAN *p, int d, int *N, int M, int *n, int m, int K, unsigned fl,          \
                   unsigned ff) {                                                             \
    PlanApi<R, PLAN>::init_lin(p, d, N, M, n, m, K, fl, ff);                                  \
  }                                                                                           \
  void P##init(PLAN *p, int d, int *N, int M) { PlanApi<R, PLAN>::init(p, d, N, M); }         \
  void P##init_1d(PLAN *p, int N1, int M) {                                                   \
    int N[1] = {N1};                                                                          \
    PlanApi<R, PLAN>::init(p, 1, N, M);                                                       \
  }                                                                                           \
  void P##init_2d(PLAN *p, int N1, int N2, int M) {                                           \
    int N[2] = {N1, N2};                                                                      \
    PlanApi<R, PLAN>::init(p, 2, N, M);                                                       \
  }                                                                                           \
  void P##init_3d(PLAN *p, int N1, int N2, int N3, int M) {                                   \
    int N[3] = {N1, N2, N3};                                                                  \
    PlanApi<R, PLAN>::init(p, 3, N, M);                                                       \
  }                                                                                           \
  void P##precompute_one_psi(PLAN *p) { PlanApi<R, PLAN>::precompute(p); }                    \
  void P##precompute_psi(PLAN *p) { PlanApi<R, PLAN>::precompute(p); }                        \
  void P##precompute_full_psi(PLAN *p) { PlanApi<R, PLAN>::precompute(p); }                   \
  void P##precompute_fg_psi(PLAN *p) { PlanApi<R, PLAN>::precompute(p); }                     \
  void P##precompute_lin_psi(PLAN *p) { PlanApi<R, PLAN>::precompute(p); }                    \
  void P##trafo(PLAN *p) { PlanApi<R, PLAN>::trafo(p); }                                      \
  void P##trafo_1d(PLAN *p) { PlanApi<R, PLAN>::trafo(p); }                                   \
  void P##trafo_2d(PLAN *p) { PlanApi<R, PLAN>::trafo(p); }                                   \
  void P##trafo_3d(PLAN *p) { PlanApi<R, PLAN>::trafo(p); }                                   \
  void P##adjoint(PLAN *p) { PlanApi<R, PLAN>::adjoint(p); }                                  \
  void P##adjoint_1d(PLAN *p) { PlanApi<R, PLAN>::adjoint(p); }                               \
  void P##adjoint_2d(PLAN *p) { PlanApi<R, PLAN>::adjoint(p); }                               \
  void P##adjoint_3d(PLAN *p) { PlanApi<R, PLAN>::adjoint(p); }                               \
  void P##trafo_direct(const PLAN *p) { PlanApi<R, PLAN>::trafo_direct(p); }                  \
  void P##adjoint_direct(const PLAN *p) { PlanApi<R, PLAN>::adjoint_direct(p); }              \
  const char *P##check(PLAN *p) { return PlanApi<R, PLAN>::check(p); }                        \
  void P##finalize(PLAN *p) { PlanApi<R, PLAN>::finalize(p); }                                \
  void P##trafo_device(PLAN *p, const void *fh, void *f) {                                    \
    PlanApi<R, PLAN>::trafo_device(p, fh, f);                                                 \
  }                                                                                           \
  void P##adjoint_device(PLAN *p, const void *f, void *fh) {                                  \
    PlanApi<R, PLAN>::adjoint_device(p, f, fh);                                               \
  }                                                                                           \
  void P##set_stream(PLAN *p, void *s) { PlanApi<R, PLAN>::set_stream(p, s); }                \
  void *P##get_stream(PLAN *p) { return PlanApi<R, PLAN>::get_stream(p); }                    \
  void P##synchronize(PLAN *p) { PlanApi<R, PLAN>::synchronize(p); }                          \
  void P##set_timing(PLAN *p, int on) { PlanApi<R, PLAN>::set_timing(p, on); }                \
  void P##timing_begin(PLAN *p) { PlanApi<R, PLAN>::eng(p)->timing_begin(); }                 \
  void P##timing_end(PLAN *p, double *out) { PlanApi<R, PLAN>::eng(p)->timing_end(out); }     \
  long P##num_subproblems(PLAN *p) { return PlanApi<R, PLAN>::num_subproblems(p); }           \
  void P##gridding_device(PLAN *p, const void *f, const R *w, void *fh) {                     \
    PlanApi<R, PLAN>::gridding_device(p, f, w, fh);                                           \
  }                                                                                           \
  void P##gridding(PLAN *p, const R *w) { PlanApi<R, PLAN>::gridding(p, w); }                 \
  void P##trafo_batch_device(PLAN *p, const void *fh, void *f, int nrhs) {                    \
    PlanApi<R, PLAN>::trafo_batch_device(p, fh, f, nrhs);                                     \
  }                                                                                           \
  void P##adjoint_batch_device(PLAN *p, const void *f, void *fh, int nrhs) {                  \
    PlanApi<R, PLAN>::adjoint_batch_device(p, f, fh, nrhs);                                   \
  }                                                                                           \
  int P##adjoint_batch_device_noalloc(PLAN *p, const void *f, void *fh, int nrhs) {           \
    return PlanApi<R, PLAN>::adjoint_batch_device_noalloc(p, f, fh, nrhs);                    \
  }                                                                                           \
  nfft_hip_graph *P##graph_trafo(PLAN **plans, int count, const void **fh, void **f,          \
                                 int nrhs) {                                                  \
    return PlanApi<R, PLAN>::capture(plans, count, fh, f, nrhs, false);                       \
  }                                                                                           \
  nfft_hip_graph *P##graph_adjoint(PLAN **plans, int count, const void **f, void **fh,        \
                                   int nrhs) {                                                \
    return PlanApi<R, PLAN>::capture(plans, count, f, fh, nrhs, true);                        \
  }                                                                                           \
  }

NFFT_HIP_EXPORT_PLAN(nfftf_, float, nfftf_plan)
NFFT_HIP_EXPORT_PLAN(nfft_, double, nfft_plan)

extern "C" {
void nfft_hip_graph_launch(nfft_hip_graph *g) {
  NFFT_HIP_CHECK(double, hipGraphLaunch(g->exec, g->stream));
}
void nfft_hip_graph_synchronize(nfft_hip_graph *g) {
  NFFT_HIP_CHECK(double, hipStreamSynchronize(g->stream));
}
void nfft_hip_graph_destroy(nfft_hip_graph *g) {
  if (!g) return;
  hipStreamSynchronize(g->stream);
  hipGraphExecDestroy(g->exec);
  delete g;
}
}

namespace nfft_hip {

template void *host_array_alloc<float>(size_t);
template void *host_array_alloc<double>(size_t);
template void host_array_free<float>(void *);
template void host_array_free<double>(void *);

// The solver reaches the plan only through the mv prefix; it is one of this
// library's plans exactly when its mv_trafo is this library's trafo.
template <>
Engine<float> *engine_of_mv<float>(void *mv) {
  nfftf_plan *p = static_cast<nfftf_plan *>(mv);
  if (p->mv_trafo != reinterpret_cast<void (*)(void *)>(nfftf_trafo)) return nullptr;
  return reinterpret_cast<Engine<float> *>(p->my_fftw_plan1);
}

template <>
Engine<double> *engine_of_mv<double>(void *mv) {
  nfft_plan *p = static_cast<nfft_plan *>(mv);
  if (p->mv_trafo != reinterpret_cast<void (*)(void *)>(nfft_trafo)) return nullptr;
  return reinterpret_cast<Engine<double> *>(p->my_fftw_plan1);
}

}  // namespace nfft_hip

namespace nfft_hip {

template <>
void ensure_nodes_of_mv<float>(void *mv) {
  nfftf_plan *p = static_cast<nfftf_plan *>(mv);
  PlanApi<float, nfftf_plan>::ensure_nodes(p, PlanApi<float, nfftf_plan>::eng(p));
}

template <>
void ensure_nodes_of_mv<double>(void *mv) {
  nfft_plan *p = static_cast<nfft_plan *>(mv);
  PlanApi<double, nfft_plan>::ensure_nodes(p, PlanApi<double, nfft_plan>::eng(p));
}

}  // namespace nfft_hip
