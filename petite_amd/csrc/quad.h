// Track-length integrals of the dark-sector tables on the device: the weight
// arrays and interaction-energy rate rows of DarkShower
// (src/PETITE/dark_shower.py:311-593: construct_*_weight_array 337-399,
// _d_rate_d_E_* 454-533) with the survival factors of the SM interaction
// integrals (src/PETITE/shower.py:298-354), and the bound-electron cross
// sections they integrate (src/PETITE/atomic_annihilation.py:10-104,
// src/PETITE/atomic_compton.py:10-109).
//
// One integral  int_lo^hi n sigma(E) / (dE/dx) * S(E, Ei) dE  per wave.  The
// reference integrates with scipy.integrate.quad; here the interval is cut
// into panels and each panel takes 8-point Gauss-Legendre:
//   * geometric edges Ei - d, d = (scale / 8) * 1.35^k with scale = 10 mean
//     free paths x dE/dx at Ei / 10 (the width of the survival factor), kept
//     where lo < Ei - d < hi;
//   * every table node strictly inside (lo, hi) (the integrand has kinks there).
// The wave lays the geometric edges out in LDS, counts the nodes inside each
// geometric panel by bisection, prefix-sums the counts, and then each lane
// takes sub-panels j = lane, lane + 64, ... of the merged partition; a
// shuffle reduction sums the lanes.
//
// The survival factor S(E, Ei) = exp(-sum_c [II_c(Ei) - II_c(E)] / (dE/dx)),
// 0 for E > Ei or a negative exponent, where II_c is the running integral of
// channel c's n sigma, exact for its piecewise-linear interpolant with fill 0
// outside the table (k_cumint: trapezoids over the union of the channel's
// table nodes and the grid II_c is kept on, summed in order).
#pragma once

namespace quad {

constexpr int kMaxGeo = 512;      // geometric panel edges per integral
constexpr int kWavesPerBlock = 4;
constexpr double kCmToM = 0.01;   // cmtom (physical_constants.py)

// 8-point Gauss-Legendre abscissae and weights on [-1, 1].
__constant__ double kGLX[8] = {-0.9602898564975362, -0.7966664774136267, -0.525532409916329, -0.18343464249564978,
                               0.18343464249564978, 0.525532409916329,  0.7966664774136267, 0.9602898564975362};
__constant__ double kGLW[8] = {0.10122853629037669, 0.22238103445337434, 0.31370664587788705, 0.36268378337836177,
                               0.36268378337836177, 0.31370664587788705, 0.22238103445337434, 0.10122853629037669};

// One job: a dark process's n sigma, the survival channels of one species,
// and the end energies of its integrals.
struct Job {
  const double* lx;          // log10 E of the log-log n sigma table
  const double* ly;          // log10 n sigma (cm^-1)
  double ns_l0, ns_inv;      // lx[0] and (n-1)/(lx[n-1]-lx[0]) when lx is near uniform (inv = 0: bisect)
  double grid_ln0[3], grid_inv[3];  // the same for ln of each survival grid
  int32_t n_ns;
  int32_t zero_small;        // n sigma < 1e-18 counts as 0 (all but bound-electron DarkAnn)
  int32_t n_loss;            // survival channels (<= 3)
  int32_t n_mfp;             // mean-free-path channels (<= 3)
  const double* loss_x[3];   // each channel's n sigma table (x, y)
  const double* loss_y[3];
  int32_t loss_n[3];
  int32_t grid_n[3];         // points of the grid its running integral is kept on
  const double* grid[3];
  double* cum[3];            // running integral on that grid (written by k_cumint)
  const double* mfp_x[3];    // n sigma tables summed for the mean free path
  const double* mfp_y[3];
  int32_t mfp_n[3];
  int32_t n_nodes;
  const double* nodes;       // panel nodes, ascending, distinct
  const double* E;           // end energies Ei
  int32_t n_E;
  int32_t mode;              // 0: weight from lo to Ei; 1: ten rate bins below Ei
  double lo;                 // weights: lower limit; rates: table floor
  double dEdx;               // GeV/m
  double* out;               // weights: n_E; rates: n_E x 10 x (centre, value)
  int64_t first;             // index of the job's first integral
};

// numpy.interp (slope form) inside [x0, x_{n-1}], 0 outside (scipy interp1d
// with fill_value 0).
__host__ __device__ inline double lin0(const double* x, const double* y, int n, double v) {
  if (n <= 0 || v < x[0] || v > x[n - 1]) return 0.0;
  if (v == x[n - 1]) return y[n - 1];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (x[mid] <= v) lo = mid; else hi = mid;
  }
  const double slope = (y[lo + 1] - y[lo]) / (x[lo + 1] - x[lo]);
  return slope * (v - x[lo]) + y[lo];
}

// Interval lo with x[lo] <= v < x[lo+1] of an ascending grid, v inside
// [x0, x_{n-1}): from the guess (g - g0) * inv on a near-uniform grid (in
// g = ln v or log10 v), corrected against the nodes; bisection when inv = 0.
__device__ inline int interval(const double* x, int n, double v, double g, double g0, double inv) {
  int lo;
  if (inv > 0.0) {
    const double t = (g - g0) * inv;
    lo = t < 0.0 ? 0 : (t > (double)(n - 2) ? n - 2 : (int)t);
    while (lo > 0 && x[lo] > v) --lo;
    while (lo < n - 2 && x[lo + 1] <= v) ++lo;
    return lo;
  }
  lo = 0;
  int hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (x[mid] <= v) lo = mid; else hi = mid;
  }
  return lo;
}

// numpy.interp inside the grid, 0 outside, with an interval guess from lnv.
__device__ inline double lin0g(const double* x, const double* y, int n, double v, double lnv, double ln0, double inv) {
  if (n <= 0 || v < x[0] || v > x[n - 1]) return 0.0;
  if (v == x[n - 1]) return y[n - 1];
  const int lo = interval(x, n, v, lnv, ln0, inv);
  const double slope = (y[lo + 1] - y[lo]) / (x[lo + 1] - x[lo]);
  return slope * (v - x[lo]) + y[lo];
}

// interpolate1d in log10-log10 with fill 1e-20 (dark_shower.py:31-46).
__device__ inline double loglog_val(const double* lx, const double* ly, int n, double E, double l0, double inv) {
  const double l = log10(E);
  if (n <= 0 || l < lx[0] || l > lx[n - 1]) return 1.0e-20;
  if (l == lx[n - 1]) return pow(10.0, ly[n - 1]);
  const int lo = interval(lx, n, l, l, l0, inv);
  const double slope = (ly[lo + 1] - ly[lo]) / (lx[lo + 1] - lx[lo]);
  return pow(10.0, slope * (l - lx[lo]) + ly[lo]);
}

// Mean free path in metres (DarkShower.get_mfp, shower.py:283-296): cmtom over
// the summed n sigma, 1e12 where it is not positive.
__device__ inline double mfp_of(const Job& J, double E) {
  double ns = 0.0;
  for (int c = 0; c < J.n_mfp; ++c) ns += lin0(J.mfp_x[c], J.mfp_y[c], J.mfp_n[c], E);
  return ns <= 0.0 ? 1.0e12 : kCmToM / ns;
}

// first index with a[i] > v (a ascending, n entries)
__device__ inline int upper_bound(const double* a, int n, double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// first index with a[i] >= v
__device__ inline int lower_bound(const double* a, int n, double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Running integral of each survival channel (shower.py:298-320): trapezoids of
// the channel's piecewise-linear n sigma over the union of its nodes and the
// grid, summed in order, read off at the grid nodes.  One thread per (job,
// channel).
__global__ void k_cumint(const Job* __restrict__ jobs, int n_jobs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = t / 3, c = t % 3;
  if (j >= n_jobs || c >= jobs[j].n_loss) return;
  const Job& J = jobs[j];
  const double* xg = J.loss_x[c];
  const double* yg = J.loss_y[c];
  const int nx = J.loss_n[c];
  const double* z = J.grid[c];
  const int nz = J.grid_n[c];
  double* cum = J.cum[c];
  if (nz <= 0) return;
  // merge z (all of it) with the channel nodes strictly inside (z0, z_{nz-1})
  int iz = 0, ix = upper_bound(xg, nx, z[0]);
  const int ix_end = lower_bound(xg, nx, z[nz - 1]);
  double acc = 0.0, p_prev = 0.0, f_prev = 0.0;
  bool first = true;
  while (iz < nz || ix < ix_end) {
    double p;
    bool is_grid;
    if (ix >= ix_end || (iz < nz && z[iz] <= xg[ix])) {
      p = z[iz];
      is_grid = true;
      if (ix < ix_end && xg[ix] == p) ++ix;  // one point for equal values
    } else {
      p = xg[ix];
      is_grid = false;
    }
    const double f = lin0(xg, yg, nx, p);
    if (!first) {
      // a segment outside the table adds nothing (its fill is 0; the jump at
      // the table's ends is not a ramp)
      const double mid = 0.5 * (p + p_prev);
      if (!(mid < xg[0] || mid > xg[nx - 1])) acc += 0.5 * (f + f_prev) * (p - p_prev);
    }
    first = false;
    p_prev = p;
    f_prev = f;
    if (is_grid) {
      // repeated grid values take the same running integral
      while (iz < nz && z[iz] == p) cum[iz++] = acc;
    } else {
      ++ix;
    }
  }
}

__device__ inline double survival_arg(const Job& J, const double* iiEi, double E) {
  const double lnE = (double)__logf((float)E);  // interval guess only
  double d = 0.0;
  for (int c = 0; c < J.n_loss; ++c)
    d += iiEi[c] - lin0g(J.grid[c], J.cum[c], J.grid_n[c], E, lnE, J.grid_ln0[c], J.grid_inv[c]);
  return d;
}

// n sigma(E) / (dE/dx) * S(E, Ei)   (dark_shower.py:311-336 integrands)
__device__ inline double integrand(const Job& J, const double* iiEi, double E, double Ei) {
  const double k = J.dEdx * kCmToM;
  double n = loglog_val(J.lx, J.ly, J.n_ns, E, J.ns_l0, J.ns_inv);
  if (J.zero_small && n < 1.0e-18) n = 0.0;
  const double d = survival_arg(J, iiEi, E);
  const double s = (d < 0.0 || E > Ei) ? 0.0 : exp(-d / J.dEdx / kCmToM);
  return n / k * s;
}

struct Lds {
  double A[kMaxGeo + 2];    // panel edges lo, geometric edges ascending, hi
  int32_t ib[kMaxGeo + 1];  // first node inside each geometric panel
  int32_t S[kMaxGeo + 2];   // prefix sums of sub-panel counts
  int32_t K;
  int32_t bad;
};

// One integral over [lo, hi] with end energy Ei, by the calling wave (all 64
// lanes; L is the wave's LDS block).  Returns the sum in every lane.
__device__ double integrate(const Job& J, const double* iiEi, double lo, double hi, double Ei, double scale, Lds& L,
                            int lane) {
  if (!(hi > lo)) return 0.0;
  if (lane == 0) {
    // geometric edges Ei - d inside (lo, hi), found in descending order
    double d = scale / 8.0;
    int K = 0;
    while (Ei - d > lo) {
      if (Ei - d < hi) {
        if (K >= kMaxGeo) {
          L.bad = 1;
          break;
        }
        L.A[1 + K] = Ei - d;
        ++K;
      }
      d *= 1.35;
    }
    for (int i = 0; i < K / 2; ++i) {  // ascending
      const double t = L.A[1 + i];
      L.A[1 + i] = L.A[K - i];
      L.A[K - i] = t;
    }
    L.A[0] = lo;
    L.A[K + 1] = hi;
    L.K = K;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const int K = L.K;
  const int n_nodes = J.n_nodes;
  for (int a = lane; a <= K; a += 64) {
    const int ib = upper_bound(J.nodes, n_nodes, L.A[a]);
    const int ie = lower_bound(J.nodes, n_nodes, L.A[a + 1]);
    L.ib[a] = ib;
    L.S[a + 1] = (ie > ib ? ie - ib : 0) + 1;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (lane == 0) {
    L.S[0] = 0;
    for (int a = 1; a <= K + 1; ++a) L.S[a] += L.S[a - 1];
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const int T = L.S[K + 1];
  double acc = 0.0;
  for (int j = lane; j < T; j += 64) {
    // geometric panel a holding sub-panel j: S[a] <= j < S[a + 1]
    int a0 = 0, a1 = K + 1;
    while (a1 - a0 > 1) {
      const int mid = (a0 + a1) >> 1;
      if (L.S[mid] <= j) a0 = mid; else a1 = mid;
    }
    const int r = j - L.S[a0];
    const int cnt = L.S[a0 + 1] - L.S[a0] - 1;  // nodes inside panel a0
    const double left = r == 0 ? L.A[a0] : J.nodes[L.ib[a0] + r - 1];
    const double right = r == cnt ? L.A[a0 + 1] : J.nodes[L.ib[a0] + r];
    const double h = 0.5 * (right - left);
    const double c = 0.5 * (right + left);
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) s += kGLW[q] * integrand(J, iiEi, c + h * kGLX[q], Ei);
    acc += h * s;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  return acc;
}

// One wave per integral: global integral g -> (job, energy, bin).
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_track_quad(const Job* __restrict__ jobs, int n_jobs,
                                                                    int64_t n_int, int* __restrict__ bad) {
  __shared__ Lds lds[kWavesPerBlock];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t g = (int64_t)blockIdx.x * kWavesPerBlock + w;
  if (g >= n_int) return;  // whole waves leave; no block-wide barrier follows
  int j = 0;
  while (j + 1 < n_jobs && jobs[j + 1].first <= g) ++j;
  const Job& J = jobs[j];
  Lds& L = lds[w];
  if (lane == 0) L.bad = 0;
  const int64_t r = g - J.first;
  const int ie = J.mode ? (int)(r / 10) : (int)r;
  const int bin = J.mode ? (int)(r % 10) : 0;
  const double Ei = J.E[ie];
  double iiEi[3];
  for (int c = 0; c < J.n_loss; ++c) iiEi[c] = lin0(J.grid[c], J.cum[c], J.grid_n[c], Ei);
  const double scale = fmax(10.0 * mfp_of(J, Ei) * J.dEdx, 1.0e-9) / 10.0;
  double lo, hi;
  if (J.mode == 0) {
    lo = J.lo;
    hi = Ei;
  } else {
    // ten equal bins over [max(Ei - 10 mfp dE/dx, floor), Ei] (numpy.linspace)
    const double start = fmax(Ei - 10.0 * mfp_of(J, Ei) * J.dEdx, J.lo);
    const double step = (Ei - start) / 10.0;
    lo = (double)bin * step + start;
    hi = bin == 9 ? Ei : (double)(bin + 1) * step + start;
  }
  const double v = integrate(J, iiEi, lo, hi, Ei, scale, L, lane);
  if (lane == 0) {
    if (J.mode == 0) {
      J.out[ie] = v;
    } else {
      J.out[(ie * 10 + bin) * 2] = 0.5 * (hi + lo);
      J.out[(ie * 10 + bin) * 2 + 1] = v;
    }
    if (L.bad) atomicOr(bad, 1);
  }
}

}  // namespace quad

// ------------------------------------------- bound-electron cross sections --
// sigma(e+ e-(bound) -> V): tree level plus radiative tail
// (atomic_annihilation.py:70-104, with that module's own alpha = 1/137 and
// m_e = 0.511 MeV), and sigma(gamma e-(bound) -> e- V) (atomic_compton.py:
// 94-109).  Terms are formed in the reference's order: near threshold the
// closed forms cancel strongly.
namespace quad {

__device__ inline double sq(double x) { return x * x; }

// radiative-tail integral (atomic_annihilation.py:10-68), series for b >= 1000
__device__ inline double fancy(double a, double b) {
  if (b >= 1000.0) {
    const double b2 = b * b, b3 = b2 * b, b6 = b3 * b3;
    return (3081.0 / 280.0 - (3081.0 * sq(a)) / 40.0) / (b6 * b2) - (709.0 * a) / (35.0 * (b6 * b)) -
           91.0 / (30.0 * b6);
  }
  const double a2 = sq(a), a4 = a2 * a2, a6 = a4 * a2;
  const double p1 = 1.0 + a2, p2 = p1 * p1, p3 = p2 * p1;
  const double c5 = 5.0 + 3.0 * a2;
  const double e1 = -1.0 + 15.0 * a2 + 12.0 * a4;
  const double b2 = b * b, b3 = b2 * b, b4 = b3 * b, b5 = b4 * b, b6 = b5 * b, b7 = b6 * b;
  const double t1 = p1 * (a * p3 * c5 - p2 * e1 * b + a * (1.0 + a2) * (-23.0 + 38.0 * a2 + 21.0 * a4) * b2 +
                          4.0 * (5.0 + a2 - 18.0 * a4 - 6.0 * a6) * b3 + 3.0 * a * (3.0 + a2) * (-1.0 + 7.0 * a2) * b4 +
                          (7.0 - 29.0 * a2 - 12.0 * a4) * b5 + a * (7.0 + 3.0 * a2) * b6);
  const double poly = 3.0 * (p3 * p3) - a * p3 * (3.0 + 26.0 * a2 + 15.0 * a4) * b +
                      3.0 * p2 * (-7.0 + 21.0 * a2 + 23.0 * a4 + 11.0 * a6) * b2 -
                      3.0 * a * p1 * (7.0 + 69.0 * a2 + 45.0 * a4 + 15.0 * a6) * b3 +
                      (-11.0 + 132.0 * a2 + 9.0 * a4 * (38.0 + 5.0 * a2 * (4.0 + a2))) * b4 -
                      3.0 * a * (11.0 + 73.0 * a2 + 41.0 * a4 + 11.0 * a6) * b5 +
                      3.0 * (-1.0 + 27.0 * a2 + 17.0 * a4 + 5.0 * a6) * b6 - a * (15.0 + 10.0 * a2 + 3.0 * a4) * b7;
  const double t4 = p3 + 6.0 * a * p2 * b - 3.0 * (1.0 + 6.0 * a2 + 5.0 * a4) * b2 + 4.0 * a * (3.0 + 5.0 * a2) * b3 -
                    3.0 * (1.0 + 5.0 * a2) * b4 + 6.0 * a * b5 - b6;
  const double num = t1 + poly * atan2(a, 1.0) + poly * atan2(1.0 + a2 - a * b, b) +
                     4.0 * b * (t4 * (log(1.0 + sq(a - b)) - 2.0 * log(b)));
  const double q = 1.0 + sq(a - b);
  return num / (8.0 * p3 * (q * q * q) * b);
}

__device__ inline double sigma_ann(double k, double mV, double Zeff) {
  const double al = 1.0 / 137.0, me = 0.511e-3, pi = 3.141592653589793;
  const double lam = Zeff * al * me;
  const double a = me / lam;
  const double b = sq(mV) / (2.0 * k * lam);
  const double q = sq(a - b) + 1.0;
  const double pref = (sq(mV) + 2.0 * sq(me)) * (2.0 / 3.0 / lam / me) * (1.0 / sq(k));
  const double pre = (4.0 * pi * al) * (pref * (1.0 / (q * q * q)));
  const double s = 2.0 * k * me + sq(me);
  const double beta = 2.0 * al / pi * (log(s / sq(me)) - 1.0);
  const double tail = (4.0 * pi * al) * (pref * (beta / 2.0) * fancy(a, b));
  return pre + tail;
}

__device__ inline double comp_inner(double a, double b) {
  const double a2 = sq(a), a4 = a2 * a2;
  const double b2 = b * b, b3 = b2 * b, b4 = b3 * b, b5 = b4 * b;
  return -a * ((a2 + 1.0) * (a2 + 1.0) * (a2 + 1.0)) * (3.0 * a2 + 5.0) + 6.0 * (a4 + 4.0 * a2 - 5.0) * b5 -
         4.0 * a * (6.0 * a4 + 23.0 * a2 - 31.0) * b4 + (a - 1.0) * (a + 1.0) * (39.0 * a4 + 190.0 * a2 + 55.0) * b3 -
         a * (a2 + 1.0) * (33.0 * a4 + 98.0 * a2 - 111.0) * b2 + sq(a2 + 1.0) * (15.0 * a4 + 32.0 * a2 - 23.0) * b;
}

__device__ inline double comp_coeff(double a, double b) {
  const double a2 = sq(a), a4 = a2 * a2, a6 = a4 * a2;
  return 3.0 * sq(sq(a2 + 1.0)) - 2.0 * a * (3.0 * a4 + 10.0 * a2 + 15.0) * (a2 + 1.0) * b +
         6.0 * (a6 + 5.0 * a4 + 15.0 * a2 - 5.0) * sq(b);
}

// combine / expr1 / expr2 of atomic_compton.py
__device__ inline double comp_combine(double a, double b) {
  const double pi = 3.141592653589793;
  const double a2 = sq(a);
  const double d = sq(a - b) + 1.0;
  const double lt = log(d) - 2.0 * log(b);
  const double p4 = sq(sq(a2 + 1.0));
  if (a2 + 1.0 < a * b) {
    const double num = 8.0 * b * (a2 - 6.0 * a * b + 1.0) * lt + (a2 + 1.0) * comp_inner(a, b) / sq(d) +
                       comp_coeff(a, b) * (atan(b / (a2 - a * b + 1.0)) + atan(1.0 / a));
    double e1 = -num / (8.0 * p4 * b);
    if (e1 < 0.0 || (a2 + 1.0 < 0.01 * a * b)) {
      const double b2 = b * b, b4 = b2 * b2;
      e1 = (-145.0 + 1015.0 * a2 + 330.0 * a * b + 64.0 * b2) / (420.0 * (b4 * b4));
    }
    return e1;
  }
  const double c = comp_coeff(a, b);
  const double t1 = (a2 + 1.0) * -comp_inner(a, b);
  const double t2 = pi * c * sq(d);
  const double t3 = c * sq(d) * atan(1.0 / a);
  const double t4 = sq(d) * (8.0 * b * (a2 - 6.0 * a * b + 1.0) * lt + c * atan(b / (a2 - a * b + 1.0)));
  return (t1 + t2 - t3 - t4) / (8.0 * p4 * b * sq(d));
}

__device__ inline double sigma_comp(double k, double mV, double Zeff, double me, double al) {
  const double pi = 3.141592653589793;
  const double lam = Zeff * al * me;
  const double a = me / lam;
  const double b = sq(mV) / (2.0 * k * lam);
  const double s = 2.0 * k * me + sq(me);
  double beta = 2.0 * al / pi * (log(s / sq(me)) - 1.0);
  if (beta < 0.0) beta = 0.0;
  return (4.0 * pi * al) *
         ((sq(mV) + 2.0 * sq(me)) * (2.0 / 3.0 / lam / me) * (1.0 / sq(k)) * (beta / 4.0) * comp_combine(a, b));
}

// ------------------------------------------ tables at m_V for the yield --
// A slot holds the bound-electron tables at one mass (dark_shower.py:254-292
// at that m_V): the annihilation grid Ea (200 points, log-spaced from
// max(E_min(e+), E_res / 1000) to the top of the dark tables), its log-log
// n sigma and the weight array W(Ea) (filled by k_track_quad), and the Compton
// grid Ec with its log-log n sigma.
constexpr int kBoundGrid = 200;
constexpr int kBoundMaxSlots = 48;  // 16 masses x (m_V, m_V (1 + h), m_V (1 - h))

struct BoundSlot {
  double mV;
  double min_ann, min_comp;       // _minimum_calculable_dark_energy of DarkAnn, DarkComp (243-245)
  double lnEa0, inv_dln_a;        // guess of the Ea interval from ln E
  double l0c, inv_dlc;            // guess of the Ec interval from log10 E
};

// per slot: Ea, lx, ly, W, lxc, lyc (kBoundGrid each), then its panel nodes
constexpr int kBoundTab = 6 * kBoundGrid;

struct BoundParams {
  double Zeff, ne_conv, m_e, alpha, ann_floor, comp_floor;
  int32_t n_static;               // nodes of the survival grids
  int32_t node_stride;            // n_static + kBoundGrid
};

struct BoundPass {
  int32_t n_slots;
  double mV[kBoundMaxSlots];
  double top_ann[kBoundMaxSlots]; // last energy of the grids (the trained mass's tables)
  double top_comp[kBoundMaxSlots];
};

// numpy.logspace(lg0, lg1, n)[i] = 10 ** (i * step + lg0), last = 10 ** lg1
__device__ inline double logspace_at(double lg0, double lg1, int n, int i) {
  const double step = (lg1 - lg0) / (double)(n - 1);
  const double e = i == n - 1 ? lg1 : (double)i * step + lg0;
  return pow(10.0, e);
}

// One block per slot, one thread per grid point.
__global__ __launch_bounds__(256) void k_bound_tables(BoundParams B, BoundPass P, BoundSlot* __restrict__ slots,
                                                      double* __restrict__ tab, Job* __restrict__ jobs) {
  const int s = blockIdx.x;
  const int i = threadIdx.x;
  const double mV = P.mV[s];
  const double me = B.m_e;
  const double E_res = (sq(mV) - 2.0 * sq(me)) / (2.0 * me);
  const double E_C = sq(mV) / (2.0 * me) + mV;
  const double la0 = log10(fmax(B.ann_floor, 0.001 * E_res)), la1 = log10(P.top_ann[s]);
  const double lc0 = log10(fmax(B.comp_floor, 0.001 * E_C)), lc1 = log10(P.top_comp[s]);
  double* t = tab + (int64_t)s * (kBoundTab + B.node_stride);
  if (i < kBoundGrid) {
    const double Ea = logspace_at(la0, la1, kBoundGrid, i);
    const double na = B.ne_conv * sigma_ann(Ea, mV, B.Zeff);
    t[i] = Ea;
    t[kBoundGrid + i] = log10(Ea + 1e-20);
    t[2 * kBoundGrid + i] = log10(na + 1e-20);
    const double Ec = logspace_at(lc0, lc1, kBoundGrid, i);
    const double nc = B.ne_conv * sigma_comp(Ec, mV, B.Zeff, me, B.alpha);
    t[4 * kBoundGrid + i] = log10(Ec + 1e-20);
    t[5 * kBoundGrid + i] = log10(nc + 1e-20);
  }
  if (i == 0) {
    BoundSlot& S = slots[s];
    S.mV = mV;
    S.min_ann = E_res / 1000.0;
    S.min_comp = E_C / 1000.0;
    const double Ea0 = pow(10.0, la0), Ea1 = pow(10.0, la1);
    S.lnEa0 = log(Ea0);
    S.inv_dln_a = (double)(kBoundGrid - 1) / (log(Ea1) - log(Ea0));
    S.l0c = lc0;
    S.inv_dlc = (double)(kBoundGrid - 1) / (lc1 - lc0);
    jobs[s].lo = Ea0;
    // lx = log10(Ea + 1e-20) of a log-spaced grid: near uniform
    jobs[s].ns_l0 = log10(Ea0 + 1e-20);
    jobs[s].ns_inv = (double)(kBoundGrid - 1) / (log10(Ea1 + 1e-20) - log10(Ea0 + 1e-20));
  }
}

// Derivative tables of the pass's masses: for point k, the slots m_k (1 + h)
// (3k + 1) and m_k (1 - h) (3k + 2) become, in place in slot 3k + 1, the
// central differences d/dm_V of the grid energies, the weight array, and the
// Compton log-log table (nodes and values) at m_k.  With them the yield
// differentiates the interpolant of slot 3k itself: the node values and
// positions move with m_V, the interval weights follow analytically.
struct BoundDeriv {
  int32_t n_pts;
  int32_t stride;
  double inv_2h[16];  // 1 / (2 h m_k)
};

__global__ void k_bound_deriv(BoundDeriv D, double* __restrict__ tab) {
  const int k = blockIdx.x;
  const int i = threadIdx.x;
  if (k >= D.n_pts || i >= kBoundGrid) return;
  double* p = tab + (int64_t)(3 * k + 1) * D.stride;
  const double* m = tab + (int64_t)(3 * k + 2) * D.stride;
  const double f = D.inv_2h[k];
  for (const int a : {0, 3, 4, 5}) {  // Ea, W, lxc, lyc
    const int j = a * kBoundGrid + i;
    p[j] = (p[j] - m[j]) * f;
  }
}

// Panel nodes of each slot: the survival grids (`fixed`, ascending) merged
// with the nodes of the slot's log-log table (its grid energies Ea; 10 ** lx
// differs from them by rounding only), duplicates once.
__global__ void k_bound_nodes(BoundParams B, int n_slots, const double* __restrict__ fixed, double* __restrict__ tab,
                              Job* __restrict__ jobs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_slots) return;
  double* t = tab + (int64_t)s * (kBoundTab + B.node_stride);
  const double* Ea = t;
  double* out = t + kBoundTab;
  int i = 0, j = 0, n = 0;
  double last = -1.0;
  while (i < B.n_static || j < kBoundGrid) {
    double v;
    if (j >= kBoundGrid || (i < B.n_static && fixed[i] <= Ea[j])) {
      v = fixed[i++];
    } else {
      v = Ea[j++];
    }
    if (n == 0 || v != last) out[n++] = v;
    last = v;
  }
  jobs[s].n_nodes = n;
}

}  // namespace quad
