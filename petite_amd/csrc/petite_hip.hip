// libpetite_hip: batched shower transport of PETITE for MI355X (gfx950).
//
// Execution model: generation-synchronous sweeps over a device particle store.
// Sweep g transports the contiguous slot range of generation g (one lane per
// particle: substeps, dE/dx, MCS, process choice), the draw kernel runs the
// accept/reject of every requested hard interaction with lane-level work
// refill, and the emission kernel builds the children, appending them to the
// store as the next contiguous range.  Every particle owns a Philox stream, so
// results do not depend on launch shapes or on the order of appends.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "petite_hip.h"
#include "physics.h"
#include "transport.h"

using namespace petite;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) return fail(PETITE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

enum StatIdx { ST_STEPS = 0, ST_DRAWS, ST_TRIALS, ST_VIOL, ST_NOSAMPLE, ST_DROPPED, ST_OVERFLOW, ST_N = 8 };

struct DrawIO {
  // explicit requests (petite_vegas_draw)
  const int32_t* proc;
  const double* E;
  const uint64_t* sid;
  // shower pipeline requests: per-process regions of req_cap entries
  const int32_t* req_slot;
  const double* req_E;
  const uint64_t* store_sid;
  int64_t req_cap;
  int64_t seg[PETITE_NPROC + 1];
  // results (explicit: index r; pipeline: p * req_cap + i)
  float* x_out;
  int32_t* trials_out;
  float* w_out;
  int64_t total;
  unsigned long long* counter;
  unsigned long long* stats;
};

__device__ inline unsigned long long shfl_u64(unsigned long long v, int src) {
  const unsigned lo = __shfl((unsigned)(v & 0xFFFFFFFFull), src);
  const unsigned hi = __shfl((unsigned)(v >> 32), src);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ inline unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += shfl_u64(v, (threadIdx.x & 63) ^ off);
  return v;
}

__device__ inline void wave_stat(unsigned long long* stats, int idx, unsigned long long v) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(stats + idx, v);
}

// ------------------------------------------------------------ draw kernel --
// Lane-level refill: a lane that accepts (or exhausts max_trials) takes the
// next request at once, so every lane evaluates candidates until the request
// list is drained.  Candidate j of a request uses Philox blocks (sid, j,
// VEGAS_A/B), which makes the accepted point equal to the first accepted
// candidate of the sequential loop of draw_sample.
template <bool PIPE>
__global__ __launch_bounds__(256) void k_draw(DevTables T, DrawIO io, uint64_t seed) {
  const int lane = threadIdx.x & 63;
  bool have = false, done = false;
  int proc = 0, ie = 0, ndim = 1;
  double E = 0.0, mf = 0.0;
  uint64_t sid = 0;
  uint32_t j = 0;
  int64_t oidx = 0;
  unsigned long long n_trials = 0, n_viol = 0, n_fail = 0, n_draw = 0;
  const uint32_t max_trials = (uint32_t)(T.d.max_trials < (int64_t)0xFFFFFFF0 ? (T.d.max_trials > 0 ? T.d.max_trials : 1)
                                                                                : (int64_t)0xFFFFFFF0);
  while (true) {
    const bool need = !have && !done;
    const unsigned long long mask = __ballot(need);
    if (mask) {
      const int n = __popcll(mask);
      const int leader = __ffsll((long long)mask) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(io.counter, (unsigned long long)n);
      base = shfl_u64(base, leader);
      if (need) {
        const unsigned long long rank = __popcll(mask & ((1ull << lane) - 1ull));
        const int64_t r = (int64_t)(base + rank);
        if (r < io.total) {
          if (PIPE) {
            int p = 0;
            while (r >= io.seg[p + 1]) ++p;
            const int64_t i = r - io.seg[p];
            oidx = (int64_t)p * io.req_cap + i;
            proc = p;
            E = io.req_E[oidx];
            sid = io.store_sid[io.req_slot[oidx]];
          } else {
            oidx = r;
            proc = io.proc[r];
            E = io.E[r];
            sid = io.sid[r];
          }
          if (proc < 0 || proc >= PETITE_NPROC || T.d.grid[proc].n_E <= 0) {
            // no grid for this process: report a failed draw without sampling
            io.trials_out[oidx] = 0;
            io.w_out[oidx] = 0.0f;
            ++n_fail;
          } else {
            ndim = T.d.grid[proc].ndim;
            ie = lu_key(T, proc, E);
            mf = T.maxf[proc * T.maxE + ie] * T.d.maxF_fudge;
            j = 0;
            have = true;
          }
        } else {
          done = true;
        }
      }
    }
    if (__ballot(have) == 0ull) break;
    if (have) {
      const U4 a = draw_block(seed, sid, j, PURPOSE_VEGAS_A);
      const uint32_t rr[4] = {a.x, a.y, a.z, a.w};
      double ua;
      if (ndim == 4) {
        const U4 b = draw_block(seed, sid, j, PURPOSE_VEGAS_B);
        ua = u01d(b.x);
      } else {
        ua = u01d(a.w);
      }
      double x[4] = {0.0, 0.0, 0.0, 0.0};
      const double J = map_point(T, proc, ie, rr, ndim, x);
      const double w = J * integrand(T, proc, E, x);
      ++j;
      const bool acc = mf * ua < w;
      if (acc || j >= max_trials) {
        float* xo = io.x_out + 4 * oidx;
        xo[0] = (float)x[0];
        xo[1] = (float)x[1];
        xo[2] = (float)x[2];
        xo[3] = (float)x[3];
        io.trials_out[oidx] = acc ? (int32_t)j : -(int32_t)j;
        io.w_out[oidx] = acc ? (float)w : 0.0f;
        n_trials += j;
        n_draw += 1;
        if (acc && w > mf) ++n_viol;
        if (!acc) ++n_fail;
        have = false;
      }
    }
  }
  if (io.stats) {
    wave_stat(io.stats, ST_TRIALS, n_trials);
    wave_stat(io.stats, ST_DRAWS, n_draw);
    wave_stat(io.stats, ST_VIOL, n_viol);
    wave_stat(io.stats, ST_NOSAMPLE, n_fail);
  }
}

// ------------------------------------------------------------- max_F scan --
// One block per grid energy; every thread evaluates a strided share of the
// n_samples candidates at the grid energy (find_maxes.py:102-110).
__global__ __launch_bounds__(256) void k_maxf_scan(DevTables T, int proc, int n_samples, uint64_t seed,
                                                   double* maxf_out, double* mean_out) {
  const int ie = blockIdx.x;
  const PetiteGridDesc& g = T.d.grid[proc];
  const double E = T.f64[g.E_off + ie];
  const uint64_t sid = ((uint64_t)proc << 32) | (uint64_t)ie;
  double mx = 0.0, sum = 0.0;
  for (int j = threadIdx.x; j < n_samples; j += blockDim.x) {
    const U4 a = draw_block(seed, sid, (uint32_t)j, PURPOSE_VEGAS_A);
    const uint32_t rr[4] = {a.x, a.y, a.z, a.w};
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    const double J = map_point(T, proc, ie, rr, g.ndim, x);
    const double w = J * integrand(T, proc, E, x);
    mx = fmax(mx, w);
    sum += w;
  }
  __shared__ double smx[256];
  __shared__ double ssum[256];
  smx[threadIdx.x] = mx;
  ssum[threadIdx.x] = sum;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + s]);
      ssum[threadIdx.x] += ssum[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    maxf_out[ie] = smx[0];
    mean_out[ie] = ssum[0] / (double)n_samples;
  }
}

// ------------------------------------------------------- point evaluation --
// J*f at given Philox words: the map transform and integrand of one candidate.
__global__ __launch_bounds__(256) void k_eval(DevTables T, const int32_t* proc, const double* E, const uint32_t* words,
                                              int64_t n, double* x_out, double* w_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = proc[i];
  double x[4] = {0.0, 0.0, 0.0, 0.0};
  double w = 0.0;
  if (p >= 0 && p < PETITE_NPROC && T.d.grid[p].n_E > 0) {
    const uint32_t rr[4] = {words[4 * i], words[4 * i + 1], words[4 * i + 2], words[4 * i + 3]};
    const int ie = lu_key(T, p, E[i]);
    const double J = map_point(T, p, ie, rr, T.d.grid[p].ndim, x);
    w = J * integrand(T, p, E[i], x);
  }
  for (int d = 0; d < 4; ++d) x_out[4 * i + d] = x[d];
  w_out[i] = w;
}

// -------------------------------------------------------------- store -----
struct DevStore {
  float4* p0;
  float4* r0;
  float4* pf;
  float4* rf;
  int32_t* pid;
  int32_t* parent;
  int32_t* meta;
  int32_t* shower;
  uint64_t* sid;
  int64_t capacity;
};

__global__ void k_init_primaries(DevStore S, const PetitePrimary* prim, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PetitePrimary p = prim[i];
  S.p0[i] = make_float4((float)p.E, (float)p.px, (float)p.py, (float)p.pz);
  S.r0[i] = make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.weight);
  S.pf[i] = S.p0[i];
  S.rf[i] = make_float4((float)p.x, (float)p.y, (float)p.z, 0.0f);
  S.pid[i] = p.pid;
  S.parent[i] = -1;
  S.meta[i] = 0;
  S.shower[i] = p.shower;
  S.sid[i] = p.sid;
}

__device__ inline double min_calc(const DevTables& T, int pid) {
  switch (pid) {
    case 11: return T.d.min_calc[0];
    case -11: return T.d.min_calc[1];
    case 22: return T.d.min_calc[2];
    case 13: return T.d.min_calc[3];
    case -13: return T.d.min_calc[4];
    default: return 0.0;
  }
}

struct Requests {
  int32_t* slot;   // [PETITE_NPROC][cap]
  double* E;
  int64_t cap;
  unsigned int* count;  // [PETITE_NPROC]
};

// Lynch-Dahl scattering of direction (nx,ny,nz) for a step of t g/cm^2.
__device__ inline void mcs_rotate(const DevTables& T, float& nx, float& ny, float& nz, double E, double p, double m_part,
                                  double m_lep, double dz, const U4& u) {
  const float beta = (float)(p / E);
  const float pMeV = (float)(m_lep / 1e-3 * p / m_part);
  const float t = (float)(T.d.rho * (dz / kCmToM));
  const float th0 = lynch_dahl_theta0(t, pMeV, beta, (float)T.d.A, (float)T.d.Z);
  const float r = sqrtf(-2.0f * logf(u01f(u.x)));
  const float sign = u01f(u.y) < 0.5f ? -1.0f : 1.0f;
  const float theta = sign * r * th0 * (float)T.d.rescale_MCS;
  float sth, cth, sph, cph;
  sincosf(theta, &sth, &cth);
  sincosf(2.0f * (float)kPi * u01f(u.z), &sph, &cph);
  const Frame f = frame_of(nx, ny, nz);
  float ox, oy, oz;
  frame_apply(f, sph * sth, -cph * sth, cth, ox, oy, oz);
  const float inv = rsqrtf(ox * ox + oy * oy + oz * oz);
  nx = ox * inv;
  ny = oy * inv;
  nz = oz * inv;
}

// ---------------------------------------------------------- transport -----
// propagate_particle (shower.py:509-601) + the process choice (665-698) for
// the store slots [first, first + n).
__global__ __launch_bounds__(256) void k_transport(DevTables T, DevStore S, int64_t first, int64_t n, uint64_t seed,
                                                   Requests R, double* shower_energy, unsigned long long* stats) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  const int lane = threadIdx.x & 63;
  int want = -1;
  double E_req = 0.0;
  int64_t slot = first + i;
  unsigned long long steps = 0;
  if (valid) {
    const float4 p0 = S.p0[slot];
    const float4 r0 = S.r0[slot];
    const int pid = S.pid[slot];
    const uint64_t sid = S.sid[slot];
    const int shower = S.shower[slot];
    const double m = pid_mass(pid);
    const double pme = fmax(fmax(min_calc(T, pid), T.d.min_energy), m);
    double E = (double)p0.x;
    float nx = p0.y, ny = p0.z, nz = p0.w;
    {
      const float pn = sqrtf(nx * nx + ny * ny + nz * nz);
      if (pn > 0.0f) {
        nx /= pn;
        ny /= pn;
        nz /= pn;
      }
    }
    double p = sqrt(fmax(E * E - m * m, 0.0));
    float rx = r0.x, ry = r0.y, rz = r0.z;
    double dep = 0.0;
    if (E < pme) {
      // ends where it starts (shower.py:532-536)
    } else if (pid == 22) {
      const double mfp = mean_free_path(T, 22, E);
      const U4 u = draw_block(seed, sid, 0, PURPOSE_STEP);
      const double dist = -mfp * log1p(-u01d(u.x));
      rx += (float)(nx * dist);
      ry += (float)(ny * dist);
      rz += (float)(nz * dist);
      steps = 1;
    } else {
      const double dEdx = T.d.dEdx;
      uint32_t k = 0;
      bool hard = false;
      double dz = 0.0;
      U4 u;
      do {
        const double mfp = mean_free_path(T, pid, E);
        u = draw_block(seed, sid, k, PURPOSE_STEP);
        dz = mfp / (6.0 + 14.0 * u01d(u.y));
        if (u01d(u.x) > exp(-dz / mfp)) {
          hard = true;
        } else {
          const double Eb = E;
          E = E - dEdx * dz;
          if (E <= m) E = m;
          p = sqrt(E * E - m * m);
          dep += Eb - E;
          if (p > 0.0) {
            rx += (float)(nx * dz);
            ry += (float)(ny * dz);
            rz += (float)(nz * dz);
            if (T.d.MS) mcs_rotate(T, nx, ny, nz, E, p, m, m, dz, draw_block(seed, sid, k, PURPOSE_MCS));
          } else {
            nx = ny = nz = 0.0f;
          }
        }
        ++k;
      } while (!hard && E >= pme);
      const double distC = u01d(u.z);
      double last;
      if (E < pme) {
        last = distC * dz;
      } else {
        const double mfp = mean_free_path(T, pid, E);
        last = -mfp * log1p(expm1(-dz / mfp) * distC);
      }
      const double Eb = E;
      E = E - dEdx * last;
      if (E <= m) E = m;
      p = sqrt(E * E - m * m);
      dep += Eb - E;
      if (p > 0.0) {
        rx += (float)(nx * last);
        ry += (float)(ny * last);
        rz += (float)(nz * last);
        if (T.d.MS) mcs_rotate(T, nx, ny, nz, E, p, m, kMe, last, draw_block(seed, sid, 0, PURPOSE_MCS_CLOSE));
      } else {
        nx = ny = nz = 0.0f;
      }
      steps = k + 1;
    }
    S.pf[slot] = make_float4((float)E, (float)(p * nx), (float)(p * ny), (float)(p * nz));
    S.rf[slot] = make_float4(rx, ry, rz, (float)steps);
    // process choice (shower.py:665-698) and the threshold of sample_scattering (469-470)
    bool interacts = false;
    if (E > pme) {
      int ch[3];
      const int nch = channels(pid, ch);
      double w[3], SC = 0.0;
      for (int c = 0; c < nch; ++c) {
        w[c] = nsigma(T, ch[c], E);
        SC += w[c];
      }
      if (nch > 0 && SC > 0.0 && !isnan(SC)) {
        const double uu = u01d(draw_block(seed, sid, 0, PURPOSE_INTERACT).x) * SC;
        double c = 0.0;
        int pick = nch - 1;
        for (int q = 0; q < nch; ++q) {
          c += w[q];
          if (uu < c) {
            pick = q;
            break;
          }
        }
        want = ch[pick];
        E_req = E;
        interacts = true;
      }
    }
    if (shower_energy) {
      atomicAdd(shower_energy + 4 * shower + 0, dep);
      if (!interacts) atomicAdd(shower_energy + 4 * shower + 1, E);
    }
  }
  if (stats) wave_stat(stats, ST_STEPS, steps);
  // wave-aggregated appends to the per-process request lists
  for (int p = 0; p < PETITE_NPROC; ++p) {
    const unsigned long long mask = __ballot(want == p);
    if (!mask) continue;
    const int leader = __ffsll((long long)mask) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(R.count + p, (unsigned)__popcll(mask));
    base = __shfl(base, leader);
    if (want == p) {
      const unsigned idx = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
      R.slot[(int64_t)p * R.cap + idx] = (int32_t)slot;
      R.E[(int64_t)p * R.cap + idx] = E_req;
    }
  }
}

// -------------------------------------------------------------- emission --
// sample_scattering's kinematics, lab rotation and child bookkeeping
// (shower.py:467-507, 702-706) for every finished draw.
__global__ __launch_bounds__(256) void k_emit(DevTables T, DevStore S, DrawIO io, uint64_t seed,
                                              unsigned long long* store_count, double* shower_energy,
                                              unsigned long long* stats) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool keep[2] = {false, false};
  V4 lab[2];
  int cpid[2] = {0, 0};
  int64_t slot = 0;
  int proc = 0;
  if (r < io.total) {
    int p = 0;
    while (r >= io.seg[p + 1]) ++p;
    proc = p;
    const int64_t oidx = (int64_t)p * io.req_cap + (r - io.seg[p]);
    slot = io.req_slot[oidx];
    const double E = io.req_E[oidx];
    const int pid = S.pid[slot];
    const int shower = S.shower[slot];
    const int32_t tr = io.trials_out[oidx];
    if (tr <= 0) {
      if (shower_energy) atomicAdd(shower_energy + 4 * shower + 1, E);
    } else {
      const float* xs = io.x_out + 4 * oidx;
      const uint64_t sid = S.sid[slot];
      const double az = u01d(draw_block(seed, sid, 0, PURPOSE_INTERACT).y);
      const double m = pid_mass(pid);
      V4 a, b;
      switch (p) {
        case BREM: case MUONBREM: kin_brem(E, m, xs, az, a, b); break;
        case PAIRPROD: kin_pairprod(E, xs, az, a, b); break;
        case COMP: kin_compton(E, 0.0, (double)xs[0], az, a, b); break;
        case ANN: kin_ann(E, 0.0, (double)xs[0], az, a, b); break;
        case MUONE: kin_mue(E, (double)xs[0], az, a, b); break;
        default: kin_ee(E, (double)xs[0], az, a, b); break;  // Moller, Bhabha
      }
      child_pids(p, pid, &cpid[0], &cpid[1]);
      const float4 pf = S.pf[slot];
      double nx = pf.y, ny = pf.z, nz = pf.w;
      const double pn = sqrt(nx * nx + ny * ny + nz * nz);
      if (pn > 0.0) {
        nx /= pn;
        ny /= pn;
        nz /= pn;
      }
      const V4 in[2] = {a, b};
      for (int k = 0; k < 2; ++k) {
        double ox, oy, oz;
        frame_apply_d(nx, ny, nz, in[k], ox, oy, oz);
        lab[k] = V4{in[k].e, ox, oy, oz};
        keep[k] = lab[k].e > T.d.min_energy;
        if (!keep[k] && shower_energy) atomicAdd(shower_energy + 4 * shower + 2, lab[k].e);
      }
      if (shower_energy && (p == COMP || p == ANN || p == MOLLER || p == BHABHA || p == MUONE))
        atomicAdd(shower_energy + 4 * shower + 3, 1.0);
      if (stats) {
        const unsigned long long nd = (!keep[0]) + (!keep[1]);
        if (nd) atomicAdd(stats + ST_DROPPED, nd);
      }
    }
  }
  const int nk = (int)keep[0] + (int)keep[1];
  // wave-aggregated slot allocation
  unsigned long long incl = nk;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long v = shfl_u64(incl, lane >= off ? lane - off : lane);
    if (lane >= off) incl += v;
  }
  const unsigned long long total = shfl_u64(incl, 63);
  unsigned long long base = 0;
  if (lane == 63 && total) base = atomicAdd(store_count, total);
  base = shfl_u64(base, 63);
  unsigned long long dst = base + incl - nk;
  if (nk) {
    const int32_t pmeta = S.meta[slot];
    const int gen = (pmeta & 0xFFFF) + 1;
    const float4 rf = S.rf[slot];
    const float4 r0 = S.r0[slot];
    const uint64_t psid = S.sid[slot];
    const int shower = S.shower[slot];
    for (int k = 0; k < 2; ++k) {
      if (!keep[k]) continue;
      if ((int64_t)dst < S.capacity) {
        const int64_t c = (int64_t)dst;
        S.p0[c] = make_float4((float)lab[k].e, (float)lab[k].x, (float)lab[k].y, (float)lab[k].z);
        S.pf[c] = S.p0[c];
        S.r0[c] = make_float4(rf.x, rf.y, rf.z, r0.w);
        S.rf[c] = make_float4(rf.x, rf.y, rf.z, 0.0f);
        S.pid[c] = cpid[k];
        S.parent[c] = (int32_t)slot;
        S.meta[c] = (gen & 0xFFFF) | ((proc + 1) << 16) | (k << 24);
        S.shower[c] = shower;
        S.sid[c] = derive_sid(seed, psid, (uint32_t)k, PURPOSE_CHILD);
      } else if (stats) {
        atomicAdd(stats + ST_OVERFLOW, 1ull);
      }
      ++dst;
    }
  }
}

}  // namespace

// ===================================================================== C ABI
struct PetiteCtx {
  PetiteTableDesc desc;
  double* d64 = nullptr;
  float* d32 = nullptr;
  double* dmaxf = nullptr;
  int32_t maxE = 0;
  DevTables dev;
  // scratch
  int32_t* req_slot = nullptr;
  double* req_E = nullptr;
  float* res_x = nullptr;
  int32_t* res_trials = nullptr;
  float* res_w = nullptr;
  int64_t req_cap = 0;
  unsigned int* req_count = nullptr;
  unsigned long long* counters = nullptr;  // [0] refill, [1] store count, [8..8+ST_N) stats
  PetitePrimary* d_prim = nullptr;
  bool timing = false;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

static int ensure_requests(PetiteCtx* c, int64_t cap) {
  if (cap <= c->req_cap) return PETITE_OK;
  cap = std::max<int64_t>(cap, 1024);
  cap = std::max<int64_t>(cap, c->req_cap + c->req_cap / 2);
  if (c->req_slot) {
    (void)hipFree(c->req_slot);
    (void)hipFree(c->req_E);
    (void)hipFree(c->res_x);
    (void)hipFree(c->res_trials);
    (void)hipFree(c->res_w);
  }
  const int64_t n = cap * PETITE_NPROC;
  HIP_TRY(hipMalloc(&c->req_slot, n * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&c->req_E, n * sizeof(double)));
  HIP_TRY(hipMalloc(&c->res_x, n * 4 * sizeof(float)));
  HIP_TRY(hipMalloc(&c->res_trials, n * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&c->res_w, n * sizeof(float)));
  c->req_cap = cap;
  return PETITE_OK;
}

extern "C" {

const char* petite_last_error(void) { return g_err.c_str(); }

int petite_ctx_create(const PetiteTableDesc* desc, const double* blob64, int64_t n64, const float* blob32,
                      int64_t n32, PetiteCtx** out) {
  if (!desc || !blob64 || !blob32 || !out || n64 <= 0 || n32 <= 0) return fail(PETITE_ERR_ARG, "petite_ctx_create: bad arguments");
  for (int p = 0; p < PETITE_NPROC; ++p) {
    const PetiteGridDesc& g = desc->grid[p];
    if (g.n_E < 0 || g.ndim < 0 || g.ndim > 4) return fail(PETITE_ERR_ARG, "petite_ctx_create: bad grid descriptor");
    if (g.n_E > 0) {
      int64_t st = 0;
      for (int d = 0; d < g.ndim; ++d) {
        if (g.ninc[d] <= 0 || g.ninc[d] > (1 << 20)) return fail(PETITE_ERR_ARG, "petite_ctx_create: bad ninc");
        st += g.ninc[d] + 1;
      }
      if (st != g.stride || g.node_off < 0 || g.node_off + (int64_t)g.n_E * g.stride > n32 || g.E_off < 0 ||
          g.E_off + g.n_E > n64)
        return fail(PETITE_ERR_ARG, "petite_ctx_create: grid descriptor outside the blobs");
    }
    const PetiteSigmaDesc& s = desc->sigma[p];
    if (s.n < 0 || (s.n > 0 && (s.E_off < 0 || s.v_off < 0 || s.E_off + s.n > n64 || s.v_off + s.n > n64)))
      return fail(PETITE_ERR_ARG, "petite_ctx_create: sigma descriptor outside the blob");
  }
  PetiteCtx* c = new PetiteCtx();
  c->desc = *desc;
  int maxE = 1;
  for (int p = 0; p < PETITE_NPROC; ++p) maxE = std::max(maxE, desc->grid[p].n_E);
  c->maxE = maxE;
  auto cleanup = [&](int code) {
    petite_ctx_destroy(c);
    return code;
  };
  if (hipMalloc(&c->d64, n64 * sizeof(double)) != hipSuccess || hipMalloc(&c->d32, n32 * sizeof(float)) != hipSuccess ||
      hipMalloc(&c->dmaxf, (size_t)PETITE_NPROC * maxE * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->req_count, PETITE_NPROC * sizeof(unsigned int)) != hipSuccess ||
      hipMalloc(&c->counters, 32 * sizeof(unsigned long long)) != hipSuccess)
    return cleanup(fail(PETITE_ERR_HIP, "petite_ctx_create: hipMalloc failed"));
  if (hipMemcpy(c->d64, blob64, n64 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d32, blob32, n32 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c->dmaxf, 0, (size_t)PETITE_NPROC * maxE * sizeof(double)) != hipSuccess)
    return cleanup(fail(PETITE_ERR_HIP, "petite_ctx_create: upload failed"));
  DevTables& T = c->dev;
  T.f64 = c->d64;
  T.f32 = c->d32;
  T.maxf = c->dmaxf;
  T.maxE = maxE;
  T.d = *desc;
  const double Z = desc->Z;
  const double a0 = 184.15 * pow(2.718, -0.5) * pow(Z, -1.0 / 3.0) / kMe;
  const double a04 = a0 * a0 * a0 * a0;
  const double Ce = 8.0 / kPi * kAlpha * (kAlpha / kMe) * (kAlpha / kMe) * pow(kMe, 4) * Z * Z * a04;
  const double Cmu = 8.0 / kPi * kAlpha * (kAlpha / kMmu) * (kAlpha / kMmu) * pow(kMmu, 4) * Z * Z * a04;
  T.brem_C_e = (float)Ce;
  T.brem_C_mu = (float)Cmu;
  T.a0sq = (float)(a0 * a0);
  *out = c;
  return PETITE_OK;
}

void petite_ctx_destroy(PetiteCtx* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  for (hipEvent_t& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (void* p : {(void*)c->d64, (void*)c->d32, (void*)c->dmaxf, (void*)c->req_slot, (void*)c->req_E, (void*)c->res_x,
                  (void*)c->res_trials, (void*)c->res_w, (void*)c->req_count, (void*)c->counters, (void*)c->d_prim})
    if (p) (void)hipFree(p);
  delete c;
}

int petite_set_timing(PetiteCtx* c, int32_t on) {
  if (!c) return fail(PETITE_ERR_ARG, "petite_set_timing: null context");
  if (on && !c->ev[0])
    for (hipEvent_t& e : c->ev) HIP_TRY(hipEventCreate(&e));
  c->timing = on != 0;
  return PETITE_OK;
}

int petite_maxf_scan(PetiteCtx* c, int32_t proc, int32_t n_samples, uint64_t seed, double* maxf_host, double* mean_host,
                     PetiteStream stream) {
  if (!c || proc < 0 || proc >= PETITE_NPROC || n_samples <= 0) return fail(PETITE_ERR_ARG, "petite_maxf_scan: bad arguments");
  const PetiteGridDesc& g = c->desc.grid[proc];
  if (g.n_E <= 0) return fail(PETITE_ERR_ARG, "petite_maxf_scan: process has no grid");
  hipStream_t s = (hipStream_t)stream;
  double* dmean = nullptr;
  HIP_TRY(hipMalloc(&dmean, g.n_E * sizeof(double)));
  double* dmax = c->dmaxf + (size_t)proc * c->maxE;
  hipLaunchKernelGGL(k_maxf_scan, dim3(g.n_E), dim3(256), 0, s, c->dev, proc, n_samples, seed, dmax, dmean);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && maxf_host) e = hipMemcpyAsync(maxf_host, dmax, g.n_E * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && mean_host) e = hipMemcpyAsync(mean_host, dmean, g.n_E * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(dmean);
  if (e != hipSuccess) return fail(PETITE_ERR_HIP, std::string("petite_maxf_scan: ") + hipGetErrorString(e));
  return PETITE_OK;
}

int petite_set_maxf(PetiteCtx* c, int32_t proc, const double* maxf_host, int32_t n_E) {
  if (!c || proc < 0 || proc >= PETITE_NPROC || !maxf_host || n_E != c->desc.grid[proc].n_E)
    return fail(PETITE_ERR_ARG, "petite_set_maxf: bad arguments");
  HIP_TRY(hipMemcpy(c->dmaxf + (size_t)proc * c->maxE, maxf_host, n_E * sizeof(double), hipMemcpyHostToDevice));
  return PETITE_OK;
}

int petite_vegas_draw(PetiteCtx* c, const int32_t* proc, const double* E, const uint64_t* sid, int64_t n, uint64_t seed,
                      float* x_out, int32_t* trials_out, float* w_out, PetiteStream stream) {
  if (!c || n < 0 || (n > 0 && (!proc || !E || !sid || !x_out || !trials_out || !w_out)))
    return fail(PETITE_ERR_ARG, "petite_vegas_draw: bad arguments");
  if (n == 0) return PETITE_OK;
  hipStream_t s = (hipStream_t)stream;
  DrawIO io{};
  io.proc = proc;
  io.E = E;
  io.sid = sid;
  io.x_out = x_out;
  io.trials_out = trials_out;
  io.w_out = w_out;
  io.total = n;
  io.counter = c->counters;
  io.stats = c->counters + 8;
  HIP_TRY(hipMemsetAsync(c->counters, 0, 32 * sizeof(unsigned long long), s));
  const int64_t waves = std::min<int64_t>((n + 63) / 64, 256 * 16);
  const int blocks = (int)((waves + 3) / 4);
  hipLaunchKernelGGL(k_draw<false>, dim3(blocks), dim3(256), 0, s, c->dev, io, seed);
  HIP_TRY(hipGetLastError());
  unsigned long long st[ST_N];
  HIP_TRY(hipMemcpyAsync(st, c->counters + 8, sizeof(st), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (st[ST_NOSAMPLE]) return fail(PETITE_ERR_NO_SAMPLE, "petite_vegas_draw: No Sample Found within max_trials");
  return PETITE_OK;
}

int petite_eval_points(PetiteCtx* c, const int32_t* proc, const double* E, const uint32_t* words, int64_t n,
                       double* x_out, double* w_out, PetiteStream stream) {
  if (!c || n < 0 || (n > 0 && (!proc || !E || !words || !x_out || !w_out)))
    return fail(PETITE_ERR_ARG, "petite_eval_points: bad arguments");
  if (n == 0) return PETITE_OK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c->dev, proc, E, words, n, x_out, w_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return PETITE_OK;
}

int petite_run_showers(PetiteCtx* c, const PetitePrimary* prim, int64_t n, uint64_t seed, PetiteStore* store,
                       double* shower_energy, PetiteStats* stats, PetiteStream stream) {
  if (!c || !prim || n <= 0 || !store || !stats) return fail(PETITE_ERR_ARG, "petite_run_showers: bad arguments");
  if (store->capacity < n) return fail(PETITE_ERR_CAPACITY, "petite_run_showers: store smaller than the primaries");
  if (store->capacity > INT32_MAX) return fail(PETITE_ERR_ARG, "petite_run_showers: store capacity above 2^31-1");
  hipStream_t s = (hipStream_t)stream;
  DevStore S{(float4*)store->p0, (float4*)store->r0, (float4*)store->pf, (float4*)store->rf, store->pid, store->parent,
             store->meta, store->shower, store->sid, store->capacity};
  unsigned long long* refill = c->counters;
  unsigned long long* scount = c->counters + 1;
  unsigned long long* st = c->counters + 8;
  HIP_TRY(hipMemsetAsync(c->counters, 0, 32 * sizeof(unsigned long long), s));
  const unsigned long long n0 = (unsigned long long)n;
  HIP_TRY(hipMemcpyAsync(scount, &n0, sizeof(n0), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_init_primaries, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S, prim, n);
  HIP_TRY(hipGetLastError());
  int64_t first = 0, count = n, written = n, iters = 0;
  double t_tr = 0.0, t_dr = 0.0, t_em = 0.0;
  int64_t l_tr = 0, l_dr = 0, l_em = 0;
  std::memset(stats, 0, sizeof(*stats));
  while (count > 0) {
    int rc = ensure_requests(c, count);
    if (rc) return rc;
    Requests R{c->req_slot, c->req_E, c->req_cap, c->req_count};
    HIP_TRY(hipMemsetAsync(c->req_count, 0, PETITE_NPROC * sizeof(unsigned int), s));
    if (c->timing) HIP_TRY(hipEventRecord(c->ev[0], s));
    hipLaunchKernelGGL(k_transport, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, c->dev, S, first, count, seed,
                       R, shower_energy, st);
    HIP_TRY(hipGetLastError());
    if (c->timing) HIP_TRY(hipEventRecord(c->ev[1], s));
    unsigned int rq[PETITE_NPROC];
    HIP_TRY(hipMemcpyAsync(rq, c->req_count, sizeof(rq), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    DrawIO io{};
    io.req_slot = c->req_slot;
    io.req_E = c->req_E;
    io.store_sid = store->sid;
    io.req_cap = c->req_cap;
    io.seg[0] = 0;
    for (int p = 0; p < PETITE_NPROC; ++p) io.seg[p + 1] = io.seg[p] + rq[p];
    io.total = io.seg[PETITE_NPROC];
    io.x_out = c->res_x;
    io.trials_out = c->res_trials;
    io.w_out = c->res_w;
    io.counter = refill;
    io.stats = st;
    if (io.total > 0) {
      HIP_TRY(hipMemsetAsync(refill, 0, sizeof(unsigned long long), s));
      const int64_t waves = std::min<int64_t>((io.total + 63) / 64, 256 * 16);
      if (c->timing) HIP_TRY(hipEventRecord(c->ev[2], s));
      hipLaunchKernelGGL(k_draw<true>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, c->dev, io, seed);
      HIP_TRY(hipGetLastError());
      if (c->timing) HIP_TRY(hipEventRecord(c->ev[3], s));
      hipLaunchKernelGGL(k_emit, dim3((unsigned)((io.total + 255) / 256)), dim3(256), 0, s, c->dev, S, io, seed, scount,
                         shower_energy, st);
      HIP_TRY(hipGetLastError());
      if (c->timing) HIP_TRY(hipEventRecord(c->ev[4], s));
    }
    unsigned long long now = 0;
    HIP_TRY(hipMemcpyAsync(&now, scount, sizeof(now), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (c->timing) {
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
      t_tr += ms;
      ++l_tr;
      if (io.total > 0) {
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        t_dr += ms;
        ++l_dr;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[3], c->ev[4]));
        t_em += ms;
        ++l_em;
      }
    }
    ++iters;
    first = written;
    count = (int64_t)now - written;
    if ((int64_t)now > store->capacity) {
      written = (int64_t)now;
      break;
    }
    written = (int64_t)now;
  }
  unsigned long long hs[ST_N];
  HIP_TRY(hipMemcpyAsync(hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  stats->n_particles = written;
  stats->n_steps = (int64_t)hs[ST_STEPS];
  stats->n_draws = (int64_t)hs[ST_DRAWS];
  stats->n_trials = (int64_t)hs[ST_TRIALS];
  stats->n_maxf_viol = (int64_t)hs[ST_VIOL];
  stats->n_no_sample = (int64_t)hs[ST_NOSAMPLE];
  stats->n_iterations = iters;
  stats->n_dropped = (int64_t)hs[ST_DROPPED];
  stats->ms_transport = t_tr;
  stats->ms_draw = t_dr;
  stats->ms_emit = t_em;
  stats->launches_transport = l_tr;
  stats->launches_draw = l_dr;
  stats->launches_emit = l_em;
  if (written > store->capacity)
    return fail(PETITE_ERR_CAPACITY, "petite_run_showers: store capacity exceeded; n_particles holds the size reached");
  if (hs[ST_NOSAMPLE]) return fail(PETITE_ERR_NO_SAMPLE, "petite_run_showers: No Sample Found within max_trials");
  return PETITE_OK;
}

}  // extern "C"
