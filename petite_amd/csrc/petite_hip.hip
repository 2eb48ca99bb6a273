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
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "petite_hip.h"
#include "physics.h"
#include "sweep.h"
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

// ST_OVERFLOW: children past the store capacity; ST_REQ_OVERFLOW: draw requests
// past their process's region (impossible while a sweep holds at most as many
// particles as a region holds requests; counted as a backstop).
enum StatIdx { ST_STEPS = 0, ST_DRAWS, ST_TRIALS, ST_VIOL, ST_NOSAMPLE, ST_DROPPED, ST_OVERFLOW, ST_REQ_OVERFLOW, ST_N = 8 };

// Layout of the context's counter block (unsigned long long).  Atomics on one
// 128-byte line are served one after another, also on different addresses:
// the per-process draw refill counters, which the concurrent draw kernels hit
// once per 64 requests each, get a line each (measured: 3.4 ms of draws per
// step of the 100k workload against one shared line).
constexpr int kCtrLine = 16;  // counters per 128-byte line
enum CounterIdx {
  C_TRANSPORT = 16,           // transport refill counter
  C_STORE = 17,               // particles written to the store
  C_SWEEPS = 18,              // sweeps with work (device-driven showers)
  C_MUONS = 19,               // primaries that are muons (muons only come from primaries)
  C_DARK_N = 20,              // dark entries written in the current pass
  C_DARK_STATS = 21,          // [3] drate rows summing to zero, vectors written, bound-electron momentum failures
  C_MESON_N = 24,             // meson-decay entries written in the current pass
  C_YIELD_OUT = 26,           // particles the current sweep's transport handed on to the next sweep
  C_RPARITY = 27,             // which half of the resume buffer the current sweep reads
  C_SPILL = 28,               // [2] depth-first walk: entries in each spill list
  C_TAIL_STEPS = 30,          // particle-steps taken in depth-first walks
  C_STATS = 32,               // [ST_N] run statistics
  C_PTRIALS = 64,             // [PETITE_NPROC] candidates per process
  C_PDRAWS = 80,              // [PETITE_NPROC] draws per process
  C_PVIOL = 96,               // [PETITE_NPROC] accepted draws above max_F per process
  C_REFILL = 112,             // [PETITE_NPROC] per-process draw refill counters, kCtrLine apart
  C_N = C_REFILL + kCtrLine * PETITE_NPROC
};
__host__ __device__ constexpr int refill_ctr(int p) { return C_REFILL + kCtrLine * p; }

// One launch of the accept/reject kernel.  Requests sit in per-process
// regions (region p holds count[p] entries from p * req_cap); each block
// serves one process, so the process, its grid descriptor and its
// integrand are wave-uniform.
struct DrawLaunch {
  const int32_t* req_slot;   // store slot (pipeline) or request index (explicit API)
  const double* req_E;
  const uint64_t* sid_src;   // stream ids, indexed by req_slot
  const uint64_t* req_sid;   // non-null: stream ids beside the requests (used instead of sid_src)
  int64_t req_cap;
  int32_t count[PETITE_NPROC];
  int32_t block_start[PETITE_NPROC + 1];
  int32_t lgG[PETITE_NPROC];  // log2 of the lanes that share one request
  int32_t explicit_out;       // write results at index req_slot instead of the region index
  float* x_out;
  int32_t* trials_out;
  float* w_out;
  unsigned long long* counters;
  const unsigned int* dcount;  // non-null: counts on the device, plan made by each block
  int32_t full_blocks;         // blocks of a whole-GPU launch (device plan)
  int32_t min_lg;              // log2 of the fewest lanes per request
  int32_t group_requests;      // requests a lane group is sized to serve (small counts)
};

// Launch plan of one process's draw kernel: groups of 2^lg lanes share a
// request, at least 2^min_lg (default 4 lanes; measured on MI355X: 10 % faster
// than single lanes, whose refills - request loads, LU_Key search, max_F -
// come four times as often), more when there are fewer than `rpg` requests per
// group, sized so the GPU stays filled and each group sees about rpg requests.
__host__ __device__ inline void draw_plan(double cnt, int full_blocks, int min_lg, int rpg, int& lg, int& blocks) {
  const double full_lanes = 256.0 * full_blocks;
  lg = min_lg;
  blocks = 0;
  if (cnt <= 0.0) return;
  if (cnt >= (double)rpg * full_lanes) {
    blocks = full_blocks;
    return;
  }
  const double want = (double)rpg * full_lanes / cnt;
  while ((1 << lg) < want && lg < 6) ++lg;
  const double lanes = cnt * (double)(1 << lg) < full_lanes ? cnt * (double)(1 << lg) : full_lanes;
  blocks = (int)ceil(lanes / 256.0);
  if (blocks < 1) blocks = 1;
}

__device__ inline unsigned long long shfl_u64(unsigned long long v, int src) {
  const unsigned lo = __shfl((unsigned)(v & 0xFFFFFFFFull), src);
  const unsigned hi = __shfl((unsigned)(v >> 32), src);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ inline double shfl_f64(double v, int src) {
  return __longlong_as_double((long long)shfl_u64((unsigned long long)__double_as_longlong(v), src));
}

__device__ inline unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += shfl_u64(v, (threadIdx.x & 63) ^ off);
  return v;
}

__device__ inline void wave_stat(unsigned long long* ctr, int idx, unsigned long long v) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(ctr + idx, v);
}

// Adds v[k], summed over the block, to *dst[k]: one atomic per block and
// counter.  Every thread of the block calls it (it synchronises the block).
// Counters that every wave of a kernel updates when it ends would otherwise
// take one atomic per wave, and atomics on one address are served one after
// another: at the end of a large draw launch, 4,096 waves of them.
template <int N>
__device__ inline void block_stat_add(const unsigned long long (&v)[N], unsigned long long* const (&dst)[N]) {
  __shared__ unsigned long long s_part[N][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const unsigned long long w = wave_sum(v[k]);
    if (lane == 0) s_part[k][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    unsigned long long t = 0;
    for (int w = 0; w < nw; ++w) t += s_part[threadIdx.x][w];
    if (t) atomicAdd(dst[threadIdx.x], t);
  }
}

// Hands out request indices to the lanes in ``need`` from wave-private
// chunks of CH consecutive indices (one atomic per chunk).
template <int CH>
struct WaveQueue {
  int64_t cur = 0, end = 0;
  __device__ inline int64_t take(unsigned long long need_mask, bool need, unsigned long long* counter) {
    const int lane = threadIdx.x & 63;
    const int n = __popcll(need_mask);
    const int64_t avail = end - cur;
    unsigned long long nb = 0;
    if (avail < n) {
      if (lane == 0) nb = atomicAdd(counter, (unsigned long long)CH);
      nb = shfl_u64(nb, 0);
    }
    int64_t r = -1;
    if (need) {
      const int64_t rank = __popcll(need_mask & ((1ull << lane) - 1ull));
      r = rank < avail ? cur + rank : (int64_t)nb + (rank - avail);
    }
    if (avail < n) {
      cur = (int64_t)nb + (n - avail);
      end = (int64_t)nb + CH;
    } else {
      cur += n;
    }
    return r;
  }
};

template <int P>
__device__ inline double integrand_p(const DevTables& T, double E, const double* x) {
  if (P == BREM) return (double)f_brem(T, E, kMe, T.brem_C_e, x);
  if (P == MUONBREM) return (double)f_brem(T, E, kMmu, T.brem_C_mu, x);
  if (P == PAIRPROD) return (double)f_pairprod(T, E, x);
  if (P == COMP) return f_compton(E, 0.0, x[0]);
  if (P == ANN) return f_annihilation(E, 0.0, T.d->Eg_min, x[0]);
  if (P == MOLLER) return f_moller(E, T.d->Ee_min, x[0]);
  if (P == BHABHA) return f_bhabha(E, T.d->Ee_min, x[0]);
  if (P == MUONE) return f_muone(E, T.d->Ee_min, x[0]);
  if (P == DARKBREM) return f_dark_brem(E, kMe, T.d->mV, T.d->mT, T.d->Z, T.d->A, x);
  if (P == DARKMUONBREM) return f_dark_brem(E, kMmu, T.d->mV, T.d->mT, T.d->Z, T.d->A, x);
  if (P == DARKANN) return f_dark_ann(E, T.d->mV, x[0]);
  if (P == DARKCOMP) return f_compton(E, T.d->mV, x[0]);
  return 0.0;
}

template <int P>
constexpr int proc_ndim() {
  return (P == BREM || P == MUONBREM || P == PAIRPROD) ? 4 : ((P == DARKBREM || P == DARKMUONBREM) ? 3 : 1);
}

// False when the integrand is NaN at every point, so that draw_sample's loop
// can only end without a sample: the radiative-return width factor
// sqrt(1 - 4 m_e^2 / m_V^2) (all_processes.py:431) for m_V <= 2 m_e.
template <int P>
__device__ inline bool draw_possible(const DevTables& T) {
  if (P == DARKANN) return T.d->mV > 2.0 * kMe;
  return true;
}

// The float64 reference-order integrand of a 4-D process (see f_brem_ref).
template <int P>
__device__ inline double integrand_ref(const DevTables& T, double E, const double* x) {
  if (P == BREM) return f_brem_ref(T, E, kMe, x[0], x[1], x[2], x[3]);
  if (P == MUONBREM) return f_brem_ref(T, E, kMmu, x[0], x[1], x[2], x[3]);
  if (P == PAIRPROD) return f_pairprod_ref(T, E, x[0], x[1], x[2], x[3]);
  return integrand_p<P>(T, E, x);
}

// Relative band around the accept threshold max_F * u in which a float32
// J*f of a 4-D process is re-evaluated in float64: the float32 bodies agree
// with the float64 reference formula to 5e-6 (99th percentile) and 2e-5
// (99.9th), so outside the band the two make the same decision.
constexpr double kRefBand = 4.0e-4;

template <int P, int ND>
__device__ inline double candidate_w(const DevTables& T, double E, const float* base, const int32_t* ninc,
                                     uint64_t seed, uint64_t sid, uint32_t j, double* x, double& ua, double& J) {
  const U4 a = draw_block(seed, sid, j, PURPOSE_VEGAS_A);
  const uint32_t rr[4] = {a.x, a.y, a.z, a.w};
  if (ND == 4) {
    // the map takes the high 23 bits of the four words; the accept uniform is
    // made of their low bits (9 of x, 9 of y, 5 of z), independent of them
    ua = u01d(((a.x & 0x1FFu) | ((a.y & 0x1FFu) << 9) | ((a.z & 0x1Fu) << 18)) << 9);
  } else {
    ua = u01d(a.w);
  }
  J = map_point_n<ND>(base, ninc, rr, x);
  return J * integrand_p<P>(T, E, x);
}

// ------------------------------------------------------------ draw kernel --
// Groups of G = 2^lgG lanes work on one request: lane g of a group tests
// candidate j + g, and the group takes the lowest accepted index, which is
// the candidate the sequential loop of draw_sample accepts (candidate j
// always uses Philox counter j).  The result does not depend on G; a group
// refills less often than a single lane, and larger G cuts the latency of
// sweeps with fewer requests than lanes.
template <int p>
__device__ void draw_loop(const DevTables& T, const DrawLaunch& io, uint64_t seed, const int64_t count, const int lgG) {
  constexpr int ND = proc_ndim<p>();
  const int lane = threadIdx.x & 63;
  const int G = 1 << lgG;
  const int gl = lane & (G - 1);
  const int g0 = lane - gl;
  const unsigned long long gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << g0);
  const PetiteGridDesc& gd = T.d->grid[p];
  const int32_t ninc[4] = {gd.ninc[0], gd.ninc[1], gd.ninc[2], gd.ninc[3]};
  const uint32_t max_trials = (uint32_t)(T.d->max_trials < (int64_t)0x7FFFFFF0 ? (T.d->max_trials > 0 ? T.d->max_trials : 1)
                                                                               : (int64_t)0x7FFFFFF0);
  WaveQueue<64> q;
  bool have = false, done = false;
  double E = 0.0, mf = 0.0;
  uint64_t sid = 0;
  const float* base = T.f32;
  uint32_t j = 0;
  int64_t oidx = 0;
  unsigned long long n_trials = 0, n_viol = 0, n_fail = 0, n_draw = 0;
  while (true) {
    // refill once half of the groups are free (or none has a request): the
    // refill's loads and table searches cost a whole-wave pass each time
    const bool need = (gl == 0) && !have && !done;
    const unsigned long long nm = __ballot(need);
    if (nm && (__popcll(nm) >= (32 >> lgG) || __ballot(have) == 0ull)) {
      const int64_t r = q.take(nm, need, io.counters + refill_ctr(p));
      if (need) {
        if (r < count) {
          const int64_t ridx = (int64_t)p * io.req_cap + r;
          const int32_t src = io.req_slot[ridx];
          oidx = io.explicit_out ? (int64_t)src : ridx;
          E = io.req_E[ridx];
          sid = io.req_sid ? io.req_sid[ridx] : io.sid_src[src];
          const int ie = lu_key(T, p, E);
          base = T.f32 + gd.node_off + (int64_t)ie * gd.stride;
          mf = T.maxf[p * T.maxE + ie] * T.d->maxF_fudge;
          // a draw that cannot accept ends after its first round, as the
          // reference's loop would after max_trials candidates
          j = draw_possible<p>(T) ? 0u : max_trials;
          have = true;
        } else {
          done = true;
        }
      }
      if (G > 1) {
        // the group leader's request becomes the group's
        have = __shfl((int)have, g0) != 0;
        done = __shfl((int)done, g0) != 0;
        E = shfl_f64(E, g0);
        mf = shfl_f64(mf, g0);
        sid = shfl_u64(sid, g0);
        oidx = (int64_t)shfl_u64((unsigned long long)oidx, g0);
        base = (const float*)shfl_u64((unsigned long long)base, g0);
        j = __shfl(j, g0);
      }
    }
    if (__ballot(have) == 0ull) break;
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    double w = 0.0, ua = 1.0, J = 0.0;
    const uint32_t jc = j + (uint32_t)gl;
    if (have) {
      w = candidate_w<p, ND>(T, E, base, ninc, seed, sid, jc, x, ua, J);
      // near the threshold the decision is the float64 reference formula's
      if (ND == 4 && fabs(mf * ua - w) <= kRefBand * w) w = J * integrand_ref<p>(T, E, x);
    }
    const bool acc = have && (jc < max_trials) && (mf * ua < w);
    const unsigned long long am = __ballot(acc) & gmask;
    if (have) {
      if (am) {
        const int win = __ffsll((long long)am) - 1;
        if (lane == win) {
          float* xo = io.x_out + 4 * oidx;
          xo[0] = (float)x[0];
          xo[1] = (float)x[1];
          xo[2] = (float)x[2];
          xo[3] = (float)x[3];
          io.trials_out[oidx] = (int32_t)(jc + 1);
          io.w_out[oidx] = (float)w;
          n_trials += jc + 1;
          n_draw += 1;
          if (w > mf) ++n_viol;
        }
        have = false;
      } else {
        j += (uint32_t)G;
        if (j >= max_trials) {
          if (gl == 0) {
            io.trials_out[oidx] = -(int32_t)max_trials;
            io.w_out[oidx] = 0.0f;
            n_trials += max_trials;
            n_draw += 1;
            ++n_fail;
          }
          have = false;
        }
      }
    }
  }
  // every block of a draw launch calls draw_loop (a block-uniform choice), so
  // the whole block reaches this point
  unsigned long long* const ctr = io.counters;
  const unsigned long long v[7] = {n_trials, n_draw, n_viol, n_fail, n_trials, n_draw, n_viol};
  unsigned long long* const dst[7] = {ctr + C_STATS + ST_TRIALS, ctr + C_STATS + ST_DRAWS, ctr + C_STATS + ST_VIOL,
                                      ctr + C_STATS + ST_NOSAMPLE, ctr + C_PTRIALS + p, ctr + C_PDRAWS + p,
                                      ctr + C_PVIOL + p};
  block_stat_add<7>(v, dst);
}

// One kernel per process: the integrand, the map dimension and the
// register allocation are specialised at compile time.  Launch p uses the
// blocks [block_start[p], block_start[p+1]) of the plan, renumbered from 0.
// The 4-D draws need 129-132 VGPRs unconstrained (3 waves per SIMD); held to
// 128 (4 waves, no spills) they hide more of their map gathers' latency: the
// Brem draw stage is 4 % faster.  Five waves (96 VGPRs) spill and are slower.
// (k_draw_small, which holds every process's draw, spills at 128: it keeps
// its 132.)
template <int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_draw(DevTables T, DrawLaunch io, uint64_t seed) {
  if (io.dcount) {
    // a count past the region (a backstop: the transport never appends one)
    // is cut to the region; the overflow is reported by the run
    const int64_t cnt = min((int64_t)io.dcount[P], io.req_cap);
    int lg, blocks;
    draw_plan((double)cnt, io.full_blocks, io.min_lg, io.group_requests, lg, blocks);
    if ((int)blockIdx.x >= blocks) return;
    draw_loop<P>(T, io, seed, cnt, lg);
  } else {
    if (io.count[P] <= 0) return;
    draw_loop<P>(T, io, seed, io.count[P], io.lgG[P]);
  }
}

typedef void (*DrawKernel)(DevTables, DrawLaunch, uint64_t);
static const DrawKernel kDrawKernels[PETITE_NPROC] = {
    k_draw<0>, k_draw<1>, k_draw<2>, k_draw<3>, k_draw<4>, k_draw<5>,
    k_draw<6>, k_draw<7>, k_draw<8>, k_draw<9>, k_draw<10>, k_draw<11>};

// One request per block: the 256 threads test candidates j .. j + 255 and
// the block takes the lowest accepted index (an LDS minimum), the candidate
// draw_sample's sequential loop accepts.  In the last sweeps of a run, where
// a process has fewer requests than blocks, a draw's latency is its slowest
// request's candidate count over the lanes that share it: 256 instead of at
// most a wave's 64.
template <int p>
__device__ void draw_block_one(const DevTables& T, const DrawLaunch& io, uint64_t seed, int64_t r) {
  constexpr int ND = proc_ndim<p>();
  __shared__ uint32_t s_win;
  const PetiteGridDesc& gd = T.d->grid[p];
  const int32_t ninc[4] = {gd.ninc[0], gd.ninc[1], gd.ninc[2], gd.ninc[3]};
  const uint32_t max_trials = (uint32_t)(T.d->max_trials < (int64_t)0x7FFFFFF0 ? (T.d->max_trials > 0 ? T.d->max_trials : 1)
                                                                               : (int64_t)0x7FFFFFF0);
  const int64_t ridx = (int64_t)p * io.req_cap + r;
  const int32_t src = io.req_slot[ridx];
  const int64_t oidx = io.explicit_out ? (int64_t)src : ridx;
  const double E = io.req_E[ridx];
  const uint64_t sid = io.req_sid ? io.req_sid[ridx] : io.sid_src[src];
  const int ie = lu_key(T, p, E);
  const float* base = T.f32 + gd.node_off + (int64_t)ie * gd.stride;
  const double mf = T.maxf[p * T.maxE + ie] * T.d->maxF_fudge;
  for (uint32_t j = draw_possible<p>(T) ? 0u : max_trials; j < max_trials; j += blockDim.x) {
    if (threadIdx.x == 0) s_win = 0xFFFFFFFFu;
    __syncthreads();
    const uint32_t jc = j + threadIdx.x;
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    double ua = 1.0, J = 0.0;
    double w = candidate_w<p, ND>(T, E, base, ninc, seed, sid, jc, x, ua, J);
    if (ND == 4 && fabs(mf * ua - w) <= kRefBand * w) w = J * integrand_ref<p>(T, E, x);
    const bool acc = (jc < max_trials) && (mf * ua < w);
    if (acc) atomicMin(&s_win, jc);
    __syncthreads();
    const uint32_t win = s_win;
    if (win != 0xFFFFFFFFu) {
      if (jc == win) {
        float* xo = io.x_out + 4 * oidx;
        xo[0] = (float)x[0];
        xo[1] = (float)x[1];
        xo[2] = (float)x[2];
        xo[3] = (float)x[3];
        io.trials_out[oidx] = (int32_t)(jc + 1);
        io.w_out[oidx] = (float)w;
        atomicAdd(io.counters + C_STATS + ST_TRIALS, (unsigned long long)(jc + 1));
        atomicAdd(io.counters + C_STATS + ST_DRAWS, 1ull);
        atomicAdd(io.counters + C_PTRIALS + p, (unsigned long long)(jc + 1));
        atomicAdd(io.counters + C_PDRAWS + p, 1ull);
        if (w > mf) {
          atomicAdd(io.counters + C_STATS + ST_VIOL, 1ull);
          atomicAdd(io.counters + C_PVIOL + p, 1ull);
        }
      }
      return;
    }
    __syncthreads();  // every thread has read s_win before the next round resets it
  }
  if (threadIdx.x == 0) {
    io.trials_out[oidx] = -(int32_t)max_trials;
    io.w_out[oidx] = 0.0f;
    atomicAdd(io.counters + C_STATS + ST_TRIALS, (unsigned long long)max_trials);
    atomicAdd(io.counters + C_STATS + ST_DRAWS, 1ull);
    atomicAdd(io.counters + C_STATS + ST_NOSAMPLE, 1ull);
    atomicAdd(io.counters + C_PTRIALS + p, (unsigned long long)max_trials);
    atomicAdd(io.counters + C_PDRAWS + p, 1ull);
  }
}

// The draws of a small sweep: every SM process in one launch on the
// caller's stream, blocks [p * per, (p + 1) * per) for process p (bit p of
// `procs` set); a process with at most `per` requests takes one block per
// request (draw_block_one).  In the last sweeps of a run, where a handful of requests
// remain, the per-process launches on their own streams cost more in
// launches, forks and joins than the draws themselves.  Same draw loop, so
// the same results.
__global__ __launch_bounds__(256) void k_draw_small(DevTables T, DrawLaunch io, uint64_t seed, int per, uint32_t procs) {
  const int p = (int)blockIdx.x / per;
  const int b = (int)blockIdx.x - p * per;
  if (p >= 8 || !((procs >> p) & 1u)) return;
  const int64_t cnt = min((int64_t)io.dcount[p], io.req_cap);
  if (cnt <= (int64_t)per) {
    if (b >= cnt) return;
    switch (p) {
      case 0: draw_block_one<0>(T, io, seed, b); break;
      case 1: draw_block_one<1>(T, io, seed, b); break;
      case 2: draw_block_one<2>(T, io, seed, b); break;
      case 3: draw_block_one<3>(T, io, seed, b); break;
      case 4: draw_block_one<4>(T, io, seed, b); break;
      case 5: draw_block_one<5>(T, io, seed, b); break;
      case 6: draw_block_one<6>(T, io, seed, b); break;
      default: draw_block_one<7>(T, io, seed, b); break;
    }
    return;
  }
  int lg, blocks;
  draw_plan((double)cnt, per, io.min_lg, io.group_requests, lg, blocks);
  if (b >= blocks) return;
  switch (p) {
    case 0: draw_loop<0>(T, io, seed, cnt, lg); break;
    case 1: draw_loop<1>(T, io, seed, cnt, lg); break;
    case 2: draw_loop<2>(T, io, seed, cnt, lg); break;
    case 3: draw_loop<3>(T, io, seed, cnt, lg); break;
    case 4: draw_loop<4>(T, io, seed, cnt, lg); break;
    case 5: draw_loop<5>(T, io, seed, cnt, lg); break;
    case 6: draw_loop<6>(T, io, seed, cnt, lg); break;
    default: draw_loop<7>(T, io, seed, cnt, lg); break;
  }
}

// Explicit requests into per-process regions (petite_vegas_draw).
__global__ void k_bucket(DevTables T, const int32_t* proc, const double* E, int64_t n, int32_t* req_slot, double* req_E,
                         int64_t req_cap, unsigned int* req_count, int32_t* trials_out, float* w_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = proc[i];
  if (p < 0 || p >= PETITE_NPROC || T.d->grid[p].n_E <= 0) {
    trials_out[i] = 0;
    w_out[i] = 0.0f;
    return;
  }
  const unsigned k = atomicAdd(req_count + p, 1u);
  req_slot[(int64_t)p * req_cap + k] = (int32_t)i;
  req_E[(int64_t)p * req_cap + k] = E[i];
}

// ----------------------------------------------------------- map training --
// VEGAS adaptation data (the reference trains its maps with the vegas package,
// utilities/generate_integrators.py:49-100, vegas_integrator_options at
// all_processes.py:1085-1096; published algorithm: Lepage 1978, 2021).  At each
// grid energy, samples uniform in y pass through the current map; per
// increment of every dimension the kernel sums (J f)^2 and counts the samples.
// The increment index is the map's integer-exact one.  Block (ie, chunk).
__global__ __launch_bounds__(256) void k_map_accumulate(DevTables T, int proc, int64_t n_per, uint64_t seed,
                                                        uint32_t iteration, double* sum_f, uint32_t* n_f) {
  const int ie = blockIdx.x;
  const PetiteGridDesc& g = T.d->grid[proc];
  const double E = T.f64[g.E_off + ie];
  const uint64_t sid = ((uint64_t)proc << 32) | (uint64_t)ie;
  const int64_t step = (int64_t)gridDim.y * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; k < n_per; k += step) {
    const uint32_t j = (uint32_t)((uint64_t)iteration * (uint64_t)n_per + (uint64_t)k);
    const U4 a = draw_block(seed, sid, j, PURPOSE_MAP_TRAIN);
    const uint32_t rr[4] = {a.x, a.y, a.z, a.w};
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    const double J = map_point(T, proc, ie, rr, g.ndim, x);
    const double f = J * integrand(T, proc, E, x);
    const double f2 = f * f;
    if (!isfinite(f2)) continue;
    int64_t off = (int64_t)ie * g.stride;
    for (int d = 0; d < g.ndim; ++d) {
      const uint32_t N = (uint32_t)g.ninc[d];
      const uint64_t q = (uint64_t)(2u * (rr[d] >> 9) + 1u) * (uint64_t)N;
      const int64_t i = off + (int64_t)(q >> 24);
      unsafeAtomicAdd(sum_f + i, f2);
      atomicAdd(n_f + i, 1u);
      off += N + 1;
    }
  }
}

// ------------------------------------------------------------- max_F scan --
// One block per grid energy; every thread evaluates a strided share of the
// n_samples candidates at the grid energy (find_maxes.py:102-110).
__global__ __launch_bounds__(256) void k_maxf_scan(DevTables T, int proc, int n_samples, uint64_t seed,
                                                   double* maxf_out, double* mean_out) {
  const int ie = blockIdx.x;
  const PetiteGridDesc& g = T.d->grid[proc];
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
  if (p >= 0 && p < PETITE_NPROC && T.d->grid[p].n_E > 0) {
    const uint32_t rr[4] = {words[4 * i], words[4 * i + 1], words[4 * i + 2], words[4 * i + 3]};
    const int ie = lu_key(T, p, E[i]);
    const double J = map_point(T, p, ie, rr, T.d->grid[p].ndim, x);
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

__global__ void k_init_primaries(DevStore S, const PetitePrimary* prim, int64_t n, unsigned long long* n_muons) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PetitePrimary p = prim[i];
  if ((p.pid == 13 || p.pid == -13) && n_muons) atomicAdd(n_muons, 1ull);
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
    case 11: return T.d->min_calc[0];
    case -11: return T.d->min_calc[1];
    case 22: return T.d->min_calc[2];
    case 13: return T.d->min_calc[3];
    case -13: return T.d->min_calc[4];
    default: return 0.0;
  }
}

struct Requests {
  int32_t* slot;   // [PETITE_NPROC][cap]
  double* E;
  int64_t cap;
  unsigned int* count;  // [PETITE_NPROC]
};

// What the draw and the emission need of the parent, written beside the
// request by the transport ([PETITE_NPROC_SM][cap], request order), so that
// neither kernel gathers from the store: requests are appended in completion
// order, which scatters their slots over the sweep's range.
struct ParentRecs {
  uint64_t* sid;
  float4* pf;    // final momentum (E, px, py, pz), as written to the store
  float4* pos;   // final position xyz, weight
  int2* info;    // (generation << 16 | pid & 0xFFFF, shower)
};
constexpr int kNprocSM = 8;

__device__ inline int info_pid(int2 v) { return (int)(int16_t)(v.x & 0xFFFF); }
__device__ inline int info_gen(int2 v) { return (v.x >> 16) & 0xFFFF; }

// n*sigma tables in LDS.  Header per process p (kLdsHeader doubles):
//   [0] float log2 E0 (low word), float nodes per log2 unit (high word; 0:
//       no node index, binary search);
//   [1] int offset of interval record -1 (low word), int n - 1 (high word);
//   [2] E_first, [3] E_last (an empty range for an empty table);
//   [4] the value of a one-point table.
// Interval i (0 <= i <= n-2) is one 32-byte record {x_i, x_i+1, y_i,
// (y_i+1 - y_i)/(x_i+1 - x_i)}; the value is slope * (E - x_i) + y_i, the
// arithmetic of scipy's interp1d (0 outside the table).  Record -1 (a copy of
// record 0) and record n-1 (a copy of record n-2) pad the interval list, so
// the node-indexed lookup reads the two records around a node without bounds
// checks (see LdsSigma::nsigma).
constexpr int kLdsHeader = 8;

// Channels of a charged species' n*sigma sum, in the order of its terms
// (shower.py:355-366), 4 bits each, 0xF = none; the sum is 0 without channels.
__device__ inline uint32_t sum_channels(int pid) {
  switch (pid) {
    case 11: return BREM | (MOLLER << 4) | (0xFu << 8);
    case -11: return BREM | (BHABHA << 4) | (ANN << 8);
    case 13: case -13: return MUONBREM | (MUONE << 4) | (0xFu << 8);
    default: return 0xFFFu;
  }
}

struct LdsSigma {
  const double* base;
  // l2 = log2f(E), shared by the lookups of one energy.  On a geometric grid
  // the node nearest E is j = round((l2 - log2 E0) * nodes per log2 unit)
  // (the host checks that every node lies within 0.4 of its index, and the
  // float32 index is off by less than 1e-4, so E lies between nodes j-1 and
  // j+1); the interval is then record j if E >= x_j, else record j-1: the
  // largest i <= n-2 with x_i <= E, as a search finds it, from one header
  // and one 64-byte record read and no branch.
  __device__ inline double nsigma(int p, double E, float l2) const {
    const double* h = base + kLdsHeader * p;
    const long long w0 = __double_as_longlong(h[0]), w1 = __double_as_longlong(h[1]);
    const float log2E0 = __int_as_float((int)w0), per = __int_as_float((int)(w0 >> 32));
    const double* rm1 = base + (int)w1;  // record -1
    const int nm1 = (int)(w1 >> 32);     // n - 1
    const bool in = (E >= h[2]) & (E <= h[3]);
    double v;
    if (per > 0.0f) {
      int j = (int)((l2 - log2E0) * per + 0.5f);
      j = min(max(j, 0), nm1);
      const double* r = rm1 + 4 * j;
      const double va = r[3] * (E - r[0]) + r[2];
      const double vb = r[7] * (E - r[4]) + r[6];
      v = E >= r[4] ? vb : va;
    } else if (nm1 <= 0) {
      v = nm1 == 0 ? h[4] : 0.0;  // a one-point table, or none
    } else {
      // a grid that is not geometric: binary search for the same interval
      const double* iv = rm1 + 4;
      int lo = 0, hi = nm1;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (iv[4 * mid] <= E) lo = mid; else hi = mid;
      }
      const double* r = iv + 4 * lo;
      v = r[3] * (E - r[0]) + r[2];
    }
    return in ? v : 0.0;
  }
  // _NSigmaElectron/Photon/Positron/Muon (shower.py:355-366), terms in their order
  __device__ inline double nsigma_sum(int pid, double E) const {
    const float lf = __log2f((float)E);
    switch (pid) {
      case 11: return nsigma(BREM, E, lf) + nsigma(MOLLER, E, lf);
      case -11: return nsigma(BREM, E, lf) + nsigma(BHABHA, E, lf) + nsigma(ANN, E, lf);
      case 22: return nsigma(PAIRPROD, E, lf) + nsigma(COMP, E, lf);
      case 13: case -13: return nsigma(MUONBREM, E, lf) + nsigma(MUONE, E, lf);
      default: return 0.0;
    }
  }
  // The same sums for a lane's charged species from its channel word (see
  // sum_channels): every lane runs the same two lookups, positrons a third,
  // so a wave that mixes species does not run each species' sum in turn.
  __device__ inline double nsigma_sum_ch(uint32_t ch, double E) const {
    if ((ch & 0xFu) == 0xFu) return 0.0;
    const float lf = __log2f((float)E);
    double s = nsigma((int)(ch & 0xFu), E, lf) + nsigma((int)((ch >> 4) & 0xFu), E, lf);
    const uint32_t c2 = (ch >> 8) & 0xFu;
    if (c2 != 0xFu) s += nsigma((int)c2, E, lf);
    return s;
  }
  __device__ inline double mfp(int pid, double E) const {
    const double s = nsigma_sum(pid, E);
    return s > 0.0 ? kCmToM / s : 1.0e12;
  }
};

// Lynch-Dahl scattering of the unit direction (nx,ny,nz) over ``dz`` metres:
// radius and azimuth from two Philox words, the sign from bit 0 of the
// azimuth word (moliere.py:265-284, 350-400).
__device__ inline void mcs_rotate(const DevTables& T, float& nx, float& ny, float& nz, double E, float fp, double m_part,
                                  double m_lep, double dz, uint32_t w_r, uint32_t w_a) {
  // float32 quotients by the hardware reciprocal: the angle is a float32
  // quantity (a float64 division per factor costs several times as much)
  const float beta = fp * __builtin_amdgcn_rcpf((float)E);
  const float pMeV = (float)(m_lep * 1.0e3) * fp * __builtin_amdgcn_rcpf((float)m_part);
  const float t = (float)(T.d->rho * (dz * (1.0 / kCmToM)));
  const float th0 = lynch_dahl_theta0(T.mcs_cA, T.mcs_cB, T.mcs_cC, t, pMeV, beta);
  const float r = sqrtf(-2.0f * logf(u01f(w_r)));
  const float theta = ((w_a & 1u) ? r : -r) * th0 * (float)T.d->rescale_MCS;
  float sth, cth;
  sincosf(theta, &sth, &cth);
  const float ua = u01f(w_a);
  const float sph = __builtin_amdgcn_sinf(ua);  // sin(2 pi ua): the hardware sine takes revolutions
  const float cph = __builtin_amdgcn_cosf(ua);
  const Frame f = frame_of(nx, ny, nz);
  float ox, oy, oz;
  frame_apply(f, sph * sth, -cph * sth, cth, ox, oy, oz);
  const float inv = rsqrtf(ox * ox + oy * oy + oz * oz);
  nx = ox * inv;
  ny = oy * inv;
  nz = oz * inv;
}

// |p| of energy E and mass m.  fmax: contracted to fma(E, E, -(m*m)), E*E -
// m*m at E == m is minus the rounding error of m*m, which can be negative (the
// reference's p is 0 there).
__device__ inline double momentum(double E, double m) { return sqrt(fmax(E * E - m * m, 0.0)); }

// Moves a charged particle by ``dist`` after losing dEdx*dist (lose_energy,
// particle.py:143-153) and scatters it (shower.py:569-581, 589-598).  The
// float64 |p| is formed once, when the particle finishes (momentum()); the
// scattering angle takes it in float32.
__device__ inline void charged_advance(const DevTables& T, double dist, double m, double m_lep, uint32_t w_r,
                                       uint32_t w_a, double& E, float& nx, float& ny, float& nz, float& rx,
                                       float& ry, float& rz, double& dep) {
  const double Eb = E;
  E = E - T.d->dEdx * dist;
  if (E <= m) E = m;
  const double q = fmax(E * E - m * m, 0.0);  // p^2, as momentum() forms it
  dep += Eb - E;
  if (q > 0.0) {
    rx += (float)(nx * dist);
    ry += (float)(ny * dist);
    rz += (float)(nz * dist);
    if (T.d->MS) mcs_rotate(T, nx, ny, nz, E, sqrtf((float)q), m, m_lep, dist, w_r, w_a);
  } else {
    nx = ny = nz = 0.0f;
  }
}

// ---------------------------------------------------------- transport -----
// propagate_particle (shower.py:509-601) and the process choice (665-698)
// for the store slots [first, first + n).  One advance per loop iteration.
// Refills (record loads, photon flights) and finishes (record stores, process
// choice, draw requests) run for the whole wave at once, so they wait until
// batch_lanes lanes need them (or no lane is stepping): each costs about as
// much as several substeps (global loads, atomics), and measured on MI355X
// 44 of 64 lanes gives the shortest sweeps (32 and 56: 1.13x longer).
// A particle that has taken yield_steps substeps in one launch hands its live
// state on to the next sweep, which starts it first: the few long-lived
// particles otherwise keep nearly empty waves running at the end of every
// launch.  Its Philox blocks are indexed by its total substep count, so the
// random numbers, and every result, are those of an uninterrupted transport.
struct TransportParams {
  int32_t batch_lanes;   // 1..64
  uint32_t yield_steps;  // >= 1
};
struct ResumeRec {
  double E, dep, dz;
  float nx, ny, nz, rx, ry, rz;
  int32_t slot;
  uint32_t k;
};
struct ResumeBuf {
  ResumeRec* rec;  // [2][cap]: the half of the current sweep's parity is read, the other written
  int64_t cap;
};
// range (non-null): the sweep's first slot and count, set on the device by
// k_sweep_setup; blocks beyond what the count needs leave at once.
__global__ __launch_bounds__(256) void k_transport(DevTables T, DevStore S, int64_t first, int64_t n, uint64_t seed,
                                                   Requests R, ParentRecs PR, double* shower_energy,
                                                   unsigned long long* counters, const double* sig_lds_src,
                                                   const int64_t* range, ResumeBuf RB, TransportParams TP) {
  // resumed particles (the previous sweep's hand-ons) come first in the queue
  int64_t nres = 0;
  const ResumeRec* rin = nullptr;
  ResumeRec* rout = nullptr;
  if (range) {
    first = range[0];
    n = range[1];
    if (RB.rec) {
      nres = range[3];
      const int64_t par = (int64_t)(counters[C_RPARITY] & 1ull);
      rin = RB.rec + par * RB.cap;
      rout = RB.rec + (1 - par) * RB.cap;
    }
    if ((int64_t)blockIdx.x * 4 >= (n + nres + 63) / 64) return;
  }
  HIP_DYNAMIC_SHARED(double, s_sig)
  for (int i = threadIdx.x; i < T.lds_doubles; i += blockDim.x) s_sig[i] = sig_lds_src[i];
  __syncthreads();
  const LdsSigma L{s_sig};
  const int lane = threadIdx.x & 63;
  WaveQueue<64> q;
  // lane states: stepping (act), finished and parked (wait), empty (neither), no work left (done)
  bool act = false, wait = false, done = false;
  int64_t slot = 0;
  int pid = 0, shower = 0, gen = 0;
  uint32_t ch = 0xFFFu;  // channels of the lane's n*sigma sum
  uint64_t sid = 0;
  double m = 0.0, pme = 0.0, E = 0.0, p = 0.0, dep = 0.0, dz = 0.0;
  float nx = 0.0f, ny = 0.0f, nz = 0.0f, rx = 0.0f, ry = 0.0f, rz = 0.0f, wgt = 0.0f;
  uint32_t k = 0;
  uint32_t kl = 0;       // substeps taken in this launch
  bool closing = false;  // the next advance is the closing increment
  unsigned long long steps = 0, req_ovf = 0;
  while (true) {
    const unsigned long long am = __ballot(act);
    // ---- refill empty lanes in batches
    const bool idle = !act && !wait && !done;
    const unsigned long long im = __ballot(idle);
    if (im && (__popcll(im) >= TP.batch_lanes || am == 0ull)) {
      const int64_t r = q.take(im, idle, counters + C_TRANSPORT);
      if (idle) {
        if (r < nres) {
          // a particle handed on by the previous sweep: constants from the
          // store, live state from its resume record
          const ResumeRec& R0 = rin[r];
          slot = R0.slot;
          pid = S.pid[slot];
          sid = S.sid[slot];
          shower = S.shower[slot];
          gen = S.meta[slot] & 0xFFFF;
          wgt = S.r0[slot].w;
          ch = sum_channels(pid);
          m = pid_mass(pid);
          pme = fmax(fmax(min_calc(T, pid), T.d->min_energy), m);
          E = R0.E;
          dep = R0.dep;
          dz = R0.dz;
          nx = R0.nx;
          ny = R0.ny;
          nz = R0.nz;
          rx = R0.rx;
          ry = R0.ry;
          rz = R0.rz;
          k = R0.k;
          kl = 0;
          closing = false;
          act = true;
        } else if (r < nres + n) {
          slot = first + (r - nres);
          kl = 0;
          const float4 p0 = S.p0[slot];
          const float4 r0 = S.r0[slot];
          pid = S.pid[slot];
          sid = S.sid[slot];
          shower = S.shower[slot];
          gen = S.meta[slot] & 0xFFFF;
          wgt = r0.w;
          ch = sum_channels(pid);
          m = pid_mass(pid);
          pme = fmax(fmax(min_calc(T, pid), T.d->min_energy), m);
          E = (double)p0.x;
          nx = p0.y;
          ny = p0.z;
          nz = p0.w;
          const float pn2 = nx * nx + ny * ny + nz * nz;
          if (pn2 > 0.0f) {
            const float inv = rsqrtf(pn2);
            nx *= inv;
            ny *= inv;
            nz *= inv;
          }
          // a species the shower does not propagate (mesons, neutrinos: no
          // channels and not a photon) keeps its |p|, and ends where it starts
          const bool still = ch == 0xFFFu && pid != 22;
          p = still ? sqrt((double)pn2) : 0.0;  // the others' |p| is formed when they finish
          rx = r0.x;
          ry = r0.y;
          rz = r0.z;
          dep = 0.0;
          k = 0;
          closing = false;
          if (E < pme || still) {
            wait = true;  // ends where it starts (shower.py:532-536, 640-653)
          } else if (pid == 22) {
            // one exponential flight, no MCS (shower.py:538-553)
            const double mfp = L.mfp(22, E);
            const double dist = -mfp * (double)log1p_f32(-u01f(draw_block(seed, sid, 0, PURPOSE_STEP).x));
            rx += (float)(nx * dist);
            ry += (float)(ny * dist);
            rz += (float)(nz * dist);
            k = 1;
            wait = true;
          } else {
            act = true;
          }
        } else {
          done = true;
        }
      }
    }
    // ---- finish parked particles in batches: write pf/rf, choose the process, queue the draw
    const unsigned long long wm = __ballot(wait);
    if (wm && (__popcll(wm) >= TP.batch_lanes || __ballot(act) == 0ull)) {
      int want = -1;
      if (wait) {
        steps += k;
        // |p| at the final energy (a species the shower does not propagate
        // keeps the |p| it came with)
        if (!(ch == 0xFFFu && pid != 22)) p = momentum(E, m);
        S.pf[slot] = make_float4((float)E, (float)(p * nx), (float)(p * ny), (float)(p * nz));
        S.rf[slot] = make_float4(rx, ry, rz, (float)k);
        // process choice (shower.py:665-698) behind sample_scattering's threshold (469-470)
        if (E > pme) {
          // channels in the order of shower.py:665-698, held in registers
          int c0 = -1, c1 = -1, c2 = -1;
          switch (pid) {
            case 11: c0 = BREM; c1 = MOLLER; break;
            case -11: c0 = BREM; c1 = ANN; c2 = BHABHA; break;
            case 22: c0 = PAIRPROD; c1 = COMP; break;
            case 13: case -13: c0 = MUONE; c1 = MUONBREM; break;
            default: break;
          }
          if (c0 >= 0) {
            const float lf = __log2f((float)E);
            const double w0 = L.nsigma(c0, E, lf), w1 = L.nsigma(c1, E, lf);
            const double w2 = c2 >= 0 ? L.nsigma(c2, E, lf) : 0.0;
            const double SC = c2 >= 0 ? w0 + w1 + w2 : w0 + w1;
            if (SC > 0.0 && !isnan(SC)) {
              const double uu = u01d(draw_block(seed, sid, 0, PURPOSE_INTERACT).x) * SC;
              if (uu < w0) want = c0;
              else if (c2 < 0 || uu < w0 + w1) want = c1;
              else want = c2;
            }
          }
        }
        if (shower_energy) {
          unsafeAtomicAdd(shower_energy + 4 * shower + 0, dep);
          if (want < 0) unsafeAtomicAdd(shower_energy + 4 * shower + 1, E);
        }
        wait = false;
      }
      // wave-aggregated appends to the per-process request regions (SM
      // processes): lane pp reserves process pp's share, so the reservations
      // of all processes are in flight together (one atomic round trip)
      if (__ballot(want >= 0)) {
        unsigned long long mine = 0ull;
        unsigned cnt = 0u;
#pragma unroll
        for (int pp = 0; pp < 8; ++pp) {
          const unsigned long long mask = __ballot(want == pp);
          if (want == pp) mine = mask;
          if (lane == pp) cnt = (unsigned)__popcll(mask);
        }
        unsigned base = 0u;
        if (cnt) base = atomicAdd(R.count + lane, cnt);
        base = __shfl(base, want >= 0 ? want : 0);
        if (want >= 0) {
          const unsigned idx = base + (unsigned)__popcll(mine & ((1ull << lane) - 1ull));
          const int64_t o = (int64_t)want * R.cap + idx;
          // a region holds as many requests as a sweep holds particles, so
          // this never fails; past the region the request is dropped and
          // counted, and the run reports an internal error
          if ((int64_t)idx >= R.cap) {
            ++req_ovf;
            want = -1;
          }
        }
        if (want >= 0) {
          const unsigned idx = base + (unsigned)__popcll(mine & ((1ull << lane) - 1ull));
          const int64_t o = (int64_t)want * R.cap + idx;
          R.slot[o] = (int32_t)slot;
          R.E[o] = E;
          PR.sid[o] = sid;
          PR.pf[o] = make_float4((float)E, (float)(p * nx), (float)(p * ny), (float)(p * nz));
          PR.pos[o] = make_float4(rx, ry, rz, wgt);
          PR.info[o] = make_int2((gen << 16) | (pid & 0xFFFF), shower);
        }
      }
    }
    if (__ballot(act) == 0ull) {
      if (__ballot(wait || !done) == 0ull) break;
      continue;
    }
    if (act) {
      // One advance per lane and iteration, so the lanes of a wave run the same
      // code: a substep (shower.py:559-581), or the closing increment
      // (583-598) in the iteration after the substep that ended the loop.
      //   substep: dz = mfp / U(6,20); hard if u > exp(-dz/mfp) = exp(-1/U(6,20)),
      //            else lose dE/dx dz, move, scatter.  exp in float32: the
      //            comparison flips only when u lies within ~1e-7 of the threshold.
      //   closing: below the cutoff dist = u dz, else the truncated exponential
      //            dist = -mfp ln(1 + (exp(-dz/mfp) - 1) u) at the current energy;
      //            its scattering uses m_e (the reference omits m_lepton there).
      const bool in_close = closing;
      const double S = L.nsigma_sum_ch(ch, E);
      const double mfp = S > 0.0 ? kCmToM / S : 1.0e12;
      const U4 u = draw_block(seed, sid, in_close ? 0u : k, in_close ? PURPOSE_CLOSE : PURPOSE_STEP);
      bool adv = true;
      double dist;
      uint32_t w_r, w_a;
      double m_lep = m;
      if (!in_close) {
        const double div = 6.0 + 14.0 * u01d(u.y);
        dz = mfp / div;
        dist = dz;
        w_r = u.z;
        w_a = u.w;
        // hard if u > exp(-dz/mfp): the float32 exponential decides unless u
        // lies within 1e-5 of it (its error is a few 1e-7), where float64 does
        const double ux = u01d(u.x);
        double thr = (double)__expf(-__builtin_amdgcn_rcpf((float)div));
        if (fabs(ux - thr) < 1.0e-5) thr = exp(-(dz / mfp));
        if (ux > thr) {
          adv = false;
          closing = true;
        }
      } else {
        const double distC = u01d(u.x);
        if (E < pme) {
          dist = distC * dz;
        } else {
          // dz / mfp = dz * S / 0.01
          const float r2 = S > 0.0 ? (float)(dz * S * (1.0 / kCmToM)) : (float)(dz * 1.0e-12);
          dist = -mfp * (double)log1p_f32(expm1f(-r2) * (float)distC);
        }
        w_r = u.y;
        w_a = u.z;
        m_lep = kMe;
      }
      if (adv) charged_advance(T, dist, m, m_lep, w_r, w_a, E, nx, ny, nz, rx, ry, rz, dep);
      ++k;
      if (in_close) {
        closing = false;
        act = false;
        wait = true;
      } else if (E < pme) {
        closing = true;  // below the cutoff after this substep: close next
      }
      // hand a long-lived particle on to the next sweep (space permitting)
      const bool give = rout && !closing && act && ++kl >= TP.yield_steps;
      const unsigned long long gm = __ballot(give);
      if (gm) {
        const int leader = __ffsll((long long)gm) - 1;
        unsigned long long gb = 0;
        if (lane == leader) gb = atomicAdd(counters + C_YIELD_OUT, (unsigned long long)__popcll(gm));
        gb = shfl_u64(gb, leader);
        const int64_t gi = (int64_t)gb + __popcll(gm & ((1ull << lane) - 1ull));
        if (give && gi < RB.cap) {
          ResumeRec& W = rout[gi];
          W.E = E;
          W.dep = dep;
          W.dz = dz;
          W.nx = nx;
          W.ny = ny;
          W.nz = nz;
          W.rx = rx;
          W.ry = ry;
          W.rz = rz;
          W.slot = (int32_t)slot;
          W.k = k;
          act = false;  // the lane is free for the next record
        }
      }
    }
  }
  wave_stat(counters, C_STATS + ST_STEPS, steps);
  wave_stat(counters, C_STATS + ST_REQ_OVERFLOW, req_ovf);
}

// -------------------------------------------------------------- emission --
// The two children of one vertex in the lab (sample_scattering,
// shower.py:471-480): the process's final state in the parent frame from the
// accepted point xs and the azimuth uniform az, rotated onto the parent's
// direction (px, py, pz; any length), and their species (process_PIDS, 87-96).
// The accepted point of a 1-D draw in float64.  The draw stores it as float32,
// but the 1-D kinematics (a cosine near 1 whose complement sets a child's
// energy, mue_to_mue_fourvecs, kinematics.py:239-265) cancel: a float32 cosine
// costs ~E * 6e-8 GeV of a knock-on's energy.  The same map transform of the
// same Philox block (candidate trials - 1 of the parent's stream) gives the
// float64 value the draw accepted.
__device__ inline double accepted_x1(const DevTables& T, int p, double E, uint64_t seed, uint64_t sid, int32_t trials,
                                     float xf) {
  const PetiteGridDesc& gd = T.d->grid[p];
  if (gd.ndim != 1 || trials <= 0 || gd.n_E <= 0) return (double)xf;
  const int ie = lu_key(T, p, E);
  const float* base = T.f32 + gd.node_off + (int64_t)ie * gd.stride;
  const U4 a = draw_block(seed, sid, (uint32_t)(trials - 1), PURPOSE_VEGAS_A);
  const uint32_t rr[4] = {a.x, a.y, a.z, a.w};
  const int32_t ninc[4] = {gd.ninc[0], gd.ninc[1], gd.ninc[2], gd.ninc[3]};
  double x[4] = {0.0, 0.0, 0.0, 0.0};
  map_point_n<1>(base, ninc, rr, x);
  return x[0];
}

// x0: the 1-D processes' variable in float64 (accepted_x1); xs: the 4-D
// processes' point.
__device__ inline void vertex_children(int p, int pid, double E, const float* xs, double x0, double az, double px,
                                       double py, double pz, V4* lab, int* cpid) {
  const double m = pid_mass(pid);
  V4 a, b;
  switch (p) {
    case BREM: case MUONBREM: kin_brem(E, m, xs, az, a, b); break;
    case PAIRPROD: kin_pairprod(E, xs, az, a, b); break;
    case COMP: kin_compton(E, 0.0, x0, az, a, b); break;
    case ANN: kin_ann(E, 0.0, x0, az, a, b); break;
    case MUONE: kin_mue(E, x0, az, a, b); break;
    default: kin_ee(E, x0, az, a, b); break;  // Moller, Bhabha
  }
  child_pids(p, pid, &cpid[0], &cpid[1]);
  const double pn = sqrt(px * px + py * py + pz * pz);
  if (pn > 0.0) {
    px /= pn;
    py /= pn;
    pz /= pn;
  }
  const V4 in[2] = {a, b};
  for (int k = 0; k < 2; ++k) {
    double ox, oy, oz;
    frame_apply_d(px, py, pz, in[k], ox, oy, oz);
    lab[k] = V4{in[k].e, ox, oy, oz};
  }
}

// sample_scattering's kinematics, lab rotation and child bookkeeping
// (shower.py:467-507, 702-706) for every finished draw.
struct EmitLaunch {
  const int32_t* req_slot;
  const double* req_E;
  int64_t req_cap;
  const float* x_out;
  const int32_t* trials_out;
  const unsigned int* dcount;  // per-process request counts (device)
  ParentRecs par;              // the parents' records, beside the requests
};

// Grid-stride over the requests of the sweep (every thread of a block takes
// part in every pass: the slot allocation below is block-aggregated).
constexpr int kEmitWaves = 4;  // waves per block
__global__ __launch_bounds__(64 * kEmitWaves) void k_emit(DevTables T, DevStore S, EmitLaunch io, uint64_t seed,
                                              unsigned long long* store_count, double* shower_energy,
                                              unsigned long long* stats) {
  // process boundaries of the request list, read once (uniform)
  int64_t seg[PETITE_NPROC + 1];
  seg[0] = 0;
#pragma unroll
  for (int q = 0; q < PETITE_NPROC; ++q) seg[q + 1] = seg[q] + min((int64_t)io.dcount[q], io.req_cap);
  const int64_t total = seg[PETITE_NPROC];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  __shared__ unsigned long long s_wtot[kEmitWaves];
  __shared__ unsigned long long s_base;
  unsigned long long n_dropped = 0;  // children below the cutoff, counted once per wave at the end
  for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x; r0 < total; r0 += stride) {
  const int64_t r = r0 + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool keep[2] = {false, false};
  V4 lab[2];
  int cpid[2] = {0, 0};
  int64_t slot = 0;
  int proc = 0;
  int64_t oidx = 0;
  int2 info{0, 0};
  uint64_t sid = 0;
  unsigned long long n_drop = 0;
  if (r < total) {
    int p = 0;
#pragma unroll
    for (int q = 1; q < PETITE_NPROC; ++q) p += r >= seg[q] ? 1 : 0;
    int64_t start = 0;
#pragma unroll
    for (int q = 0; q < PETITE_NPROC; ++q) start = q == p ? seg[q] : start;
    proc = p;
    oidx = (int64_t)p * io.req_cap + (r - start);
    slot = io.req_slot[oidx];
    const double E = io.req_E[oidx];
    info = io.par.info[oidx];
    const int pid = info_pid(info);
    const int shower = info.y;
    const int32_t tr = io.trials_out[oidx];
    if (tr <= 0) {
      if (shower_energy) unsafeAtomicAdd(shower_energy + 4 * shower + 1, E);
    } else {
      const float* xs = io.x_out + 4 * oidx;
      sid = io.par.sid[oidx];
      const double az = u01d(draw_block(seed, sid, 0, PURPOSE_INTERACT).y);
      const float4 pf = io.par.pf[oidx];
      const double x0 = accepted_x1(T, p, E, seed, sid, tr, xs[0]);
      vertex_children(p, pid, E, xs, x0, az, pf.y, pf.z, pf.w, lab, cpid);
      for (int k = 0; k < 2; ++k) {
        keep[k] = lab[k].e > T.d->min_energy;
        if (!keep[k] && shower_energy) unsafeAtomicAdd(shower_energy + 4 * shower + 2, lab[k].e);
      }
      if (shower_energy && (p == COMP || p == ANN || p == MOLLER || p == BHABHA || p == MUONE))
        unsafeAtomicAdd(shower_energy + 4 * shower + 3, 1.0);
      n_drop = (!keep[0]) + (!keep[1]);
    }
  }
  n_dropped += n_drop;
  const int nk = (int)keep[0] + (int)keep[1];
  // block-aggregated slot allocation: one atomic on the store count per block
  // and pass.  Atomics on one address are served one after another (~10 ns
  // each, measured): a wave's atomic per pass on the count, and another on the
  // dropped-children statistic beside it, held the emission at 41 ms per step
  // of the 100k workload, against 16 ms with one per block and pass.
  unsigned long long incl = nk;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long v = shfl_u64(incl, lane >= off ? lane - off : lane);
    if (lane >= off) incl += v;
  }
  const int wid = threadIdx.x >> 6;
  if (lane == 63) s_wtot[wid] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sum = 0;
    for (int w = 0; w < kEmitWaves; ++w) sum += s_wtot[w];
    s_base = sum ? atomicAdd(store_count, sum) : 0ull;
  }
  __syncthreads();
  unsigned long long dst = s_base + incl - nk;
  for (int w = 0; w < wid; ++w) dst += s_wtot[w];
  __syncthreads();  // every wave has read s_wtot and s_base before the next pass writes them
  if (nk) {
    const int gen = info_gen(info) + 1;
    const float4 pos = io.par.pos[oidx];
    const int shower = info.y;
    for (int k = 0; k < 2; ++k) {
      if (!keep[k]) continue;
      if ((int64_t)dst < S.capacity) {
        const int64_t c = (int64_t)dst;
        // pf and rf are left to the transport, which writes them for every
        // stored particle when it finishes it
        S.p0[c] = make_float4((float)lab[k].e, (float)lab[k].x, (float)lab[k].y, (float)lab[k].z);
        S.r0[c] = make_float4(pos.x, pos.y, pos.z, pos.w);
        S.pid[c] = cpid[k];
        S.parent[c] = (int32_t)slot;
        S.meta[c] = (gen & 0xFFFF) | ((proc + 1) << 16) | (k << 24);
        S.shower[c] = shower;
        S.sid[c] = derive_sid(seed, sid, (uint32_t)k, PURPOSE_CHILD);
      } else if (stats) {
        atomicAdd(stats + ST_OVERFLOW, 1ull);
      }
      ++dst;
    }
  }
  }
  if (stats) {
    const unsigned long long v[1] = {n_dropped};
    unsigned long long* const dst[1] = {stats + ST_DROPPED};
    block_stat_add<1>(v, dst);
  }
}

// ------------------------------------------------------- single vertices --
// sample_scattering for given particles and processes (shower.py:467-507):
// requests into the per-process regions (slot = the vertex index, stream id
// beside it), then, after the draws, both children of every vertex.
__global__ void k_vertex_requests(DevTables T, const PetiteVertex* v, const int32_t* proc, int64_t n, Requests R,
                                  uint64_t* req_sid, int32_t* trials_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = proc[i];
  if (p < 0 || p >= kNprocSM || T.d->grid[p].n_E <= 0) {
    trials_out[i] = 0;
    return;
  }
  const unsigned k = atomicAdd(R.count + p, 1u);
  const int64_t o = (int64_t)p * R.cap + k;
  R.slot[o] = (int32_t)i;
  R.E[o] = v[i].E;
  req_sid[o] = v[i].sid;
}

// Children 2i and 2i+1 of vertex i (child index k = 0, 1), every child kept
// (sample_scattering returns both; generate_shower's cut is the caller's).
__global__ void k_vertex_emit(DevTables T, const PetiteVertex* v, const int32_t* proc, int64_t n, const float* x_out,
                              const int32_t* trials, uint64_t seed, DevStore out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PetiteVertex q = v[i];
  const int p = proc[i];
  const bool ok = trials[i] > 0;
  V4 lab[2] = {V4{0.0, 0.0, 0.0, 0.0}, V4{0.0, 0.0, 0.0, 0.0}};
  int cpid[2] = {0, 0};
  if (ok) {
    const double az = u01d(draw_block(seed, q.sid, 0, PURPOSE_INTERACT).y);
    const double x0 = accepted_x1(T, p, q.E, seed, q.sid, trials[i], x_out[4 * i]);
    vertex_children(p, q.pid, q.E, x_out + 4 * i, x0, az, q.px, q.py, q.pz, lab, cpid);
  }
  for (int k = 0; k < 2; ++k) {
    const int64_t c = 2 * i + k;
    const float4 p4 = make_float4((float)lab[k].e, (float)lab[k].x, (float)lab[k].y, (float)lab[k].z);
    const float4 r4 = make_float4((float)q.x, (float)q.y, (float)q.z, (float)q.weight);
    out.p0[c] = p4;
    out.pf[c] = p4;
    out.r0[c] = r4;
    out.rf[c] = make_float4((float)q.x, (float)q.y, (float)q.z, 0.0f);
    out.pid[c] = cpid[k];
    out.parent[c] = (int32_t)i;
    out.meta[c] = ok ? (((q.generation + 1) & 0xFFFF) | ((p + 1) << 16) | (k << 24)) : 0;
    out.shower[c] = 0;
    out.sid[c] = derive_sid(seed, q.sid, (uint32_t)k, PURPOSE_CHILD);
  }
}

// Start of a device-driven sweep: the range is the particles appended since
// the last sweep (at most `bound`, the rest waits for the next sweep), and the
// per-sweep counters are cleared.  range = {first, n, end}.
__global__ void k_sweep_setup(unsigned long long* ctr, int64_t* range, unsigned int* req_count, int64_t capacity,
                              int64_t bound, int64_t resume_cap) {
  const int t = threadIdx.x;
  if (t < PETITE_NPROC) req_count[t] = 0u;
  if (t < PETITE_NPROC) ctr[refill_ctr(t)] = 0ull;  // draw refill counters
  if (t == 0) ctr[C_TRANSPORT] = 0ull;              // transport refill counter
  if (t == 0) {
    const SweepRange sr = sweep_range(range[2], (int64_t)ctr[C_STORE], capacity, bound, (int64_t)ctr[C_YIELD_OUT],
                                      resume_cap);
    range[0] = sr.first;
    range[1] = sr.n;
    range[2] = sr.first + sr.n;
    ctr[C_YIELD_OUT] = 0ull;
    ctr[C_RPARITY] ^= 1ull;
    range[3] = sr.nres;
    if (sr.n > 0 || sr.nres > 0) ctr[C_SWEEPS] += 1ull;
  }
}

#include "tail.h"
#include "quad.h"
#include "dark.h"

// draw_pe_sample (dark_shower.py:706-719) for n stream ids: the bound
// electron's momentum [GeV], or -1 when no sample came within 65536 pairs.
__global__ void k_draw_pe(const uint64_t* sid, int64_t n, uint64_t seed, double Zeff, double* pe_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double pe = -1.0;
  if (!dark::draw_pe(seed, sid[i], Zeff, pe)) pe = -1.0;
  pe_out[i] = pe;
}


// ------------------------------------------------- detector acceptance --
// transverse_position and detector_cut (shower.py:815-865): straight-line
// extrapolation of each record from its production point along its
// production momentum to the planes z = z_d, an annulus cut on the transverse
// radius, and the per-plane sums of the passing weights.
struct DetectorPlanes {
  int32_t n;
  int32_t has_ecut;
  double radius, inner, emin, emax;
  double z[PETITE_MAX_DETECTORS];
};

__device__ inline double wave_sum_f64(double v) {
  for (int off = 32; off > 0; off >>= 1) v += shfl_f64(v, (threadIdx.x & 63) ^ off);
  return v;
}

// Pass bits of one record for planes [0, n_planes): T = (z - z0) / pz,
// r_T = |(x0 + T px, y0 + T py)|, inner < r_T < radius.
__device__ inline bool plane_pass(const DetectorPlanes& D, int d, double x0, double y0, double z0, double px, double py,
                                  double pz) {
#pragma clang fp contract(off)
  const double T = (D.z[d] - z0) / pz;
  const double xf = x0 + T * px, yf = y0 + T * py;
  const double rT = sqrt(xf * xf + yf * yf);
  return rT > D.inner && rT < D.radius;
}

// Grid-stride over the records.  With at most kRegPlanes planes every thread
// keeps its weight sums in registers and they are reduced once per wave at
// the end; otherwise each block pass reduces them per wave.  Per block the
// wave sums meet in LDS, then one float64 atomic per block and plane.
constexpr int kRegPlanes = 8;
template <bool REG>
__global__ __launch_bounds__(256) void k_detector_cut(const float4* __restrict__ p, const float4* __restrict__ r,
                                                      int64_t n, DetectorPlanes D, uint64_t* __restrict__ mask,
                                                      double* __restrict__ wsum) {
  __shared__ double s_w[PETITE_MAX_DETECTORS + 1];
  for (int d = threadIdx.x; d <= D.n; d += blockDim.x) s_w[d] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  double acc[kRegPlanes + 1];
#pragma unroll
  for (int d = 0; d <= kRegPlanes; ++d) acc[d] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
    const int64_t i = i0 + threadIdx.x;
    bool inwin = false;
    double w = 0.0, x0 = 0.0, y0 = 0.0, z0 = 0.0, px = 0.0, py = 0.0, pz = 1.0;
    if (i < n) {
      const float4 pi = p[i], ri = r[i];
      const double E = (double)pi.x;
      inwin = !D.has_ecut || (E < D.emax && E > D.emin);
      w = (double)ri.w;
      x0 = (double)ri.x;
      y0 = (double)ri.y;
      z0 = (double)ri.z;
      px = (double)pi.y;
      py = (double)pi.z;
      pz = (double)pi.w;
    }
    uint64_t bits = 0;
    if (REG) {
#pragma unroll
      for (int d = 0; d < kRegPlanes; ++d) {
        if (d < D.n) {
          const bool pass = inwin && plane_pass(D, d, x0, y0, z0, px, py, pz);
          bits |= (uint64_t)pass << d;
          acc[d] += pass ? w : 0.0;
        }
      }
      acc[kRegPlanes] += inwin ? w : 0.0;
    } else {
      for (int d = 0; d < D.n; ++d) {
        const bool pass = inwin && plane_pass(D, d, x0, y0, z0, px, py, pz);
        bits |= (uint64_t)pass << d;
        const double v = wave_sum_f64(pass ? w : 0.0);
        if (lane == 0 && v != 0.0) atomicAdd(&s_w[d], v);
      }
      const double t = wave_sum_f64(inwin ? w : 0.0);
      if (lane == 0 && t != 0.0) atomicAdd(&s_w[D.n], t);
    }
    if (mask && i < n) mask[i] = bits;
  }
  if (REG) {
#pragma unroll
    for (int d = 0; d <= kRegPlanes; ++d) {
      const int slot = d == kRegPlanes ? D.n : d;
      if (d < D.n || d == kRegPlanes) {
        const double v = wave_sum_f64(acc[d]);
        if (lane == 0 && v != 0.0) atomicAdd(&s_w[slot], v);
      }
    }
  }
  __syncthreads();
  if (wsum)
    for (int d = threadIdx.x; d <= D.n; d += blockDim.x)
      if (s_w[d] != 0.0) unsafeAtomicAdd(wsum + d, s_w[d]);
}

// --------------------------------------------------- allocation-free draws --
// petite_vegas_draw_noalloc keeps its request regions and counters in one
// file-scope device array instead of context scratch: [PETITE_NPROC][kNoAllocCap]
// request slots (int32) and energies (float64), then the counter block
// (C_N uint64) and the per-process request counts (uint32).  Every GPU the
// library runs on holds one copy until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns on that device.
constexpr int64_t kNoAllocCap = PETITE_NOALLOC_MAX_DRAWS;
constexpr size_t kNoAllocSlotOff = 0;
constexpr size_t kNoAllocEOff = kNoAllocSlotOff + sizeof(int32_t) * PETITE_NPROC * kNoAllocCap;
constexpr size_t kNoAllocCtrOff = kNoAllocEOff + sizeof(double) * PETITE_NPROC * kNoAllocCap;
constexpr size_t kNoAllocCountOff = kNoAllocCtrOff + sizeof(unsigned long long) * C_N;
constexpr size_t kNoAllocBytes = (kNoAllocCountOff + sizeof(unsigned int) * PETITE_NPROC + 255) / 256 * 256;
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Clears the counter block and the request counts of the allocation-free
// draws (a kernel rather than a memset, so the call stays capturable).
__global__ void k_noalloc_clear(unsigned long long* ctr, unsigned int* counts) {
  for (int t = threadIdx.x; t < C_N; t += blockDim.x) ctr[t] = 0ull;
  if (threadIdx.x < PETITE_NPROC) counts[threadIdx.x] = 0u;
}

}  // namespace

// ===================================================================== C ABI
struct PetiteCtx {
  PetiteTableDesc desc;
  double* d64 = nullptr;
  float* d32 = nullptr;
  double* dmaxf = nullptr;
  PetiteTableDesc* ddesc = nullptr;
  double* dsig_lds = nullptr;
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
  ParentRecs par{};  // [kNprocSM][req_cap], shower pipeline only
  unsigned long long* counters = nullptr;  // C_N entries, layout CounterIdx
  bool timing = false;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // one stream per process for the draw kernels (default priority), forked
  // from and joined back into the caller's stream
  hipStream_t dstream[PETITE_NPROC] = {};
  hipEvent_t dfork = nullptr;
  hipEvent_t djoin[PETITE_NPROC] = {};
  // device-driven sweeps: {first, n, end} of the current sweep, and timing events
  int64_t* drange = nullptr;
  ResumeRec* dresume = nullptr;  // [2][resume_cap]
  int64_t resume_cap = 0;
  int32_t* tail_spill = nullptr;  // [2][kTailSpill]: depth-first walk spill lists
  std::vector<hipEvent_t> tev;
  int n_cu = 256;
  PetiteTuning tune{};  // launch and queue sizes in effect (defaults filled in)
  unsigned char* noalloc = nullptr;  // this device's copy of `scratch` (hipGetSymbolAddress)
  double tbar[PETITE_NPROC];  // running mean of candidates per draw, per process
  // dark sector
  bool has_dark = false;
  PetiteDarkDesc dark{};
  double* dark_blob = nullptr;
  dark::Entries ent{};
  float* dres_x = nullptr;
  int32_t* dres_trials = nullptr;
  float* dres_w = nullptr;
  double* dyield = nullptr;
  // trained masses for the m_V-continuous yield
  int32_t n_mass = 0;
  double masses[PETITE_MAX_MASSES] = {};
  dark::YieldTables ytab{};
  dark::YieldTables* ytab_dev = nullptr;  // device copy read by the kernel
  double* yield_blob = nullptr;
  double* yacc = nullptr;
  int32_t* ylist = nullptr;               // [2][ylist_cap]: the records of k_dark_yield, of k_bound_yield
  int64_t ylist_cap = 0;
  unsigned long long* ylist_cnt = nullptr;  // [2]
  // bound-electron tables at each m_V of the yield (petite_yield_bound)
  bool bnd_on = false;
  quad::BoundParams bnd{};
  double bnd_h = 0.0;
  double bnd_top_ann[PETITE_MAX_MASSES] = {};
  double bnd_top_comp[PETITE_MAX_MASSES] = {};
  double* bnd_blob = nullptr;          // survival and mean-free-path tables, running integrals, fixed nodes
  double* bnd_tab = nullptr;           // per slot: grids, log-log tables, weights, panel nodes
  quad::BoundSlot* bnd_slots = nullptr;
  quad::Job* bnd_jobs = nullptr;
  int* bnd_bad = nullptr;
  const double* bnd_fixed = nullptr;   // survival-grid panel nodes (in bnd_blob)
  hipStream_t cstream = nullptr;       // capture stream of the sweep-group graphs (default priority)
};

static void free_entries(PetiteCtx* c) {
  for (void* p : {(void*)c->ent.parent, (void*)c->ent.flag, (void*)c->ent.weight, (void*)c->ent.E, (void*)c->ent.dir,
                  (void*)c->ent.dsid, (void*)c->dres_x, (void*)c->dres_trials, (void*)c->dres_w})
    if (p) (void)hipFree(p);
  c->ent = dark::Entries{};
  c->dres_x = nullptr;
  c->dres_trials = nullptr;
  c->dres_w = nullptr;
}

static int ensure_entries(PetiteCtx* c, int64_t cap) {
  if (cap <= c->ent.cap) return PETITE_OK;
  free_entries(c);
  dark::Entries& q = c->ent;
  HIP_TRY(hipMalloc(&q.parent, cap * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&q.flag, cap * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&q.weight, cap * sizeof(double)));
  HIP_TRY(hipMalloc(&q.E, cap * sizeof(double)));
  HIP_TRY(hipMalloc(&q.dir, cap * sizeof(float4)));
  HIP_TRY(hipMalloc(&q.dsid, cap * sizeof(uint64_t)));
  HIP_TRY(hipMalloc(&c->dres_x, cap * 4 * sizeof(float)));
  HIP_TRY(hipMalloc(&c->dres_trials, cap * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&c->dres_w, cap * sizeof(float)));
  q.cap = cap;
  return PETITE_OK;
}

// Default launch and queue sizes (PetiteTuning), measured on MI355X with the
// 100k-primary lead workload.  Device-driven sweeps: at most 2^23 particles
// per sweep (request scratch scales with it) and 16 sweeps queued per host
// check.
static PetiteTuning default_tuning() {
  PetiteTuning t{};
  t.batch_lanes = 44;
  t.yield_steps = 40;
  t.transport_waves = 20;
  t.emit_waves = 32;
  for (int p = 0; p < 8; ++p) t.draw_waves[p] = 2;
  t.draw_waves[BREM] = 16;
  t.draw_waves[PAIRPROD] = 8;
  t.min_group_lanes = 4;
  t.sweep_group = 16;
  t.group_requests = 4;
  t.small_sweep = 1 << 20;
  t.small_yield_steps = 40;
  t.graph = 2;
  t.tail_sweep = 16384;
  t.tail_waves = 8;
  t.tail_depth = kTailMaxDepth;
  t.tail_budget = 16;
  t.sweep_bound = (int64_t)1 << 23;
  t.resume_cap = (int64_t)1 << 20;
  return t;
}

// log2 of a power-of-two lane count.
static int lanes_log2(int lanes) {
  int lg = 0;
  while ((1 << lg) < lanes && lg < 6) ++lg;
  return lg;
}

// Blocks of 4 waves for `waves` waves per CU (at least one block).
static unsigned blocks_for(const PetiteCtx* c, int waves) {
  const int64_t b = ((int64_t)c->n_cu * waves + 3) / 4;
  return (unsigned)std::max<int64_t>(1, b);
}

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
    c->req_slot = nullptr;
    c->req_cap = 0;
  }
  if (c->par.sid) {
    for (void* p : {(void*)c->par.sid, (void*)c->par.pf, (void*)c->par.pos, (void*)c->par.info}) (void)hipFree(p);
    c->par = ParentRecs{};
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

// the parents' records beside the requests of the SM processes
static int ensure_parent_recs(PetiteCtx* c) {
  if (c->par.sid) return PETITE_OK;
  const int64_t n = c->req_cap * kNprocSM;
  HIP_TRY(hipMalloc(&c->par.sid, n * sizeof(uint64_t)));
  HIP_TRY(hipMalloc(&c->par.pf, n * sizeof(float4)));
  HIP_TRY(hipMalloc(&c->par.pos, n * sizeof(float4)));
  HIP_TRY(hipMalloc(&c->par.info, n * sizeof(int2)));
  return PETITE_OK;
}

static void plan_draw(PetiteCtx* c, const unsigned int* rq, DrawLaunch& io, int64_t& blocks_total) {
  io.block_start[0] = 0;
  io.full_blocks = c->n_cu * 4;  // 4-wave blocks at 4 waves per SIMD
  io.min_lg = lanes_log2(c->tune.min_group_lanes);
  io.group_requests = c->tune.group_requests;
  for (int p = 0; p < PETITE_NPROC; ++p) {
    int b, lg;
    io.count[p] = (int32_t)rq[p];
    draw_plan((double)rq[p], io.full_blocks, io.min_lg, io.group_requests, lg, b);
    io.lgG[p] = lg;
    io.block_start[p + 1] = io.block_start[p] + b;
  }
  blocks_total = io.block_start[PETITE_NPROC];
}

static int ensure_draw_streams(PetiteCtx* c);

// The per-process kernels are independent: with more than one, each runs on
// its own stream so small processes overlap the large ones.
static int launch_draws(PetiteCtx* c, const DrawLaunch& io, uint64_t seed, hipStream_t s) {
  int active[PETITE_NPROC], na = 0;
  for (int p = 0; p < PETITE_NPROC; ++p) {
    const int b = io.block_start[p + 1] - io.block_start[p];
    if (b > 0 && io.count[p] > 0 && c->desc.grid[p].n_E > 0) active[na++] = p;
  }
  if (na > 1) {
    const int rc = ensure_draw_streams(c);
    if (rc) return rc;
  }
  if (na > 1) HIP_TRY(hipEventRecord(c->dfork, s));
  for (int i = 0; i < na; ++i) {
    const int p = active[i];
    const int b = io.block_start[p + 1] - io.block_start[p];
    hipStream_t sp = s;
    if (na > 1) {
      sp = c->dstream[p];
      HIP_TRY(hipStreamWaitEvent(sp, c->dfork, 0));
    }
    hipLaunchKernelGGL(kDrawKernels[p], dim3((unsigned)b), dim3(256), 0, sp, c->dev, io, seed);
    HIP_TRY(hipGetLastError());
    if (na > 1) HIP_TRY(hipEventRecord(c->djoin[p], sp));
  }
  for (int i = 0; na > 1 && i < na; ++i) HIP_TRY(hipStreamWaitEvent(s, c->djoin[active[i]], 0));
  return PETITE_OK;
}

static int ensure_draw_streams(PetiteCtx* c) {
  if (c->dfork) return PETITE_OK;
  HIP_TRY(hipEventCreateWithFlags(&c->dfork, hipEventDisableTiming));
  for (int p = 0; p < PETITE_NPROC; ++p) {
    HIP_TRY(hipStreamCreateWithFlags(&c->dstream[p], hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c->djoin[p], hipEventDisableTiming));
  }
  return PETITE_OK;
}

// Device-planned draws of the shower pipeline: every SM process with a grid
// gets a whole-GPU launch on its stream; blocks the count does not need leave.
// Blocks of a device-planned draw launch of process p, the processes running
// side by side: Brem and pair production hold most of a shower's draws (and
// the longest accept/reject loops), so they get the most of the GPU; the
// others finish early, after which Brem runs alone at full occupancy
// (4 blocks of 4 waves per CU).  Measured on MI355X against n_cu * 2 blocks
// for every process: draws 13 % shorter.
static int draw_blocks(const PetiteCtx* c, int p) { return (int)blocks_for(c, c->tune.draw_waves[p & 7]); }

static int launch_draws_device(PetiteCtx* c, const DrawLaunch& io, uint64_t seed, hipStream_t s, bool muons) {
  int rc = ensure_draw_streams(c);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(c->dfork, s));
  auto used = [&](int p) { return c->desc.grid[p].n_E > 0 && (muons || (p != MUONE && p != MUONBREM)); };
  for (int p = 0; p < 8; ++p) {
    if (!used(p)) continue;
    HIP_TRY(hipStreamWaitEvent(c->dstream[p], c->dfork, 0));
    DrawLaunch iop = io;
    iop.full_blocks = draw_blocks(c, p);
    hipLaunchKernelGGL(kDrawKernels[p], dim3((unsigned)iop.full_blocks), dim3(256), 0, c->dstream[p], c->dev, iop, seed);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->djoin[p], c->dstream[p]));
  }
  for (int p = 0; p < 8; ++p)
    if (used(p)) HIP_TRY(hipStreamWaitEvent(s, c->djoin[p], 0));
  return PETITE_OK;
}

// Device address of `scratch` on each device, looked up once per device with
// that device current (the array's host-side name is not a device address).
static std::mutex g_noalloc_mu;
static unsigned char* g_noalloc_ptr[64] = {};

static int noalloc_scratch(int dev, unsigned char** out) {
  if (dev < 0 || dev >= 64) return fail(PETITE_ERR_ARG, "noalloc scratch: device index above 63");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(PETITE_ERR_HIP, "noalloc scratch: hipGetSymbolAddress returned a null address");
    g_noalloc_ptr[dev] = (unsigned char*)p;
  }
  *out = g_noalloc_ptr[dev];
  return PETITE_OK;
}

// Draw kernels of the active processes one after another on the caller's
// stream (no per-process streams: creating them would allocate).
static int launch_draws_serial(PetiteCtx* c, const DrawLaunch& io, uint64_t seed, hipStream_t s) {
  for (int p = 0; p < PETITE_NPROC; ++p) {
    const int b = io.block_start[p + 1] - io.block_start[p];
    if (b <= 0 || io.count[p] <= 0 || c->desc.grid[p].n_E <= 0) continue;
    hipLaunchKernelGGL(kDrawKernels[p], dim3((unsigned)b), dim3(256), 0, s, c->dev, io, seed);
    HIP_TRY(hipGetLastError());
  }
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
      hipMalloc(&c->counters, C_N * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->ddesc, sizeof(PetiteTableDesc)) != hipSuccess)
    return cleanup(fail(PETITE_ERR_HIP, "petite_ctx_create: hipMalloc failed"));
  if (hipMemcpy(c->d64, blob64, n64 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d32, blob32, n32 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c->dmaxf, 0, (size_t)PETITE_NPROC * maxE * sizeof(double)) != hipSuccess ||
      hipMemcpy(c->ddesc, desc, sizeof(PetiteTableDesc), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(PETITE_ERR_HIP, "petite_ctx_create: upload failed"));
  DevTables& T = c->dev;
  T.f64 = c->d64;
  T.f32 = c->d32;
  T.maxf = c->dmaxf;
  T.maxE = maxE;
  T.d = c->ddesc;
  const double Z = desc->Z;
  const double a0 = 184.15 * pow(2.718, -0.5) * pow(Z, -1.0 / 3.0) / kMe;
  const double a04 = a0 * a0 * a0 * a0;
  const double Ce = 8.0 / kPi * kAlpha * (kAlpha / kMe) * (kAlpha / kMe) * pow(kMe, 4) * Z * Z * a04;
  const double Cmu = 8.0 / kPi * kAlpha * (kAlpha / kMmu) * (kAlpha / kMmu) * pow(kMmu, 4) * Z * Z * a04;
  T.brem_C_e = (float)Ce;
  T.brem_C_mu = (float)Cmu;
  T.a0sq = (float)(a0 * a0);
  T.brem_a0 = a0;
  for (int p = 0; p < PETITE_NPROC; ++p) c->tbar[p] = 16.0;
  T.mcs_cA = (float)(0.157 * Z * (Z + 1.0) / desc->A);
  T.mcs_cB = (float)(2.007e-5 * pow(Z, 2.0 / 3.0));
  T.mcs_cC = (float)(3.34 * (Z * kAlpha) * (Z * kAlpha));
  {
    // n*sigma tables of the transport kernel, laid out for LDS (LdsSigma)
    auto pack = [](uint32_t lo, uint32_t hi) {
      const uint64_t b = (uint64_t)lo | ((uint64_t)hi << 32);
      double d;
      std::memcpy(&d, &b, sizeof d);
      return d;
    };
    auto fbits = [](float f) {
      uint32_t u;
      std::memcpy(&u, &f, sizeof u);
      return u;
    };
    std::vector<double> lds(kLdsHeader * PETITE_NPROC, 0.0);
    for (int p = 0; p < PETITE_NPROC; ++p) {
      const PetiteSigmaDesc& sd = desc->sigma[p];
      const int n = sd.n > 0 ? sd.n : 0;
      double* h = &lds[kLdsHeader * p];
      h[0] = pack(0u, 0u);
      h[1] = pack(0u, (uint32_t)(n - 1));
      h[2] = 1.0;  // empty range
      h[3] = 0.0;
      if (n == 0) continue;
      const double* x = blob64 + sd.E_off;
      const double* y = blob64 + sd.v_off;
      h[2] = x[0];
      h[3] = x[n - 1];
      h[4] = y[0];
      if (n == 1) continue;
      // records -1 .. n-1: record -1 and n-1 repeat records 0 and n-2
      const size_t rec_m1 = lds.size();
      for (int i = -1; i <= n - 1; ++i) {
        const int k = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
        lds.push_back(x[k]);
        lds.push_back(x[k + 1]);
        lds.push_back(y[k]);
        lds.push_back((y[k + 1] - y[k]) / (x[k + 1] - x[k]));
      }
      // the node index of a geometric grid, kept only when every node lies
      // within 0.4 of its index by the float32 constants the kernel uses
      float log2E0 = 0.0f, per = 0.0f;
      if (sd.inv_dlog > 0.0) {
        log2E0 = (float)(sd.lnE0 / std::log(2.0));
        per = (float)(sd.inv_dlog * std::log(2.0));
        for (int k = 0; k < n && per > 0.0f; ++k)
          if (!(x[k] > 0.0) || !(std::fabs((std::log2(x[k]) - (double)log2E0) * (double)per - k) < 0.4)) per = 0.0f;
      }
      // (the appends above may have moved the vector: index it afresh)
      lds[kLdsHeader * p + 0] = pack(fbits(log2E0), fbits(per));
      lds[kLdsHeader * p + 1] = pack((uint32_t)rec_m1, (uint32_t)(n - 1));
    }
    T.lds_doubles = (int32_t)lds.size();
    if (hipMalloc(&c->dsig_lds, lds.size() * sizeof(double)) != hipSuccess ||
        hipMemcpy(c->dsig_lds, lds.data(), lds.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(PETITE_ERR_HIP, "petite_ctx_create: LDS table upload failed"));
  }
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
    c->n_cu = prop.multiProcessorCount;
  c->tune = default_tuning();
  {
    const int rc = noalloc_scratch(dev, &c->noalloc);
    if (rc) return cleanup(rc);
  }
  *out = c;
  return PETITE_OK;
}

void petite_ctx_destroy(PetiteCtx* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  for (hipEvent_t& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->dfork) (void)hipEventDestroy(c->dfork);
  if (c->cstream) (void)hipStreamDestroy(c->cstream);
  for (hipEvent_t e : c->tev)
    if (e) (void)hipEventDestroy(e);
  if (c->drange) (void)hipFree(c->drange);
  if (c->dresume) (void)hipFree(c->dresume);
  for (int p = 0; p < PETITE_NPROC; ++p) {
    if (c->djoin[p]) (void)hipEventDestroy(c->djoin[p]);
    if (c->dstream[p]) (void)hipStreamDestroy(c->dstream[p]);
  }
  for (void* p : {(void*)c->d64, (void*)c->d32, (void*)c->dmaxf, (void*)c->req_slot, (void*)c->req_E, (void*)c->res_x,
                  (void*)c->res_trials, (void*)c->res_w, (void*)c->req_count, (void*)c->counters, (void*)c->ddesc,
                  (void*)c->dsig_lds, (void*)c->dark_blob, (void*)c->dyield, (void*)c->yield_blob, (void*)c->yacc, (void*)c->ylist, (void*)c->ylist_cnt, (void*)c->ytab_dev,
                  (void*)c->bnd_blob, (void*)c->bnd_tab, (void*)c->bnd_slots, (void*)c->bnd_jobs, (void*)c->bnd_bad,
                  (void*)c->par.sid, (void*)c->par.pf, (void*)c->par.pos, (void*)c->par.info, (void*)c->tail_spill})
    if (p) (void)hipFree(p);
  free_entries(c);
  delete c;
}

int petite_set_tuning(PetiteCtx* c, const PetiteTuning* in) {
  if (!c) return fail(PETITE_ERR_ARG, "petite_set_tuning: null context");
  const PetiteTuning d = default_tuning();
  if (!in) {
    c->tune = d;
    return PETITE_OK;
  }
  PetiteTuning t = *in;
  auto pick = [](auto v, auto def) { return v > 0 ? v : def; };
  t.batch_lanes = pick(t.batch_lanes, d.batch_lanes);
  t.yield_steps = pick(t.yield_steps, d.yield_steps);
  t.transport_waves = pick(t.transport_waves, d.transport_waves);
  t.emit_waves = pick(t.emit_waves, d.emit_waves);
  for (int p = 0; p < 8; ++p) t.draw_waves[p] = pick(t.draw_waves[p], d.draw_waves[p]);
  t.min_group_lanes = pick(t.min_group_lanes, d.min_group_lanes);
  t.sweep_group = pick(t.sweep_group, d.sweep_group);
  t.group_requests = pick(t.group_requests, d.group_requests);
  t.sweep_bound = pick(t.sweep_bound, d.sweep_bound);
  t.resume_cap = pick(t.resume_cap, std::min<int64_t>(d.resume_cap, t.sweep_bound));
  t.small_sweep = pick(t.small_sweep, d.small_sweep);
  t.small_yield_steps = pick(t.small_yield_steps, d.small_yield_steps);
  t.graph = pick(t.graph, d.graph);
  t.tail_sweep = pick(t.tail_sweep, d.tail_sweep);
  t.tail_waves = pick(t.tail_waves, d.tail_waves);
  t.tail_depth = pick(t.tail_depth, d.tail_depth);
  t.tail_budget = pick(t.tail_budget, d.tail_budget);
  if (t.tail_waves > 64) return fail(PETITE_ERR_ARG, "petite_set_tuning: more than 64 tail waves per CU");
  if (t.tail_depth > kTailMaxDepth) return fail(PETITE_ERR_ARG, "petite_set_tuning: tail_depth above 128");
  if (t.graph > 2) return fail(PETITE_ERR_ARG, "petite_set_tuning: graph is 1 (capture) or 2 (direct launches)");
  if (t.batch_lanes > 64) return fail(PETITE_ERR_ARG, "petite_set_tuning: batch_lanes above 64");
  if (t.transport_waves > 64 || t.emit_waves > 64) return fail(PETITE_ERR_ARG, "petite_set_tuning: more than 64 waves per CU");
  for (int p = 0; p < 8; ++p)
    if (t.draw_waves[p] > 64) return fail(PETITE_ERR_ARG, "petite_set_tuning: more than 64 draw waves per CU");
  if (t.min_group_lanes > 64 || (t.min_group_lanes & (t.min_group_lanes - 1)) != 0)
    return fail(PETITE_ERR_ARG, "petite_set_tuning: min_group_lanes is not a power of two up to 64");
  if (t.sweep_group > 64) return fail(PETITE_ERR_ARG, "petite_set_tuning: sweep_group above 64");
  if (t.group_requests > 16) return fail(PETITE_ERR_ARG, "petite_set_tuning: group_requests above 16");
  if (t.sweep_bound < 64 || t.sweep_bound > ((int64_t)1 << 26))
    return fail(PETITE_ERR_ARG, "petite_set_tuning: sweep_bound outside 64..2^26");
  if (t.resume_cap > t.sweep_bound) return fail(PETITE_ERR_ARG, "petite_set_tuning: resume_cap above sweep_bound");
  c->tune = t;
  return PETITE_OK;
}

int petite_get_tuning(const PetiteCtx* c, PetiteTuning* t) {
  if (!c || !t) return fail(PETITE_ERR_ARG, "petite_get_tuning: bad arguments");
  *t = c->tune;
  return PETITE_OK;
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

int petite_map_accumulate(PetiteCtx* c, int32_t proc, int64_t n_per_energy, uint64_t seed, uint32_t iteration,
                          double* sum_f, uint32_t* n_f, PetiteStream stream) {
  if (!c || proc < 0 || proc >= PETITE_NPROC || n_per_energy <= 0 || !sum_f || !n_f)
    return fail(PETITE_ERR_ARG, "petite_map_accumulate: bad arguments");
  const PetiteGridDesc& g = c->desc.grid[proc];
  if (g.n_E <= 0) return fail(PETITE_ERR_ARG, "petite_map_accumulate: process has no grid");
  if ((uint64_t)(iteration + 1ull) * (uint64_t)n_per_energy > 0xFFFFFFFFull)
    return fail(PETITE_ERR_ARG, "petite_map_accumulate: iteration * n_per_energy exceeds the 32-bit sample index");
  hipStream_t s = (hipStream_t)stream;
  const int64_t chunks = std::max<int64_t>(1, std::min<int64_t>((n_per_energy + 1023) / 1024, 64));
  hipLaunchKernelGGL(k_map_accumulate, dim3((unsigned)g.n_E, (unsigned)chunks), dim3(256), 0, s, c->dev, proc,
                     n_per_energy, seed, iteration, sum_f, n_f);
  HIP_TRY(hipGetLastError());
  return PETITE_OK;
}

int petite_set_grid(PetiteCtx* c, int32_t proc, const float* nodes_host, int64_t n) {
  if (!c || proc < 0 || proc >= PETITE_NPROC || !nodes_host) return fail(PETITE_ERR_ARG, "petite_set_grid: bad arguments");
  const PetiteGridDesc& g = c->desc.grid[proc];
  if (g.n_E <= 0 || n != (int64_t)g.n_E * g.stride)
    return fail(PETITE_ERR_ARG, "petite_set_grid: node count differs from the process's grid");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(c->d32 + g.node_off, nodes_host, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
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
  if (n > INT32_MAX) return fail(PETITE_ERR_ARG, "petite_vegas_draw: more than 2^31-1 requests");
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_requests(c, n);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(c->counters, 0, C_N * sizeof(unsigned long long), s));
  HIP_TRY(hipMemsetAsync(c->req_count, 0, PETITE_NPROC * sizeof(unsigned int), s));
  hipLaunchKernelGGL(k_bucket, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c->dev, proc, E, n, c->req_slot, c->req_E,
                     c->req_cap, c->req_count, trials_out, w_out);
  HIP_TRY(hipGetLastError());
  unsigned int rq[PETITE_NPROC];
  HIP_TRY(hipMemcpyAsync(rq, c->req_count, sizeof(rq), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  DrawLaunch io{};
  io.req_slot = c->req_slot;
  io.req_E = c->req_E;
  io.sid_src = sid;
  io.req_cap = c->req_cap;
  io.explicit_out = 1;
  io.x_out = x_out;
  io.trials_out = trials_out;
  io.w_out = w_out;
  io.counters = c->counters;
  int64_t blocks = 0;
  plan_draw(c, rq, io, blocks);
  rc = launch_draws(c, io, seed, s);
  if (rc) return rc;
  unsigned long long st[ST_N];
  HIP_TRY(hipMemcpyAsync(st, c->counters + C_STATS, sizeof(st), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (st[ST_NOSAMPLE]) return fail(PETITE_ERR_NO_SAMPLE, "petite_vegas_draw: No Sample Found within max_trials");
  return PETITE_OK;
}

int petite_vegas_draw_noalloc(PetiteCtx* c, const int32_t* proc, const double* E, const uint64_t* sid, int64_t n,
                              uint64_t seed, float* x_out, int32_t* trials_out, float* w_out, PetiteStream stream) {
  if (!c || n < 0 || (n > 0 && (!proc || !E || !sid || !x_out || !trials_out || !w_out)))
    return fail(PETITE_ERR_ARG, "petite_vegas_draw_noalloc: bad arguments");
  if (n > kNoAllocCap)
    return fail(PETITE_ERR_ARG, "petite_vegas_draw_noalloc: more than PETITE_NOALLOC_MAX_DRAWS (16384) requests");
  if (n == 0) return PETITE_OK;
  if (!c->noalloc) return fail(PETITE_ERR_HIP, "petite_vegas_draw_noalloc: no scratch address for this device");
  hipStream_t s = (hipStream_t)stream;
  int32_t* req_slot = (int32_t*)(c->noalloc + kNoAllocSlotOff);
  double* req_E = (double*)(c->noalloc + kNoAllocEOff);
  unsigned long long* ctr = (unsigned long long*)(c->noalloc + kNoAllocCtrOff);
  unsigned int* counts = (unsigned int*)(c->noalloc + kNoAllocCountOff);
  // the previous call's counters are still there: clear what must start at zero
  hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(128), 0, s, ctr, counts);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bucket, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c->dev, proc, E, n, req_slot, req_E,
                     kNoAllocCap, counts, trials_out, w_out);
  HIP_TRY(hipGetLastError());
  unsigned int rq[PETITE_NPROC];
  HIP_TRY(hipMemcpyAsync(rq, counts, sizeof(rq), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  DrawLaunch io{};
  io.req_slot = req_slot;
  io.req_E = req_E;
  io.sid_src = sid;
  io.req_cap = kNoAllocCap;
  io.explicit_out = 1;
  io.x_out = x_out;
  io.trials_out = trials_out;
  io.w_out = w_out;
  io.counters = ctr;
  int64_t blocks = 0;
  plan_draw(c, rq, io, blocks);
  int rc = launch_draws_serial(c, io, seed, s);
  if (rc) return rc;
  unsigned long long st[ST_N];
  HIP_TRY(hipMemcpyAsync(st, ctr + C_STATS, sizeof(st), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (st[ST_NOSAMPLE]) return fail(PETITE_ERR_NO_SAMPLE, "petite_vegas_draw_noalloc: No Sample Found within max_trials");
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
  std::memset(stats, 0, sizeof(*stats));
  // A sweep transports the particles appended since the previous one (at
  // most `bound` of them), draws their interactions and appends the children.
  // Sizes live on the device: the host queues tune.sweep_group sweeps at a
  // time and reads the counters once per group.
  const PetiteTuning& tn = c->tune;
  const int64_t bound = std::min<int64_t>(store->capacity, tn.sweep_bound);
  int rc = ensure_requests(c, bound);
  if (rc) return rc;
  rc = ensure_parent_recs(c);
  if (rc) return rc;
  if (!c->drange) HIP_TRY(hipMalloc(&c->drange, 4 * sizeof(int64_t)));
  // hand-on records per sweep: the buffer keeps its allocation, the cap in
  // use is the tuning's (at most the sweep bound)
  const int64_t rcap = std::max<int64_t>(1, std::min<int64_t>(bound, tn.resume_cap));
  if (c->resume_cap < rcap) {
    if (c->dresume) HIP_TRY(hipFree(c->dresume));
    c->dresume = nullptr;
    c->resume_cap = 0;
    HIP_TRY(hipMalloc(&c->dresume, 2 * rcap * sizeof(ResumeRec)));
    c->resume_cap = rcap;
  }
  unsigned long long* ctr = c->counters;
  HIP_TRY(hipMemsetAsync(ctr, 0, C_N * sizeof(unsigned long long), s));
  HIP_TRY(hipMemsetAsync(c->drange, 0, 4 * sizeof(int64_t), s));
  const unsigned long long n0 = (unsigned long long)n;
  HIP_TRY(hipMemcpyAsync(ctr + C_STORE, &n0, sizeof(n0), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_init_primaries, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S, prim, n, ctr + C_MUONS);
  HIP_TRY(hipGetLastError());
  // muon processes can only be requested when a primary is a muon
  unsigned long long n_muons = 0;
  HIP_TRY(hipMemcpyAsync(&n_muons, ctr + C_MUONS, sizeof(n_muons), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const bool muons = n_muons > 0;
  Requests R{c->req_slot, c->req_E, c->req_cap, c->req_count};
  DrawLaunch io{};
  io.req_slot = c->req_slot;
  io.req_E = c->req_E;
  io.sid_src = store->sid;
  io.req_sid = c->par.sid;
  io.req_cap = c->req_cap;
  io.explicit_out = 0;
  io.x_out = c->res_x;
  io.trials_out = c->res_trials;
  io.w_out = c->res_w;
  io.counters = ctr;
  io.dcount = c->req_count;
  io.full_blocks = 0;  // per process, set by launch_draws_device
  EmitLaunch em{};
  em.req_slot = c->req_slot;
  em.req_E = c->req_E;
  em.req_cap = c->req_cap;
  em.x_out = c->res_x;
  em.trials_out = c->res_trials;
  em.dcount = c->req_count;
  em.par = c->par;
  io.min_lg = lanes_log2(tn.min_group_lanes);
  io.group_requests = tn.group_requests;
  const unsigned tr_blocks = blocks_for(c, tn.transport_waves);
  const unsigned em_blocks = blocks_for(c, tn.emit_waves);
  const TransportParams TP{tn.batch_lanes, (uint32_t)tn.yield_steps};
  const TransportParams TP_small{tn.batch_lanes, (uint32_t)tn.small_yield_steps};
  const int group = tn.sweep_group;
  // small sweeps (the deep generations at the end of a run): one block per CU
  // for the transport and the emission, and the draws of every process in
  // one launch (k_draw_small); set from the last sweep of each host check
  bool small = false;
  // small-sweep launch sizes, from the size of the last sweep the host saw:
  // about one transport / emission lane per particle (at least a block per
  // CU), and draw blocks per process for about 32 requests per block (at
  // least 64, at most a big sweep's Brem blocks)
  unsigned small_tr = (unsigned)c->n_cu, small_em = (unsigned)c->n_cu;
  int small_per = 64;
  uint32_t small_procs = 0;
  for (int p = 0; p < 8; ++p)
    if (c->desc.grid[p].n_E > 0 && (muons || (p != MUONE && p != MUONBREM))) small_procs |= 1u << p;
  const size_t lds = (size_t)c->dev.lds_doubles * sizeof(double);
  // the depth-first walk of the last particles (k_tail): entered once fewer
  // than tn.tail_sweep particles wait at a host check, then one launch per
  // check until no spill is left
  bool tail = false;
  int tpar = 0;  // spill list the next walk reads
  int64_t tail_launches = 0;
  double t_tail = 0.0;
  const unsigned tail_blocks = blocks_for(c, tn.tail_waves);
  if (c->timing) {
    while ((int)c->tev.size() < 4 * group + 2 * kTailGroup) {
      hipEvent_t e;
      HIP_TRY(hipEventCreate(&e));
      c->tev.push_back(e);
    }
  }
  double t_tr = 0.0, t_dr = 0.0, t_em = 0.0;
  int64_t queued = 0;
  unsigned long long h_store = n0;
  int64_t h_range[4] = {0, 0, 0, 0};
  // One group of sweeps.  Its launches take the same arguments every time
  // (sizes live on the device), so with tn.graph a group that uses the
  // per-process draw streams is captured once per run as a HIP graph, fork
  // and join events included, and replayed for the later such groups.
  auto enqueue_group = [&](bool sm, hipStream_t s) -> int {
    for (int g = 0; g < group; ++g) {
      hipLaunchKernelGGL(k_sweep_setup, dim3(1), dim3(64), 0, s, ctr, c->drange, c->req_count, store->capacity, bound,
                         rcap);
      HIP_TRY(hipGetLastError());
      if (c->timing) HIP_TRY(hipEventRecord(c->tev[4 * g + 0], s));
      hipLaunchKernelGGL(k_transport, dim3(sm ? small_tr : tr_blocks), dim3(256), lds, s, c->dev, S,
                         (int64_t)0, (int64_t)0, seed, R,
                         c->par, shower_energy, ctr, c->dsig_lds, (const int64_t*)c->drange,
                         ResumeBuf{c->dresume, rcap}, sm ? TP_small : TP);
      HIP_TRY(hipGetLastError());
      if (c->timing) HIP_TRY(hipEventRecord(c->tev[4 * g + 1], s));
      if (sm) {
        hipLaunchKernelGGL(k_draw_small, dim3((unsigned)(8 * small_per)), dim3(256), 0, s, c->dev, io, seed, small_per,
                           small_procs);
        HIP_TRY(hipGetLastError());
      } else {
        const int rd = launch_draws_device(c, io, seed, s, muons);
        if (rd) return rd;
      }
      if (c->timing) HIP_TRY(hipEventRecord(c->tev[4 * g + 2], s));
      hipLaunchKernelGGL(k_emit, dim3(sm ? small_em : em_blocks), dim3(64 * kEmitWaves), 0, s, c->dev, S, em, seed,
                         ctr + C_STORE, shower_energy, ctr + C_STATS);
      HIP_TRY(hipGetLastError());
      if (c->timing) HIP_TRY(hipEventRecord(c->tev[4 * g + 3], s));
    }
    return PETITE_OK;
  };
  // The capture runs on an internal stream (the caller's may be the null
  // stream, which cannot capture); the graph is then launched on the caller's.
  const bool use_graph = tn.graph == 1 && !c->timing;
  hipGraphExec_t gexec = nullptr;
  auto drop_graph = [&]() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
  };
  if (use_graph) {
    rc = ensure_draw_streams(c);
    if (rc) return rc;
    if (!c->cstream) HIP_TRY(hipStreamCreate(&c->cstream));
  }
  while (true) {
    if (tail) {
      if (!c->tail_spill) HIP_TRY(hipMalloc(&c->tail_spill, 2 * kTailSpill * sizeof(int32_t)));
      // the first walk alone (it often leaves no spill), then kTailGroup per
      // host check: a launch whose inputs are empty ends at once
      const int ng = tail_launches == 0 ? 1 : kTailGroup;
      for (int g = 0; g < ng; ++g) {
        hipLaunchKernelGGL(k_tail_setup, dim3(1), dim3(64), 0, s, ctr, c->drange, store->capacity, rcap, tpar);
        HIP_TRY(hipGetLastError());
        if (c->timing) HIP_TRY(hipEventRecord(c->tev[4 * group + 2 * g], s));
        const TailParams tp{c->tail_spill + (int64_t)tpar * kTailSpill, ctr + C_SPILL + tpar,
                            c->tail_spill + (int64_t)(tpar ^ 1) * kTailSpill, ctr + C_SPILL + (tpar ^ 1), kTailSpill,
                            tn.tail_depth, tn.tail_budget};
        hipLaunchKernelGGL(k_tail, dim3(tail_blocks), dim3(256), lds, s, c->dev, S, seed, shower_energy, ctr,
                           c->dsig_lds, (const int64_t*)c->drange, ResumeBuf{c->dresume, rcap}, tp);
        HIP_TRY(hipGetLastError());
        if (c->timing) HIP_TRY(hipEventRecord(c->tev[4 * group + 2 * g + 1], s));
        hipLaunchKernelGGL(k_tail_finish, dim3(1), dim3(64), 0, s, ctr, c->drange, store->capacity);
        HIP_TRY(hipGetLastError());
        tpar ^= 1;
        ++tail_launches;
      }
      unsigned long long h_spill = 0;
      HIP_TRY(hipMemcpyAsync(&h_store, ctr + C_STORE, sizeof(h_store), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipMemcpyAsync(&h_spill, ctr + C_SPILL + tpar, sizeof(h_spill), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      if (c->timing) {
        for (int g = 0; g < ng; ++g) {
          float ms = 0.0f;
          HIP_TRY(hipEventElapsedTime(&ms, c->tev[4 * group + 2 * g], c->tev[4 * group + 2 * g + 1]));
          t_tail += ms;
        }
      }
      if ((int64_t)h_store > store->capacity) break;
      if (h_spill == 0) break;  // every particle walked
      if (tail_launches > 100000) return fail(PETITE_ERR_HIP, "petite_run_showers: tail launch runaway");
      continue;
    }
    if (use_graph && !small) {
      if (!gexec) {
        HIP_TRY(hipStreamBeginCapture(c->cstream, hipStreamCaptureModeRelaxed));
        const int rq = enqueue_group(false, c->cstream);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->cstream, &graph);
        if (rq) {
          if (graph) (void)hipGraphDestroy(graph);
          return rq;
        }
        if (ec != hipSuccess) return fail(PETITE_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
          gexec = nullptr;
          return fail(PETITE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
      }
      const hipError_t el = hipGraphLaunch(gexec, s);
      if (el != hipSuccess) {
        drop_graph();
        return fail(PETITE_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(el));
      }
    } else {
      rc = enqueue_group(small, s);
      if (rc) {
        drop_graph();
        return rc;
      }
    }
    queued += group;
    unsigned long long h_yield = 0;
    HIP_TRY(hipMemcpyAsync(&h_store, ctr + C_STORE, sizeof(h_store), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&h_yield, ctr + C_YIELD_OUT, sizeof(h_yield), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_range, c->drange, sizeof(h_range), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (c->timing) {
      for (int g = 0; g < group; ++g) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, c->tev[4 * g + 0], c->tev[4 * g + 1]));
        t_tr += ms;
        HIP_TRY(hipEventElapsedTime(&ms, c->tev[4 * g + 1], c->tev[4 * g + 2]));
        t_dr += ms;
        HIP_TRY(hipEventElapsedTime(&ms, c->tev[4 * g + 2], c->tev[4 * g + 3]));
        t_em += ms;
      }
    }
    {
      const int64_t last = h_range[1] + h_range[3];
      small = last < (int64_t)tn.small_sweep;
      const int64_t lanes_blocks = (last + 255) / 256;
      small_tr = (unsigned)std::min<int64_t>(tr_blocks, std::max<int64_t>(c->n_cu, lanes_blocks));
      small_em = (unsigned)std::min<int64_t>(em_blocks, std::max<int64_t>(c->n_cu, lanes_blocks));
      small_per = (int)std::min<int64_t>(std::max(64, draw_blocks(c, BREM)), std::max<int64_t>(64, (last + 31) / 32));
    }
    if ((int64_t)h_store > store->capacity) break;  // children were dropped: report the size reached
    if (h_range[1] == 0 && h_range[3] == 0 && (int64_t)h_store == h_range[2]) break;  // the last sweep found nothing to do
    {
      // particles waiting for the next sweep: children of the last one and its hand-ons
      const int64_t waiting = (int64_t)h_store - h_range[2] + std::min<int64_t>((int64_t)h_yield, rcap);
      if (waiting > 0 && waiting < (int64_t)tn.tail_sweep) {
        drop_graph();
        tail = true;
      }
    }
    if (queued > 1000000) {
      drop_graph();
      return fail(PETITE_ERR_HIP, "petite_run_showers: sweep count runaway");
    }
  }
  drop_graph();
  unsigned long long hs[ST_N], sweeps = 0;
  HIP_TRY(hipMemcpyAsync(hs, ctr + C_STATS, sizeof(hs), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(&sweeps, ctr + C_SWEEPS, sizeof(sweeps), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const int64_t written = (int64_t)h_store;
  stats->n_particles = written;
  stats->n_steps = (int64_t)hs[ST_STEPS];
  stats->n_draws = (int64_t)hs[ST_DRAWS];
  stats->n_trials = (int64_t)hs[ST_TRIALS];
  stats->n_maxf_viol = (int64_t)hs[ST_VIOL];
  stats->n_no_sample = (int64_t)hs[ST_NOSAMPLE];
  stats->n_iterations = (int64_t)sweeps;
  stats->n_dropped = (int64_t)hs[ST_DROPPED];
  {
    unsigned long long pt[PETITE_NPROC], pd[PETITE_NPROC], pv[PETITE_NPROC];
    HIP_TRY(hipMemcpyAsync(pt, ctr + C_PTRIALS, sizeof(pt), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(pd, ctr + C_PDRAWS, sizeof(pd), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(pv, ctr + C_PVIOL, sizeof(pv), hipMemcpyDeviceToHost, s));
    unsigned long long ts = 0;
    HIP_TRY(hipMemcpyAsync(&ts, ctr + C_TAIL_STEPS, sizeof(ts), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    stats->n_steps_tail = (int64_t)ts;
    for (int p = 0; p < PETITE_NPROC; ++p) {
      stats->n_viol_proc[p] = (int64_t)pv[p];
      stats->n_trials_proc[p] = (int64_t)pt[p];
      stats->n_draws_proc[p] = (int64_t)pd[p];
    }
  }
  stats->ms_transport = t_tr;
  stats->ms_draw = t_dr;
  stats->ms_emit = t_em;
  // every queued sweep launched each stage once (n_iterations counts the
  // sweeps with work; the trailing empty ones of the last group exit at once)
  stats->launches_transport = queued;
  stats->launches_draw = queued;
  stats->launches_emit = queued;
  stats->ms_tail = t_tail;
  stats->launches_tail = tail_launches;

  if (hs[ST_REQ_OVERFLOW])
    return fail(PETITE_ERR_INTERNAL, "petite_run_showers: draw requests overflowed their region (library bug)");
  if (written > store->capacity)
    return fail(PETITE_ERR_CAPACITY, "petite_run_showers: store capacity exceeded; n_particles holds the size reached");
  if (hs[ST_NOSAMPLE]) return fail(PETITE_ERR_NO_SAMPLE, "petite_run_showers: No Sample Found within max_trials");
  return PETITE_OK;
}

int petite_sample_scattering(PetiteCtx* c, const PetiteVertex* v, const int32_t* proc, int64_t n, uint64_t seed,
                             PetiteStore* out, int32_t* trials_out, PetiteStream stream) {
  if (!c || n < 0 || (n > 0 && (!v || !proc || !out || !trials_out)))
    return fail(PETITE_ERR_ARG, "petite_sample_scattering: bad arguments");
  if (n == 0) return PETITE_OK;
  if (n > INT32_MAX / 2) return fail(PETITE_ERR_ARG, "petite_sample_scattering: more than 2^30 vertices");
  if (out->capacity < 2 * n) return fail(PETITE_ERR_CAPACITY, "petite_sample_scattering: out holds fewer than 2n entries");
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_requests(c, n);
  if (rc) return rc;
  rc = ensure_parent_recs(c);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(c->counters, 0, C_N * sizeof(unsigned long long), s));
  HIP_TRY(hipMemsetAsync(c->req_count, 0, PETITE_NPROC * sizeof(unsigned int), s));
  Requests R{c->req_slot, c->req_E, c->req_cap, c->req_count};
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_vertex_requests, dim3(nb), dim3(256), 0, s, c->dev, v, proc, n, R, c->par.sid, trials_out);
  HIP_TRY(hipGetLastError());
  unsigned int rq[PETITE_NPROC];
  HIP_TRY(hipMemcpyAsync(rq, c->req_count, sizeof(rq), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  DrawLaunch io{};
  io.req_slot = c->req_slot;
  io.req_E = c->req_E;
  io.req_sid = c->par.sid;
  io.req_cap = c->req_cap;
  io.explicit_out = 1;
  io.x_out = c->res_x;
  io.trials_out = trials_out;
  io.w_out = c->res_w;
  io.counters = c->counters;
  int64_t blocks = 0;
  plan_draw(c, rq, io, blocks);
  rc = launch_draws(c, io, seed, s);
  if (rc) return rc;
  DevStore O{(float4*)out->p0, (float4*)out->r0, (float4*)out->pf, (float4*)out->rf, out->pid, out->parent, out->meta,
             out->shower, out->sid, out->capacity};
  hipLaunchKernelGGL(k_vertex_emit, dim3(nb), dim3(256), 0, s, c->dev, v, proc, n, (const float*)c->res_x,
                     (const int32_t*)trials_out, seed, O);
  HIP_TRY(hipGetLastError());
  unsigned long long st[ST_N];
  HIP_TRY(hipMemcpyAsync(st, c->counters + C_STATS, sizeof(st), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (st[ST_NOSAMPLE]) return fail(PETITE_ERR_NO_SAMPLE, "petite_sample_scattering: No Sample Found within max_trials");
  return PETITE_OK;
}

int petite_propagate(PetiteCtx* c, PetiteStore* recs, int64_t n, uint64_t seed, PetiteStream stream) {
  if (!c || !recs || n < 0) return fail(PETITE_ERR_ARG, "petite_propagate: bad arguments");
  if (n == 0) return PETITE_OK;
  if (n > recs->capacity || n > INT32_MAX) return fail(PETITE_ERR_ARG, "petite_propagate: n above the store capacity");
  hipStream_t s = (hipStream_t)stream;
  // the transport also queues the draw requests of the particles it
  // finishes; they go to the context's regions and are not used
  int rc = ensure_requests(c, n);
  if (rc) return rc;
  rc = ensure_parent_recs(c);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(c->counters, 0, C_N * sizeof(unsigned long long), s));
  HIP_TRY(hipMemsetAsync(c->req_count, 0, PETITE_NPROC * sizeof(unsigned int), s));
  DevStore S{(float4*)recs->p0, (float4*)recs->r0, (float4*)recs->pf, (float4*)recs->rf, recs->pid, recs->parent,
             recs->meta, recs->shower, recs->sid, recs->capacity};
  Requests R{c->req_slot, c->req_E, c->req_cap, c->req_count};
  const PetiteTuning& tn = c->tune;
  // no hand-on (no resume buffer): every particle is transported to its end in this launch
  const TransportParams TP{tn.batch_lanes, 0xFFFFFFFFu};
  const int64_t waves = (n + 63) / 64;
  const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((waves + 3) / 4, blocks_for(c, tn.transport_waves)));
  const size_t lds = (size_t)c->dev.lds_doubles * sizeof(double);
  hipLaunchKernelGGL(k_transport, dim3(blocks), dim3(256), lds, s, c->dev, S, (int64_t)0, n, seed, R, c->par,
                     (double*)nullptr, c->counters, c->dsig_lds, (const int64_t*)nullptr, ResumeBuf{nullptr, 0}, TP);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return PETITE_OK;
}

int petite_draw_pe(const uint64_t* sid, int64_t n, uint64_t seed, double Zeff, double* pe_out, PetiteStream stream) {
  if (n < 0 || (n > 0 && (!sid || !pe_out)) || !(Zeff > 0.0)) return fail(PETITE_ERR_ARG, "petite_draw_pe: bad arguments");
  if (n == 0) return PETITE_OK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_draw_pe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sid, n, seed, Zeff, pe_out);
  HIP_TRY(hipGetLastError());
  return PETITE_OK;
}

static int check_dark_desc(const PetiteDarkDesc* desc, const double* blob, int64_t n, const char* who) {
  if (desc->n_active < 0 || desc->n_active > 8) return fail(PETITE_ERR_ARG, std::string(who) + ": n_active outside 0..8");
  for (int a = 0; a < desc->n_active; ++a)
    if ((desc->active[a] < DARKBREM || desc->active[a] > DARKMUONBREM) && desc->active[a] != dark::TWO_BODY)
      return fail(PETITE_ERR_ARG, std::string(who) + ": active process is not a dark process");
  for (int f = 0; f < PETITE_DARK_FAMILIES; ++f) {
    if (desc->w_n[f] < 0 || (desc->w_n[f] > 0 && (desc->w_off[f] < 0 || desc->w_off[f] + 2 * (int64_t)desc->w_n[f] > n)))
      return fail(PETITE_ERR_ARG, std::string(who) + ": weight table outside the blob");
    const int64_t need = (int64_t)desc->d_n[f] * (2 + 2 * PETITE_DARK_DRATE_BINS);
    if (desc->d_n[f] < 0 || (desc->d_n[f] > 0 && (desc->d_off[f] < 0 || desc->d_off[f] + need > n)))
      return fail(PETITE_ERR_ARG, std::string(who) + ": drate table outside the blob");
    const double* cnt = blob + desc->d_off[f] + desc->d_n[f];
    for (int i = 0; i < desc->d_n[f]; ++i)
      if (!(cnt[i] >= 0.0 && cnt[i] <= PETITE_DARK_DRATE_BINS))
        return fail(PETITE_ERR_ARG, std::string(who) + ": drate bin count outside 0..10");
  }
  if (desc->comp_n < 0 || (desc->comp_n > 0 && (desc->comp_off < 0 || desc->comp_off + 2 * (int64_t)desc->comp_n > n)))
    return fail(PETITE_ERR_ARG, std::string(who) + ": DarkComp table outside the blob");
  return PETITE_OK;
}

// (g0, inv) of an interval guess (g - g0) * inv for an ascending grid in g =
// f(x) (ln or log10): used when every node lies within one interval of its
// guessed place, else inv = 0 (bisection).
static void grid_guess(const double* x, int n, bool log10_values, double& g0, double& inv) {
  g0 = 0.0;
  inv = 0.0;
  if (n < 3) return;
  auto g = [&](double v) { return log10_values ? v : std::log(v); };
  if (!log10_values && !(x[0] > 0.0)) return;
  const double a = g(x[0]), b = g(x[n - 1]);
  if (!(b > a)) return;
  const double iv = (double)(n - 1) / (b - a);
  for (int i = 0; i < n; ++i)
    if (std::fabs((g(x[i]) - a) * iv - (double)i) > 0.5) return;
  g0 = a;
  inv = iv;
}

int petite_dark_quadrature(const PetiteDarkQuadJob* jobs, int32_t n_jobs, const double* blob, int64_t n_blob,
                           double* out, int64_t n_out) {
  if (n_jobs < 0 || (n_jobs > 0 && (!jobs || !blob || !out)) || n_blob < 0 || n_out < 0)
    return fail(PETITE_ERR_ARG, "petite_dark_quadrature: bad arguments");
  if (n_jobs == 0) return PETITE_OK;
  auto in_blob = [&](const PetiteQTable& t, int cols) {
    return t.n >= 0 && t.off >= 0 && t.off + (int64_t)cols * t.n <= n_blob;
  };
  // device layout: blob, then the running integrals (one per job channel)
  int64_t n_cum = 0, n_int = 0;
  for (int j = 0; j < n_jobs; ++j) {
    const PetiteDarkQuadJob& q = jobs[j];
    if (q.n_loss < 0 || q.n_loss > 3 || q.n_mfp < 0 || q.n_mfp > 3 || (q.mode != 0 && q.mode != 1) || q.ns.n < 1 ||
        !in_blob(q.ns, 2) || !in_blob(q.nodes, 1) || !in_blob(q.energies, 1) || !(q.dEdx > 0.0))
      return fail(PETITE_ERR_ARG, "petite_dark_quadrature: bad job " + std::to_string(j));
    for (int c = 0; c < q.n_loss; ++c)
      if (!in_blob(q.loss[c], 2) || !in_blob(q.grid[c], 1) || q.grid[c].n < 1)
        return fail(PETITE_ERR_ARG, "petite_dark_quadrature: bad survival channel in job " + std::to_string(j));
    for (int c = 0; c < q.n_mfp; ++c)
      if (!in_blob(q.mfp[c], 2)) return fail(PETITE_ERR_ARG, "petite_dark_quadrature: bad mfp table in job " + std::to_string(j));
    const double* nd = blob + q.nodes.off;
    for (int i = 1; i < q.nodes.n; ++i)
      if (!(nd[i] > nd[i - 1])) return fail(PETITE_ERR_ARG, "petite_dark_quadrature: nodes must ascend strictly");
    const int64_t per = q.mode ? 20 : 1;
    if (q.out_off < 0 || q.out_off + per * q.energies.n > n_out)
      return fail(PETITE_ERR_ARG, "petite_dark_quadrature: results outside out in job " + std::to_string(j));
    for (int c = 0; c < q.n_loss; ++c) n_cum += q.grid[c].n;
    n_int += (q.mode ? 10 : 1) * (int64_t)q.energies.n;
  }
  double* d_blob = nullptr;
  double* d_out = nullptr;
  quad::Job* d_jobs = nullptr;
  int* d_bad = nullptr;
  auto release = [&](int rc) {
    (void)hipFree(d_blob);
    (void)hipFree(d_out);
    (void)hipFree(d_jobs);
    (void)hipFree(d_bad);
    return rc;
  };
#define QTRY(expr)                                                                                     \
  do {                                                                                                 \
    hipError_t _e = (expr);                                                                            \
    if (_e != hipSuccess) return release(fail(PETITE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e))); \
  } while (0)
  QTRY(hipMalloc(&d_blob, sizeof(double) * (size_t)(n_blob + n_cum + 1)));
  QTRY(hipMalloc(&d_out, sizeof(double) * (size_t)(n_out + 1)));
  QTRY(hipMalloc(&d_jobs, sizeof(quad::Job) * (size_t)n_jobs));
  QTRY(hipMalloc(&d_bad, sizeof(int)));
  std::vector<quad::Job> hj((size_t)n_jobs);
  int64_t cum_at = n_blob, first = 0;
  for (int j = 0; j < n_jobs; ++j) {
    const PetiteDarkQuadJob& q = jobs[j];
    quad::Job& J = hj[(size_t)j];
    J.lx = d_blob + q.ns.off;
    J.ly = d_blob + q.ns.off + q.ns.n;
    J.n_ns = q.ns.n;
    grid_guess(blob + q.ns.off, q.ns.n, true, J.ns_l0, J.ns_inv);
    J.zero_small = q.zero_small ? 1 : 0;
    J.n_loss = q.n_loss;
    J.n_mfp = q.n_mfp;
    for (int c = 0; c < 3; ++c) {
      const bool on = c < q.n_loss;
      J.loss_x[c] = on ? d_blob + q.loss[c].off : nullptr;
      J.loss_y[c] = on ? d_blob + q.loss[c].off + q.loss[c].n : nullptr;
      J.loss_n[c] = on ? q.loss[c].n : 0;
      J.grid[c] = on ? d_blob + q.grid[c].off : nullptr;
      J.grid_n[c] = on ? q.grid[c].n : 0;
      J.grid_ln0[c] = J.grid_inv[c] = 0.0;
      if (on) grid_guess(blob + q.grid[c].off, q.grid[c].n, false, J.grid_ln0[c], J.grid_inv[c]);
      J.cum[c] = on ? d_blob + cum_at : nullptr;
      if (on) cum_at += q.grid[c].n;
      const bool mon = c < q.n_mfp;
      J.mfp_x[c] = mon ? d_blob + q.mfp[c].off : nullptr;
      J.mfp_y[c] = mon ? d_blob + q.mfp[c].off + q.mfp[c].n : nullptr;
      J.mfp_n[c] = mon ? q.mfp[c].n : 0;
    }
    J.n_nodes = q.nodes.n;
    J.nodes = d_blob + q.nodes.off;
    J.E = d_blob + q.energies.off;
    J.n_E = q.energies.n;
    J.mode = q.mode;
    J.lo = q.lo;
    J.dEdx = q.dEdx;
    J.out = d_out + q.out_off;
    J.first = first;
    first += (q.mode ? 10 : 1) * (int64_t)q.energies.n;
  }
  hipStream_t s = nullptr;
  QTRY(hipMemcpyAsync(d_blob, blob, sizeof(double) * (size_t)n_blob, hipMemcpyHostToDevice, s));
  QTRY(hipMemcpyAsync(d_jobs, hj.data(), sizeof(quad::Job) * (size_t)n_jobs, hipMemcpyHostToDevice, s));
  QTRY(hipMemsetAsync(d_bad, 0, sizeof(int), s));
  QTRY(hipMemsetAsync(d_out, 0, sizeof(double) * (size_t)n_out, s));
  hipLaunchKernelGGL(quad::k_cumint, dim3((unsigned)((3 * n_jobs + 63) / 64)), dim3(64), 0, s, d_jobs, n_jobs);
  QTRY(hipGetLastError());
  if (n_int > 0) {
    hipLaunchKernelGGL(quad::k_track_quad, dim3((unsigned)((n_int + quad::kWavesPerBlock - 1) / quad::kWavesPerBlock)),
                       dim3(64 * quad::kWavesPerBlock), 0, s, d_jobs, n_jobs, n_int, d_bad);
    QTRY(hipGetLastError());
  }
  int bad = 0;
  QTRY(hipMemcpyAsync(out, d_out, sizeof(double) * (size_t)n_out, hipMemcpyDeviceToHost, s));
  QTRY(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
  QTRY(hipStreamSynchronize(s));
#undef QTRY
  if (bad) return release(fail(PETITE_ERR_INTERNAL, "petite_dark_quadrature: an integral needs more than 512 geometric panel edges"));
  return release(PETITE_OK);
}

int petite_dark_setup(PetiteCtx* c, const PetiteDarkDesc* desc, const double* blob, int64_t n) {
  if (!c || !desc || !blob || n <= 0) return fail(PETITE_ERR_ARG, "petite_dark_setup: bad arguments");
  if (desc->n_active < 0 || desc->n_active > 8) return fail(PETITE_ERR_ARG, "petite_dark_setup: n_active outside 0..8");
  for (int a = 0; a < desc->n_active; ++a)
    if (desc->active[a] >= DARKBREM && desc->active[a] <= DARKMUONBREM && c->desc.grid[desc->active[a]].n_E <= 0)
      return fail(PETITE_ERR_ARG, "petite_dark_setup: active process has no grid in the context");
  {
    const int rc = check_dark_desc(desc, blob, n, "petite_dark_setup");
    if (rc) return rc;
  }
  if (c->dark_blob) {
    (void)hipFree(c->dark_blob);
    c->dark_blob = nullptr;
  }
  HIP_TRY(hipMalloc(&c->dark_blob, n * sizeof(double)));
  HIP_TRY(hipMemcpy(c->dark_blob, blob, n * sizeof(double), hipMemcpyHostToDevice));
  if (!c->dyield) HIP_TRY(hipMalloc(&c->dyield, 5 * sizeof(double)));
  c->dark = *desc;
  c->has_dark = true;
  return PETITE_OK;
}

int petite_dark_reweight(PetiteCtx* c, const PetiteStore* sm, int64_t n_sm, uint64_t seed, PetiteDarkOut* out,
                         PetiteDarkStats* stats, PetiteStream stream) {
  if (!c || !sm || n_sm < 0 || !out || !stats) return fail(PETITE_ERR_ARG, "petite_dark_reweight: bad arguments");
  if (!c->has_dark) return fail(PETITE_ERR_ARG, "petite_dark_reweight: no dark tables (petite_dark_setup)");
  if (n_sm > sm->capacity) return fail(PETITE_ERR_ARG, "petite_dark_reweight: n_sm above the store capacity");
  std::memset(stats, 0, sizeof(*stats));
  if (n_sm == 0 || c->dark.n_active == 0) return PETITE_OK;
  hipStream_t s = (hipStream_t)stream;
  DevStore S{(float4*)sm->p0, (float4*)sm->r0, (float4*)sm->pf, (float4*)sm->rf, sm->pid, sm->parent, sm->meta,
             sm->shower, sm->sid, sm->capacity};
  const int na = c->dark.n_active;
  // SM records per pass: entry scratch stays below 2^27 entries
  const int64_t chunk = std::min<int64_t>(n_sm, std::max<int64_t>(1, ((int64_t)1 << 27) / na));
  int rc = ensure_entries(c, chunk * na);
  if (rc) return rc;
  rc = ensure_requests(c, chunk * na);
  if (rc) return rc;
  unsigned long long* ctr = c->counters;
  HIP_TRY(hipMemsetAsync(c->dyield, 0, 5 * sizeof(double), s));
  HIP_TRY(hipMemsetAsync(ctr, 0, C_N * sizeof(unsigned long long), s));
  int64_t out_base = 0;
  const bool mesons = c->dark.meson_w[0] > 0.0 || c->dark.meson_w[1] > 0.0 || c->dark.meson_w[2] > 0.0;
  // the select / draw / emit pass serves the processes of SM particles (codes
  // 8..11); with TwoBody_BSMDecay alone only the meson pass runs
  bool sm_dark = false;
  for (int a = 0; a < c->dark.n_active; ++a) sm_dark = sm_dark || c->dark.active[a] != dark::TWO_BODY;
  double t_sel = 0.0, t_dr = 0.0, t_em = 0.0, t_me = 0.0;
  int64_t launches = 0;
  for (int64_t first = 0; first < n_sm; first += chunk) {
    const int64_t n = std::min(chunk, n_sm - first);
    unsigned long long n_e = 0;
    if (sm_dark) {
      HIP_TRY(hipMemsetAsync(c->req_count, 0, PETITE_NPROC * sizeof(unsigned int), s));
      HIP_TRY(hipMemsetAsync(ctr + C_DARK_N, 0, sizeof(unsigned long long), s));
      HIP_TRY(hipMemsetAsync(ctr + C_REFILL, 0, kCtrLine * PETITE_NPROC * sizeof(unsigned long long), s));  // draw refill
      Requests R{c->req_slot, c->req_E, c->req_cap, c->req_count};
      if (c->timing) HIP_TRY(hipEventRecord(c->ev[0], s));
      hipLaunchKernelGGL(dark::k_dark_select, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c->dev, c->dark,
                         (const double*)c->dark_blob, S, first, n, seed, c->ent, R, ctr + C_DARK_N, ctr + C_DARK_STATS);
      HIP_TRY(hipGetLastError());
      if (c->timing) HIP_TRY(hipEventRecord(c->ev[1], s));
      unsigned int rq[PETITE_NPROC];
      HIP_TRY(hipMemcpyAsync(rq, c->req_count, sizeof(rq), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipMemcpyAsync(&n_e, ctr + C_DARK_N, sizeof(n_e), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      if ((int64_t)n_e > c->ent.cap) return fail(PETITE_ERR_HIP, "petite_dark_reweight: entry scratch overflow");
      stats->n_entries += (int64_t)n_e;
      if (out_base + (int64_t)n_e > out->capacity) {
        return fail(PETITE_ERR_CAPACITY, "petite_dark_reweight: output smaller than the entries (n_entries holds the count so far)");
      }
      DrawLaunch io{};
      io.req_slot = c->req_slot;
      io.req_E = c->req_E;
      io.sid_src = c->ent.dsid;
      io.req_cap = c->req_cap;
      io.explicit_out = 1;
      io.x_out = c->dres_x;
      io.trials_out = c->dres_trials;
      io.w_out = c->dres_w;
      io.counters = ctr;
      int64_t blocks = 0;
      plan_draw(c, rq, io, blocks);
      if (c->timing) HIP_TRY(hipEventRecord(c->ev[2], s));
      rc = launch_draws(c, io, seed, s);
      if (rc) return rc;
      if (c->timing) HIP_TRY(hipEventRecord(c->ev[3], s));
      if (n_e > 0) {
        dark::EmitOut O{(float4*)out->p, (float4*)out->r, out->parent, out->meta};
        hipLaunchKernelGGL(dark::k_dark_emit, dim3((unsigned)((n_e + 255) / 256)), dim3(256), 0, s, c->dev, c->dark, S,
                           c->ent, (int64_t)n_e, (const float*)c->dres_x, (const int32_t*)c->dres_trials, O, out_base,
                           seed, c->dyield, ctr + C_DARK_STATS);
        HIP_TRY(hipGetLastError());
      }
      if (c->timing) {
        HIP_TRY(hipEventRecord(c->ev[4], s));
        HIP_TRY(hipEventSynchronize(c->ev[4]));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        t_sel += ms;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        t_dr += ms;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[3], c->ev[4]));
        t_em += ms;
      }
    }
    ++launches;
    out_base += (int64_t)n_e;
    if (mesons) {
      HIP_TRY(hipMemsetAsync(ctr + C_MESON_N, 0, sizeof(unsigned long long), s));
      dark::EmitOut O{(float4*)out->p, (float4*)out->r, out->parent, out->meta};
      if (c->timing) HIP_TRY(hipEventRecord(c->ev[0], s));
      hipLaunchKernelGGL(dark::k_meson_decay, dim3((unsigned)((n + dark::kMesonBlock - 1) / dark::kMesonBlock)),
                         dim3(dark::kMesonBlock), 0, s, c->dev, c->dark, S,
                         first, n, seed, O, out_base, out->capacity, ctr + C_MESON_N, c->dyield, ctr + C_DARK_STATS);
      HIP_TRY(hipGetLastError());
      if (c->timing) HIP_TRY(hipEventRecord(c->ev[1], s));
      unsigned long long n_m = 0;
      HIP_TRY(hipMemcpyAsync(&n_m, ctr + C_MESON_N, sizeof(n_m), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      if (c->timing) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        t_me += ms;
      }
      stats->n_entries += (int64_t)n_m;
      if (out_base + (int64_t)n_m > out->capacity)
        return fail(PETITE_ERR_CAPACITY, "petite_dark_reweight: output smaller than the entries (n_entries holds the count so far)");
      out_base += (int64_t)n_m;
    }
  }
  unsigned long long hs[ST_N], dst[3];
  HIP_TRY(hipMemcpyAsync(hs, ctr + C_STATS, sizeof(hs), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(dst, ctr + C_DARK_STATS, sizeof(dst), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(stats->yield, c->dyield, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  stats->n_vectors = (int64_t)dst[1];
  stats->n_zero_rate = (int64_t)dst[0];
  stats->n_pe_fail = (int64_t)dst[2];
  stats->n_draws = (int64_t)hs[ST_DRAWS];
  stats->n_trials = (int64_t)hs[ST_TRIALS];
  stats->n_no_sample = (int64_t)hs[ST_NOSAMPLE];
  stats->ms_select = t_sel;
  stats->ms_meson = t_me;
  stats->ms_draw = t_dr;
  stats->ms_emit = t_em;
  stats->launches = launches;
  return PETITE_OK;
}

int petite_yield_setup(PetiteCtx* c, int32_t n_mass, const double* masses, const PetiteDarkDesc* descs,
                       const double* blob, int64_t n) {
  if (!c || !masses || !descs || !blob || n <= 0 || n_mass < 1 || n_mass > PETITE_MAX_MASSES)
    return fail(PETITE_ERR_ARG, "petite_yield_setup: bad arguments");
  for (int j = 0; j < n_mass; ++j) {
    if (!(masses[j] > 0.0) || (j > 0 && !(masses[j] > masses[j - 1])))
      return fail(PETITE_ERR_ARG, "petite_yield_setup: masses must be positive and ascending");
    int rc = check_dark_desc(descs + j, blob, n, "petite_yield_setup");
    if (rc) return rc;
    if (descs[j].n_active != descs[0].n_active ||
        std::memcmp(descs[j].active, descs[0].active, sizeof(descs[0].active)) != 0)
      return fail(PETITE_ERR_ARG, "petite_yield_setup: the masses must share the active processes");
  }
  if (c->yield_blob) {
    (void)hipFree(c->yield_blob);
    c->yield_blob = nullptr;
  }
  HIP_TRY(hipMalloc(&c->yield_blob, n * sizeof(double)));
  HIP_TRY(hipMemcpy(c->yield_blob, blob, n * sizeof(double), hipMemcpyHostToDevice));
  if (!c->yacc) HIP_TRY(hipMalloc(&c->yacc, 2 * dark::kYieldBatch * sizeof(double)));
  c->n_mass = n_mass;
  for (int j = 0; j < n_mass; ++j) {
    c->masses[j] = masses[j];
    c->ytab.d[j] = descs[j];
  }
  c->ytab.n_mass = n_mass;
  c->ytab.b = c->yield_blob;
  c->bnd_on = false;  // a new model brings its own bound-electron inputs (petite_yield_bound)
  if (!c->ytab_dev) HIP_TRY(hipMalloc(&c->ytab_dev, sizeof(dark::YieldTables)));
  HIP_TRY(hipMemcpy(c->ytab_dev, &c->ytab, sizeof(dark::YieldTables), hipMemcpyHostToDevice));
  return PETITE_OK;
}

int petite_yield_bound(PetiteCtx* c, const PetiteBoundDesc* d, const double* blob, int64_t n) {
  if (!c || !d || !blob || n <= 0) return fail(PETITE_ERR_ARG, "petite_yield_bound: bad arguments");
  if (c->n_mass < 1) return fail(PETITE_ERR_ARG, "petite_yield_bound: set the trained masses first (petite_yield_setup)");
  auto in_blob = [&](const PetiteQTable& t, int cols) { return t.n >= 1 && t.off >= 0 && t.off + (int64_t)cols * t.n <= n; };
  for (int q = 0; q < 3; ++q)
    if (!in_blob(d->loss[q], 2) || !in_blob(d->mfp[q], 2)) return fail(PETITE_ERR_ARG, "petite_yield_bound: bad table");
  if (!in_blob(d->nodes, 1) || !(d->rel_step > 0.0 && d->rel_step < 0.01) || !(d->dEdx > 0.0) || !(d->Zeff > 0.0) ||
      !(d->ne_conv > 0.0) || !(d->m_e > 0.0) || !(d->alpha > 0.0))
    return fail(PETITE_ERR_ARG, "petite_yield_bound: bad parameters");
  for (int j = 0; j < c->n_mass; ++j)
    if (!(d->top_ann[j] > 0.0) || !(d->top_comp[j] > 0.0)) return fail(PETITE_ERR_ARG, "petite_yield_bound: bad grid tops");
  const double* nd = blob + d->nodes.off;
  for (int i = 1; i < d->nodes.n; ++i)
    if (!(nd[i] > nd[i - 1])) return fail(PETITE_ERR_ARG, "petite_yield_bound: nodes must ascend strictly");
  for (void* p : {(void*)c->bnd_blob, (void*)c->bnd_tab, (void*)c->bnd_slots, (void*)c->bnd_jobs, (void*)c->bnd_bad})
    if (p) (void)hipFree(p);
  c->bnd_blob = c->bnd_tab = nullptr;
  c->bnd_slots = nullptr;
  c->bnd_jobs = nullptr;
  c->bnd_bad = nullptr;
  c->bnd_on = false;
  int64_t n_cum = 0;
  for (int q = 0; q < 3; ++q) n_cum += d->loss[q].n;
  const int stride = quad::kBoundTab + d->nodes.n + quad::kBoundGrid;
  HIP_TRY(hipMalloc(&c->bnd_blob, sizeof(double) * (size_t)(n + n_cum)));
  HIP_TRY(hipMalloc(&c->bnd_tab, sizeof(double) * (size_t)stride * quad::kBoundMaxSlots));
  HIP_TRY(hipMalloc(&c->bnd_slots, sizeof(quad::BoundSlot) * quad::kBoundMaxSlots));
  HIP_TRY(hipMalloc(&c->bnd_jobs, sizeof(quad::Job) * quad::kBoundMaxSlots));
  HIP_TRY(hipMalloc(&c->bnd_bad, sizeof(int)));
  HIP_TRY(hipMemset(c->bnd_bad, 0, sizeof(int)));
  HIP_TRY(hipMemcpy(c->bnd_blob, blob, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
  // one job per slot; every slot shares the survival channels and their
  // running integrals (computed once here by the first job)
  std::vector<quad::Job> jobs(quad::kBoundMaxSlots);
  int64_t cum_at = n;
  double* cum[3];
  for (int q = 0; q < 3; ++q) {
    cum[q] = c->bnd_blob + cum_at;
    cum_at += d->loss[q].n;
  }
  for (int sl = 0; sl < quad::kBoundMaxSlots; ++sl) {
    quad::Job& J = jobs[(size_t)sl];
    double* t = c->bnd_tab + (int64_t)sl * stride;
    J.lx = t + quad::kBoundGrid;
    J.ly = t + 2 * quad::kBoundGrid;
    J.n_ns = quad::kBoundGrid;
    J.zero_small = 0;  // bound-electron DarkAnn keeps n sigma below 1e-18 (dark_shower.py:311-336)
    J.n_loss = 3;
    J.n_mfp = 3;
    for (int q = 0; q < 3; ++q) {
      J.loss_x[q] = c->bnd_blob + d->loss[q].off;
      J.loss_y[q] = c->bnd_blob + d->loss[q].off + d->loss[q].n;
      J.loss_n[q] = d->loss[q].n;
      J.grid[q] = J.loss_x[q];
      J.grid_n[q] = d->loss[q].n;
      grid_guess(blob + d->loss[q].off, d->loss[q].n, false, J.grid_ln0[q], J.grid_inv[q]);
      J.cum[q] = cum[q];
      J.mfp_x[q] = c->bnd_blob + d->mfp[q].off;
      J.mfp_y[q] = c->bnd_blob + d->mfp[q].off + d->mfp[q].n;
      J.mfp_n[q] = d->mfp[q].n;
    }
    J.n_nodes = 0;  // k_bound_nodes
    J.nodes = t + quad::kBoundTab;
    J.E = t;
    J.n_E = quad::kBoundGrid;
    J.mode = 0;
    J.lo = 0.0;  // k_bound_tables
    J.dEdx = d->dEdx;
    J.out = t + 3 * quad::kBoundGrid;
    J.first = (int64_t)sl * quad::kBoundGrid;
  }
  HIP_TRY(hipMemcpy(c->bnd_jobs, jobs.data(), sizeof(quad::Job) * jobs.size(), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(quad::k_cumint, dim3(1), dim3(64), 0, nullptr, c->bnd_jobs, 1);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  c->bnd.Zeff = d->Zeff;
  c->bnd.ne_conv = d->ne_conv;
  c->bnd.m_e = d->m_e;
  c->bnd.alpha = d->alpha;
  c->bnd.ann_floor = d->ann_floor;
  c->bnd.comp_floor = d->comp_floor;
  c->bnd.n_static = d->nodes.n;
  c->bnd.node_stride = d->nodes.n + quad::kBoundGrid;
  c->bnd_h = d->rel_step;
  for (int j = 0; j < c->n_mass; ++j) {
    c->bnd_top_ann[j] = d->top_ann[j];
    c->bnd_top_comp[j] = d->top_comp[j];
  }
  // the fixed nodes live in the blob at nodes.off
  c->bnd_on = true;
  c->bnd_fixed = c->bnd_blob + d->nodes.off;
  return PETITE_OK;
}

int petite_dark_yield(PetiteCtx* c, const PetiteStore* sm, int64_t n_sm, int32_t n_points, const double* mV,
                      double* yield_host, double* dyield_host, PetiteStream stream) {
  if (!c || !sm || n_sm < 0 || n_points < 0 || (n_points > 0 && (!mV || !yield_host)))
    return fail(PETITE_ERR_ARG, "petite_dark_yield: bad arguments");
  if (c->n_mass < 1) return fail(PETITE_ERR_ARG, "petite_dark_yield: no trained masses (petite_yield_setup)");
  if (n_sm > sm->capacity) return fail(PETITE_ERR_ARG, "petite_dark_yield: n_sm above the store capacity");
  const int nm = c->n_mass;
  for (int k = 0; k < n_points; ++k) {
    if (!(mV[k] > 0.0)) return fail(PETITE_ERR_ARG, "petite_dark_yield: m_V must be positive");
    if (mV[k] < c->masses[0] || mV[k] > c->masses[nm - 1]) {
      char msg[200];
      std::snprintf(msg, sizeof(msg), "petite_dark_yield: m_V = %.6g GeV outside the trained masses [%.6g, %.6g] GeV",
                    mV[k], c->masses[0], c->masses[nm - 1]);
      return fail(PETITE_ERR_ARG, msg);
    }
  }
  hipStream_t s = (hipStream_t)stream;
  DevStore S{(float4*)sm->p0, (float4*)sm->r0, (float4*)sm->pf, (float4*)sm->rf, sm->pid, sm->parent, sm->meta,
             sm->shower, sm->sid, sm->capacity};
  if (n_sm > INT32_MAX) return fail(PETITE_ERR_ARG, "petite_dark_yield: more than 2^31-1 records");
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n_sm + 255) / 256, (int64_t)c->n_cu * 8));
  bool use_bound = false;
  for (int a = 0; a < c->ytab.d[0].n_active; ++a)
    use_bound |= c->ytab.d[0].active[a] == DARKANN || c->ytab.d[0].active[a] == DARKCOMP;
  // the records each kernel visits, by species (bits: e-, e+, photon, mu-, mu+):
  // k_dark_yield takes the interpolated processes' species, k_bound_yield
  // positrons (DarkAnn) and photons (DarkComp) when the tables at each m_V
  // carry those processes
  const bool bound_all = c->bnd_on && use_bound;
  uint32_t mask_a = 0u, mask_b = 0u;
  for (int a = 0; a < c->ytab.d[0].n_active; ++a) {
    const int code = c->ytab.d[0].active[a];
    const bool bnd = code == DARKANN || code == DARKCOMP;
    if (bound_all && bnd) {
      mask_b |= code == DARKANN ? 2u : 4u;
      continue;
    }
    if (code == DARKBREM) mask_a |= 1u | 2u;
    else if (code == DARKANN) mask_a |= 2u;
    else if (code == DARKMUONBREM) mask_a |= 8u | 16u;
    else if (code == DARKCOMP) mask_a |= 4u;
  }
  if (n_sm > 0) {
    if (c->ylist_cap < n_sm) {
      if (c->ylist) HIP_TRY(hipFree(c->ylist));
      c->ylist = nullptr;
      c->ylist_cap = 0;
      HIP_TRY(hipMalloc(&c->ylist, 2 * n_sm * sizeof(int32_t)));
      c->ylist_cap = n_sm;
    }
    if (!c->ylist_cnt) HIP_TRY(hipMalloc(&c->ylist_cnt, 2 * sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(c->ylist_cnt, 0, 2 * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(dark::k_yield_lists, dim3((unsigned)blocks), dim3(256), 0, s, S, n_sm, mask_a, mask_b, c->ylist,
                       c->ylist + c->ylist_cap, c->ylist_cnt);
    HIP_TRY(hipGetLastError());
  }
  const dark::RecordList list_a{c->ylist, c->ylist_cnt}, list_b{c->ylist + c->ylist_cap, c->ylist_cnt + 1};
  // passes over the points in ascending m_V, so that a pass needs few trained masses
  std::vector<int> order(n_points);
  for (int k = 0; k < n_points; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return mV[a] < mV[b]; });
  // segment and interpolation parameter of every point, ascending m_V
  std::vector<int> seg(n_points);
  std::vector<double> tt(n_points), dd(n_points);
  for (int q = 0; q < n_points; ++q) {
    const double m = mV[order[q]];
    int j = 0;
    while (j + 1 < nm - 1 && c->masses[j + 1] <= m) ++j;
    double t = 0.0, dtdm = 0.0;
    if (nm > 1) {
      const double dl = std::log(c->masses[j + 1] / c->masses[j]);
      t = std::log(m / c->masses[j]) / dl;
      dtdm = 1.0 / (m * dl);
      // inside the trained range (checked above) t leaves [0, 1] only by rounding
      t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    }
    seg[q] = j;
    tt[q] = t;
    dd[q] = dtdm;
  }
  // passes of at most kYieldBatch points whose segments fit one window of
  // kYieldWindow consecutive trained masses
  for (int k0 = 0; k0 < n_points;) {
    dark::YieldPoints P{};
    P.j0 = seg[k0];
    P.n = 0;
    while (P.n < dark::kYieldBatch && k0 + P.n < n_points && seg[k0 + P.n] + 1 < P.j0 + dark::kYieldWindow) ++P.n;
    for (int q = 0; q < dark::kYieldWindow; ++q)
      P.ln_eps2[q] = P.j0 + q < nm ? std::log(c->ytab.d[P.j0 + q].eps2) : 0.0;
    for (int k = 0; k < P.n; ++k) {
      const int j = seg[k0 + k] - P.j0;
      P.j[k] = j;
      P.t[k] = tt[k0 + k];
      P.dtdm[k] = dd[k0 + k];
      P.need |= 1u << j;
      if (P.j0 + j + 1 < nm) P.need |= 1u << (j + 1);
    }
    dark::BoundView Bv{};
    if (c->bnd_on && use_bound) {
      // tables at each point's m_V (and m_V (1 +- h) for the derivative),
      // the grid tops of the trained mass at or below it
      const int nv = dyield_host ? 3 : 1;
      quad::BoundPass BP{};
      BP.n_slots = P.n * nv;
      for (int k = 0; k < P.n; ++k) {
        const double m = mV[order[k0 + k]];
        int j = seg[k0 + k];
        if (j + 1 < nm && m >= c->masses[j + 1]) ++j;
        for (int v = 0; v < nv; ++v) {
          const int sl = k * nv + v;
          BP.mV[sl] = v == 0 ? m : (v == 1 ? m * (1.0 + c->bnd_h) : m * (1.0 - c->bnd_h));
          BP.top_ann[sl] = c->bnd_top_ann[j];
          BP.top_comp[sl] = c->bnd_top_comp[j];
        }
      }
      hipLaunchKernelGGL(quad::k_bound_tables, dim3((unsigned)BP.n_slots), dim3(256), 0, s, c->bnd, BP, c->bnd_slots,
                         c->bnd_tab, c->bnd_jobs);
      HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(quad::k_bound_nodes, dim3(1), dim3(64), 0, s, c->bnd, BP.n_slots, c->bnd_fixed, c->bnd_tab,
                         c->bnd_jobs);
      HIP_TRY(hipGetLastError());
      const int64_t n_int = (int64_t)BP.n_slots * quad::kBoundGrid;
      hipLaunchKernelGGL(quad::k_track_quad, dim3((unsigned)((n_int + quad::kWavesPerBlock - 1) / quad::kWavesPerBlock)),
                         dim3(64 * quad::kWavesPerBlock), 0, s, (const quad::Job*)c->bnd_jobs, BP.n_slots, n_int,
                         c->bnd_bad);
      HIP_TRY(hipGetLastError());
      if (nv == 3) {
        quad::BoundDeriv D{};
        D.n_pts = P.n;
        D.stride = quad::kBoundTab + c->bnd.node_stride;
        for (int k = 0; k < P.n; ++k) D.inv_2h[k] = 1.0 / (2.0 * c->bnd_h * mV[order[k0 + k]]);
        hipLaunchKernelGGL(quad::k_bound_deriv, dim3((unsigned)P.n), dim3(256), 0, s, D, c->bnd_tab);
        HIP_TRY(hipGetLastError());
      }
      Bv.slot = c->bnd_slots;
      Bv.tab = c->bnd_tab;
      Bv.stride = quad::kBoundTab + c->bnd.node_stride;
      Bv.nv = nv;
    }
    HIP_TRY(hipMemsetAsync(c->yacc, 0, 2 * dark::kYieldBatch * sizeof(double), s));
    if (n_sm > 0) {
      const bool bound_on = Bv.slot != nullptr;
      if (dyield_host)
        hipLaunchKernelGGL(dark::k_dark_yield<true>, dim3((unsigned)blocks), dim3(256), 0, s, c->dev,
                           (const dark::YieldTables*)c->ytab_dev, S, list_a, P, bound_on, c->yacc);
      else
        hipLaunchKernelGGL(dark::k_dark_yield<false>, dim3((unsigned)blocks), dim3(256), 0, s, c->dev,
                           (const dark::YieldTables*)c->ytab_dev, S, list_a, P, bound_on, c->yacc);
      HIP_TRY(hipGetLastError());
      if (bound_on) {
        int ann = 0, comp = 0;
        for (int a = 0; a < c->ytab.d[0].n_active; ++a) {
          ann |= c->ytab.d[0].active[a] == DARKANN;
          comp |= c->ytab.d[0].active[a] == DARKCOMP;
        }
        if (dyield_host)
          hipLaunchKernelGGL(dark::k_bound_yield<true>, dim3((unsigned)blocks), dim3(256), 0, s, c->dev, S, list_b, P.n,
                             ann, comp, P.ln_eps2[0], Bv, c->yacc);
        else
          hipLaunchKernelGGL(dark::k_bound_yield<false>, dim3((unsigned)blocks), dim3(256), 0, s, c->dev, S, list_b, P.n,
                             ann, comp, P.ln_eps2[0], Bv, c->yacc);
        HIP_TRY(hipGetLastError());
      }
    }
    double h[2 * dark::kYieldBatch];
    HIP_TRY(hipMemcpyAsync(h, c->yacc, sizeof(h), hipMemcpyDeviceToHost, s));
    int bad = 0;
    if (Bv.slot) HIP_TRY(hipMemcpyAsync(&bad, c->bnd_bad, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (bad) return fail(PETITE_ERR_INTERNAL, "petite_dark_yield: a bound-electron weight needs more than 512 panel edges");
    for (int k = 0; k < P.n; ++k) {
      yield_host[order[k0 + k]] = h[2 * k];
      if (dyield_host) dyield_host[order[k0 + k]] = h[2 * k + 1];
    }
    k0 += P.n;
  }
  return PETITE_OK;
}

int petite_detector_cut(const float* p, const float* r, int64_t n, const double* det_z, int32_t n_det, double radius,
                        double inner_radius, int32_t has_ecut, double emin, double emax, uint64_t* mask, double* wsum,
                        PetiteStream stream) {
  if (n < 0 || n_det < 1 || n_det > PETITE_MAX_DETECTORS || !det_z || (n > 0 && (!p || !r)))
    return fail(PETITE_ERR_ARG, "petite_detector_cut: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  DetectorPlanes D{};
  D.n = n_det;
  D.has_ecut = has_ecut ? 1 : 0;
  D.radius = radius;
  D.inner = inner_radius;
  D.emin = emin;
  D.emax = emax;
  for (int d = 0; d < n_det; ++d) D.z[d] = det_z[d];
  if (wsum) HIP_TRY(hipMemsetAsync(wsum, 0, (size_t)(n_det + 1) * sizeof(double), s));
  if (n > 0) {
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    if (n_det <= kRegPlanes)
      hipLaunchKernelGGL(k_detector_cut<true>, dim3((unsigned)blocks), dim3(256), 0, s, (const float4*)p,
                         (const float4*)r, n, D, mask, wsum);
    else
      hipLaunchKernelGGL(k_detector_cut<false>, dim3((unsigned)blocks), dim3(256), 0, s, (const float4*)p,
                         (const float4*)r, n, D, mask, wsum);
    HIP_TRY(hipGetLastError());
  }
  return PETITE_OK;
}

}  // extern "C"
