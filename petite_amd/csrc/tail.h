// The last generations of a run as depth-first walks (k_tail).
//
// Sweeps advance a batch one generation at a time, so a run takes as many
// sweeps as its deepest chain of generations, and every sweep costs a fixed
// latency however few particles it holds: a transport launch whose slowest
// particle takes its substeps one after another, a draw launch whose slowest
// request takes its candidate rounds, an emission launch.  A 400 GeV muon's
// ~2,000 MuonE vertices are 2,000 generations.  Once few particles remain, a
// wave instead takes one particle and walks its whole subtree: transport to
// the next hard interaction (propagate_particle, shower.py:509-601), the
// process choice (665-698), the accept/reject draw with the wave's 64 lanes
// on candidates j .. j + 63 (draw_sample, 401-465; the lowest accepted index
// is the sequential loop's candidate), and the children (sample_scattering,
// 467-507) appended to the store.  The kept children go on the wave's stack,
// the lower-energy one on top, so it is walked first: its energy is at most
// half the parent's (plus m_e), so the stack never holds more than about
// log2(E_primary / min_energy) + 1 entries.  A full stack (a depth limit
// below that) passes the child to a spill list, which the next launch walks.
// Every random number comes from the particle's own Philox stream, as in the
// sweeps, so the records are the sweeps' records, bit for bit.
//
// The walk is wave-uniform: every lane computes the same transport (the
// loads are the same addresses) and lane 0 writes; only the draw spreads the
// candidates over the lanes.
#pragma once

constexpr int kTailMaxDepth = 128;  // two stack entries per lane
constexpr int kTailGroup = 8;       // walk launches queued per host check after the first
constexpr int64_t kTailSpill = 1 << 20;  // entries per spill list

struct TailParams {
  const int32_t* spill_in;            // slots left by the previous launch
  const unsigned long long* n_spill_in;
  int32_t* spill_out;                 // slots past a full stack
  unsigned long long* n_spill_out;
  int64_t spill_cap;
  int32_t depth;   // stack entries per wave (1..kTailMaxDepth)
  int32_t budget;  // particles a wave walks from one claimed particle before it passes its stack on
};

// The wave's stack in registers: entry e sits in lane e % 64, in `lo` for
// e < 64 and in `hi` above; sp is wave-uniform.
struct WaveStack {
  int32_t lo = 0, hi = 0;
  int sp = 0;
  __device__ inline void push(int32_t v, int lane) {
    if (sp < 64) {
      if (lane == sp) lo = v;
    } else if (lane == sp - 64) {
      hi = v;
    }
    ++sp;
  }
  __device__ inline int32_t pop() {
    --sp;
    return sp < 64 ? __shfl(lo, sp) : __shfl(hi, sp - 64);
  }
};

// One accept/reject draw of process P by the whole wave (results uniform):
// the candidate count as the draw stores it (negative when max_trials passed
// without an acceptance), the accepted point and its J*f, and max_F.
template <int P>
__device__ inline int32_t wave_draw(const DevTables& T, double E, uint64_t seed, uint64_t sid, float* xs, double& w_acc,
                                    double& mf) {
  constexpr int ND = proc_ndim<P>();
  const int lane = threadIdx.x & 63;
  const PetiteGridDesc& gd = T.d->grid[P];
  const int32_t ninc[4] = {gd.ninc[0], gd.ninc[1], gd.ninc[2], gd.ninc[3]};
  const uint32_t max_trials = (uint32_t)(T.d->max_trials < (int64_t)0x7FFFFFF0 ? (T.d->max_trials > 0 ? T.d->max_trials : 1)
                                                                               : (int64_t)0x7FFFFFF0);
  const int ie = lu_key(T, P, E);
  const float* base = T.f32 + gd.node_off + (int64_t)ie * gd.stride;
  mf = T.maxf[P * T.maxE + ie] * T.d->maxF_fudge;
  for (uint32_t j = draw_possible<P>(T) ? 0u : max_trials; j < max_trials; j += 64u) {
    const uint32_t jc = j + (uint32_t)lane;
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    double ua = 1.0, J = 0.0;
    double w = candidate_w<P, ND>(T, E, base, ninc, seed, sid, jc, x, ua, J);
    if (ND == 4 && fabs(mf * ua - w) <= kRefBand * w) w = J * integrand_ref<P>(T, E, x);
    const bool acc = (jc < max_trials) && (mf * ua < w);
    const unsigned long long am = __ballot(acc);
    if (am) {
      const int win = __ffsll((long long)am) - 1;
      for (int d = 0; d < 4; ++d) xs[d] = (float)shfl_f64(x[d], win);
      w_acc = shfl_f64(w, win);
      return (int32_t)(j + (uint32_t)win + 1u);
    }
  }
  w_acc = 0.0;
  return -(int32_t)max_trials;
}

__device__ inline int32_t wave_draw_any(int p, const DevTables& T, double E, uint64_t seed, uint64_t sid, float* xs,
                                        double& w_acc, double& mf) {
  switch (p) {
    case 0: return wave_draw<0>(T, E, seed, sid, xs, w_acc, mf);
    case 1: return wave_draw<1>(T, E, seed, sid, xs, w_acc, mf);
    case 2: return wave_draw<2>(T, E, seed, sid, xs, w_acc, mf);
    case 3: return wave_draw<3>(T, E, seed, sid, xs, w_acc, mf);
    case 4: return wave_draw<4>(T, E, seed, sid, xs, w_acc, mf);
    case 5: return wave_draw<5>(T, E, seed, sid, xs, w_acc, mf);
    case 6: return wave_draw<6>(T, E, seed, sid, xs, w_acc, mf);
    default: return wave_draw<7>(T, E, seed, sid, xs, w_acc, mf);
  }
}

// Walks the items [0, nres + n + n_spill_in): the particles the previous
// sweep handed on (live state in its resume records), the store slots
// [first, first + n), then the previous launch's spills.  range: the tail
// setup's {first, n, end, nres}.
__global__ __launch_bounds__(256) void k_tail(DevTables T, DevStore S, uint64_t seed, double* shower_energy,
                                              unsigned long long* counters, const double* sig_lds_src,
                                              const int64_t* range, ResumeBuf RB, TailParams TPm) {
  const int64_t first = range[0], n = range[1];
  const int64_t nres = RB.rec ? range[3] : 0;
  const ResumeRec* rin = RB.rec ? RB.rec + (int64_t)(counters[C_RPARITY] & 1ull) * RB.cap : nullptr;
  // a count past the list (children that did not fit, an internal error the
  // run reports) is cut to the list
  const int64_t n_in = (int64_t)*TPm.n_spill_in;
  const int64_t total = nres + n + (n_in < TPm.spill_cap ? n_in : TPm.spill_cap);
  if ((int64_t)blockIdx.x * 4 >= total) return;
  HIP_DYNAMIC_SHARED(double, s_sig)
  for (int i = threadIdx.x; i < T.lds_doubles; i += blockDim.x) s_sig[i] = sig_lds_src[i];
  __syncthreads();
  const LdsSigma L{s_sig};
  const int lane = threadIdx.x & 63;
  const int depth = TPm.depth < 1 ? 1 : (TPm.depth > kTailMaxDepth ? kTailMaxDepth : TPm.depth);
  WaveStack st;
  unsigned long long steps = 0, n_trials = 0, n_draws = 0, n_viol = 0, n_fail = 0, n_drop = 0, n_ovf = 0, n_lost = 0;
  int walked = 0;  // particles walked since the wave's last claim
  while (true) {
    // a claimed particle's subtree beyond `budget` particles is left to the
    // next launch, which spreads the stack's entries over all of its waves:
    // one large subtree no longer holds a single wave while the others idle
    if (TPm.budget > 0 && walked >= TPm.budget && st.sp > 0) {
      unsigned long long o = 0;
      if (lane == 0) o = atomicAdd(TPm.n_spill_out, (unsigned long long)st.sp);
      o = shfl_u64(o, 0);
      if (lane < st.sp && (int64_t)(o + lane) < TPm.spill_cap) TPm.spill_out[o + lane] = st.lo;
      if (lane + 64 < st.sp && (int64_t)(o + lane + 64) < TPm.spill_cap) TPm.spill_out[o + lane + 64] = st.hi;
      const int64_t past = (int64_t)o + st.sp - TPm.spill_cap;
      if (past > 0) n_lost += (unsigned long long)(past < st.sp ? past : st.sp);  // reported as an internal error
      st.sp = 0;
    }
    int64_t slot;
    int64_t rr = -1;  // the claimed item (-1: popped from the stack)
    if (st.sp > 0) {
      slot = st.pop();
    } else {
      walked = 0;
      unsigned long long c = 0;
      if (lane == 0) c = atomicAdd(counters + C_TRANSPORT, 1ull);
      rr = (int64_t)shfl_u64(c, 0);
      if (rr >= total) break;
      if (rr < nres) slot = rin[rr].slot;
      else if (rr < nres + n) slot = first + (rr - nres);
      else slot = TPm.spill_in[rr - nres - n];
    }
    ++walked;
    const bool resumed = rr >= 0 && rr < nres;
    // ---- the particle: constants from the store, the start state from the
    // store or the resume record (k_transport's refill)
    const int pid = S.pid[slot];
    const uint64_t sid = S.sid[slot];
    const int shower = S.shower[slot];
    const int gen = S.meta[slot] & 0xFFFF;
    const float wgt = S.r0[slot].w;
    const uint32_t ch = sum_channels(pid);
    const double m = pid_mass(pid);
    const double pme = fmax(fmax(min_calc(T, pid), T.d->min_energy), m);
    double E, p = 0.0, dep = 0.0, dz = 0.0;
    float nx, ny, nz, rx, ry, rz;
    uint32_t k = 0;
    bool act = false;
    if (resumed) {
      const ResumeRec& R0 = rin[rr];
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
      act = true;
    } else {
      const float4 p0 = S.p0[slot];
      const float4 r0 = S.r0[slot];
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
      const bool still = ch == 0xFFFu && pid != 22;
      p = still ? sqrt((double)pn2) : 0.0;  // the others' |p| is formed when they finish
      rx = r0.x;
      ry = r0.y;
      rz = r0.z;
      if (E < pme || still) {
        // ends where it starts (shower.py:532-536, 640-653)
      } else if (pid == 22) {
        // one exponential flight, no MCS (shower.py:538-553)
        const double mfp = L.mfp(22, E);
        const double dist = -mfp * (double)log1p_f32(-u01f(draw_block(seed, sid, 0, PURPOSE_STEP).x));
        rx += (float)(nx * dist);
        ry += (float)(ny * dist);
        rz += (float)(nz * dist);
        k = 1;
      } else {
        act = true;
      }
    }
    // ---- substeps, then the closing increment (k_transport's advance)
    bool closing = false;
    while (act) {
      const bool in_close = closing;
      const double Ssum = L.nsigma_sum_ch(ch, E);
      const double mfp = Ssum > 0.0 ? kCmToM / Ssum : 1.0e12;
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
          const float r2 = Ssum > 0.0 ? (float)(dz * Ssum * (1.0 / kCmToM)) : (float)(dz * 1.0e-12);
          dist = -mfp * (double)log1p_f32(expm1f(-r2) * (float)distC);
        }
        w_r = u.y;
        w_a = u.z;
        m_lep = kMe;
      }
      if (adv) charged_advance(T, dist, m, m_lep, w_r, w_a, E, nx, ny, nz, rx, ry, rz, dep);
      ++k;
      if (in_close) act = false;
      else if (E < pme) closing = true;
    }
    // ---- the final state and the process choice (k_transport's finish)
    steps += k;
    if (!(ch == 0xFFFu && pid != 22)) p = momentum(E, m);
    const float4 pf = make_float4((float)E, (float)(p * nx), (float)(p * ny), (float)(p * nz));
    if (lane == 0) {
      S.pf[slot] = pf;
      S.rf[slot] = make_float4(rx, ry, rz, (float)k);
    }
    int want = -1;
    if (E > pme) {
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
    if (shower_energy && lane == 0) {
      unsafeAtomicAdd(shower_energy + 4 * shower + 0, dep);
      if (want < 0) unsafeAtomicAdd(shower_energy + 4 * shower + 1, E);
    }
    if (want < 0) continue;
    // ---- the draw
    float xs[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    double w_acc = 0.0, mf = 0.0;
    const int32_t tr = wave_draw_any(want, T, E, seed, sid, xs, w_acc, mf);
    const unsigned long long ntr = tr > 0 ? (unsigned long long)tr : (unsigned long long)(-(int64_t)tr);
    const bool viol = tr > 0 && w_acc > mf;
    n_draws += 1;
    n_trials += ntr;
    n_viol += viol ? 1ull : 0ull;
    n_fail += tr > 0 ? 0ull : 1ull;
    if (lane == 0) {
      atomicAdd(counters + C_PTRIALS + want, ntr);
      atomicAdd(counters + C_PDRAWS + want, 1ull);
      if (viol) atomicAdd(counters + C_PVIOL + want, 1ull);
    }
    if (tr <= 0) {
      if (shower_energy && lane == 0) unsafeAtomicAdd(shower_energy + 4 * shower + 1, E);
      continue;
    }
    // ---- the children (k_emit)
    const double az = u01d(draw_block(seed, sid, 0, PURPOSE_INTERACT).y);
    const double x0 = accepted_x1(T, want, E, seed, sid, tr, xs[0]);
    V4 lab[2];
    int cpid[2];
    vertex_children(want, pid, E, xs, x0, az, pf.y, pf.z, pf.w, lab, cpid);
    bool keep[2];
    for (int c = 0; c < 2; ++c) {
      keep[c] = lab[c].e > T.d->min_energy;
      if (!keep[c] && shower_energy && lane == 0) unsafeAtomicAdd(shower_energy + 4 * shower + 2, lab[c].e);
    }
    if (shower_energy && lane == 0 &&
        (want == COMP || want == ANN || want == MOLLER || want == BHABHA || want == MUONE))
      unsafeAtomicAdd(shower_energy + 4 * shower + 3, 1.0);
    n_drop += (unsigned long long)(!keep[0]) + (unsigned long long)(!keep[1]);
    const int nk = (int)keep[0] + (int)keep[1];
    if (nk == 0) continue;
    unsigned long long cb = 0;
    if (lane == 0) cb = atomicAdd(counters + C_STORE, (unsigned long long)nk);
    cb = shfl_u64(cb, 0);
    int32_t cslot[2] = {-1, -1};
    for (int c = 0; c < 2; ++c) {
      if (!keep[c]) continue;
      if ((int64_t)cb < S.capacity) {
        const int64_t cs = (int64_t)cb;
        if (lane == 0) {
          S.p0[cs] = make_float4((float)lab[c].e, (float)lab[c].x, (float)lab[c].y, (float)lab[c].z);
          S.r0[cs] = make_float4(rx, ry, rz, wgt);
          S.pid[cs] = cpid[c];
          S.parent[cs] = (int32_t)slot;
          S.meta[cs] = ((gen + 1) & 0xFFFF) | ((want + 1) << 16) | (c << 24);
          S.shower[cs] = shower;
          S.sid[cs] = derive_sid(seed, sid, (uint32_t)c, PURPOSE_CHILD);
        }
        cslot[c] = (int32_t)cs;
      } else {
        ++n_ovf;
      }
      ++cb;
    }
    // lane 0's writes above are read by the whole wave when a child is popped
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    // the higher-energy child below, the lower-energy one on top
    const int lower = (cslot[0] >= 0 && cslot[1] >= 0 && lab[1].e < lab[0].e) ? 1 : 0;
    for (int t = 0; t < 2; ++t) {
      const int c = t == 0 ? 1 - lower : lower;
      if (cslot[c] < 0) continue;
      if (st.sp < depth) {
        st.push(cslot[c], lane);
      } else {
        unsigned long long o = 0;
        if (lane == 0) o = atomicAdd(TPm.n_spill_out, 1ull);
        o = shfl_u64(o, 0);
        if ((int64_t)o < TPm.spill_cap) {
          if (lane == 0) TPm.spill_out[o] = cslot[c];
        } else {
          ++n_lost;  // reported by the run as an internal error
        }
      }
    }
  }
  if (lane == 0) {
    if (steps) {
      atomicAdd(counters + C_STATS + ST_STEPS, steps);
      atomicAdd(counters + C_TAIL_STEPS, steps);
    }
    if (n_draws) atomicAdd(counters + C_STATS + ST_DRAWS, n_draws);
    if (n_trials) atomicAdd(counters + C_STATS + ST_TRIALS, n_trials);
    if (n_viol) atomicAdd(counters + C_STATS + ST_VIOL, n_viol);
    if (n_fail) atomicAdd(counters + C_STATS + ST_NOSAMPLE, n_fail);
    if (n_drop) atomicAdd(counters + C_STATS + ST_DROPPED, n_drop);
    if (n_ovf) atomicAdd(counters + C_STATS + ST_OVERFLOW, n_ovf);
    if (n_lost) atomicAdd(counters + C_STATS + ST_REQ_OVERFLOW, n_lost);
  }
}

// Start of a tail launch: the items are the particles handed on by the last
// sweep, every store entry appended since the last sweep or tail launch (no
// bound: the walk files no draw requests), and the previous launch's spills
// (in the list of parity `par`; the other list is cleared for this launch's).
__global__ void k_tail_setup(unsigned long long* ctr, int64_t* range, int64_t capacity, int64_t resume_cap, int par) {
  if (threadIdx.x == 0) {
    const int64_t big = (int64_t)1 << 62;
    const SweepRange sr = sweep_range(range[2], (int64_t)ctr[C_STORE], capacity, big, (int64_t)ctr[C_YIELD_OUT],
                                      resume_cap);
    range[0] = sr.first;
    range[1] = sr.n;
    range[2] = sr.first + sr.n;
    range[3] = sr.nres;
    ctr[C_YIELD_OUT] = 0ull;
    ctr[C_RPARITY] ^= 1ull;
    ctr[C_TRANSPORT] = 0ull;
    ctr[C_SPILL + (par ^ 1)] = 0ull;
    if (sr.n > 0 || sr.nres > 0 || ctr[C_SPILL + par] > 0ull) ctr[C_SWEEPS] += 1ull;
  }
}

// After a tail launch: everything it appended was walked or spilled, so the
// next launch (or sweep) starts past it.
__global__ void k_tail_finish(unsigned long long* ctr, int64_t* range, int64_t capacity) {
  if (threadIdx.x == 0) {
    const int64_t now = (int64_t)ctr[C_STORE];
    range[2] = now < capacity ? now : capacity;
  }
}
