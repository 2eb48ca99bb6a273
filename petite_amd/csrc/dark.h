// Dark-vector production over a shower record (DarkShower.generate_dark_shower,
// src/PETITE/dark_shower.py:806-849).
//
// k_dark_select reads each SM record once and, for every active process with
// a positive weight (GetBSMWeights, 595-647), writes an entry: the weight, the
// pair's dark stream, and the state produce_bsm_particle (721-804) reaches
// before its draw: interaction energy from the drate table, multiple
// scattering over the distance travelled and the energy lost.  Entries that
// need draw_dark_sample go to the per-process draw regions; k_dark_emit turns
// every entry into its vector (threshold shortcuts, bound-electron Compton,
// radiative return, dark bremsstrahlung) and rotates it into the lab frame.
//
// Random numbers of one pair come from dsid = derive_sid(seed, sid, code,
// PURPOSE_DARK): INTERACT block 0 (x drate bin, y MCS radius, z MCS azimuth
// with its sign in bit 0, w kinematic azimuth), INTERACT block 1 (x the bound
// electron's cos(theta), y the azimuth of its frame), BOUND_E blocks (pairs of
// bound-electron momentum accept/reject uniforms) and the VEGAS candidates.
#pragma once

namespace dark {

enum Family { F_BREM_E = 0, F_BREM_P = 1, F_MUBREM = 2, F_ANN = 3 };
enum State { S_NONE = 0, S_DRAW = 1, S_DIRECT = 2 };
constexpr int kPePairs = 1 << 16;

struct Entries {
  int32_t* parent;   // SM store slot
  int32_t* flag;     // code | state << 8
  double* weight;    // wg * parent weight
  double* E;         // energy after the drate step (pf[0] of produce_bsm_particle)
  float4* dir;       // unit direction of pf
  uint64_t* dsid;    // dark stream
  int64_t cap;
};

// Interval lo with x[lo] <= v < x[lo+1] (v inside [x0, x_{n-1})): an O(1)
// guess g0 + (v' - v0) * inv on uniform grids (in v' = ln v or v itself),
// corrected against the stored nodes; bisection when inv = 0.
__device__ inline int find_interval(const double* x, int n, double v, float guess_arg, double v0, double inv) {
  int lo;
  if (inv > 0.0) {
    lo = (int)((guess_arg - (float)v0) * (float)inv);
    lo = lo < 0 ? 0 : (lo > n - 2 ? n - 2 : lo);
    while (lo > 0 && x[lo] > v) --lo;
    while (lo < n - 2 && x[lo + 1] <= v) ++lo;
  } else {
    lo = 0;
    int hi = n - 1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (x[mid] <= v) lo = mid; else hi = mid;
    }
  }
  return lo;
}

// numpy.interp on an ascending grid (slope form), 0 outside [x0, x_{n-1}]
// (scipy interp1d with fill_value=0, dark_shower.py:447-452).
__device__ inline double interp_fill0(const double* x, const double* y, int n, double E, double lnE0 = 0.0,
                                      double inv_dlog = 0.0) {
  if (n <= 0 || !(E >= x[0]) || E > x[n - 1]) return 0.0;
  if (E == x[n - 1]) return y[n - 1];
  const int lo = find_interval(x, n, E, __logf((float)E), lnE0, inv_dlog);
  const double slope = (y[lo + 1] - y[lo]) / (x[lo + 1] - x[lo]);
  return slope * (E - x[lo]) + y[lo];
}

// interpolate1d in log10-log10 with fill -20 (dark_shower.py:31-46, 294-309);
// l = log10(E).
__device__ inline double loglog(const double* lx, const double* ly, int n, double l, double l0 = 0.0,
                                double inv_dl = 0.0) {
  if (n <= 0 || l < lx[0] || l > lx[n - 1]) return 1.0e-20;
  if (l == lx[n - 1]) return pow(10.0, ly[n - 1]);
  const int lo = find_interval(lx, n, l, (float)l, l0, inv_dl);
  const double slope = (ly[lo + 1] - ly[lo]) / (lx[lo + 1] - lx[lo]);
  return pow(10.0, slope * (l - lx[lo]) + ly[lo]);
}

__device__ inline int family_of(int code, int pid) {
  if (code == DARKBREM) return pid == 11 ? F_BREM_E : (pid == -11 ? F_BREM_P : -1);
  if (code == DARKANN) return pid == -11 ? F_ANN : -1;
  if (code == DARKMUONBREM) return (pid == 13 || pid == -13) ? F_MUBREM : -1;
  return -1;
}

// Quantities of one SM particle shared by the weights of every mass.
struct WeightArgs {
  double E0;
  float lf;    // logf(E0): guess of the geometric-grid interval
  double l10;  // log10(E0) for the log-log DarkComp table
  double den;  // n*sigma_PairProd + n*sigma_Comp at E0 (photons)
};

__device__ inline WeightArgs weight_args(const DevTables& T, int pid, double E0) {
  WeightArgs a;
  a.E0 = E0;
  a.lf = __logf((float)E0);
  a.l10 = pid == 22 ? log10(E0) : 0.0;
  a.den = pid == 22 ? nsigma(T, PAIRPROD, E0) + nsigma(T, COMP, E0) : 0.0;
  return a;
}

// GetBSMWeights (dark_shower.py:595-647), eps^2 included.
__device__ inline double bsm_weight(const DevTables& T, const PetiteDarkDesc& D, const double* b, int pid, int code,
                                    const WeightArgs& a) {
  const double E0 = a.E0;
  if (code == DARKCOMP) {
    if (pid != 22 || E0 < D.min_comp || E0 < T.d->min_calc[2]) return 0.0;
    const double* lx = b + D.comp_off;
    const double num = loglog(lx, lx + D.comp_n, D.comp_n, a.l10, D.comp_l0, D.comp_inv_dl);
    return D.eps2 * (num / a.den);
  }
  const int f = family_of(code, pid);
  if (f < 0) return 0.0;
  const double thr = (code == DARKBREM) ? D.min_brem : (code == DARKANN ? D.min_ann : D.min_muon_brem);
  if (E0 < thr) return 0.0;
  const double* x = b + D.w_off[f];
  const int n = D.w_n[f];
  if (n <= 0 || !(E0 >= x[0]) || E0 > x[n - 1]) return 0.0;
  const double* y = x + n;
  if (E0 == x[n - 1]) return D.eps2 * y[n - 1];
  const int lo = find_interval(x, n, E0, a.lf, D.w_lnE0[f], D.w_inv_dlog[f]);
  const double slope = (y[lo + 1] - y[lo]) / (x[lo + 1] - x[lo]);
  return D.eps2 * (slope * (E0 - x[lo]) + y[lo]);
}

// numpy's sum of a short row: sequential below 8 terms, else eight partial
// sums combined pairwise and the remainder added in order.
__device__ inline double np_sum(const double* r, int n, int stride) {
  if (n < 8) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += r[i * stride];
    return s;
  }
  double s = ((r[0] + r[stride]) + (r[2 * stride] + r[3 * stride])) +
             ((r[4 * stride] + r[5 * stride]) + (r[6 * stride] + r[7 * stride]));
  for (int i = 8; i < n; ++i) s += r[i * stride];
  return s;
}

// Interaction energy of produce_bsm_particle (dark_shower.py:741-756): the
// saved energy Ei (the largest <= E0, clamped to the table), a bin drawn with
// probability proportional to its rate, shifted by E0 - Ei.  Returns false
// when the row's rates sum to zero (the reference returns None).
__device__ inline bool drate_energy(const PetiteDarkDesc& D, const double* b, int f, double E0, double u,
                                    double& E_int) {
  const int n = D.d_n[f];
  if (n <= 0) return false;
  const double* Es = b + D.d_off[f];
  const double* cnt = Es + n;
  const double* rows = cnt + n;
  int i;
  if (E0 < Es[0]) {
    i = 0;
  } else if (E0 > Es[n - 1]) {
    i = n - 1;
  } else {
    int lo = 0, hi = n;  // largest i with Es[i] <= E0
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (Es[mid] <= E0) lo = mid; else hi = mid;
    }
    i = lo;
  }
  const int c = (int)cnt[i];
  const double* row = rows + (int64_t)i * 2 * PETITE_DARK_DRATE_BINS;
  const double tot = np_sum(row + 1, c, 2);
  if (!(tot != 0.0) || c <= 0) return false;
  double cdf[PETITE_DARK_DRATE_BINS];
  double acc = 0.0;
  for (int k = 0; k < c; ++k) {
    acc += row[2 * k + 1] / tot;
    cdf[k] = acc;
  }
  int k = c - 1;
  for (int j = 0; j < c; ++j) {
    if (cdf[j] / acc > u) {
      k = j;
      break;
    }
  }
  E_int = row[2 * k] + (E0 - Es[i]);
  return true;
}

__global__ __launch_bounds__(256) void k_dark_select(DevTables T, PetiteDarkDesc D, const double* b, DevStore S,
                                                     int64_t first, int64_t n, uint64_t seed, Entries Q,
                                                     Requests R, unsigned long long* n_entries,
                                                     unsigned long long* stats) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool live = t < n;
  const int64_t i = first + t;
  int pid = 0;
  float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), pf = p0;
  double wpar = 0.0;
  uint64_t sid = 0;
  if (live) {
    pid = S.pid[i];
    p0 = S.p0[i];
    pf = S.pf[i];
    wpar = (double)S.r0[i].w;
    sid = S.sid[i];
  }
  const double E0 = (double)p0.x;
  const WeightArgs wa = weight_args(T, pid, E0);
  unsigned long long n_zero = 0;
  for (int a = 0; a < D.n_active; ++a) {
    const int code = D.active[a];  // wave-uniform
    const double wg = live ? bsm_weight(T, D, b, pid, code, wa) : 0.0;
    const bool want = wg > 0.0;
    const unsigned long long m = __ballot(want);
    if (!m) continue;
    unsigned long long base = 0;
    if (lane == __ffsll((long long)m) - 1) base = atomicAdd(n_entries, (unsigned long long)__popcll(m));
    base = shfl_u64(base, __ffsll((long long)m) - 1);
    int state = S_NONE;
    int64_t e = -1;
    double E = 0.0;
    float nx = 0.f, ny = 0.f, nz = 0.f;
    if (want) {
      e = (int64_t)(base + __popcll(m & ((1ull << lane) - 1ull)));
      const uint64_t dsid = derive_sid(seed, sid, (uint32_t)code, PURPOSE_DARK);
      const int f = (code == DARKCOMP) ? -1 : family_of(code, pid);
      const double m_part = pid_mass(pid);
      state = S_DIRECT;
      if (f >= 0) {
        const U4 w0 = draw_block(seed, dsid, 0, PURPOSE_INTERACT);
        double E_int = 0.0;
        if (!drate_energy(D, b, f, E0, u01d(w0.x), E_int)) {
          state = S_NONE;
          ++n_zero;
        } else {
          // scatter the initial momentum over the distance (m_lepton = m_e,
          // dark_shower.py:753-755), then lose E0 - E_int.  beta comes from the
          // on-shell |p| = sqrt(E0^2 - m^2): a float32 record's |p| can round to E0.
          const double dist = (E0 - E_int) / T.d->dEdx;
          const double px = p0.y, py = p0.z, pz = p0.w;
          const double pn = sqrt(px * px + py * py + pz * pz);
          const double pon = sqrt(fmax(E0 * E0 - m_part * m_part, 0.0));
          if (pn > 0.0 && pon > 0.0) {
            nx = (float)(px / pn);
            ny = (float)(py / pn);
            nz = (float)(pz / pn);
            mcs_rotate(T, nx, ny, nz, E0, pon, m_part, kMe, dist, w0.y, w0.z);
          }
          double Eu = E0 - (E0 - E_int);
          if (Eu <= m_part) Eu = m_part;
          if (!(sqrt(Eu * Eu - m_part * m_part) > 0.0)) nx = ny = nz = 0.0f;
          E = Eu;
        }
      } else {
        E = (double)pf.x;
        const double pn = sqrt((double)pf.y * pf.y + (double)pf.z * pf.z + (double)pf.w * pf.w);
        if (pn > 0.0) {
          nx = (float)(pf.y / pn);
          ny = (float)(pf.z / pn);
          nz = (float)(pf.w / pn);
        }
      }
      if (state == S_DIRECT) {
        const bool shortcut = (code == DARKANN && E <= D.E_res) || (code == DARKCOMP && E <= D.E_C);
        if (!shortcut) state = S_DRAW;
      }
      if (e < Q.cap) {
        Q.parent[e] = (int32_t)i;
        Q.flag[e] = code | (state << 8);
        Q.weight[e] = wg * wpar;
        Q.E[e] = E;
        Q.dir[e] = make_float4(nx, ny, nz, 0.0f);
        Q.dsid[e] = dsid;
      }
    }
    // draw requests of this process, one atomic per wave
    const bool drq = want && state == S_DRAW && e < Q.cap;
    const unsigned long long dm = __ballot(drq);
    if (dm) {
      const int lead = __ffsll((long long)dm) - 1;
      unsigned int rb = 0;
      if (lane == lead) rb = atomicAdd(R.count + code, (unsigned int)__popcll(dm));
      rb = __shfl(rb, lead);
      if (drq) {
        const int64_t k = (int64_t)code * R.cap + rb + __popcll(dm & ((1ull << lane) - 1ull));
        R.slot[k] = (int32_t)e;
        R.E[k] = E;
      }
    }
  }
  wave_stat(stats, 0, n_zero);
}

// draw_pe_sample (dark_shower.py:706-719): pe uniform in [0, 1 MeV) accepted
// against the 1s momentum density normalised at its maximum.
__device__ inline double electron_wf(double pe, double lam) {
  const double d = pe * pe + lam * lam;
  return 32.0 / kPi * (lam * lam * lam * lam * lam) * pe * pe / (d * d * d * d);
}
__device__ inline bool draw_pe(uint64_t seed, uint64_t dsid, double Zeff, double& pe) {
  const double lam = kAlpha * Zeff * kMe;
  const double c = electron_wf(lam / sqrt(3.0), lam);
  for (int k = 0; k < kPePairs / 2; ++k) {
    const U4 w = draw_block(seed, dsid, (uint32_t)k, PURPOSE_BOUND_E);
    double x = 1.0e-3 * u01d(w.x);
    if (u01d(w.y) < electron_wf(x, lam) / c) {
      pe = x;
      return true;
    }
    x = 1.0e-3 * u01d(w.z);
    if (u01d(w.w) < electron_wf(x, lam) / c) {
      pe = x;
      return true;
    }
  }
  return false;
}

// compton_fourvecs_boundelectron (kinematics.py:134-183): the vector.
__device__ inline V4 kin_compton_bound(double Eg, double mV, double ct, double Pe, double cte, double u_ph,
                                       double u_phie) {
  const double m = kMe;
  const double s = m * m + 2.0 * Eg * (sqrt(m * m + Pe * Pe) - cte * Pe);
  const double rs = sqrt(s);
  const double Ee = (s - mV * mV + m * m) / (2.0 * rs);
  const double EV = (s + mV * mV - m * m) / (2.0 * rs);
  const double pF = sqrt(Ee * Ee - m * m);
  const double bnum = sqrt(Eg * Eg + 2.0 * cte * Eg * Pe + Pe * Pe);
  const double b0 = bnum / (Eg + sqrt(m * m + Pe * Pe));
  const double g0 = 1.0 / sqrt(1.0 - b0 * b0);
  double sph, cph;
  sincos(2.0 * kPi * u_ph, &sph, &cph);
  const double st = sqrt(1.0 - ct * ct);
  const double vx = pF * st * sph, vy = pF * st * cph, vz = b0 * g0 * EV - g0 * pF * ct;
  const double ctz = (Eg + cte * Pe) / bnum;
  const double stz = sqrt(1.0 - ctz * ctz);
  double spe, cpe;
  sincos(2.0 * kPi * u_phie, &spe, &cpe);
  return V4{g0 * EV - b0 * g0 * pF * ct, ctz * cpe * vx - spe * vy + stz * cpe * vz,
            ctz * spe * vx + cpe * vy + stz * spe * vz, -stz * vx + ctz * vz};
}

// radiative_return_fourvecs (kinematics.py:267-299) with the boost of
// radiative_return.py:18-24 into the frame of the electron at rest.
__device__ inline V4 kin_radiative_return(double E, double u, double mV) {
  const double m = kMe;
  const double s = 2.0 * m * (m + E);
  const double beta = (2.0 * kAlpha / kPi) * (log(s / (m * m)) - 1.0);
  const double umax = pow(1.0 - mV * mV / s, beta / 2.0);
  const double x1 = 1.0 - pow(u * umax, 2.0 / beta);
  const double x2 = mV * mV / (x1 * s);
  const double rs = sqrt(s);
  const double E1 = x1 * rs / 2.0, E2 = x2 * rs / 2.0;
  const double v0 = E1 + E2, v3 = E1 - E2;
  // p = (rs/2, 0, 0, -sqrt(s/4 - m^2))
  const double p0 = rs / 2.0, p3 = -sqrt(s / 4.0 - m * m);
  const double rsq = sqrt(p0 * p0 - p3 * p3);
  const double w0 = (p0 * v0 - p3 * v3) / rsq;
  const double c1 = (v0 + w0) / (rsq + p0);
  return V4{w0, 0.0, 0.0, v3 - c1 * p3};
}

struct EmitOut {
  float4* p;
  float4* r;
  int32_t* parent;
  int32_t* meta;
};

__global__ __launch_bounds__(256) void k_dark_emit(DevTables T, PetiteDarkDesc D, DevStore S, Entries Q, int64_t n_e,
                                                   const float* x_res, const int32_t* trials_res, EmitOut O,
                                                   int64_t out_base, uint64_t seed, double* yield,
                                                   unsigned long long* stats) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double wsum[4] = {0.0, 0.0, 0.0, 0.0};
  unsigned long long n_vec = 0, n_pe_fail = 0;
  if (e < n_e) {
    const int flag = Q.flag[e];
    const int code = flag & 0xFF;
    const int state = flag >> 8;
    const int32_t slot = Q.parent[e];
    double w = Q.weight[e];
    const double E = Q.E[e];
    const double mV = T.d->mV;
    bool ok = state != S_NONE;
    V4 v{0.0, 0.0, 0.0, 0.0};
    if (ok && state == S_DIRECT) {
      if (code == DARKANN) {
        v = V4{sqrt(E * E - kMe * kMe + mV * mV), 0.0, 0.0, sqrt(E * E - kMe * kMe)};
      } else {
        v = V4{sqrt(E * E + mV * mV), 0.0, 0.0, E};
      }
    } else if (ok) {
      ok = trials_res[e] > 0;
      if (ok) {
        const float* xs = x_res + 4 * e;
        const uint64_t dsid = Q.dsid[e];
        const U4 w0 = draw_block(seed, dsid, 0, PURPOSE_INTERACT);
        if (code == DARKCOMP) {
          const U4 w1 = draw_block(seed, dsid, 1, PURPOSE_INTERACT);
          double pe = 0.0;
          if (!draw_pe(seed, dsid, T.d->Zeff, pe)) {
            ok = false;
            ++n_pe_fail;
          } else {
            const double c0 = -1.0 + 2.0 * u01d(w1.x);
            const double s = kMe * kMe + 2.0 * E * (sqrt(kMe * kMe + pe * pe) - c0 * pe);
            const double Ee = (s - mV * mV + kMe * kMe) / (2.0 * sqrt(s));
            if (Ee < kMe) {
              v = V4{mV, 0.0, 0.0, 0.0};
              w = 0.0;
            } else {
              v = kin_compton_bound(E, mV, (double)xs[0], pe, c0, u01d(w0.w), u01d(w1.y));
            }
          }
        } else if (code == DARKANN) {
          v = kin_radiative_return(E, (double)xs[0], mV);
        } else {
          // l_to_lV_fourvecs (kinematics.py:43-68): the vector
          const double wv = (double)xs[0] * E;
          const double ct = 1.0 - pow(10.0, (double)xs[1]);
          const double k = sqrt(wv * wv - mV * mV);
          double sa, ca;
          sincos(2.0 * kPi * u01d(w0.w), &sa, &ca);
          const double st = sqrt(1.0 - ct * ct);
          v = V4{wv, k * ca * st, k * sa * st, k * ct};
        }
      }
    }
    float4 pv = make_float4(0.f, 0.f, 0.f, 0.f), rv = pv;
    if (ok) {
      const float4 d = Q.dir[e];
      double ox, oy, oz;
      frame_apply_d((double)d.x, (double)d.y, (double)d.z, v, ox, oy, oz);
      pv = make_float4((float)v.e, (float)ox, (float)oy, (float)oz);
      const float4 rf = S.rf[slot];
      rv = make_float4(rf.x, rf.y, rf.z, (float)w);
      ++n_vec;
      wsum[code - DARKBREM] = w;
    }
    const int64_t o = out_base + e;
    O.p[o] = pv;
    O.r[o] = rv;
    O.parent[o] = slot;
    O.meta[o] = code | (ok ? 0x100 : 0);
  }
  for (int k = 0; k < 4; ++k) {
    double s = wsum[k];
    for (int off = 32; off > 0; off >>= 1) s += shfl_f64(s, (threadIdx.x & 63) ^ off);
    if ((threadIdx.x & 63) == 0 && s != 0.0) unsafeAtomicAdd(yield + k, s);
  }
  wave_stat(stats, 1, n_vec);
  wave_stat(stats, 2, n_pe_fail);
}

// ------------------------------------------------------ meson decays ------
constexpr int TWO_BODY = 12;  // TwoBody_BSMDecay: dark process code of the output entries

// TwoBody_BSMDecay of pi0 / eta / eta' records (dark_shower.py:633-640,
// 836-844): weight meson_w (1 - r^2)^3 with r = m_V / m_M, and the vector of
// the decay M -> gamma V, isotropic in the meson's rest frame (cos theta and
// phi uniform; particle.py:209-254 with V second) and boosted along the
// meson's final momentum with the boost_matrix of particle.py:187-207; the
// vector sits at the meson's end point.  Meson masses are the PDG values of
// mass_dict.  Entries are appended after out_base through *count; an entry
// beyond capacity is counted but not written.
// Slots, the yield sum and the vector count are aggregated per block (one
// atomic of each per 1024 records): per-wave atomics on the three shared
// addresses serialised the kernel (845 us for 4e6 records).
constexpr int kMesonBlock = 1024;
__global__ __launch_bounds__(kMesonBlock) void k_meson_decay(DevTables T, PetiteDarkDesc D, DevStore S, int64_t first,
                                                             int64_t n, uint64_t seed, EmitOut O, int64_t out_base,
                                                             int64_t capacity, unsigned long long* count,
                                                             double* yield, unsigned long long* stats) {
  __shared__ unsigned int wave_off[kMesonBlock / 64];
  __shared__ double wave_w[kMesonBlock / 64];
  __shared__ unsigned long long block_base;
  const int64_t i = first + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  bool have = false;
  double w = 0.0;
  float4 pv = make_float4(0.f, 0.f, 0.f, 0.f), rv = pv;
  if (i < first + n) {
    const int pid = S.pid[i];
    const int k = pid == 111 ? 0 : (pid == 221 ? 1 : (pid == 331 ? 2 : -1));
    const double mM = k == 0 ? 0.1349768 : (k == 1 ? 0.547862 : 0.95778);
    const double mV = T.d->mV;
    if (k >= 0 && D.meson_w[k] > 0.0 && mV / mM < 1.0) {
      const double r = mV / mM;
      const double q = 1.0 - r * r;
      const double wg = D.meson_w[k] * q * q * q;
      w = wg * (double)S.r0[i].w;
      const uint64_t dsid = derive_sid(seed, S.sid[i], (uint32_t)TWO_BODY, PURPOSE_DARK);
      const U4 u = draw_block(seed, dsid, 0, PURPOSE_INTERACT);
      const double ct = -1.0 + 2.0 * u01d(u.x);
      const double phi = 2.0 * kPi * u01d(u.y);
      const double st = sqrt(1.0 - ct * ct);
      const double E2 = (mM * mM + mV * mV) / (2.0 * mM);
      const double pF = (mM * mM - mV * mV) / (2.0 * mM);  // the photon's energy = |p| (m_gamma = 0)
      double sp, cp;
      sincos(phi, &sp, &cp);
      const double v[4] = {E2, pF * st * sp, pF * st * cp, pF * ct};
      const float4 pf = S.pf[i];
      const double E0 = pf.x, px = pf.y, py = pf.z, pz = pf.w;
      const double gam = E0 / mM;
      const double beta = gam == 1.0 ? 1.0 : sqrt(1.0 - 1.0 / (gam * gam));
      const double pm = sqrt(px * px + py * py + pz * pz);
      double o[4] = {v[0], v[1], v[2], v[3]};
      if (pm != 0.0) {
        const double b[3] = {beta * px / pm, beta * py / pm, beta * pz / pm};
        const double g1 = (gam - 1.0) / (beta * beta);
        const double bv = b[0] * v[1] + b[1] * v[2] + b[2] * v[3];
        o[0] = gam * v[0] + gam * bv;
        for (int a = 0; a < 3; ++a) o[a + 1] = gam * b[a] * v[0] + v[a + 1] + g1 * b[a] * bv;
      }
      pv = make_float4((float)o[0], (float)o[1], (float)o[2], (float)o[3]);
      const float4 rf = S.rf[i];
      rv = make_float4(rf.x, rf.y, rf.z, (float)w);
      have = true;
    }
  }
  const unsigned long long m = __ballot(have);
  double s = have ? w : 0.0;
  for (int off = 32; off > 0; off >>= 1) s += shfl_f64(s, lane ^ off);
  if (lane == 0) {
    wave_off[wave] = (unsigned int)__popcll(m);
    wave_w[wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int tot = 0;
    double sw = 0.0;
    for (int k = 0; k < kMesonBlock / 64; ++k) {
      const unsigned int c = wave_off[k];
      wave_off[k] = tot;  // exclusive prefix: the wave's first slot in the block
      tot += c;
      sw += wave_w[k];
    }
    block_base = tot ? atomicAdd(count, (unsigned long long)tot) : 0ull;
    if (tot) {
      unsafeAtomicAdd(yield + 4, sw);
      atomicAdd(stats + 1, (unsigned long long)tot);
    }
  }
  __syncthreads();
  if (have) {
    const int64_t o = out_base + (int64_t)(block_base + wave_off[wave]) + __popcll(m & ((1ull << lane) - 1ull));
    if (o < capacity) {
      O.p[o] = pv;
      O.r[o] = rv;
      O.parent[o] = (int32_t)i;
      O.meta[o] = TWO_BODY | 0x100;
    }
  }
}

// ------------------------------------------------------ m_V-continuous yield --
constexpr int kYieldBatch = 16;
constexpr int kYieldWindow = 8;  // trained masses a pass may use (consecutive)
constexpr double kLn10 = 2.302585092994045684;

struct YieldTables {
  PetiteDarkDesc d[PETITE_MAX_MASSES];
  int32_t n_mass;
  const double* b;
};

// Mass points of one pass (ascending m_V): segment j0 + j[k] (trained masses
// j0 + j[k] and the next), t and dt/dm_V; `need` marks which of the window's
// trained masses j0 .. j0 + kYieldWindow - 1 the segments use.
struct YieldPoints {
  int32_t n;
  int32_t j0;
  uint32_t need;
  int32_t j[kYieldBatch];
  double t[kYieldBatch];
  double dtdm[kYieldBatch];
  double ln_eps2[kYieldWindow];
};

// interpolate1d in log10-log10 (dark_shower.py:31-46), returning the log10 of
// the value: -20 outside the table (the fill 1e-20); l = log10(E).
__device__ inline double loglog10(const double* lx, const double* ly, int n, double l, double l0, double inv_dl) {
  if (n <= 0 || l < lx[0] || l > lx[n - 1]) return -20.0;
  if (l == lx[n - 1]) return ly[n - 1];
  const int lo = find_interval(lx, n, l, (float)l, l0, inv_dl);
  const double slope = (ly[lo + 1] - ly[lo]) / (lx[lo + 1] - lx[lo]);
  return slope * (l - lx[lo]) + ly[lo];
}

// Weight of one trained mass for the yield, as (w, ln w): GetBSMWeights with
// that mass's tables (dark_shower.py:595-647).  DarkComp's log-log table gives
// ln w = ln eps^2 - ln(n sigma_PP + n sigma_Comp) + ln 10 * L directly (w
// itself, eps^2 10^L / den, is formed only where the linear branch needs it;
// here w = 1 marks it positive); the other families take the log of the
// interpolated weight.
__device__ inline void yield_weight(const DevTables& T, const PetiteDarkDesc& D, const double* b, int pid, int code,
                                    const WeightArgs& a, double ln_eps2, double ln_den, double& wv, double& lam) {
  if (code == DARKCOMP) {
    if (pid != 22 || a.E0 < D.min_comp || a.E0 < T.d->min_calc[2]) {
      wv = 0.0;
      lam = 0.0;
      return;
    }
    const double* lx = b + D.comp_off;
    lam = ln_eps2 - ln_den + kLn10 * loglog10(lx, lx + D.comp_n, D.comp_n, a.l10, D.comp_l0, D.comp_inv_dl);
    wv = 1.0;
    return;
  }
  wv = bsm_weight(T, D, b, pid, code, a);
  lam = wv > 0.0 ? log(wv) : 0.0;
}

// Bound-electron tables at the pass's masses (quad.h): slot k * nv + v holds
// m_k (v = 0), m_k (1 + h) (v = 1) and m_k (1 - h) (v = 2, with the
// derivative).  slot == nullptr: DarkAnn and DarkComp interpolate between the
// trained masses like the other processes.
struct BoundView {
  const quad::BoundSlot* slot;
  const double* tab;
  int32_t stride;  // doubles per slot
  int32_t nv;  // slots per point: 3 with the derivative (m_k, and its node derivatives after k_bound_deriv), else 1
};

// W and dW/dm_V of DarkAnn / DarkComp at point k: the interpolant of slot
// k * nv at E0, and (GRAD) its derivative from the node derivatives in slot
// k * nv + 1 (k_bound_deriv).  With s the interval weight,
//   d/dm [y_lo + s (y_hi - y_lo)] = y'_lo + s (y'_hi - y'_lo) + (y_hi - y_lo) s',
//   s' = -[x'_lo (1 - s) + x'_hi s] / (x_hi - x_lo);
// DarkComp in log10 space: W = eps^2 10^L / den, dW = W ln 10 dL.  Below the
// thresholds or outside the grid W = dW = 0.
template <bool GRAD>
__device__ inline void bound_wd(const DevTables& T, const BoundView& B, int k, int code, const WeightArgs& a,
                                double ln_den, double ln_eps2, double& W, double& dW) {
  const int s0 = k * B.nv;
  const quad::BoundSlot& S = B.slot[s0];
  const double* t = B.tab + (int64_t)s0 * B.stride;
  const double* d = t + B.stride;
  constexpr int n = quad::kBoundGrid;
  W = dW = 0.0;
  if (code == DARKANN) {
    const double E0 = a.E0;
    if (E0 < S.min_ann || !(E0 >= t[0]) || E0 > t[n - 1]) return;
    const double* y = t + 3 * n;
    const double e2 = exp(ln_eps2);
    if (E0 == t[n - 1]) {
      W = e2 * y[n - 1];
      if (GRAD) dW = e2 * d[3 * n + n - 1];
      return;
    }
    const int lo = find_interval(t, n, E0, a.lf, S.lnEa0, S.inv_dln_a);
    const double xl = t[lo], xh = t[lo + 1], yl = y[lo], yh = y[lo + 1];
    const double slope = (yh - yl) / (xh - xl);
    W = e2 * (slope * (E0 - xl) + yl);
    if (GRAD) {
      const double sw = (E0 - xl) / (xh - xl);
      const double sd = -(d[lo] * (1.0 - sw) + d[lo + 1] * sw) / (xh - xl);
      const double dyl = d[3 * n + lo], dyh = d[3 * n + lo + 1];
      dW = e2 * (dyl + sw * (dyh - dyl) + (yh - yl) * sd);
    }
    return;
  }
  if (a.E0 < S.min_comp || a.E0 < T.d->min_calc[2]) return;
  const double* lx = t + 4 * n;
  const double* ly = t + 5 * n;
  const double l = a.l10;
  if (l < lx[0] || l > lx[n - 1]) {
    W = exp(ln_eps2 - ln_den - 20.0 * kLn10);  // the fill 1e-20
    return;
  }
  double L, dL = 0.0;
  if (l == lx[n - 1]) {
    L = ly[n - 1];
    if (GRAD) dL = d[5 * n + n - 1];
  } else {
    const int lo = find_interval(lx, n, l, (float)l, S.l0c, S.inv_dlc);
    const double xl = lx[lo], xh = lx[lo + 1], yl = ly[lo], yh = ly[lo + 1];
    L = (yh - yl) / (xh - xl) * (l - xl) + yl;
    if (GRAD) {
      const double sw = (l - xl) / (xh - xl);
      const double sd = -(d[4 * n + lo] * (1.0 - sw) + d[4 * n + lo + 1] * sw) / (xh - xl);
      const double dyl = d[5 * n + lo], dyh = d[5 * n + lo + 1];
      dL = dyl + sw * (dyh - dyl) + (yh - yl) * sd;
    }
  }
  W = exp(ln_eps2 - ln_den + kLn10 * L);
  if (GRAD) dW = W * kLn10 * dL;
}

// Record indices one yield kernel visits (the rest contribute nothing to it).
struct RecordList {
  const int32_t* idx;
  const unsigned long long* count;  // on the device
};

// Species bit of a record's pid (0 e-, 1 e+, 2 photon, 3 mu-, 4 mu+; 5 other).
__device__ inline uint32_t species_bit(int pid) {
  switch (pid) {
    case 11: return 1u;
    case -11: return 2u;
    case 22: return 4u;
    case 13: return 8u;
    case -13: return 16u;
    default: return 32u;
  }
}

// The records of each yield kernel (species in mask_a / mask_b): the scans
// then run over the records that carry weight only (in a lead shower about
// half of them for each kernel) instead of masking the other lanes in every
// pass.  Each wave lists a run of kListRun x 64 consecutive records in record
// order: a first look counts its entries, one atomic per block and list
// reserves the block's share (atomics on one address are served one after
// another, so one per wave and 64 records cost more than the scans save), a
// second look writes them.
constexpr int kListRun = 16;
__global__ __launch_bounds__(256) void k_yield_lists(DevStore S, int64_t n, uint32_t mask_a, uint32_t mask_b,
                                                     int32_t* list_a, int32_t* list_b, unsigned long long* counts) {
  __shared__ unsigned long long s_cnt[2][4], s_base[2];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  constexpr int64_t kWaveRun = 64 * kListRun, kBlockRun = 4 * kWaveRun;
  for (int64_t b0 = (int64_t)blockIdx.x * kBlockRun; b0 < n; b0 += (int64_t)gridDim.x * kBlockRun) {
    const int64_t w0 = b0 + wid * kWaveRun;
    unsigned long long ca = 0, cb = 0;
    for (int j = 0; j < kListRun; ++j) {
      const int64_t i = w0 + j * 64 + lane;
      const uint32_t bit = i < n ? species_bit(S.pid[i]) : 0u;
      ca += (unsigned long long)__popcll(__ballot((bit & mask_a) != 0u));
      cb += (unsigned long long)__popcll(__ballot((bit & mask_b) != 0u));
    }
    if (lane == 0) {
      s_cnt[0][wid] = ca;
      s_cnt[1][wid] = cb;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
      const int l = threadIdx.x;
      const unsigned long long tot = s_cnt[l][0] + s_cnt[l][1] + s_cnt[l][2] + s_cnt[l][3];
      s_base[l] = tot ? atomicAdd(counts + l, tot) : 0ull;
    }
    __syncthreads();
    unsigned long long oa = s_base[0], ob = s_base[1];
    for (int w = 0; w < wid; ++w) {
      oa += s_cnt[0][w];
      ob += s_cnt[1][w];
    }
    for (int j = 0; j < kListRun; ++j) {
      const int64_t i = w0 + j * 64 + lane;
      const uint32_t bit = i < n ? species_bit(S.pid[i]) : 0u;
      const bool in_a = (bit & mask_a) != 0u, in_b = (bit & mask_b) != 0u;
      const unsigned long long ma = __ballot(in_a), mb = __ballot(in_b);
      if (in_a) list_a[oa + __popcll(ma & below)] = (int32_t)i;
      if (in_b) list_b[ob + __popcll(mb & below)] = (int32_t)i;
      oa += (unsigned long long)__popcll(ma);
      ob += (unsigned long long)__popcll(mb);
    }
    __syncthreads();  // s_cnt and s_base are read before the next run writes them
  }
}

// One float64 pass over the record: acc[2k] += w W(m_k), acc[2k+1] += w dW/dm(m_k)
// (GRAD = false: the yield alone).  Between trained masses j and j+1 a
// particle's weight is interpolated geometrically, W = exp(ln w_j + t ln(w_j+1 /
// w_j)), where both are positive, linearly otherwise.
template <bool GRAD>
__global__ __launch_bounds__(256) void k_dark_yield(DevTables T, const YieldTables* __restrict__ Yp, DevStore S,
                                                    RecordList L, YieldPoints P, bool bound_on, double* acc) {
  const YieldTables& Y = *Yp;
  double y[kYieldBatch], dy[kYieldBatch];
#pragma unroll
  for (int k = 0; k < kYieldBatch; ++k) y[k] = dy[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const PetiteDarkDesc& d0 = Y.d[0];
  const int64_t n = (int64_t)*L.count;
  for (int64_t ii = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ii < n; ii += stride) {
    const int64_t i = L.idx[ii];
    const int pid = S.pid[i];
    const double E0 = (double)S.p0[i].x;
    const double w = (double)S.r0[i].w;
    const WeightArgs wa = weight_args(T, pid, E0);
    const double ln_den = pid == 22 ? log(wa.den) : 0.0;
    for (int a = 0; a < d0.n_active; ++a) {
      const int code = d0.active[a];
      if (code == DARKCOMP ? pid != 22 : family_of(code, pid) < 0) continue;
      if (bound_on && (code == DARKANN || code == DARKCOMP)) continue;  // k_bound_yield
      double wv[kYieldWindow], lam[kYieldWindow];
#pragma unroll
      for (int q = 0; q < kYieldWindow; ++q) {
        wv[q] = lam[q] = 0.0;
        if (P.need & (1u << q))
          yield_weight(T, Y.d[P.j0 + q], Y.b, pid, code, wa, P.ln_eps2[q], ln_den, wv[q], lam[q]);
      }
#pragma unroll
      for (int k = 0; k < kYieldBatch; ++k) {
        if (k >= P.n) break;
        double lo = 0.0, hi = 0.0, llo = 0.0, lhi = 0.0;
        // segment P.j[k] of the window (uniform over the grid); selects
        // without indexing the register arrays
#pragma unroll
        for (int q = 0; q + 1 < kYieldWindow; ++q)
          if (P.j[k] == q) {
            lo = wv[q];
            llo = lam[q];
            hi = wv[q + 1];
            lhi = lam[q + 1];
          }
        if (P.j[k] >= kYieldWindow - 1) {
          lo = hi = wv[kYieldWindow - 1];
          llo = lhi = lam[kYieldWindow - 1];
        }
        if (P.j0 + P.j[k] + 1 >= Y.n_mass) {
          hi = lo;
          lhi = llo;
        }
        const double t = P.t[k];
        double W, dW;
        if (lo > 0.0 && hi > 0.0) {
          const double r = lhi - llo;
          W = exp(llo + t * r);
          dW = GRAD ? W * r * P.dtdm[k] : 0.0;
        } else {
          if (code == DARKCOMP) {
            lo = lo > 0.0 ? exp(llo) : 0.0;
            hi = hi > 0.0 ? exp(lhi) : 0.0;
          }
          W = lo + t * (hi - lo);
          dW = GRAD ? (hi - lo) * P.dtdm[k] : 0.0;
        }
        y[k] += w * W;
        if (GRAD) dy[k] += w * dW;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kYieldBatch; ++k) {
    double a = y[k], b = dy[k];
    for (int off = 32; off > 0; off >>= 1) {
      a += shfl_f64(a, (threadIdx.x & 63) ^ off);
      if (GRAD) b += shfl_f64(b, (threadIdx.x & 63) ^ off);
    }
    if ((threadIdx.x & 63) == 0 && k < P.n) {
      unsafeAtomicAdd(acc + 2 * k, a);
      if (GRAD) unsafeAtomicAdd(acc + 2 * k + 1, b);
    }
  }
}

// The bound-electron processes' part of the yield (DarkAnn of positrons,
// DarkComp of photons) from the tables at the pass's masses: one float64 pass
// over the record beside k_dark_yield, which leaves them out, into the same
// accumulators.  A kernel of its own keeps the interpolation kernel's
// registers (its three table lookups per mass would double them).
template <bool GRAD>
__global__ __launch_bounds__(256) void k_bound_yield(DevTables T, DevStore S, RecordList L, int32_t n_pts, int32_t ann,
                                                     int32_t comp, double ln_eps2, BoundView Bv, double* acc) {
  double y[kYieldBatch], dy[kYieldBatch];
#pragma unroll
  for (int k = 0; k < kYieldBatch; ++k) y[k] = dy[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t n = (int64_t)*L.count;
  for (int64_t ii = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ii < n; ii += stride) {
    const int64_t i = L.idx[ii];
    const int pid = S.pid[i];
    const bool is_ann = ann && pid == -11, is_comp = comp && pid == 22;
    if (!is_ann && !is_comp) continue;
    const double E0 = (double)S.p0[i].x;
    const double w = (double)S.r0[i].w;
    const WeightArgs wa = weight_args(T, pid, E0);
    const double ln_den = is_comp ? log(wa.den) : 0.0;
    const int code = is_ann ? DARKANN : DARKCOMP;
#pragma unroll
    for (int k = 0; k < kYieldBatch; ++k) {
      if (k >= n_pts) break;
      double W, dW;
      bound_wd<GRAD>(T, Bv, k, code, wa, ln_den, ln_eps2, W, dW);
      y[k] += w * W;
      if (GRAD) dy[k] += w * dW;
    }
  }
#pragma unroll
  for (int k = 0; k < kYieldBatch; ++k) {
    double a = y[k], b = dy[k];
    for (int off = 32; off > 0; off >>= 1) {
      a += shfl_f64(a, (threadIdx.x & 63) ^ off);
      if (GRAD) b += shfl_f64(b, (threadIdx.x & 63) ^ off);
    }
    if ((threadIdx.x & 63) == 0 && k < n_pts) {
      unsafeAtomicAdd(acc + 2 * k, a);
      if (GRAD) unsafeAtomicAdd(acc + 2 * k + 1, b);
    }
  }
}

}  // namespace dark
