// Device physics of the PETITE shower-transport path for gfx950.
//
// Cross sections follow src/PETITE/all_processes.py of the reference, with
// the 4-D integrands (Brem, MuonBrem, PairProd) evaluated in float32 after a
// float64 prologue that forms the cancellation-prone variables (1-x, x2-x3,
// ep-w) exactly, and with PF*FF folded so that no q^4/q^4 ratio is formed in
// float32.  The 1-D and dark integrands run in float64 with the reference's
// own formulas.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "petite_hip.h"
#include "philox.h"

namespace petite {

constexpr double kAlpha = 1.0 / 137.035999;
constexpr double kMe = 510.998950e-6;
constexpr double kMmu = 105.6583755e-3;
constexpr double kMp = 938.272088e-3;
constexpr double kPi = 3.14159265358979323846;
constexpr double kCmToM = 0.01;

enum Proc : int32_t {
  BREM = 0, ANN = 1, PAIRPROD = 2, COMP = 3, MOLLER = 4, BHABHA = 5, MUONE = 6, MUONBREM = 7,
  DARKBREM = 8, DARKANN = 9, DARKCOMP = 10, DARKMUONBREM = 11
};

struct DevTables {
  const double* f64;
  const float* f32;
  const double* maxf;  // [PETITE_NPROC][maxE]
  int32_t maxE;
  float brem_C_e;      // 8/pi alpha (alpha/m)^2 m^4 Z^2 a0^4 for m = m_e
  float brem_C_mu;     // same for m = m_mu (a0 keeps m_e, all_processes.py:92-103)
  float a0sq;          // a0^2 [GeV^-2]
  double brem_a0;      // a0 = 184.15 e^-1/2 Z^-1/3 / m_e (all_processes.py:92-94), float64
  float mcs_cA;        // Lynch-Dahl constants of the material: 0.157 Z (Z+1) / A,
  float mcs_cB;        //   2.007e-5 Z^(2/3),
  float mcs_cC;        //   3.34 (Z alpha)^2  (moliere.py:196-219)
  int32_t lds_doubles; // size of the n*sigma tables staged in LDS by the transport kernel
  const PetiteTableDesc* d;  // device copy of the descriptor (read through the scalar cache)
};

// ---------------------------------------------------------------- tables ---

// Interval index lo (0 <= lo <= n-2) with xe[lo] <= E, the largest such
// (E inside [xe[0], xe[n-1]]).  Geometric grids (inv_dlog > 0) take an O(1)
// guess from ln E and correct it against the stored nodes; others bisect.
__device__ inline int grid_interval(const double* xe, int n, double lnE0, double inv_dlog, double E) {
  int lo;
  if (inv_dlog > 0.0) {
    lo = (int)((logf((float)E) - (float)lnE0) * (float)inv_dlog);
    lo = lo < 0 ? 0 : (lo > n - 2 ? n - 2 : lo);
    while (lo > 0 && xe[lo] > E) --lo;
    while (lo < n - 2 && xe[lo + 1] <= E) ++lo;
  } else {
    lo = 0;
    int hi = n - 1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (xe[mid] <= E) lo = mid; else hi = mid;
    }
  }
  return lo;
}

// Linear interpolation with zero outside [E0, E_{n-1}] (scipy interp1d with
// fill_value=0, bounds_error=False as in src/PETITE/shower.py:280-295).
__device__ inline double nsigma(const DevTables& T, int proc, double E) {
  const PetiteSigmaDesc& s = T.d->sigma[proc];
  if (s.n <= 0) return 0.0;
  const double* xe = T.f64 + s.E_off;
  const double* ye = T.f64 + s.v_off;
  if (!(E >= xe[0]) || E > xe[s.n - 1]) return 0.0;
  if (s.n == 1) return ye[0];
  const int lo = grid_interval(xe, s.n, s.lnE0, s.inv_dlog, E);
  const double t = (E - xe[lo]) / (xe[lo + 1] - xe[lo]);
  return ye[lo] + t * (ye[lo + 1] - ye[lo]);
}

// Channels per species in the reference's order (shower.py:665-698).
__device__ inline int channels(int pid, int* ch) {
  switch (pid) {
    case 11: ch[0] = BREM; ch[1] = MOLLER; return 2;
    case -11: ch[0] = BREM; ch[1] = ANN; ch[2] = BHABHA; return 3;
    case 22: ch[0] = PAIRPROD; ch[1] = COMP; return 2;
    case 13: case -13: ch[0] = MUONE; ch[1] = MUONBREM; return 2;
    default: return 0;
  }
}

// get_mfp (shower.py:370-389): metres, 1e12 when n*sigma <= 0.
__device__ inline double mean_free_path(const DevTables& T, int pid, double E) {
  int ch[3];
  const int n = channels(pid, ch);
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += nsigma(T, ch[i], E);
  return s > 0.0 ? kCmToM / s : 1.0e12;
}

// LU_Key = argmin|E_i - E| + 1, clamped to n_E - 1 (shower.py:416-426).
__device__ inline int lu_key(const DevTables& T, int proc, double E) {
  const PetiteGridDesc& g = T.d->grid[proc];
  const double* Eg = T.f64 + g.E_off;
  const int n = g.n_E;
  if (n < 2) return 0;
  if (E <= Eg[0]) return 1;
  if (E >= Eg[n - 1]) return n - 1;
  const int lo = grid_interval(Eg, n, g.lnE0, g.inv_dlog, E);
  // argmin takes the first index on ties (numpy.argmin)
  const int best = (fabs(Eg[lo] - E) <= fabs(Eg[lo + 1] - E)) ? lo : lo + 1;
  const int k = best + 1;
  return k > n - 1 ? n - 1 : k;
}

// VEGAS map transform with an integer-exact increment index: the uniform
// y = (2m+1) 2^-24 (m = the top 23 bits of the Philox word) times N is formed
// as an exact 34-bit integer, so the bin and the fraction equal what float64
// gives on the host.  ``base`` points at the nodes of one grid energy; the
// node pairs of all dimensions are gathered before they are used.
template <int ND>
__device__ inline double map_point_n(const float* base, const int32_t* ninc, const uint32_t* r, double* x) {
  uint32_t idx[ND];
  double fr[ND];
  float g0[ND], g1[ND];
  int64_t off = 0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const uint32_t N = (uint32_t)ninc[d];
    const uint64_t q = (uint64_t)(2u * (r[d] >> 9) + 1u) * (uint64_t)N;
    idx[d] = (uint32_t)(q >> 24) + (uint32_t)off;
    fr[d] = (double)(uint32_t)(q & 0xFFFFFFu) * 0x1.0p-24;
    off += N + 1;
  }
  // each node pair in one 8-byte load (4-byte aligned, which global loads
  // allow): half the gather instructions of two 4-byte loads
  typedef float pair_t __attribute__((ext_vector_type(2), aligned(4)));
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const pair_t g = *(const pair_t*)(base + idx[d]);
    g0[d] = g.x;
    g1[d] = g.y;
  }
  double J = 1.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const double a = (double)g0[d];
    const double dg = (double)g1[d] - a;
    x[d] = a + dg * fr[d];
    J *= (double)ninc[d] * dg;
  }
  return J;
}

__device__ inline double map_point(const DevTables& T, int proc, int ie, const uint32_t* r, int ndim, double* x) {
  const PetiteGridDesc& g = T.d->grid[proc];
  const float* base = T.f32 + g.node_off + (int64_t)ie * g.stride;
  switch (ndim) {
    case 1: return map_point_n<1>(base, g.ninc, r, x);
    case 3: return map_point_n<3>(base, g.ninc, r, x);
    case 4: return map_point_n<4>(base, g.ninc, r, x);
    default: return map_point_n<2>(base, g.ninc, r, x);
  }
}

// -------------------------------------------------------------- integrands --

// sin(pi t) and cos(pi t) in float32 with the argument reduced in float64
// first (t - rint(t) is exact), so the float32 sine sees |arg| <= pi/2 with
// full relative precision near the zeros.
__device__ inline float sinpi_f(double t) {
  const double n = rint(t);
  const float s = sinf((float)((t - n) * kPi));
  return ((int64_t)n & 1) ? -s : s;
}
__device__ inline float cospi_f(double t) {
  // cos(pi t) = sin(pi (1/2 - t))
  return sinpi_f(0.5 - t);
}

// 1/x by the hardware reciprocal (1 ulp).  The float32 integrand bodies below
// only decide candidates outside the band kRefBand (4e-4 relative) around the
// accept threshold, inside which the float64 reference formula decides; their
// error against it (a few 1e-6) leaves the decisions unchanged, and the
// correctly rounded division costs several times the instructions.
__device__ inline float rcp_f(float x) { return __builtin_amdgcn_rcpf(x); }

// Brem / MuonBrem, all_processes.py:139-206.  PF*FF = C epp d dp /(w ep (1+a0^2 q^2)^2).
__device__ inline float f_brem(const DevTables& T, double ep, double m, float C, const double* x) {
  const double egmin = T.d->Eg_min;
  const double range = ep - m - egmin;
  const double w = egmin + x[0] * range;
  const double epp = ep - w;
  const double k = ep * (0.5 / m);
  const double d = k * (x[1] + x[2]);
  const double dp = k * (x[1] - x[2]);
  if (!((egmin < w) && (w < ep - m) && (m < epp) && (epp < ep) && (d > 0.0) && (dp > 0.0))) return 0.0f;
  const float fd = (float)d, fdp = (float)dp, fw = (float)w, fepp = (float)epp, fep = (float)ep, fm = (float)m;
  const float dmd = (float)(2.0 * k * x[2]);                 // d - dp
  const float Bn = (float)(-w + d * d * epp - dp * dp * ep);  // (1+d^2) epp - (1+dp^2) ep
  const float s2 = sinpi_f(x[3] - 0.5);                      // sin(ph/2)
  const float r_pp = rcp_f(fep * fepp);                      // 1/(ep epp)
  const float B = 0.5f * Bn * r_pp;
  const float qm2 = dmd * dmd + 4.0f * fd * fdp * s2 * s2 + fm * fm * B * B;
  const float qsq = fm * fm * qm2;
  const float rden = rcp_f(1.0f + T.a0sq * qsq);
  const float pfff = C * fepp * fd * fdp * rcp_f(fw * fep) * rden * rden;
  const float jac = (float)kPi * fep * fep * (float)range * (float)(1.0 / (m * m));
  // T1+T2+T3+T4 of the reference as a sum of non-negative terms: the direct
  // sum cancels to O((w/ep)^2) for soft photons.  With a=d, b=dp, A=1/(1+a^2),
  // B=1/(1+b^2):  (a-b)^2 (1-ab)^2 A^2 B^2 + 4 a b sin^2(ph/2) A B
  //               + w^2/(2 ep epp) A B ((a-b)^2 + 4 a b sin^2(ph/2))
  const float i12 = rcp_f((1.0f + fd * fd) * (1.0f + fdp * fdp));
  const float ab = fd * fdp;
  const float q1 = dmd * (1.0f - ab) * i12;
  const float ang = 4.0f * ab * s2 * s2;
  const float bracket = q1 * q1 + ang * i12 + 0.5f * fw * fw * r_pp * i12 * (dmd * dmd + ang);
  return pfff * bracket * jac;
}

// PairProd, all_processes.py:531-622.
__device__ inline float f_pairprod(const DevTables& T, double w, const double* x) {
  const double m = kMe;
  const double r2 = w - 2.0 * m;
  const double epp = m + x[0] * r2;
  const double epm = m + (1.0 - x[0]) * r2;
  const double k = w * (0.5 / m);
  const double dp = k * (x[1] + x[2]);
  const double dm = k * (x[1] - x[2]);
  if (!((m < epm) && (epm < w) && (m < epp) && (epp < w) && (dm > 0.0) && (dp > 0.0))) return 0.0f;
  const float fdp = (float)dp, fdm = (float)dm, fw = (float)w, fepp = (float)epp, fepm = (float)epm, fm = (float)m;
  const float dpm = (float)(2.0 * k * x[2]);                 // dp - dm
  const float c2 = cospi_f(x[3]);                            // cos(ph/2)
  const float dp2 = fdp * fdp, dm2 = fdm * fdm;
  const float r_pm = rcp_f(fepp * fepm);                     // 1/(epp epm)
  const float S = 0.5f * ((1.0f + dp2) * fepm + (1.0f + dm2) * fepp) * r_pm;
  const float qq = dpm * dpm + 4.0f * fdp * fdm * c2 * c2 + fm * fm * S * S;
  const float rden = rcp_f(1.0f + T.a0sq * fm * fm * qq);
  const float pfff = T.brem_C_e * fepp * fepm * fdp * fdm * rcp_f(fw * fw * fw) * rden * rden;
  const float jac = (float)kPi * fw * fw * (float)r2 * (float)(1.0 / (m * m));
  // T1+T2+T3+T4 of the reference as a sum of non-negative terms (a=dp, b=dm):
  //   (a^2-b^2)^2 A^2 B^2 + (epp^2+epm^2)/(2 epp epm) A B ((a-b)^2 + 4 a b cos^2(ph/2))
  const float i12 = rcp_f((1.0f + dp2) * (1.0f + dm2));
  const float t1 = dpm * (fdp + fdm) * i12;
  const float bracket = t1 * t1 + 0.5f * (fepp * fepp + fepm * fepm) * r_pm * i12 *
                                      (dpm * dpm + 4.0f * fdp * fdm * c2 * c2);
  return pfff * bracket * jac;
}

// The Brem and PairProd integrands in float64 with the reference's own term
// order (all_processes.py:139-206, 531-622): T1+T2+T3+T4 summed directly and
// the form factor g2_elastic (99-103) applied as a separate factor.  The
// draws take these only for a candidate whose float32 J*f lies within a
// small band of the accept threshold, so the accept decision is the one a
// float64 evaluation of the reference formula makes.  Not inlined (and the
// point passed by value): the rare call keeps the draw loop's registers at
// the float32 body's count.
__device__ __noinline__ double f_brem_ref(const DevTables& T, double ep, double m, double x0, double x1, double x2,
                                          double x3) {
  const double x[4] = {x0, x1, x2, x3};
  const double egmin = T.d->Eg_min;
  const double w = egmin + x[0] * (ep - m - egmin);
  const double d = ep / (2.0 * m) * (x[1] + x[2]);
  const double dp = ep / (2.0 * m) * (x[1] - x[2]);
  const double ph = (x[3] - 0.5) * 2.0 * kPi;
  const double epp = ep - w;
  if (!((egmin < w) && (w < ep - m) && (m < epp) && (epp < ep) && (d > 0.0) && (dp > 0.0))) return 0.0;
  const double cph = cos(ph);
  const double bb = (1.0 + d * d) / (2.0 * ep) - (1.0 + dp * dp) / (2.0 * epp);
  const double qsq = m * m * ((d * d + dp * dp - 2.0 * d * dp * cph) + m * m * bb * bb);
  const double a = kAlpha / m;
  const double PF = 8.0 / kPi * kAlpha * (a * a) * (epp * (m * m * m * m)) / (w * ep * (qsq * qsq)) * d * dp;
  const double jac = kPi * ep * ep * (ep - m - egmin) / (m * m);
  const double a0 = T.brem_a0;
  const double Z = T.d->Z;
  const double a0t = 1.0 + a0 * a0 * qsq;
  const double FF = Z * Z * (a0 * a0 * a0 * a0) * (qsq * qsq) / (a0t * a0t);
  const double A = 1.0 + d * d, B = 1.0 + dp * dp;
  const double T1 = d * d / (A * A);
  const double T2 = dp * dp / (B * B);
  const double T3 = w * w / (2.0 * ep * epp) * (d * d + dp * dp) / (A * B);
  const double T4 = -(epp / ep + ep / epp) * (d * dp * cph) / (A * B);
  return PF * (T1 + T2 + T3 + T4) * jac * FF;
}

__device__ __noinline__ double f_pairprod_ref(const DevTables& T, double w, double x0, double x1, double x2,
                                              double x3) {
  const double x[4] = {x0, x1, x2, x3};
  const double m = kMe;
  const double epp = m + x[0] * (w - 2.0 * m);
  const double dp = w / (2.0 * m) * (x[1] + x[2]);
  const double dm = w / (2.0 * m) * (x[1] - x[2]);
  const double ph = x[3] * 2.0 * kPi;
  const double epm = w - epp;
  if (!((m < epm) && (epm < w) && (m < epp) && (epp < w) && (dm > 0.0) && (dp > 0.0))) return 0.0;
  const double cph = cos(ph);
  const double S = (1.0 + dp * dp) / (2.0 * epp) + (1.0 + dm * dm) / (2.0 * epm);
  const double qq = (dp * dp + dm * dm + 2.0 * dp * dm * cph) + m * m * S * S;
  const double a = kAlpha / m;
  const double PF = 8.0 / kPi * kAlpha * (a * a) * epp * epm / (w * w * w * (qq * qq)) * dp * dm;
  const double jac = kPi * w * w * (w - 2.0 * m) / (m * m);
  const double a0 = T.brem_a0;
  const double Z = T.d->Z;
  const double t = m * m * qq;
  const double a0t = 1.0 + a0 * a0 * t;
  const double FF = Z * Z * (a0 * a0 * a0 * a0) * (t * t) / (a0t * a0t);
  const double A = 1.0 + dp * dp, B = 1.0 + dm * dm;
  const double T1 = -1.0 * dp * dp / (A * A);
  const double T2 = -1.0 * dm * dm / (B * B);
  const double T3 = w * w / (2.0 * epp * epm) * (dp * dp + dm * dm) / (A * B);
  const double T4 = (epp / epm + epm / epp) * (dp * dm * cph) / (A * B);
  return PF * (T1 + T2 + T3 + T4) * jac * FF;
}

// Compton / DarkComp on a free electron at rest, all_processes.py:624-742.
__device__ inline double f_compton(double Eg, double mV, double ct) {
  const double m = kMe;
  const double m2 = m * m, mV2 = mV * mV;
  const double s = m2 + 2.0 * Eg * m;
  if (s < (m + mV) * (m + mV)) return 0.0;
  const double R = sqrt(m2 * m2 + (mV2 - s) * (mV2 - s) - 2.0 * m2 * (mV2 + s));
  const double jac = (s - m2) / (2.0 * s) * sqrt((s - mV2) * (s - mV2) - 2.0 * m2 * (s + mV2) + m2 * m2);
  const double t = -0.5 * (m2 * m2 + s * (-mV2 + s + ct * R) - m2 * (mV2 + 2.0 * s + ct * R)) / s;
  const double PF = 2.0 * kPi * kAlpha * kAlpha / ((s - m2) * (s - m2));
  const double T1 = (2.0 * m2 * (mV2 - 3.0 * s) - 3.0 * m2 * m2 - 2.0 * mV2 * s + 2.0 * mV2 * mV2 + s * s) /
                    ((m2 - s) * (m2 + mV2 - s - t));
  const double D = m2 + mV2 - s - t;
  const double T2 = (2.0 * m2 * (2.0 * m2 + mV2)) / (D * D);
  const double T3 = ((m2 + s) * (m2 + mV2 + s) + t * (s - m2)) / ((m2 - s) * (m2 - s));
  return PF * jac * (T1 + T2 + T3);
}

// Ann (e+ e- -> gamma gamma), all_processes.py:468-529 (mV = 0, Eg_min cut).
__device__ inline double f_annihilation(double Ee, double mV, double egmin, double ct) {
  const double m = kMe;
  const double s = 2.0 * m * (Ee + m);
  const double ctMax = sqrt((Ee + m) / (Ee - m)) * (2.0 * m * (Ee - 2.0 * egmin + m) - mV * mV) /
                       (2.0 * m * (Ee + m) - mV * mV);
  if (s < mV * mV) return 0.0;
  if (ct > ctMax) return 0.0;
  const double b2 = 1.0 - 4.0 * m * m / s;
  return 4.0 * kPi * kAlpha * kAlpha / (s * (1.0 - b2 * ct * ct)) *
         ((s - mV * mV) / (2.0 * s) * (1.0 + ct * ct) + 2.0 * mV * mV / (s - mV * mV));
}

// Moller, all_processes.py:786-840.
__device__ inline double f_moller(double Ee, double DE, double ct) {
  const double m = kMe;
  const double dct = 2.0 * DE / (Ee - m);
  if (!((ct > -1.0 + dct) && (ct < 1.0 - dct))) return 0.0;
  const double s = m * m + 2.0 * Ee * m;
  const double c2 = ct * ct;
  const double num = s * s * (3.0 + c2) * (3.0 + c2) - 8.0 * m * m * s * (7.0 + c2 * c2) +
                     16.0 * m * m * m * m * (6.0 - 3.0 * c2 + c2 * c2);
  const double sm = s - 4.0 * m * m;
  return 16.0 * kPi * kPi * kAlpha * kAlpha * num /
         (8.0 * kPi * s * sm * sm * (1.0 - ct) * (1.0 - ct) * (1.0 + ct) * (1.0 + ct));
}

// Bhabha, all_processes.py:947-1007.
__device__ inline double f_bhabha(double Ee, double DE, double ct) {
  const double m = kMe;
  const double dct = 2.0 * DE / (Ee - m);
  if (!((ct > -1.0 + dct) && (ct < 1.0 - dct))) return 0.0;
  const double s = m * m + 2.0 * Ee * m;
  const double m2 = m * m, m4 = m2 * m2, m6 = m4 * m2, m8 = m4 * m4;
  const double cm1 = -1.0 + ct;
  const double num = 256.0 * cm1 * cm1 * ct * ct * m8 -
                     128.0 * cm1 * (1.0 + ct * (1.0 + ct) * (-3.0 + 2.0 * ct)) * m6 * s +
                     16.0 * (7.0 + ct * (2.0 + ct * (-5.0 + 6.0 * cm1 * ct))) * m4 * s * s -
                     8.0 * (7.0 + ct * (-3.0 + ct * (3.0 + ct * (-1.0 + 2.0 * ct)))) * m2 * s * s * s +
                     (3.0 + ct * ct) * (3.0 + ct * ct) * s * s * s * s;
  const double sm = -4.0 * m2 + s;
  return kAlpha * kAlpha * kPi * num / (2.0 * cm1 * cm1 * s * s * s * sm * sm);
}

// mu e -> mu e, all_processes.py:842-886.
__device__ inline double f_muone(double Emu, double DE, double ct) {
  const double me = kMe, mm = kMmu;
  const double s = me * me + mm * mm + 2.0 * me * Emu;
  const double t_limit = 2.0 * me * (me - DE);
  const double a = s + me * me - mm * mm;
  const double t = -2.0 * (1.0 - ct) * (a * a / (4.0 * s) - me * me);
  if (!(t < t_limit)) return 0.0;
  const double b = s + t - 4.0 * me * me;
  return 16.0 * kPi * kPi * kAlpha * kAlpha *
         (s * s + 2.0 * (me * me + mm * mm) * (2.0 * t + mm * mm - 3.0 * me * me) + b * b) / (16.0 * kPi * s * t * t);
}

// Gninenko et al. elastic+inelastic form factor over t^2, all_processes.py:111-133.
__device__ inline double g_el_inel_over_t2(double Z, double A, double t) {
  const double mu_p = 2.79;
  const double c1 = pow(111.0 * pow(Z, -1.0 / 3.0) / kMe, 2.0);
  const double c2 = 0.164 * pow(A, -2.0 / 3.0);
  const double a1 = 1.0 / (1.0 + c1 * t);
  const double a2 = 1.0 + t / c2;
  const double Gel = a1 * a1 / (a2 * a2);
  const double ap2 = pow(773.0 * pow(Z, -2.0 / 3.0) / kMe, 2.0);
  const double b1 = ap2 / (1.0 + ap2 * t);
  const double b2 = 1.0 + t / 0.71;
  const double Ginel = Z / (c1 * c1 * Z * Z) * b1 * b1 *
                       ((1.0 + (mu_p * mu_p - 1.0) * t / (4.0 * kMp * kMp)) / (b2 * b2 * b2 * b2));
  return Z * Z * c1 * c1 * (Gel + Ginel);
}

// DarkBrem / DarkMuonBrem exact tree level, "Log" variables, all_processes.py:208-372.
__device__ inline double f_dark_brem(double Ebeam, double m_lepton, double mV, double MT, double Z, double A,
                                     const double* xv) {
  const double x = xv[0];
  const double one_minus_ct = pow(10.0, xv[1]);
  const double ct = 1.0 - one_minus_ct;
  const double ttilde = pow(10.0, xv[2]);
  const double ln10 = 2.302585092994046;
  const double Jac = one_minus_ct * ttilde * ln10 * ln10;
  const double k = sqrt(fabs((x * Ebeam) * (x * Ebeam) - mV * mV));
  const double p = sqrt(Ebeam * Ebeam - m_lepton * m_lepton);
  const double V = sqrt(p * p + k * k - 2.0 * p * k * ct);
  const double ut = -2.0 * (x * Ebeam * Ebeam - k * p * ct) + mV * mV;
  const double E1 = (1.0 - x) * Ebeam + MT;
  const double discr = ut * ut + 4.0 * MT * ut * E1 + 4.0 * MT * MT * V * V;
  const double sq = sqrt(fabs(discr));
  const double den = 2.0 * E1 * E1 - 2.0 * V * V;
  const double Qp = fabs((V * (ut + 2.0 * MT * E1) + E1 * sq) / den);
  const double Qm = fabs((V * (ut + 2.0 * MT * E1) - E1 * sq) / den);
  const double tplus = 2.0 * MT * (sqrt(MT * MT + Qp * Qp) - MT);
  const double tminus = 2.0 * MT * (sqrt(MT * MT + Qm * Qm) - MT);
  const double tc = 2.0 * MT * (MT + Ebeam) * sqrt(Ebeam * Ebeam + m_lepton * m_lepton) /
                    (MT * (MT + 2.0 * Ebeam) + m_lepton * m_lepton);
  const double tconv = tc * tc;
  const double t = ttilde * tconv;
  const double q0 = -t / (2.0 * MT);
  const double q = sqrt(t * t / (4.0 * MT * MT) + t);
  const double cq = -(V * V + q * q + m_lepton * m_lepton - (Ebeam + q0 - x * Ebeam) * (Ebeam + q0 - x * Ebeam)) /
                    (2.0 * V * q);
  const double mv2 = mV * mV + 2.0 * m_lepton * m_lepton;
  const double Am2 = -8.0 * MT * (4.0 * Ebeam * Ebeam * MT - t * (2.0 * Ebeam + MT)) * mv2;
  const double A1 = 8.0 * MT * MT / ut;
  const double Am1 = (8.0 / ut) *
                     (MT * MT *
                          (2.0 * t * ut + ut * ut +
                           4.0 * Ebeam * Ebeam * (2.0 * (x - 1.0) * mv2 - t * ((x - 2.0) * x + 2.0)) +
                           2.0 * t * (-(mV * mV) + 2.0 * m_lepton * m_lepton + t)) -
                      2.0 * Ebeam * MT * t * ((1.0 - x) * ut + (x - 2.0) * (mv2 + t)) + t * t * (ut - mV * mV));
  const double A0 = (8.0 / (ut * ut)) * (MT * MT * (2.0 * t * ut + (t - 4.0 * Ebeam * Ebeam * (x - 1.0) * (x - 1.0)) * mv2) +
                                        2.0 * Ebeam * MT * t * (ut - (x - 1.0) * mv2));
  const double Y = -t + 2.0 * q0 * Ebeam - 2.0 * q * p * (p - k * ct) * cq / V;
  const double W = fabs(Y * Y - 4.0 * q * q * p * p * k * k * (1.0 - ct * ct) * (1.0 - cq * cq) / (V * V));
  const bool ok = (x * Ebeam >= mV) && (discr >= 0.0) && (tplus > tminus) && (t > tminus) && (t < tplus) &&
                  (fabs(cq) <= 1.0) && (W > 0.0);
  if (!ok) return 0.0;
  const double phi_int = (A0 + Y * A1 + Am1 / sqrt(W) + Y * Am2 / (W * sqrt(W))) / (8.0 * MT * MT);
  const double ff = g_el_inel_over_t2(Z, A, t);
  const double ans = ff * kAlpha * kAlpha * kAlpha * k * Ebeam * phi_int / (p * sqrt(k * k + p * p - 2.0 * p * k * ct));
  const double out = ans * tconv * Jac;
  return isfinite(out) ? out : 0.0;
}

// Kuraev-Fadin structure functions, radiative_return.py:26-44, 61-79.
__device__ inline double fl_kf(double x, double s) {
  const double beta = (2.0 * kAlpha / kPi) * (log(s / (kMe * kMe)) - 1.0);
  if (x >= 1.0) x = 1.0 - 1e-10;
  return (beta / 16.0) * ((8.0 + 3.0 * beta) * pow(1.0 - x, beta / 2.0 - 1.0) - 4.0 * (1.0 + x));
}
__device__ inline double fl_kf_scaled(double x, double s) {
  const double beta = (2.0 * kAlpha / kPi) * (log(s / (kMe * kMe)) - 1.0);
  return (beta / 16.0) * ((8.0 + 3.0 * beta) - 4.0 * (1.0 + x) * pow(1.0 - x, 1.0 - beta / 2.0));
}

// DarkAnn radiative return in u, all_processes.py:400-466.
__device__ inline double f_dark_ann(double Ee, double mV, double u0) {
  const double s = 2.0 * kMe * (Ee + kMe);
  if (s < mV * mV) return 0.0;
  const double beta = (2.0 * kAlpha / kPi) * (log(s / (kMe * kMe)) - 1.0);
  const double umax = pow(1.0 - mV * mV / s, beta / 2.0);
  const double betaf = sqrt(1.0 - 4.0 * kMe * kMe / (mV * mV));
  const double prefac = (4.0 * kPi * kPi) * kAlpha * betaf * (1.5 - betaf * betaf / 2.0) / s * umax;
  const double x1 = 1.0 - pow(u0 * umax, 2.0 / beta);
  const double x2 = mV * mV / (x1 * s);
  if (!((x2 < 1.0) && (x1 > 0.0) && (u0 < 1.0))) return 0.0;
  const double y = mV * mV / s;
  const double xx = 1.0 - pow(u0 * umax, 2.0 / beta);
  const double lumi = fl_kf(y / xx, s) * fl_kf_scaled(xx, s) * (1.0 / xx) * (2.0 / beta);
  const double out = 2.0 * prefac * lumi;
  return isfinite(out) ? out : 0.0;
}

// J*f of one candidate; x holds the mapped point.  E is the true energy.
__device__ inline double integrand(const DevTables& T, int proc, double E, const double* x) {
  switch (proc) {
    case BREM: return (double)f_brem(T, E, kMe, T.brem_C_e, x);
    case MUONBREM: return (double)f_brem(T, E, kMmu, T.brem_C_mu, x);
    case PAIRPROD: return (double)f_pairprod(T, E, x);
    case COMP: return f_compton(E, 0.0, x[0]);
    case ANN: return f_annihilation(E, 0.0, T.d->Eg_min, x[0]);
    case MOLLER: return f_moller(E, T.d->Ee_min, x[0]);
    case BHABHA: return f_bhabha(E, T.d->Ee_min, x[0]);
    case MUONE: return f_muone(E, T.d->Ee_min, x[0]);
    case DARKBREM: return f_dark_brem(E, kMe, T.d->mV, T.d->mT, T.d->Z, T.d->A, x);
    case DARKMUONBREM: return f_dark_brem(E, kMmu, T.d->mV, T.d->mT, T.d->Z, T.d->A, x);
    case DARKANN: return f_dark_ann(E, T.d->mV, x[0]);
    case DARKCOMP: return f_compton(E, T.d->mV, x[0]);
    default: return 0.0;
  }
}

}  // namespace petite
