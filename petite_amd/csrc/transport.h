// Kinematics, multiple scattering and the particle record for the transport
// kernels.  Kinematics run in float64 (once per hard interaction); the
// per-substep MCS runs in float32 with momentum formed without 1 - beta^2.
#pragma once
#include "physics.h"

namespace petite {

struct V4 {
  double e, x, y, z;
};

// ---------------------------------------------------------- kinematics ----
// Final states in the parent frame (parent along +z), src/PETITE/kinematics.py.
// ``az`` is the azimuth uniform.  Returns the two outgoing four-vectors in the
// reference's order.
__device__ inline void kin_brem(double ep, double m, const float* xs, double az, V4& l, V4& g) {
  const double egmin = 0.001;  // kinematics.py:9
  const double x1 = xs[0], x2 = xs[1], x3 = xs[2], x4 = xs[3];
  const double w = egmin + x1 * (ep - m - egmin);
  double st, ct, stp, ctp;
  sincos((x2 + x3) / 2.0, &st, &ct);
  sincos((x2 - x3) * ep / (2.0 * (ep - w)), &stp, &ctp);
  st = fabs(st);
  stp = fabs(stp);
  const double ph = (x4 - 0.5) * 2.0 * kPi;
  const double epp = ep - w;
  const double pp = sqrt(epp * epp - m * m);
  double sal, cal, sp, cp;
  sincos(2.0 * kPi * az, &sal, &cal);
  sincos(ph, &sp, &cp);
  g = V4{w, w * cal * st, w * sal * st, w * ct};
  l = V4{epp, pp * (sal * sp * stp + cal * (ctp * st - cp * ct * stp)),
         pp * (ctp * sal * st - (cp * ct * sal + cal * sp) * stp), pp * (ct * ctp + cp * st * stp)};
}

__device__ inline void kin_pairprod(double w, const float* xs, double az, V4& pos, V4& ele) {
  const double m = kMe;
  const double x1 = xs[0], x2 = xs[1], x3 = xs[2], x4 = xs[3];
  const double epp = m + x1 * (w - 2.0 * m);
  double stp, ctp, stm, ctm;
  sincos(w * (x2 + x3) / (2.0 * epp), &stp, &ctp);
  sincos(w * (x2 - x3) / (2.0 * (w - epp)), &stm, &ctm);
  stp = fabs(stp);
  stm = fabs(stm);
  const double ph = x4 * 2.0 * kPi;
  const double epm = w - epp;
  const double pm = sqrt(epm * epm - m * m), pp = sqrt(epp * epp - m * m);
  const double al = 2.0 * kPi * az;
  double sal, cal, spal, cpal;
  sincos(al, &sal, &cal);
  sincos(ph + al, &spal, &cpal);
  pos = V4{epp, pp * stp * cal, pp * stp * sal, pp * ctp};
  ele = V4{epm, pm * stm * cpal, pm * stm * spal, pm * ctm};
}

// 2 -> 2 on an electron at rest, boosted from the CM frame (compton, ee_to_ee,
// mue_to_mue, annihilation of kinematics.py).  sgn selects the reference's
// sign convention of each function.
__device__ inline void kin_compton(double Eg, double mV, double ct, double az, V4& e, V4& v) {
  const double m = kMe;
  const double s = m * m + 2.0 * Eg * m;
  const double rs = sqrt(s);
  const double Ee0 = (s + m * m) / (2.0 * rs);
  const double Ee = (s - mV * mV + m * m) / (2.0 * rs);
  const double EV = (s + mV * mV - m * m) / (2.0 * rs);
  const double pF = sqrt(Ee * Ee - m * m);
  const double g0 = Ee0 / m;
  const double b0 = sqrt(g0 * g0 - 1.0) / g0;
  double sp, cp;
  sincos(2.0 * kPi * az, &sp, &cp);
  const double st = sqrt(1.0 - ct * ct);
  e = V4{g0 * Ee + b0 * g0 * pF * ct, -pF * st * sp, -pF * st * cp, b0 * g0 * Ee + g0 * pF * ct};
  v = V4{g0 * EV - b0 * g0 * pF * ct, pF * st * sp, pF * st * cp, b0 * g0 * EV - g0 * pF * ct};
}

__device__ inline void kin_ee(double E, double ct, double az, V4& a, V4& b) {
  const double m = kMe;
  const double s = 2.0 * m * m + 2.0 * E * m;
  const double Ee0 = sqrt(s) / 2.0;
  const double pF = sqrt(Ee0 * Ee0 - m * m);
  const double g0 = Ee0 / m;
  const double b0 = sqrt(g0 * g0 - 1.0) / g0;
  double sp, cp;
  sincos(2.0 * kPi * az, &sp, &cp);
  const double st = sqrt(1.0 - ct * ct);
  a = V4{g0 * Ee0 + b0 * g0 * pF * ct, -pF * st * sp, -pF * st * cp, b0 * g0 * Ee0 + g0 * pF * ct};
  b = V4{g0 * Ee0 - b0 * g0 * pF * ct, pF * st * sp, pF * st * cp, b0 * g0 * Ee0 - g0 * pF * ct};
}

__device__ inline void kin_mue(double E, double ct, double az, V4& mu, V4& el) {
  const double me = kMe, mm = kMmu;
  const double s = me * me + mm * mm + 2.0 * E * me;
  const double rs = sqrt(s);
  const double Ee0 = (s + me * me - mm * mm) / (2.0 * rs);
  const double Em0 = (s + mm * mm - me * me) / (2.0 * rs);
  const double pe = sqrt(Ee0 * Ee0 - me * me);
  const double pm = sqrt(Em0 * Em0 - mm * mm);
  const double g0 = Ee0 / me;
  const double b0 = sqrt(g0 * g0 - 1.0) / g0;
  double sp, cp;
  sincos(2.0 * kPi * az, &sp, &cp);
  const double st = sqrt(1.0 - ct * ct);
  mu = V4{g0 * Em0 + b0 * g0 * pm * ct, pm * st * sp, pm * st * cp, b0 * g0 * Em0 + g0 * pm * ct};
  el = V4{g0 * Ee0 - b0 * g0 * pe * ct, -pe * st * sp, -pe * st * cp, b0 * g0 * Ee0 - g0 * pe * ct};
}

__device__ inline void kin_ann(double E, double mV, double ct, double az, V4& g, V4& v) {
  const double m = kMe;
  const double s = 2.0 * m * (E + m);
  const double rs = sqrt(s);
  const double EeCM = rs / 2.0;
  const double Eg = (s - mV * mV) / (2.0 * rs);
  const double EV = (s + mV * mV) / (2.0 * rs);
  const double pF = Eg;
  const double g0 = EeCM / m;
  const double b0 = sqrt(g0 * g0 - 1.0) / g0;
  double sp, cp;
  sincos(2.0 * kPi * az, &sp, &cp);
  const double st = sqrt(1.0 - ct * ct);
  g = V4{g0 * Eg - b0 * g0 * pF * ct, -pF * st * sp, -pF * st * cp, b0 * g0 * Eg - g0 * pF * ct};
  v = V4{g0 * EV + b0 * g0 * pF * ct, pF * st * sp, pF * st * cp, b0 * g0 * EV + g0 * pF * ct};
}

// Children PIDs (0 = parent PID), shower.py:87-96.
__device__ inline void child_pids(int proc, int parent, int* a, int* b) {
  switch (proc) {
    case PAIRPROD: *a = -11; *b = 11; return;
    case BREM: case MUONBREM: *a = parent; *b = 22; return;
    case COMP: *a = 11; *b = 22; return;
    case ANN: *a = 22; *b = 22; return;
    default: *a = parent; *b = 11; return;  // Moller, Bhabha, MuonE
  }
}

__device__ inline double pid_mass(int pid) {
  const int a = pid < 0 ? -pid : pid;
  if (a == 11) return kMe;
  if (a == 13) return kMmu;
  return 0.0;
}

// Rotation that takes +z to the unit direction n (Particle.rotation_matrix,
// particle.py:176-185, with the angles' sines and cosines formed directly).
struct Frame {
  float c_t, s_t, c_p, s_p;
};
__device__ inline Frame frame_of(float nx, float ny, float nz) {
  const float rho = sqrtf(nx * nx + ny * ny);
  Frame f;
  f.c_t = nz;
  f.s_t = rho;
  if (rho > 0.0f) {
    const float ir = __builtin_amdgcn_rcpf(rho);
    f.c_p = nx * ir;
    f.s_p = ny * ir;
  } else {
    f.c_p = 1.0f;
    f.s_p = 0.0f;
  }
  return f;
}
__device__ inline void frame_apply(const Frame& f, float lx, float ly, float lz, float& ox, float& oy, float& oz) {
  ox = f.c_t * f.c_p * lx - f.s_p * ly + f.s_t * f.c_p * lz;
  oy = f.c_t * f.s_p * lx + f.c_p * ly + f.s_t * f.s_p * lz;
  oz = -f.s_t * lx + f.c_t * lz;
}

// Same rotation in float64 for the children of an interaction.
__device__ inline void frame_apply_d(double nx, double ny, double nz, const V4& v, double& ox, double& oy, double& oz) {
  const double rho = sqrt(nx * nx + ny * ny);
  const double ct = nz, st = rho;
  double cp = 1.0, sp = 0.0;
  if (rho > 0.0) {
    cp = nx / rho;
    sp = ny / rho;
  }
  ox = ct * cp * v.x - sp * v.y + st * cp * v.z;
  oy = ct * sp * v.x + cp * v.y + st * sp * v.z;
  oz = -st * v.x + ct * v.z;
}

// -------------------------------------------------------------- MCS -----
// Lynch-Dahl Gaussian width (moliere.py:196-219, 265-284) with the material
// constants folded: chi_c^2 = cA t / (p beta)^2 and
// v = chi_c^2 / (2 (1-F) chi_a^2) = cA t / (2 (1-F) cB (beta^2 + cC)).
// ln(1 + y) for y > -1 to a few ulp: Goldberg's correction of the rounded
// sum u = 1 + y (D. Goldberg, What every computer scientist should know about
// floating-point arithmetic, ACM Comput. Surv. 23 (1991), Theorem 4), with the
// hardware logarithm and reciprocal.
__device__ inline float log1p_f32(float y) {
  const float u = 1.0f + y;
  const float d = u - 1.0f;
  return d == 0.0f ? y : __logf(u) * (y * __builtin_amdgcn_rcpf(d));
}

__device__ inline float lynch_dahl_theta0(float cA, float cB, float cC, float t_gcm2, float pMeV, float beta) {
  // quotients with the hardware reciprocal (about 1 ulp): the angle is a float32 quantity
  const float F = 0.98f;
  const float pb = pMeV * beta;
  const float ct = cA * t_gcm2;
  const float chic2 = ct * __builtin_amdgcn_rcpf(pb * pb);
  const float v = ct * __builtin_amdgcn_rcpf(2.0f * (1.0f - F) * cB * (beta * beta + cC));
  float g;  // (1+v) ln(1+v)/v - 1
  if (v < 0.1f) {
    // sum_{n>=1} (-1)^{n+1} v^n / (n (n+1))
    g = v * (0.5f + v * (-1.0f / 6.0f + v * (1.0f / 12.0f + v * (-1.0f / 20.0f + v * (1.0f / 30.0f + v * (-1.0f / 42.0f + v * (1.0f / 56.0f)))))));
  } else {
    g = (1.0f + v) * log1p_f32(v) * __builtin_amdgcn_rcpf(v) - 1.0f;
  }
  return sqrtf(chic2 * g * (1.0f / (1.0f + F * F)));
}

}  // namespace petite
