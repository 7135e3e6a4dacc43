// twobody.hpp -- the softened two-body propagator of the Sakura kernel,
// compiled for the device (sakura.hip) and, for CPU-side checks of the same
// code, for the host.
//
// For one pair the relative orbit (r, v) with total mass m is advanced by
// the flag-selected composition of a free drift and the exact softened
// two-body motion (tupan/lib/src/sakura_kernel_common.h:94-191): a softened
// leapfrog when r > 64 m / v^2, otherwise the universal-variable Kepler
// solution -- Laguerre-Conway iteration on the universal Kepler equation
// with Stumpff functions (tupan/lib/src/universal_kepler_solver.h:17-357),
// retried on n = 2, 4, ... equal sub-steps until every sub-step converges
// (:466-507) and, with softening, sub-stepped further until the softened
// two-body energy is conserved to 64 * 2^-42 (:510-614).
//
// The reference retries without bound.  Here a pair carries a budget of
// Kepler solves (default kBudget; the softened tight binaries of the test
// set need up to 2^19 in the reference itself); a pair that exhausts it
// keeps its last state and pair_evolve returns false.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TUPAN_HD __host__ __device__
#else
#define TUPAN_HD
#endif

namespace tupan {


constexpr double kKepTol = 2.2737367544323205948e-13;  // 2^-42
constexpr int kKepMaxIter = 64;
constexpr int kLaguerreOrder = 5;
constexpr int kBudget = 1 << 20;  // default Kepler solves per pair

TUPAN_HD inline int sgn(double x) { return (x > 0) - (x < 0); }

TUPAN_HD inline double stumpff_c0(double z) {
  if (z < 0) return cos(sqrt(-z));
  if (z > 0) return cosh(sqrt(z));
  return 1.0;
}
TUPAN_HD inline double stumpff_c1(double z) {
  if (z < 0) { double s = sqrt(-z); return sin(s) / s; }
  if (z > 0) { double s = sqrt(z); return sinh(s) / s; }
  return 1.0;
}
TUPAN_HD inline double stumpff_c2(double z) {
  if (z < 0) { double s = sqrt(-z); return (cos(s) - 1) / z; }
  if (z > 0) { double s = sqrt(z); return (cosh(s) - 1) / z; }
  return 0.5;
}
TUPAN_HD inline double stumpff_c3(double z) {
  if (z < 0) { double s = sqrt(-z); return (sin(s) / s - 1) / z; }
  if (z > 0) { double s = sqrt(z); return (sinh(s) / s - 1) / z; }
  return 1.0 / 6.0;
}

struct KepEq {
  double dt, r0, rv, m, alpha;
};

// Universal Kepler equation f(s) = r0 s + r0.v0 S2 + (m + alpha r0) S3 - dt
// and its first two s-derivatives (universal_kepler_solver.h:166-198).
TUPAN_HD inline void kep_eval(double s, const KepEq &k, double *f,
                                         double *df, double *ddf) {
  double s2 = s * s;
  double z = k.alpha * s2;
  double S0 = stumpff_c0(z);
  double S1 = s * stumpff_c1(z);
  double S2 = s2 * stumpff_c2(z);
  double S3 = s * s2 * stumpff_c3(z);
  double mu = k.m + k.alpha * k.r0;
  *f = k.r0 * s + k.rv * S2 + mu * S3 - k.dt;
  *df = k.r0 + k.rv * S1 + mu * S2;
  *ddf = k.rv * S0 + mu * S1;
}

// Laguerre-Conway iteration, order 5 (universal_kepler_solver.h:225-254):
// 0 = converged, -1 zero denominator, -2 iteration cap, -3 sign flip.
TUPAN_HD inline int laguerre(double s0, double *s, const KepEq &k) {
  double x = s0, delta;
  int it = 0;
  do {
    double f, df, ddf;
    kep_eval(x, k, &f, &df, &ddf);
    double a2 = df * df;
    double b = a2 - f * ddf;
    double h = df + sgn(df) * sqrt(fabs((kLaguerreOrder - 1) *
                                        (kLaguerreOrder * b - a2)));
    if (h == 0) return -1;
    delta = -(kLaguerreOrder * f) / h;
    x += delta;
    if (++it > kKepMaxIter) return -2;
  } while (fabs(delta) > kKepTol);
  *s = x;
  if (sgn(x) != sgn(s0)) return -3;
  return 0;
}

// One Kepler step of the softened problem (universal_kepler_solver.h:
// 312-463); st = (r, v) in and out, untouched on failure.
TUPAN_HD inline int kepler_once(double dt0, double m, double e2, double *st) {
  double rx = st[0], ry = st[1], rz = st[2];
  double vx = st[3], vy = st[4], vz = st[5];
  double r2 = rx * rx + ry * ry + rz * rz;
  if (!(r2 > 0)) return 0;
  r2 += e2;
  double r = sqrt(r2);
  double v2 = vx * vx + vy * vy + vz * vz;
  double rv = rx * vx + ry * vy + rz * vz;
  double beta = 2 - e2 / r2;
  double alpha = v2 - beta * m / r;
  double dt = dt0;
  if (alpha < 0) {  // bound: reduce dt modulo the period
    double a = m / fabs(alpha);
    double T = 2 * M_PI * a * sqrt(a / m);
    double ratio = dt0 / T;
    dt = (ratio - (double)(long long)ratio) * T;
  }
  double s0 = dt / r;
  if (alpha > 0) {  // hyperbolic first guess (Bate, Mueller & White 4.5.11)
    double salpha = sqrt(alpha);
    double ss = fabs(2 * alpha * dt / (rv + (m + alpha * r) / salpha));
    if (ss > 1) s0 = sgn(dt) * log(ss) / salpha;
  }
  KepEq k{dt, r, rv, m, alpha};
  double s = 0;
  int err = laguerre(s0, &s, k);
  if (err) return err;
  double s2 = s * s, z = alpha * s2;
  double S1 = s * stumpff_c1(z);
  double S2 = s2 * stumpff_c2(z);
  double S3 = s * s2 * stumpff_c3(z);
  double lf = 1 - m * S2 / r;
  double lg = dt - m * S3;
  double r1x = rx * lf + vx * lg;
  double r1y = ry * lf + vy * lg;
  double r1z = rz * lf + vz * lg;
  double r1 = sqrt(r1x * r1x + r1y * r1y + r1z * r1z + e2);
  double ldf = -m * S1 / (r * r1);
  double ldg = (1 + lg * ldf) / lf;
  st[0] = r1x;
  st[1] = r1y;
  st[2] = r1z;
  st[3] = rx * ldf + vx * ldg;
  st[4] = ry * ldf + vy * ldg;
  st[5] = rz * ldf + vz * ldg;
  return 0;
}

// n = 1, 2, 4, ... equal sub-steps until all converge (:466-507).
// Returns false when the budget ran out (out = last attempt's start).
TUPAN_HD inline bool kepler_substeps(double dt, double m, double e2,
                                const double *in, double *out, int &budget) {
  for (long n = 1;; n *= 2) {
    double st[6] = {in[0], in[1], in[2], in[3], in[4], in[5]};
    bool ok = true;
    for (long i = 0; i < n; ++i) {
      if (--budget < 0) {
        for (int q = 0; q < 6; ++q) out[q] = st[q];
        return false;
      }
      if (kepler_once(dt / n, m, e2, st) != 0) {
        ok = false;
        break;
      }
    }
    if (ok) {
      for (int q = 0; q < 6; ++q) out[q] = st[q];
      return true;
    }
  }
}

TUPAN_HD inline double kep_energy(const double *st, double m,
                                             double e2, double *u) {
  double r = sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2] + e2);
  *u = 2 * m / r;
  return (st[3] * st[3] + st[4] * st[4] + st[5] * st[5]) - *u;
}

// Softened Kepler drift with the energy-conservation sub-stepping
// (universal_kepler_solver.h:510-614).
TUPAN_HD inline bool kepler_drift(double dt, double m, double e2, double *st,
                             int &budget) {
  double in[6] = {st[0], st[1], st[2], st[3], st[4], st[5]};
  double cur[6];
  bool ok = kepler_substeps(dt, m, e2, in, cur, budget);
  if (!ok || e2 == 0) {
    for (int q = 0; q < 6; ++q) st[q] = cur[q];
    return ok;
  }
  if (!(in[0] * in[0] + in[1] * in[1] + in[2] * in[2] > 0)) return true;
  double u0, u1;
  double E0 = kep_energy(in, m, e2, &u0);
  double E1 = kep_energy(cur, m, e2, &u1);
  const double tol = 64 * kKepTol;
  if (2 * fabs(E1 - E0) < tol * (u1 + u0)) {
    for (int q = 0; q < 6; ++q) st[q] = cur[q];
    return true;
  }
  for (long n = 2;; n *= 2) {
    for (int q = 0; q < 6; ++q) cur[q] = in[q];
    bool redo = false;
    for (long i = 0; i < n; ++i) {
      double nxt[6];
      if (!kepler_substeps(dt / n, m, e2, cur, nxt, budget)) {
        for (int q = 0; q < 6; ++q) st[q] = nxt[q];
        return false;
      }
      for (int q = 0; q < 6; ++q) cur[q] = nxt[q];
      E1 = kep_energy(cur, m, e2, &u1);
      if (2 * fabs(E1 - E0) > tol * (u1 + u0)) {
        redo = true;
        break;
      }
    }
    if (!redo) break;
  }
  for (int q = 0; q < 6; ++q) st[q] = cur[q];
  return true;
}

// Softened kick-drift-kick for wide pairs (sakura_kernel_common.h:24-91).
TUPAN_HD inline void pair_leapfrog(double dt, double m, double e2,
                                              double *st) {
  double h = dt / 2;
  double rx = st[0] + st[3] * h, ry = st[1] + st[4] * h,
         rz = st[2] + st[5] * h;
  double r2 = rx * rx + ry * ry + rz * rz;
  double inv_r2 = r2 > 0 ? 1.0 / (r2 + e2) : 0.0;
  double k = m * (inv_r2 * sqrt(inv_r2));
  double vx = st[3] - k * rx * dt, vy = st[4] - k * ry * dt,
         vz = st[5] - k * rz * dt;
  st[0] = rx + vx * h;
  st[1] = ry + vy * h;
  st[2] = rz + vz * h;
  st[3] = vx;
  st[4] = vy;
  st[5] = vz;
}

TUPAN_HD inline bool pair_twobody(double dt, double m, double e2, double *st,
                             int &budget) {
  double r2 = st[0] * st[0] + st[1] * st[1] + st[2] * st[2];
  double v2 = st[3] * st[3] + st[4] * st[4] + st[5] * st[5];
  double R = 64 * (m / v2);
  if (r2 > R * R) {
    pair_leapfrog(dt, m, e2, st);
    return true;
  }
  return kepler_drift(dt, m, e2, st, budget);
}

TUPAN_HD inline void free_drift(double *st, double t) {
  st[0] -= st[3] * t;
  st[1] -= st[4] * t;
  st[2] -= st[5] * t;
}

// flag: -1 drift then K(dt); +1 K(dt) then drift; -2 half drift, K(dt),
// half drift; +2 K(dt/2), drift, K(dt/2); otherwise identity
// (sakura_kernel_common.h:151-183).  "drift" moves by -v dt.
TUPAN_HD inline bool pair_evolve(double dt, int flag, double m, double e2,
                            double *st, int budget = kBudget) {
  bool ok = true;
  switch (flag) {
    case -1:
      free_drift(st, dt);
      ok = pair_twobody(dt, m, e2, st, budget);
      break;
    case 1:
      ok = pair_twobody(dt, m, e2, st, budget);
      free_drift(st, dt);
      break;
    case -2:
      free_drift(st, dt / 2);
      ok = pair_twobody(dt, m, e2, st, budget);
      free_drift(st, dt / 2);
      break;
    case 2:
      ok = pair_twobody(dt / 2, m, e2, st, budget);
      free_drift(st, dt);
      ok = pair_twobody(dt / 2, m, e2, st, budget) && ok;
      break;
    default:
      break;
  }
  return ok;
}

}  // namespace tupan
