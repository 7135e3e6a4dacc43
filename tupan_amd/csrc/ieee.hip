// ieee.hip -- the reference's arithmetic, bit for bit (TUPAN_MODE_FP64_IEEE
// and TUPAN_MODE_FP32_IEEE: its FP64 and FP32 builds).
//
// Phi, Acc, AccJerk, SnapCrackle and Tstep evaluated exactly as the
// reference's C kernels evaluate them, in double or in float (T, the
// reference's REAL with or without CONFIG_USE_DOUBLE, c_common.h:7-43): the
// same operations in the same order, one IEEE operation at a time (no FMA
// contraction), the softened inverse square as a correctly rounded division
// 1 / (r^2 + e2) and square roots correctly rounded (smoothing.h:382-414;
// <tgmath.h> makes sqrt of a float sqrtf), and every j folded into one
// accumulator per i in ascending order.  On the same inputs the results are
// the reference C's bits (tupan/lib/src/*_kernel_common.h and
// tstep_kernel.c, compiled as ISO C99 without contraction).  Float inputs
// are the double fields rounded to float at load; the float results are
// widened exactly.

#include "ieee_engine.hpp"

namespace tupan {
namespace {

using namespace ieee;

template <class T>
struct PhiIeee : IeeeBase<1, T> {
  using typename IeeeBase<1, T>::Params;
  static constexpr int kFields = 5;
  // 48 B (double) or 32 B (float) records
  struct JRec { T m, x, y, z, e2, pad[sizeof(T) == 8 ? 1 : 3]; };
  struct I { T x, y, z, e2; };
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {(T)s.f[M][j], (T)s.f[RX][j], (T)s.f[RY][j], (T)s.f[RZ][j],
            (T)s.f[E2][j], {}};
  }
  __device__ static I load_i(const Soa &s, int i) {
    return {(T)s.f[RX][i], (T)s.f[RY][i], (T)s.f[RZ][i], (T)s.f[E2][i]};
  }
  // phi_kernel_common.h:7-31
  __device__ static void terms(const I &q, const JRec &r, T *t,
                               const Params &) {
#pragma clang fp contract(off)
    T rx = q.x - r.x;
    T ry = q.y - r.y;
    T rz = q.z - r.z;
    T e2 = q.e2 + r.e2;
    T r2 = rx * rx + ry * ry + rz * rz;
    T inv_r1 = sqrt(ieee_inv_r2(r2, e2));
    t[0] = r.m * inv_r1;
  }
};

template <class T>
struct AccIeee : IeeeBase<3, T> {
  using typename IeeeBase<3, T>::Params;
  static constexpr int kFields = 5;
  using JRec = typename PhiIeee<T>::JRec;
  using I = typename PhiIeee<T>::I;
  __device__ static JRec pack(const Soa &s, int j, double c) {
    return PhiIeee<T>::pack(s, j, c);
  }
  __device__ static I load_i(const Soa &s, int i) {
    return PhiIeee<T>::load_i(s, i);
  }
  // acc_kernel_common.h:7-37
  __device__ static void terms(const I &q, const JRec &r, T *t,
                               const Params &) {
#pragma clang fp contract(off)
    T rx = q.x - r.x;
    T ry = q.y - r.y;
    T rz = q.z - r.z;
    T e2 = q.e2 + r.e2;
    T r2 = rx * rx + ry * ry + rz * rz;
    T inv_r2 = ieee_inv_r2(r2, e2);
    T k = inv_r2 * sqrt(inv_r2);
    k *= r.m;
    t[0] = k * rx;
    t[1] = k * ry;
    t[2] = k * rz;
  }
};

template <class T>
struct AccJerkIeee : IeeeBase<6, T> {
  using typename IeeeBase<6, T>::Params;
  static constexpr int kFields = 8;
  struct JRec { T m, x, y, z, e2, vx, vy, vz; };
  struct I { T x, y, z, e2, vx, vy, vz; };
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {(T)s.f[M][j],  (T)s.f[RX][j], (T)s.f[RY][j], (T)s.f[RZ][j],
            (T)s.f[E2][j], (T)s.f[VX][j], (T)s.f[VY][j], (T)s.f[VZ][j]};
  }
  __device__ static I load_i(const Soa &s, int i) {
    return {(T)s.f[RX][i], (T)s.f[RY][i], (T)s.f[RZ][i], (T)s.f[E2][i],
            (T)s.f[VX][i], (T)s.f[VY][i], (T)s.f[VZ][i]};
  }
  // acc_jerk_kernel_common.h:7-55
  __device__ static void terms(const I &q, const JRec &r, T *t,
                               const Params &) {
#pragma clang fp contract(off)
    T rx = q.x - r.x;
    T ry = q.y - r.y;
    T rz = q.z - r.z;
    T e2 = q.e2 + r.e2;
    T vx = q.vx - r.vx;
    T vy = q.vy - r.vy;
    T vz = q.vz - r.vz;
    T r2 = rx * rx + ry * ry + rz * rz;
    T rv = rx * vx + ry * vy + rz * vz;
    T inv_r2 = ieee_inv_r2(r2, e2);
    T k = inv_r2 * sqrt(inv_r2);
    T alpha = 3 * rv * inv_r2;
    vx -= alpha * rx;
    vy -= alpha * ry;
    vz -= alpha * rz;
    k *= r.m;
    t[0] = k * rx;
    t[1] = k * ry;
    t[2] = k * rz;
    t[3] = k * vx;
    t[4] = k * vy;
    t[5] = k * vz;
  }
};

template <class T>
struct SnapCrackleIeee : IeeeBase<6, T> {
  using typename IeeeBase<6, T>::Params;
  static constexpr int kFields = 14;
  // 128 B (double) or 64 B (float) records
  struct JRec { T m, x, y, z, e2, vx, vy, vz, ax, ay, az, jx, jy, jz, pad[2]; };
  struct I { T x, y, z, e2, vx, vy, vz, ax, ay, az, jx, jy, jz; };
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {(T)s.f[M][j],  (T)s.f[RX][j], (T)s.f[RY][j], (T)s.f[RZ][j],
            (T)s.f[E2][j], (T)s.f[VX][j], (T)s.f[VY][j], (T)s.f[VZ][j],
            (T)s.f[AX][j], (T)s.f[AY][j], (T)s.f[AZ][j], (T)s.f[JX][j],
            (T)s.f[JY][j], (T)s.f[JZ][j], {}};
  }
  __device__ static I load_i(const Soa &s, int i) {
    return {(T)s.f[RX][i], (T)s.f[RY][i], (T)s.f[RZ][i], (T)s.f[E2][i],
            (T)s.f[VX][i], (T)s.f[VY][i], (T)s.f[VZ][i], (T)s.f[AX][i],
            (T)s.f[AY][i], (T)s.f[AZ][i], (T)s.f[JX][i], (T)s.f[JY][i],
            (T)s.f[JZ][i]};
  }
  // snap_crackle_kernel_common.h:7-94
  __device__ static void terms(const I &q, const JRec &r, T *t,
                               const Params &) {
#pragma clang fp contract(off)
    T rx = q.x - r.x;
    T ry = q.y - r.y;
    T rz = q.z - r.z;
    T e2 = q.e2 + r.e2;
    T vx = q.vx - r.vx;
    T vy = q.vy - r.vy;
    T vz = q.vz - r.vz;
    T Ax = q.ax - r.ax;
    T Ay = q.ay - r.ay;
    T Az = q.az - r.az;
    T Jx = q.jx - r.jx;
    T Jy = q.jy - r.jy;
    T Jz = q.jz - r.jz;
    T r2 = rx * rx + ry * ry + rz * rz;
    T rv = rx * vx + ry * vy + rz * vz;
    T v2 = vx * vx + vy * vy + vz * vz;
    T rJ = rx * Jx + ry * Jy + rz * Jz;
    T rA = rx * Ax + ry * Ay + rz * Az;
    T vA = vx * Ax + vy * Ay + vz * Az;
    T inv_r2 = ieee_inv_r2(r2, e2);
    T k = inv_r2 * sqrt(inv_r2);
    T alpha = rv * inv_r2;
    T alpha2 = alpha * alpha;
    T beta = 3 * ((v2 + rA) * inv_r2 + alpha2);
    T gamma = (3 * vA + rJ) * inv_r2 + alpha * (beta - 4 * alpha2);
    alpha *= 3;
    gamma *= 3;
    vx -= alpha * rx;
    vy -= alpha * ry;
    vz -= alpha * rz;
    T two_alpha = 2 * alpha;
    Ax -= two_alpha * vx + beta * rx;
    Ay -= two_alpha * vy + beta * ry;
    Az -= two_alpha * vz + beta * rz;
    alpha *= 3;
    beta *= 3;
    Jx -= alpha * Ax + beta * vx + gamma * rx;
    Jy -= alpha * Ay + beta * vy + gamma * ry;
    Jz -= alpha * Az + beta * vz + gamma * rz;
    k *= r.m;
    t[0] = k * Ax;
    t[1] = k * Ay;
    t[2] = k * Az;
    t[3] = k * Jx;
    t[4] = k * Jy;
    t[5] = k * Jz;
  }
};

// tstep_kernel_common.h:7-56 and the finalisation of tstep_kernel.c:46-47:
// per pair w^2 (masked to 0 unless r^2 > 0), summed and maximised (fmax)
// per i; outputs eta / sqrt(1 + sum) and eta / sqrt(1 + max).
template <class T>
struct TstepIeee {
  using Real = T;
  static constexpr bool kTwoLevel = false;
  static constexpr int kOut = 2, kTerms = 1, kAcc = 2, kFields = 8;
  struct Params { T eta; };
  using JRec = typename AccJerkIeee<T>::JRec;
  struct I { T m, x, y, z, e2, vx, vy, vz; };
  __device__ static JRec pack(const Soa &s, int j, double c) {
    return AccJerkIeee<T>::pack(s, j, c);
  }
  __device__ static I load_i(const Soa &s, int i) {
    return {(T)s.f[M][i],  (T)s.f[RX][i], (T)s.f[RY][i], (T)s.f[RZ][i],
            (T)s.f[E2][i], (T)s.f[VX][i], (T)s.f[VY][i], (T)s.f[VZ][i]};
  }
  __device__ static void terms(const I &q, const JRec &r, T *t,
                               const Params &prm) {
#pragma clang fp contract(off)
    T rx = q.x - r.x;
    T ry = q.y - r.y;
    T rz = q.z - r.z;
    T e2 = q.e2 + r.e2;
    T vx = q.vx - r.vx;
    T vy = q.vy - r.vy;
    T vz = q.vz - r.vz;
    T m = q.m + r.m;
    T r2 = rx * rx + ry * ry + rz * rz;
    T rv = rx * vx + ry * vy + rz * vz;
    T v2 = vx * vx + vy * vy + vz * vz;
    bool pair = r2 > 0;
    T inv_r2 = ieee_inv_r2(r2, e2);
    T inv_r1 = sqrt(inv_r2);
    T a = 2;
    T b = (1 + a / 2) * inv_r2;
    T w2 = inv_r2;
    T gam = inv_r2;
    T phi = m * inv_r1;
    w2 *= (v2 + a * phi);
    gam *= (w2 + b * phi);
    gam *= (prm.eta / sqrt(w2));
    w2 -= gam * rv;
    t[0] = pair ? w2 : (T)0;  // a masked pair may hold 0*inf = NaN
  }
  __device__ static void init(T *acc) { acc[0] = acc[1] = 0; }
  __device__ static void fold(T *acc, const T *t) {
#pragma clang fp contract(off)
    acc[0] = acc[0] + t[0];
    acc[1] = fmax(t[0], acc[1]);
  }
  __device__ static void finish(const T *acc, Out out, int i,
                                const Params &prm) {
#pragma clang fp contract(off)
    const T a = prm.eta / sqrt(1 + acc[0]);
    const T b = prm.eta / sqrt(1 + acc[1]);
    out.f[0][i] = (double)a;
    out.f[1][i] = (double)b;
  }
};

}  // namespace

// Entry for gravity.hip's mode dispatch: kernel ids TUPAN_K_PHI ..
// TUPAN_K_SNAP_CRACKLE; fp32: the reference's FP32 build (float
// arithmetic), else its FP64 build.
int run_ieee(int kid, bool fp32, size_t ni, const double *const *ip,
             size_t nj, const double *const *jp, double *const *op,
             hipStream_t st) {
  if (fp32) {
    switch (kid) {
      case TUPAN_K_PHI:
        return run_ieee_engine<PhiIeee<float>, 4>(ni, ip, nj, jp, op, st);
      case TUPAN_K_ACC:
        return run_ieee_engine<AccIeee<float>, 4>(ni, ip, nj, jp, op, st);
      case TUPAN_K_ACC_JERK:
        return run_ieee_engine<AccJerkIeee<float>, 4>(ni, ip, nj, jp, op, st);
      case TUPAN_K_SNAP_CRACKLE:
        return run_ieee_engine<SnapCrackleIeee<float>, 2>(ni, ip, nj, jp, op,
                                                          st);
    }
    return set_error(TUPAN_ERR_ARG, "no FP32 IEEE mode for kernel %d", kid);
  }
  switch (kid) {
    case TUPAN_K_PHI:
      return run_ieee_engine<PhiIeee<double>, 4>(ni, ip, nj, jp, op, st);
    case TUPAN_K_ACC:
      return run_ieee_engine<AccIeee<double>, 4>(ni, ip, nj, jp, op, st);
    case TUPAN_K_ACC_JERK:
      return run_ieee_engine<AccJerkIeee<double>, 4>(ni, ip, nj, jp, op, st);
    case TUPAN_K_SNAP_CRACKLE:
      return run_ieee_engine<SnapCrackleIeee<double>, 2>(ni, ip, nj, jp, op,
                                                         st);
  }
  return set_error(TUPAN_ERR_ARG, "no IEEE mode for kernel %d", kid);
}

int run_ieee_tstep(bool fp32, size_t ni, const double *const *ip, size_t nj,
                   const double *const *jp, double eta, double *const *op,
                   hipStream_t st) {
  if (fp32)
    return run_ieee_engine<TstepIeee<float>, 4>(
        ni, ip, nj, jp, op, st, TstepIeee<float>::Params{(float)eta});
  return run_ieee_engine<TstepIeee<double>, 4>(
      ni, ip, nj, jp, op, st, TstepIeee<double>::Params{eta});
}

}  // namespace tupan
