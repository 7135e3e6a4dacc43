// ieee.hip -- the reference's arithmetic, bit for bit (TUPAN_MODE_FP64_IEEE).
//
// Phi, Acc, AccJerk, SnapCrackle and Tstep evaluated exactly as the
// reference's C kernels evaluate them: the same operations in the same
// order, one IEEE operation at a time (no FMA contraction), the softened
// inverse square as a correctly rounded division 1 / (r^2 + e2) and square
// roots correctly rounded (smoothing.h:382-414), and every j folded into one
// accumulator per i in ascending order.  On the same inputs the results are
// the reference C's bits (tupan/lib/src/*_kernel_common.h and
// tstep_kernel.c, compiled as ISO C99 without contraction).

#include "ieee_engine.hpp"

namespace tupan {
namespace {

using namespace ieee;

struct PhiIeee : IeeeBase<1> {
  static constexpr int kFields = 5;
  struct JRec { double m, x, y, z, e2, pad; };
  struct I { double x, y, z, e2; };
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j], s.f[RX][j], s.f[RY][j], s.f[RZ][j], s.f[E2][j], 0.0};
  }
  __device__ static I load_i(const Soa &s, int i) {
    return {s.f[RX][i], s.f[RY][i], s.f[RZ][i], s.f[E2][i]};
  }
  // phi_kernel_common.h:7-31
  __device__ static void terms(const I &q, const JRec &r, double *t,
                               const Params &) {
#pragma clang fp contract(off)
    double rx = q.x - r.x;
    double ry = q.y - r.y;
    double rz = q.z - r.z;
    double e2 = q.e2 + r.e2;
    double r2 = rx * rx + ry * ry + rz * rz;
    double inv_r1 = sqrt(ieee_inv_r2(r2, e2));
    t[0] = r.m * inv_r1;
  }
};

struct AccIeee : IeeeBase<3> {
  static constexpr int kFields = 5;
  using JRec = PhiIeee::JRec;
  using I = PhiIeee::I;
  __device__ static JRec pack(const Soa &s, int j, double c) {
    return PhiIeee::pack(s, j, c);
  }
  __device__ static I load_i(const Soa &s, int i) {
    return PhiIeee::load_i(s, i);
  }
  // acc_kernel_common.h:7-37
  __device__ static void terms(const I &q, const JRec &r, double *t,
                               const Params &) {
#pragma clang fp contract(off)
    double rx = q.x - r.x;
    double ry = q.y - r.y;
    double rz = q.z - r.z;
    double e2 = q.e2 + r.e2;
    double r2 = rx * rx + ry * ry + rz * rz;
    double inv_r2 = ieee_inv_r2(r2, e2);
    double k = inv_r2 * sqrt(inv_r2);
    k *= r.m;
    t[0] = k * rx;
    t[1] = k * ry;
    t[2] = k * rz;
  }
};

struct AccJerkIeee : IeeeBase<6> {
  static constexpr int kFields = 8;
  struct JRec { double m, x, y, z, e2, vx, vy, vz; };
  struct I { double x, y, z, e2, vx, vy, vz; };
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j],  s.f[RX][j], s.f[RY][j], s.f[RZ][j],
            s.f[E2][j], s.f[VX][j], s.f[VY][j], s.f[VZ][j]};
  }
  __device__ static I load_i(const Soa &s, int i) {
    return {s.f[RX][i], s.f[RY][i], s.f[RZ][i], s.f[E2][i],
            s.f[VX][i], s.f[VY][i], s.f[VZ][i]};
  }
  // acc_jerk_kernel_common.h:7-55
  __device__ static void terms(const I &q, const JRec &r, double *t,
                               const Params &) {
#pragma clang fp contract(off)
    double rx = q.x - r.x;
    double ry = q.y - r.y;
    double rz = q.z - r.z;
    double e2 = q.e2 + r.e2;
    double vx = q.vx - r.vx;
    double vy = q.vy - r.vy;
    double vz = q.vz - r.vz;
    double r2 = rx * rx + ry * ry + rz * rz;
    double rv = rx * vx + ry * vy + rz * vz;
    double inv_r2 = ieee_inv_r2(r2, e2);
    double k = inv_r2 * sqrt(inv_r2);
    double alpha = 3 * rv * inv_r2;
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

struct SnapCrackleIeee : IeeeBase<6> {
  static constexpr int kFields = 14;
  struct JRec { double m, x, y, z, e2, vx, vy, vz, ax, ay, az, jx, jy, jz; };
  struct I { double x, y, z, e2, vx, vy, vz, ax, ay, az, jx, jy, jz; };
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j],  s.f[RX][j], s.f[RY][j], s.f[RZ][j], s.f[E2][j],
            s.f[VX][j], s.f[VY][j], s.f[VZ][j], s.f[AX][j], s.f[AY][j],
            s.f[AZ][j], s.f[JX][j], s.f[JY][j], s.f[JZ][j]};
  }
  __device__ static I load_i(const Soa &s, int i) {
    return {s.f[RX][i], s.f[RY][i], s.f[RZ][i], s.f[E2][i], s.f[VX][i],
            s.f[VY][i], s.f[VZ][i], s.f[AX][i], s.f[AY][i], s.f[AZ][i],
            s.f[JX][i], s.f[JY][i], s.f[JZ][i]};
  }
  // snap_crackle_kernel_common.h:7-94
  __device__ static void terms(const I &q, const JRec &r, double *t,
                               const Params &) {
#pragma clang fp contract(off)
    double rx = q.x - r.x;
    double ry = q.y - r.y;
    double rz = q.z - r.z;
    double e2 = q.e2 + r.e2;
    double vx = q.vx - r.vx;
    double vy = q.vy - r.vy;
    double vz = q.vz - r.vz;
    double Ax = q.ax - r.ax;
    double Ay = q.ay - r.ay;
    double Az = q.az - r.az;
    double Jx = q.jx - r.jx;
    double Jy = q.jy - r.jy;
    double Jz = q.jz - r.jz;
    double r2 = rx * rx + ry * ry + rz * rz;
    double rv = rx * vx + ry * vy + rz * vz;
    double v2 = vx * vx + vy * vy + vz * vz;
    double rJ = rx * Jx + ry * Jy + rz * Jz;
    double rA = rx * Ax + ry * Ay + rz * Az;
    double vA = vx * Ax + vy * Ay + vz * Az;
    double inv_r2 = ieee_inv_r2(r2, e2);
    double k = inv_r2 * sqrt(inv_r2);
    double alpha = rv * inv_r2;
    double alpha2 = alpha * alpha;
    double beta = 3 * ((v2 + rA) * inv_r2 + alpha2);
    double gamma = (3 * vA + rJ) * inv_r2 + alpha * (beta - 4 * alpha2);
    alpha *= 3;
    gamma *= 3;
    vx -= alpha * rx;
    vy -= alpha * ry;
    vz -= alpha * rz;
    double two_alpha = 2 * alpha;
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
struct TstepIeee {
  static constexpr bool kTwoLevel = false;
  static constexpr int kOut = 2, kTerms = 1, kAcc = 2, kFields = 8;
  struct Params { double eta; };
  using JRec = AccJerkIeee::JRec;
  struct I { double m, x, y, z, e2, vx, vy, vz; };
  __device__ static JRec pack(const Soa &s, int j, double c) {
    return AccJerkIeee::pack(s, j, c);
  }
  __device__ static I load_i(const Soa &s, int i) {
    return {s.f[M][i],  s.f[RX][i], s.f[RY][i], s.f[RZ][i],
            s.f[E2][i], s.f[VX][i], s.f[VY][i], s.f[VZ][i]};
  }
  __device__ static void terms(const I &q, const JRec &r, double *t,
                               const Params &prm) {
#pragma clang fp contract(off)
    double rx = q.x - r.x;
    double ry = q.y - r.y;
    double rz = q.z - r.z;
    double e2 = q.e2 + r.e2;
    double vx = q.vx - r.vx;
    double vy = q.vy - r.vy;
    double vz = q.vz - r.vz;
    double m = q.m + r.m;
    double r2 = rx * rx + ry * ry + rz * rz;
    double rv = rx * vx + ry * vy + rz * vz;
    double v2 = vx * vx + vy * vy + vz * vz;
    bool pair = r2 > 0;
    double inv_r2 = ieee_inv_r2(r2, e2);
    double inv_r1 = sqrt(inv_r2);
    double a = 2;
    double b = (1 + a / 2) * inv_r2;
    double w2 = inv_r2;
    double gam = inv_r2;
    double phi = m * inv_r1;
    w2 *= (v2 + a * phi);
    gam *= (w2 + b * phi);
    gam *= (prm.eta / sqrt(w2));
    w2 -= gam * rv;
    t[0] = pair ? w2 : 0.0;  // a masked pair may hold 0*inf = NaN
  }
  __device__ static void init(double *acc) { acc[0] = acc[1] = 0.0; }
  __device__ static void fold(double *acc, const double *t) {
#pragma clang fp contract(off)
    acc[0] = acc[0] + t[0];
    acc[1] = fmax(t[0], acc[1]);
  }
  __device__ static void finish(const double *acc, Out out, int i,
                                const Params &prm) {
#pragma clang fp contract(off)
    out.f[0][i] = prm.eta / sqrt(1 + acc[0]);
    out.f[1][i] = prm.eta / sqrt(1 + acc[1]);
  }
};

}  // namespace

// Entry for gravity.hip's mode dispatch: kernel ids TUPAN_K_PHI ..
// TUPAN_K_SNAP_CRACKLE.
int run_ieee(int kid, size_t ni, const double *const *ip, size_t nj,
             const double *const *jp, double *const *op, hipStream_t st) {
  switch (kid) {
    case TUPAN_K_PHI:
      return run_ieee_engine<PhiIeee, 4>(ni, ip, nj, jp, op, st);
    case TUPAN_K_ACC:
      return run_ieee_engine<AccIeee, 4>(ni, ip, nj, jp, op, st);
    case TUPAN_K_ACC_JERK:
      return run_ieee_engine<AccJerkIeee, 4>(ni, ip, nj, jp, op, st);
    case TUPAN_K_SNAP_CRACKLE:
      return run_ieee_engine<SnapCrackleIeee, 2>(ni, ip, nj, jp, op, st);
  }
  return set_error(TUPAN_ERR_ARG, "no IEEE mode for kernel %d", kid);
}

int run_ieee_tstep(size_t ni, const double *const *ip, size_t nj,
                   const double *const *jp, double eta, double *const *op,
                   hipStream_t st) {
  return run_ieee_engine<TstepIeee, 4>(ni, ip, nj, jp, op, st,
                                       TstepIeee::Params{eta});
}

}  // namespace tupan
