// ieee.hip -- the reference's arithmetic, bit for bit (TUPAN_MODE_FP64_IEEE).
//
// Phi, Acc, AccJerk and SnapCrackle evaluated exactly as the reference's C
// kernels evaluate them: the same operations in the same order, one IEEE
// operation at a time (no FMA contraction), the softened inverse square as
// a correctly rounded division 1 / (r^2 + e2) and its square root as a
// correctly rounded sqrt (smoothing.h:382-414), and every j summed in
// ascending order into one accumulator per i (the mode runs with a single
// j split).  On the same inputs the results are the reference C's bits
// (tupan/lib/src/*_kernel_common.h, compiled as ISO C99 without
// contraction).  This is the reproducibility mode, not the fast one: the
// division and square root cost ~30 FP64 instructions per pair and a single
// split leaves the chip underfilled below a few hundred thousand particles.
#include "launch.hpp"

namespace tupan {
namespace {

struct IeeeBase {
  static constexpr bool kTwoLevel = false;
  using Params = NoParams;
};

// smoothing.h:382-390 / :405-414 -- 1/(r^2+e2), zero unless r^2 > 0
__device__ __forceinline__ double ieee_inv_r2(double r2, double e2) {
#pragma clang fp contract(off)
  double inv = 1 / (r2 + e2);
  return r2 > 0 ? inv : 0.0;
}

struct PhiIeee : IeeeBase, SumOut<1> {
  static constexpr int kFields = 5, kMaxIpl = 2, kPrefIpl = 1;
  struct JRec { double m, x, y, z, e2; };
  struct I { double x, y, z, e2; };
  using A64 = F64Acc<1>;
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j], s.f[RX][j], s.f[RY][j], s.f[RZ][j], s.f[E2][j]};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return {s.f[RX][i], s.f[RY][i], s.f[RZ][i], s.f[E2][i]};
  }
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  // phi_kernel_common.h:7-31
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &) {
#pragma clang fp contract(off)
    double rx = q.x - r.x;
    double ry = q.y - r.y;
    double rz = q.z - r.z;
    double e2 = q.e2 + r.e2;
    double r2 = rx * rx + ry * ry + rz * rz;
    double inv_r1 = sqrt(ieee_inv_r2(r2, e2));
    a.v[0] -= r.m * inv_r1;
  }
};

struct AccIeee : IeeeBase, SumOut<3> {
  static constexpr int kFields = 5, kMaxIpl = 2, kPrefIpl = 1;
  using JRec = PhiIeee::JRec;
  using I = PhiIeee::I;
  using A64 = F64Acc<3>;
  __device__ static JRec pack(const Soa &s, int j, double c) {
    return PhiIeee::pack(s, j, c);
  }
  __device__ static I load_i(const Soa &s, int i, const Params &p) {
    return PhiIeee::load_i(s, i, p);
  }
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  // acc_kernel_common.h:7-37
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
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
    a.v[0] -= k * rx;
    a.v[1] -= k * ry;
    a.v[2] -= k * rz;
  }
};

struct AccJerkIeee : IeeeBase, SumOut<6> {
  static constexpr int kFields = 8, kMaxIpl = 2, kPrefIpl = 1;
  struct JRec { double m, x, y, z, e2, vx, vy, vz; };
  struct I { double x, y, z, e2, vx, vy, vz; };
  using A64 = F64Acc<6>;
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j],  s.f[RX][j], s.f[RY][j], s.f[RZ][j],
            s.f[E2][j], s.f[VX][j], s.f[VY][j], s.f[VZ][j]};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return {s.f[RX][i], s.f[RY][i], s.f[RZ][i], s.f[E2][i],
            s.f[VX][i], s.f[VY][i], s.f[VZ][i]};
  }
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  // acc_jerk_kernel_common.h:7-55
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
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
    a.v[0] -= k * rx;
    a.v[1] -= k * ry;
    a.v[2] -= k * rz;
    a.v[3] -= k * vx;
    a.v[4] -= k * vy;
    a.v[5] -= k * vz;
  }
};

struct SnapCrackleIeee : IeeeBase, SumOut<6> {
  static constexpr int kFields = 14, kMaxIpl = 1, kPrefIpl = 1;
  struct JRec { double m, x, y, z, e2, vx, vy, vz, ax, ay, az, jx, jy, jz; };
  struct I { double x, y, z, e2, vx, vy, vz, ax, ay, az, jx, jy, jz; };
  using A64 = F64Acc<6>;
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j],  s.f[RX][j], s.f[RY][j], s.f[RZ][j], s.f[E2][j],
            s.f[VX][j], s.f[VY][j], s.f[VZ][j], s.f[AX][j], s.f[AY][j],
            s.f[AZ][j], s.f[JX][j], s.f[JY][j], s.f[JZ][j]};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return {s.f[RX][i], s.f[RY][i], s.f[RZ][i], s.f[E2][i], s.f[VX][i],
            s.f[VY][i], s.f[VZ][i], s.f[AX][i], s.f[AY][i], s.f[AZ][i],
            s.f[JX][i], s.f[JY][i], s.f[JZ][i]};
  }
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  // snap_crackle_kernel_common.h:7-94
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
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
    a.v[0] -= k * Ax;
    a.v[1] -= k * Ay;
    a.v[2] -= k * Az;
    a.v[3] -= k * Jx;
    a.v[4] -= k * Jy;
    a.v[5] -= k * Jz;
  }
};

}  // namespace

// Entry for gravity.hip's mode dispatch: kernel ids TUPAN_K_PHI ..
// TUPAN_K_SNAP_CRACKLE.
int run_ieee(int kid, size_t ni, const double *const *ip, size_t nj,
             const double *const *jp, double *const *op, hipStream_t st) {
  switch (kid) {
    case TUPAN_K_PHI:
      return run_engine<PhiIeee>(kid, TUPAN_MODE_FP64_IEEE, ni, ip, nj, jp,
                                 op, {}, st);
    case TUPAN_K_ACC:
      return run_engine<AccIeee>(kid, TUPAN_MODE_FP64_IEEE, ni, ip, nj, jp,
                                 op, {}, st);
    case TUPAN_K_ACC_JERK:
      return run_engine<AccJerkIeee>(kid, TUPAN_MODE_FP64_IEEE, ni, ip, nj,
                                     jp, op, {}, st);
    case TUPAN_K_SNAP_CRACKLE:
      return run_engine<SnapCrackleIeee>(kid, TUPAN_MODE_FP64_IEEE, ni, ip,
                                         nj, jp, op, {}, st);
  }
  return set_error(TUPAN_ERR_ARG, "no IEEE mode for kernel %d", kid);
}

}  // namespace tupan
