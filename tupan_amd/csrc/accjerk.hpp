// accjerk.hpp -- the AccJerk pair policies (FP64 and FP32C), shared by the
// single-system kernels (gravity.hip) and the ensemble (ensemble.hip).
// Pair formula: tupan/lib/src/acc_jerk_kernel_common.h:7-55, with the
// softening of tupan/lib/src/smoothing.h:10-184.
#pragma once
#include "f32c.hpp"

namespace tupan {

// ============================================================ AccJerk
struct AccJerkF64 : SumOut<6> {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 8, kMaxIpl = 4, kPrefIpl = 1;
  using Params = NoParams;
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
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    double vx = q.vx - r.vx, vy = q.vy - r.vy, vz = q.vz - r.vz;
    double r2 = rx * rx + ry * ry + rz * rz;
    double rv = rx * vx + ry * vy + rz * vz;
    double y = rsq64(r2 + (q.e2 + r.e2));
    y = r2 > 0.0 ? y : 0.0;
    double y2 = y * y;
    double k = r.m * y * y2;
    double alpha = 3.0 * rv * y2;
    vx -= alpha * rx;
    vy -= alpha * ry;
    vz -= alpha * rz;
    a.v[0] -= k * rx;
    a.v[1] -= k * ry;
    a.v[2] -= k * rz;
    a.v[3] -= k * vx;
    a.v[4] -= k * vy;
    a.v[5] -= k * vz;
  }
};

// kDebias: the stored mass carries kMassY3 and the velocity offsets
// kVelY2, which cancel the mean bias of v_rsq_f32 in k = m y^3 and in
// alpha = 3 (r.v) y^2 at no cost per pair (the single-system kernels);
// without it, the plain values (the ensemble, which builds its records
// itself).
template <bool kDebias>
struct AccJerkF32cT : F32cBase<AccJerkF32cT<kDebias>, AccJerkF64, 6> {
  using Base = F32cBase<AccJerkF32cT<kDebias>, AccJerkF64, 6>;
  using typename Base::A32;
  static constexpr int kFields = 8, kMaxIpl = 4, kPrefIpl = 1;
  // Velocity differences from FP32 offsets to the j tile's FP64 velocity
  // origin: the j offset is rounded once in pack, the i offset once per
  // tile (tile_begin).  Rounding v_i itself to FP32 would add the same error
  // to every pair of particle i, and that coherent term
  // (dv_i * sum_j m_j/s^3) would not average out in the jerk; per tile it
  // is shared by 32 pairs only.
  static constexpr int kVelOrigins = 1;
  // The velocity origins are stored multiplied by kVelY2 (kDebias), and so
  // are the i velocities of I: the per-tile i offsets are then one FP64
  // subtraction each, with no FP64 multiplication per tile.
  static constexpr double kOrgScale = kDebias ? kVelY2 : 1.0;
  struct JRec { double x, y, z; float dvx, dvy, dvz, m, e2, near2; };
  struct I {
    double x, y, z, vx, vy, vz;  // vx, vy, vz times kOrgScale
    float e2, dvx, dvy, dvz;
    int i;
    float px, py, pz;  // position offsets to the tile's origin (far tiles)
  };
  // far-tile record: one 32-byte scalar load per pair
  static constexpr bool kFarRec = true;
  struct alignas(32) FarRec {
    float x, y, z, dvx, dvy, dvz, m, e2;
  };
  // the far record of r, whose tile has the position origin po
  __device__ static FarRec far_rec(const JRec &r, const double *po) {
    return {(float)(r.x - po[0]), (float)(r.y - po[1]), (float)(r.z - po[2]),
            r.dvx, r.dvy, r.dvz, r.m, r.e2};
  }
  // record of j with the (scaled) velocity origin o (3 doubles) of its tile
  __device__ static JRec pack_o(const Soa &s, int j, const PackCtx &c,
                                const double *o) {
    constexpr double vs = kOrgScale;
    constexpr double ms = kDebias ? kMassY3 : 1.0;
    return {s.f[RX][j],
            s.f[RY][j],
            s.f[RZ][j],
            (float)(s.f[VX][j] * vs - o[0]),
            (float)(s.f[VY][j] * vs - o[1]),
            (float)(s.f[VZ][j] * vs - o[2]),
            (float)(s.f[M][j] * ms),
            (float)s.f[E2][j],
            near_threshold(s.f[M][j], s.f[E2][j], c)};
  }
  __device__ static JRec pack(const Soa &s, int j, const PackCtx &c) {
    return pack_o(s, j, c, c.origins + 4 * (j / kTile));
  }
  __device__ static I load_i(const Soa &s, int i, const NoParams &) {
    constexpr double vs = kOrgScale;
    return {s.f[RX][i],      s.f[RY][i],      s.f[RZ][i],
            s.f[VX][i] * vs, s.f[VY][i] * vs, s.f[VZ][i] * vs,
            (float)s.f[E2][i], 0.0f, 0.0f, 0.0f, i};
  }
  // the FP64 fields the sorted engine keeps in LDS
  static constexpr int kI64 = 6;
  __device__ static void save64(const I &q, double *f) {
    f[0] = q.x, f[1] = q.y, f[2] = q.z, f[3] = q.vx, f[4] = q.vy, f[5] = q.vz;
  }
  __device__ static void load64(I &q, const double *f) {
    q.x = f[0], q.y = f[1], q.z = f[2], q.vx = f[3], q.vy = f[4], q.vz = f[5];
  }
  // AccJerkF64's i record from q's positions, and e2 and (when q's are
  // scaled) the velocities from the caller's arrays
  __device__ static AccJerkF64::I exact_i(const I &q, const Soa &is) {
    if constexpr (kDebias)
      return {q.x, q.y, q.z, is.f[E2][q.i], is.f[VX][q.i], is.f[VY][q.i],
              is.f[VZ][q.i]};
    else
      return {q.x, q.y, q.z, is.f[E2][q.i], q.vx, q.vy, q.vz};
  }
  // the i-side velocity offsets to tile t's (scaled) origin
  __device__ static void tile_begin(I &q, const double *origins, int t) {
    const double *o = origins + 4 * t;
    q.dvx = (float)(q.vx - o[0]);
    q.dvy = (float)(q.vy - o[1]);
    q.dvz = (float)(q.vz - o[2]);
  }
  template <class R>  // JRec or FarRec
  __device__ static void body32(const I &q, const R &r, float rx, float ry,
                                float rz, float s2, bool far, A32 &a) {
    float vx = q.dvx - r.dvx, vy = q.dvy - r.dvy, vz = q.dvz - r.dvz;
    float rv = rx * vx + ry * vy + rz * vz;
    float y = rsq32(s2);
    y = far ? y : 0.0f;
    float y2 = y * y;
    float k = r.m * y * y2;
    float alpha = 3.0f * rv * y2;
    vx -= alpha * rx;
    vy -= alpha * ry;
    vz -= alpha * rz;
    a.v[0] -= k * rx;
    a.v[1] -= k * ry;
    a.v[2] -= k * rz;
    a.v[3] -= k * vx;
    a.v[4] -= k * vy;
    a.v[5] -= k * vz;
  }
};
using AccJerkF32c = AccJerkF32cT<true>;
using AccJerkF32cTile = AccJerkF32cT<false>;

}  // namespace tupan
