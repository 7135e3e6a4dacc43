// gravity.hip -- Phi, Acc, AccJerk, SnapCrackle and Tstep on the pair engine.
//
// Each kernel has an FP64 policy (FP64 arithmetic, hardware rsq refined by
// one Newton-Householder step) and, except Tstep, an FP32C policy
// (f32c.hpp):
//   * r_ij = x_i - x_j is formed in FP64 and rounded once to FP32
//     (dsub32), so the separation is accurate to 1/2 ulp of itself;
//   * the rest of the pair arithmetic is FP32 with v_rsq_f32; Phi, Acc and
//     AccJerk fold its measured mean bias into the stored masses (and
//     velocity offsets), SnapCrackle refines it by a Newton step;
//   * sums run in FP32 over a tile of kTile j's and are folded into FP64
//     accumulators after every tile (block-compensated summation);
//   * near field: a pair inside its near-field radius (see ctx_kernel)
//     contributes nothing in FP32; after each tile the engine re-evaluates
//     exactly those pairs with the FP64 policy from the caller's FP64
//     arrays.  The few largest pair forces of each particle -- the ones
//     whose FP32 rounding would dominate its error when they nearly cancel
//     -- are thereby exact to FP64.
//
// Pair formulas and semantics follow the reference cores:
//   Phi         tupan/lib/src/phi_kernel_common.h:7-31
//   Acc         tupan/lib/src/acc_kernel_common.h:7-37
//   AccJerk     tupan/lib/src/acc_jerk_kernel_common.h:7-55
//   SnapCrackle tupan/lib/src/snap_crackle_kernel_common.h:7-94
//   Tstep       tupan/lib/src/tstep_kernel_common.h:7-56 (+ finalisation
//               tupan/lib/src/tstep_kernel.c:46-47)
// with the softening of tupan/lib/src/smoothing.h:10-184: s^2 = r^2 + e2_i +
// e2_j, and a pair contributes nothing unless the *unsoftened* r^2 > 0.
#include <mutex>

#include "accjerk.hpp"
#include "f32c.hpp"
#include "launch.hpp"

namespace tupan {

// ================================================================ Phi
struct PhiF64 : SumOut<1> {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 5, kMaxIpl = 4, kPrefIpl = 2;
  using Params = NoParams;
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
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    double r2 = rx * rx + ry * ry + rz * rz;
    double y = rsq64(r2 + (q.e2 + r.e2));
    y = r2 > 0.0 ? y : 0.0;
    a.v[0] -= r.m * y;
  }
};

// j record / i state shared by the Phi and Acc FP32C policies
struct Pos32cJ { double x, y, z; float m, e2, near2, pad; };
struct Pos32cI {
  double x, y, z;
  float e2;
  int i;
  float px, py, pz;  // position offsets to the tile's origin (far tiles)
};
// far-tile record (F32cBase::pair32_farrec)
struct alignas(32) Pos32cFar { float x, y, z, m, e2, pad[3]; };
__device__ __forceinline__ Pos32cFar far_pos32c(const Pos32cJ &r,
                                                const double *po) {
  return {(float)(r.x - po[0]), (float)(r.y - po[1]), (float)(r.z - po[2]),
          r.m, r.e2, {0.0f, 0.0f, 0.0f}};
}

// ms: the mass factor that cancels v_rsq_f32's mean bias (kMassY1 for
// Phi's m y, kMassY3 for Acc's m y^3)
__device__ __forceinline__ Pos32cJ pack_pos32c(const Soa &s, int j,
                                               const PackCtx &c, double ms) {
  return {s.f[RX][j], s.f[RY][j], s.f[RZ][j], (float)(s.f[M][j] * ms),
          (float)s.f[E2][j], near_threshold(s.f[M][j], s.f[E2][j], c), 0.0f};
}
__device__ __forceinline__ Pos32cI load_pos32c(const Soa &s, int i) {
  return {s.f[RX][i], s.f[RY][i], s.f[RZ][i], (float)s.f[E2][i], i};
}

struct PhiF32c : F32cBase<PhiF32c, PhiF64, 1> {
  static constexpr int kFields = 5, kMaxIpl = 4, kPrefIpl = 2;
  using JRec = Pos32cJ;
  using I = Pos32cI;
  __device__ static JRec pack(const Soa &s, int j, const PackCtx &c) {
    return pack_pos32c(s, j, c, kMassY1);
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return load_pos32c(s, i);
  }
  static constexpr bool kFarRec = true;
  // far records two per scalar load: the short pair body no longer waits
  // for its record (4 per load: slower); on tiles that share one e2 the
  // 16-byte records, four per load
  static constexpr int kFarGroup = 2;
  static constexpr bool kFar16 = true;
  using FarRec = Pos32cFar;
  __device__ static FarRec far_rec(const JRec &r, const double *po) {
    return far_pos32c(r, po);
  }
  template <class R>  // JRec or FarRec
  __device__ static void body32(const I &, const R &r, float, float,
                                float, float s2, bool far, A32 &a) {
    float y = rsq32(s2);
    y = far ? y : 0.0f;
    a.v[0] -= r.m * y;
  }
};

// ================================================================ Acc
struct AccF64 : SumOut<3> {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 5, kMaxIpl = 4, kPrefIpl = 1;
  using Params = NoParams;
  using JRec = PhiF64::JRec;
  using I = PhiF64::I;
  using A64 = F64Acc<3>;
  __device__ static JRec pack(const Soa &s, int j, double sc) {
    return PhiF64::pack(s, j, sc);
  }
  __device__ static I load_i(const Soa &s, int i, const Params &p) {
    return PhiF64::load_i(s, i, p);
  }
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    double r2 = rx * rx + ry * ry + rz * rz;
    double y = rsq64(r2 + (q.e2 + r.e2));
    y = r2 > 0.0 ? y : 0.0;
    double k = r.m * y * (y * y);
    a.v[0] -= k * rx;
    a.v[1] -= k * ry;
    a.v[2] -= k * rz;
  }
};

// Acc restricted to the pairs whose classification positions (fields
// 5-7: cx, cy, cz) are at least rcut apart: the far field of a hybrid
// scheme whose close pairs are integrated by the Sakura Kepler drift
// (SakuraNearF64 takes the complementary set of pairs).  The scheme passes
// positions fixed for the whole step, so that every kick and drift of the
// step splits the pairs the same way; passing the positions themselves
// selects by the current separation.  FP64, AccF64's pair terms.
struct AccFarF64 : SumOut<3> {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 8, kMaxIpl = 4, kPrefIpl = 1;
  struct Params { double rc2; };
  struct JRec { double m, x, y, z, e2, cx, cy, cz; };
  struct I { double x, y, z, e2, cx, cy, cz; };
  using A64 = F64Acc<3>;
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j],  s.f[RX][j], s.f[RY][j], s.f[RZ][j],
            s.f[E2][j], s.f[5][j],  s.f[6][j],  s.f[7][j]};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return {s.f[RX][i], s.f[RY][i], s.f[RZ][i], s.f[E2][i],
            s.f[5][i],  s.f[6][i],  s.f[7][i]};
  }
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &prm) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    double r2 = rx * rx + ry * ry + rz * rz;
    double cx = q.cx - r.cx, cy = q.cy - r.cy, cz = q.cz - r.cz;
    double c2 = cx * cx + cy * cy + cz * cz;
    double y = rsq64(r2 + (q.e2 + r.e2));
    y = (r2 > 0.0 && c2 >= prm.rc2) ? y : 0.0;
    double k = r.m * y * (y * y);
    a.v[0] -= k * rx;
    a.v[1] -= k * ry;
    a.v[2] -= k * rz;
  }
};

// AccFarF64 whose finishing step also applies kick_pn's first half kick,
// v += (a * tau + w) / 2 (tupan_kick_pn_half_dev's operations, in its
// order), to the particle whose acceleration it has just written.  The pair
// terms read masses, positions and softenings only, so updating v where
// the sums finish (inside the pair kernel when there is one block row)
// races with no other block.
struct AccFarKickF64 : AccFarF64 {
  struct Params : AccFarF64::Params {
    double tau;
    double *v[3];
    const double *w[3];
  };
  __device__ static void finish(const double *a, Out out, int i,
                                const Params &prm) {
#pragma clang fp contract(off)
    for (int k = 0; k < 3; ++k) {
      out.f[k][i] = a[k];
      prm.v[k][i] = prm.v[k][i] + (a[k] * prm.tau + prm.w[k][i]) / 2;
    }
  }
};

struct AccF32c : F32cBase<AccF32c, AccF64, 3> {
  static constexpr int kFields = 5, kMaxIpl = 4, kPrefIpl = 2;
  using JRec = Pos32cJ;
  using I = Pos32cI;
  __device__ static JRec pack(const Soa &s, int j, const PackCtx &c) {
    return pack_pos32c(s, j, c, kMassY3);
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return load_pos32c(s, i);
  }
  static constexpr bool kFarRec = true;
  // far records two per scalar load: the short pair body no longer waits
  // for its record (4 per load: slower); on tiles that share one e2 the
  // 16-byte records, four per load
  static constexpr int kFarGroup = 2;
  static constexpr bool kFar16 = true;
  using FarRec = Pos32cFar;
  __device__ static FarRec far_rec(const JRec &r, const double *po) {
    return far_pos32c(r, po);
  }
  template <class R>  // JRec or FarRec
  __device__ static void body32(const I &q, const R &r, float rx, float ry,
                                float rz, float s2, bool far, A32 &a) {
    float y = rsq32(s2);
    y = far ? y : 0.0f;
    float k = r.m * y * (y * y);
    a.v[0] -= k * rx;
    a.v[1] -= k * ry;
    a.v[2] -= k * rz;
  }
};

// ========================================================= SnapCrackle
// Snap and crackle of the softened pair force from the relative state
// (r, v, A = a_i - a_j, J = j_i - j_j): the second and third time
// derivatives of -m r / s^3 via the alpha, beta, gamma recurrences
// (snap_crackle_kernel_common.h:63-93).
template <class T>
__device__ __forceinline__ void snap_crackle_pair(
    T rx, T ry, T rz, T vx, T vy, T vz, T Ax, T Ay, T Az, T Jx, T Jy, T Jz,
    T y, T m, T *acc) {
  T r2y = y * y;  // 1/s^2
  T rv = rx * vx + ry * vy + rz * vz;
  T v2 = vx * vx + vy * vy + vz * vz;
  T rJ = rx * Jx + ry * Jy + rz * Jz;
  T rA = rx * Ax + ry * Ay + rz * Az;
  T vA = vx * Ax + vy * Ay + vz * Az;
  T alpha = rv * r2y;
  T alpha2 = alpha * alpha;
  T beta = T(3) * ((v2 + rA) * r2y + alpha2);
  T gamma = (T(3) * vA + rJ) * r2y + alpha * (beta - T(4) * alpha2);
  alpha *= T(3);
  gamma *= T(3);
  vx -= alpha * rx;
  vy -= alpha * ry;
  vz -= alpha * rz;
  T two_alpha = T(2) * alpha;
  Ax -= two_alpha * vx + beta * rx;
  Ay -= two_alpha * vy + beta * ry;
  Az -= two_alpha * vz + beta * rz;
  alpha *= T(3);
  beta *= T(3);
  Jx -= alpha * Ax + beta * vx + gamma * rx;
  Jy -= alpha * Ay + beta * vy + gamma * ry;
  Jz -= alpha * Az + beta * vz + gamma * rz;
  T k = m * y * r2y;
  acc[0] -= k * Ax;
  acc[1] -= k * Ay;
  acc[2] -= k * Az;
  acc[3] -= k * Jx;
  acc[4] -= k * Jy;
  acc[5] -= k * Jz;
}

struct SnapCrackleF64 : SumOut<6> {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 14, kMaxIpl = 2, kPrefIpl = 1;
  using Params = NoParams;
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
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    double r2 = rx * rx + ry * ry + rz * rz;
    double y = rsq64(r2 + (q.e2 + r.e2));
    y = r2 > 0.0 ? y : 0.0;
    snap_crackle_pair<double>(rx, ry, rz, q.vx - r.vx, q.vy - r.vy,
                              q.vz - r.vz, q.ax - r.ax, q.ay - r.ay,
                              q.az - r.az, q.jx - r.jx, q.jy - r.jy,
                              q.jz - r.jz, y, r.m, a.v);
  }
};

// SnapCrackle over Morton-sorted particles (the sorted engine's policy of
// SnapCrackleF32c calls): velocity, acceleration and jerk differences from
// FP32 offsets to per-tile FP64 origins (the j offsets rounded once when
// packed, the i offsets once per tile, as AccJerk's velocities), positions
// from far records on far tiles and rounded FP64 differences (dsub32) on
// the others.  A pair then forms its twelve differences with FP32
// subtractions instead of twelve FP64 subtractions and conversions, and
// the i's FP64 fields are needed at tile starts only (they wait in LDS).
struct SnapCrackleF32s : F32cBase<SnapCrackleF32s, SnapCrackleF64, 6> {
  static constexpr int kFields = 14, kMaxIpl = 1, kPrefIpl = 1;
  static constexpr int kVelOrigins = 3;  // v, a, j: 12 doubles per tile
  static constexpr int kOrg = 4 * kVelOrigins;
  struct JRec {
    double x, y, z;
    float d[9];  // offsets of v, a, j to the tile's origins
    float m, e2, near2;
  };
  struct I {
    double x, y, z, vx, vy, vz, ax, ay, az, jx, jy, jz;
    float e2;
    float d[9];  // the i offsets to the current tile's origins
    int i;
    float px, py, pz;  // position offsets to the tile's origin (far tiles)
  };
  static constexpr bool kFarRec = true;
  struct alignas(32) FarRec {
    float x, y, z, d[9], m, e2;
  };
  __device__ static FarRec far_rec(const JRec &r, const double *po) {
    FarRec f;
    f.x = (float)(r.x - po[0]);
    f.y = (float)(r.y - po[1]);
    f.z = (float)(r.z - po[2]);
#pragma unroll
    for (int k = 0; k < 9; ++k) f.d[k] = r.d[k];
    f.m = r.m;
    f.e2 = r.e2;
    return f;
  }
  // record of j with the tile's origins o (kOrg doubles)
  __device__ static JRec pack_o(const Soa &s, int j, const PackCtx &c,
                                const double *o) {
    JRec r;
    r.x = s.f[RX][j];
    r.y = s.f[RY][j];
    r.z = s.f[RZ][j];
#pragma unroll
    for (int g = 0; g < 3; ++g)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        r.d[3 * g + k] = (float)(s.f[VX + 3 * g + k][j] - o[4 * g + k]);
    r.m = (float)s.f[M][j];
    r.e2 = (float)s.f[E2][j];
    r.near2 = near_threshold(s.f[M][j], s.f[E2][j], c);
    return r;
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    I q;
    q.x = s.f[RX][i], q.y = s.f[RY][i], q.z = s.f[RZ][i];
    q.vx = s.f[VX][i], q.vy = s.f[VY][i], q.vz = s.f[VZ][i];
    q.ax = s.f[AX][i], q.ay = s.f[AY][i], q.az = s.f[AZ][i];
    q.jx = s.f[JX][i], q.jy = s.f[JY][i], q.jz = s.f[JZ][i];
    q.e2 = (float)s.f[E2][i];
#pragma unroll
    for (int k = 0; k < 9; ++k) q.d[k] = 0.0f;
    q.i = i;
    q.px = q.py = q.pz = 0.0f;
    return q;
  }
  // the FP64 fields the sorted engine keeps in LDS
  static constexpr int kI64 = 12;
  __device__ static void save64(const I &q, double *f) {
    f[0] = q.x, f[1] = q.y, f[2] = q.z, f[3] = q.vx, f[4] = q.vy;
    f[5] = q.vz, f[6] = q.ax, f[7] = q.ay, f[8] = q.az, f[9] = q.jx;
    f[10] = q.jy, f[11] = q.jz;
  }
  __device__ static void load64(I &q, const double *f) {
    q.x = f[0], q.y = f[1], q.z = f[2], q.vx = f[3], q.vy = f[4];
    q.vz = f[5], q.ax = f[6], q.ay = f[7], q.az = f[8], q.jx = f[9];
    q.jy = f[10], q.jz = f[11];
  }
  __device__ static SnapCrackleF64::I exact_i(const I &q, const Soa &is) {
    return {q.x,  q.y,  q.z,  is.f[E2][q.i], q.vx, q.vy, q.vz,
            q.ax, q.ay, q.az, q.jx,          q.jy, q.jz};
  }
  // the i-side offsets to tile t's origins
  __device__ static void tile_begin(I &q, const double *origins, int t) {
    const double *o = origins + kOrg * t;
    const double f[9] = {q.vx, q.vy, q.vz, q.ax, q.ay, q.az, q.jx, q.jy, q.jz};
#pragma unroll
    for (int g = 0; g < 3; ++g)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        q.d[3 * g + k] = (float)(f[3 * g + k] - o[4 * g + k]);
  }
  template <class R>  // JRec or FarRec
  __device__ static void body32(const I &q, const R &r, float rx, float ry,
                                float rz, float s2, bool far, A32 &a) {
    float y = rsq32n(s2);
    y = far ? y : 0.0f;
    snap_crackle_pair<float>(rx, ry, rz, q.d[0] - r.d[0], q.d[1] - r.d[1],
                             q.d[2] - r.d[2], q.d[3] - r.d[3],
                             q.d[4] - r.d[4], q.d[5] - r.d[5],
                             q.d[6] - r.d[6], q.d[7] - r.d[7],
                             q.d[8] - r.d[8], y, r.m, a.v);
  }
};

struct SnapCrackleF32c : F32cBase<SnapCrackleF32c, SnapCrackleF64, 6> {
  static constexpr int kFields = 14, kMaxIpl = 2, kPrefIpl = 1;
  // calls with more than 4096 j's go to the sorted engine with its own
  // policy
  using Sorted = SnapCrackleF32s;
  // 5 waves per SIMD in the sorted engine (96 VGPRs, 116 B per lane
  // spilled, 188 B at 7 waves): 6% faster at N = 65536 than 7 waves, as
  // fast as 4
  static constexpr int kWavesPerEu = 5;
  // all relative quantities are formed in FP64 and rounded once (see
  // AccJerkF32c)
  struct JRec {
    double x, y, z, vx, vy, vz, ax, ay, az, jx, jy, jz;
    float m, e2, near2, pad;
  };
  struct I {
    double x, y, z, vx, vy, vz, ax, ay, az, jx, jy, jz;
    float e2;
    int i;
  };
  __device__ static JRec pack(const Soa &s, int j, const PackCtx &c) {
    return {s.f[RX][j],       s.f[RY][j],        s.f[RZ][j], s.f[VX][j],
            s.f[VY][j],       s.f[VZ][j],        s.f[AX][j], s.f[AY][j],
            s.f[AZ][j],       s.f[JX][j],        s.f[JY][j], s.f[JZ][j],
            (float)s.f[M][j], (float)s.f[E2][j],
            near_threshold(s.f[M][j], s.f[E2][j], c), 0.0f};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return {s.f[RX][i], s.f[RY][i], s.f[RZ][i], s.f[VX][i],
            s.f[VY][i], s.f[VZ][i], s.f[AX][i], s.f[AY][i],
            s.f[AZ][i], s.f[JX][i], s.f[JY][i], s.f[JZ][i],
            (float)s.f[E2][i], i};
  }
  __device__ static void body32(const I &q, const JRec &r, float rx, float ry,
                                float rz, float s2, bool far, A32 &a) {
    float y = rsq32n(s2);
    y = far ? y : 0.0f;
    snap_crackle_pair<float>(
        rx, ry, rz, dsub32(q.vx, r.vx), dsub32(q.vy, r.vy),
        dsub32(q.vz, r.vz), dsub32(q.ax, r.ax), dsub32(q.ay, r.ay),
        dsub32(q.az, r.az), dsub32(q.jx, r.jx), dsub32(q.jy, r.jy),
        dsub32(q.jz, r.jz), y, r.m, a.v);
  }
};

// ============================================================== Tstep
// Per pair w^2 = (v^2 + 2 phi)/s^2 with phi = (m_i+m_j)/s, corrected by the
// time derivative term -gamma (r.v) with gamma = (w^2 + 2 phi/s^2)/s^2 *
// eta/sqrt(w^2); the i-particle keeps sum_j w^2 and max_j w^2 and returns
// eta/sqrt(1 + .) of each (tstep_kernel_common.h:41-55, tstep_kernel.c:46-47).
struct TstepF64 {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 8, kMaxIpl = 2, kPrefIpl = 1, kOut = 2;
  struct Params { double eta; };
  using JRec = AccJerkF64::JRec;
  struct I { double m, x, y, z, e2, vx, vy, vz; };
  struct A64 { double sum, mx; };
  __device__ static JRec pack(const Soa &s, int j, double sc) {
    return AccJerkF64::pack(s, j, sc);
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return {s.f[M][i],  s.f[RX][i], s.f[RY][i], s.f[RZ][i],
            s.f[E2][i], s.f[VX][i], s.f[VY][i], s.f[VZ][i]};
  }
  __device__ static void zero(A64 &a) { a.sum = 0.0; a.mx = 0.0; }
  __device__ static void values(const A64 &a, double *v) {
    v[0] = a.sum;
    v[1] = a.mx;
  }
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &prm) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    double vx = q.vx - r.vx, vy = q.vy - r.vy, vz = q.vz - r.vz;
    double m = q.m + r.m;
    double r2 = rx * rx + ry * ry + rz * rz;
    double rv = rx * vx + ry * vy + rz * vz;
    double v2 = vx * vx + vy * vy + vz * vz;
    const bool real_pair = r2 > 0.0;
    double inv_r1 = rsq64(r2 + (q.e2 + r.e2));
    inv_r1 = real_pair ? inv_r1 : 0.0;
    double inv_r2 = inv_r1 * inv_r1;
    double phi = m * inv_r1;
    double w2 = inv_r2 * (v2 + 2.0 * phi);
    double gam = inv_r2 * (w2 + 2.0 * inv_r2 * phi);
    gam *= prm.eta * rsq64(w2);
    w2 -= gam * rv;
    // a masked pair may hold 0*inf = NaN here: select, do not multiply
    w2 = real_pair ? w2 : 0.0;
    a.sum += w2;
    a.mx = fmax(w2, a.mx);
  }
  __device__ static void combine(double *v, const double *w) {
    v[0] += w[0];
    v[1] = fmax(w[1], v[1]);
  }
  __device__ static void finish(const double *v, Out out, int i,
                                const Params &prm) {
    out.f[0][i] = prm.eta / sqrt(1.0 + v[0]);
    out.f[1][i] = prm.eta / sqrt(1.0 + v[1]);
  }
};

// ------------------------------------------------ static scratch
// Device memory of tupan_acc_jerk_static_dev: the j records and block-row
// partials of its FP64 AccJerk live in this file-scope array instead of
// the scratch cache, so that the call allocates nothing.  Each device holds
// its own copy for the life of the process.  In host code the array's name
// is not a device address: its address on each device comes from
// hipGetSymbolAddress (static_scratch), and only that is passed to kernels.
constexpr size_t kStaticScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_static_mu;
static void *g_static_ptr[kMaxDevices];

// The array's address on the current device, looked up once per device.
static int static_scratch(void **ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(TUPAN_ERR_ARG, "device %d beyond the static table", dev);
  std::lock_guard<std::mutex> lk(g_static_mu);
  if (!g_static_ptr[dev]) {
    void *q = nullptr;
    e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_static_ptr[dev] = q;
  }
  *ptr = g_static_ptr[dev];
  return TUPAN_OK;
}

}  // namespace tupan

using namespace tupan;

static hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

template <class P64, class P32c>
static int dispatch(int kid, int mode, size_t ni, const double *const *ip,
                    size_t nj, const double *const *jp, double *const *op,
                    hipStream_t st) {
  if (mode == TUPAN_MODE_FP64)
    return run_engine<P64>(kid, mode, ni, ip, nj, jp, op, {}, st);
  if (mode == TUPAN_MODE_FP32C)
    return run_engine<P32c>(kid, mode, ni, ip, nj, jp, op, {}, st);
  if (mode == TUPAN_MODE_FP64_IEEE || mode == TUPAN_MODE_FP32_IEEE)
    return run_ieee(kid, mode == TUPAN_MODE_FP32_IEEE, ni, ip, nj, jp, op, st);
  return set_error(TUPAN_ERR_ARG, "unknown mode %d", mode);
}

template <class P64, class P32c>
static int dispatch_host(int kid, int mode, int nfields, int nout, size_t ni,
                         const double *const *ip, size_t nj,
                         const double *const *jp, double *const *op) {
  return host_call(ni, ip, nj, jp, nfields, op, nout,
                   [&](const double *const *di, const double *const *dj,
                       double *const *dout, hipStream_t st) {
                     return dispatch<P64, P32c>(kid, mode, ni, di, nj, dj,
                                                dout, st);
                   });
}

extern "C" {

int tupan_phi_dev(int mode, size_t ni, const double *const *ip, size_t nj,
                  const double *const *jp, double *const *op, void *stream) {
  return dispatch<PhiF64, PhiF32c>(TUPAN_K_PHI, mode, ni, ip, nj, jp, op,
                                   as_stream(stream));
}
int tupan_phi_host(int mode, size_t ni, const double *const *ip, size_t nj,
                   const double *const *jp, double *const *op) {
  return dispatch_host<PhiF64, PhiF32c>(TUPAN_K_PHI, mode, 5, 1, ni, ip, nj,
                                        jp, op);
}

int tupan_acc_dev(int mode, size_t ni, const double *const *ip, size_t nj,
                  const double *const *jp, double *const *op, void *stream) {
  return dispatch<AccF64, AccF32c>(TUPAN_K_ACC, mode, ni, ip, nj, jp, op,
                                   as_stream(stream));
}
int tupan_acc_host(int mode, size_t ni, const double *const *ip, size_t nj,
                   const double *const *jp, double *const *op) {
  return dispatch_host<AccF64, AccF32c>(TUPAN_K_ACC, mode, 5, 3, ni, ip, nj,
                                        jp, op);
}

int tupan_acc_far_dev(size_t ni, const double *const *ip, size_t nj,
                      const double *const *jp, double rcut, double *const *op,
                      void *stream) {
  if (!(rcut >= 0.0))
    return set_error(TUPAN_ERR_ARG, "rcut must be >= 0 (got %g)", rcut);
  AccFarF64::Params prm{rcut * rcut};
  return run_engine<AccFarF64>(TUPAN_K_ACC, TUPAN_MODE_FP64, ni, ip, nj, jp,
                               op, prm, as_stream(stream));
}

int tupan_acc_far_kick_pn_half_dev(size_t n, const double *const *ip,
                                   double rcut, double *const *acc,
                                   double *const *v, const double *const *w,
                                   double tau, void *stream) {
  if (!(rcut >= 0.0))
    return set_error(TUPAN_ERR_ARG, "rcut must be >= 0 (got %g)", rcut);
  if (!v || !w) return set_error(TUPAN_ERR_ARG, "NULL v or w");
  AccFarKickF64::Params prm{};
  prm.rc2 = rcut * rcut;
  prm.tau = tau;
  for (int k = 0; k < 3; ++k) {
    if (!v[k] || !w[k])
      return set_error(TUPAN_ERR_ARG, "NULL v or w field %d", k);
    prm.v[k] = v[k];
    prm.w[k] = w[k];
  }
  return run_engine<AccFarKickF64>(TUPAN_K_ACC, TUPAN_MODE_FP64, n, ip, n, ip,
                                   acc, prm, as_stream(stream));
}

int tupan_acc_jerk_dev(int mode, size_t ni, const double *const *ip,
                       size_t nj, const double *const *jp, double *const *op,
                       void *stream) {
  return dispatch<AccJerkF64, AccJerkF32c>(TUPAN_K_ACC_JERK, mode, ni, ip, nj,
                                           jp, op, as_stream(stream));
}
int tupan_acc_jerk_host(int mode, size_t ni, const double *const *ip,
                        size_t nj, const double *const *jp,
                        double *const *op) {
  return dispatch_host<AccJerkF64, AccJerkF32c>(TUPAN_K_ACC_JERK, mode, 8, 6,
                                                ni, ip, nj, jp, op);
}

int tupan_acc_jerk_static_dev(size_t ni, const double *const *ip,
                              size_t nj, const double *const *jp,
                              double *const *op, void *stream) {
  // the j records alone must fit (nj above the direct-record bound)
  if (nj > (size_t)kDirectMaxJ &&
      sizeof(AccJerkF64::JRec) * nj > kStaticScratchBytes)
    return set_error(TUPAN_ERR_ARG,
                     "nj = %zu needs %zu bytes of j records; the static "
                     "scratch holds %zu", nj, sizeof(AccJerkF64::JRec) * nj,
                     kStaticScratchBytes);
  if (ni == 0) return TUPAN_OK;
  void *ptr = nullptr;
  const int rc = static_scratch(&ptr);
  if (rc) return rc;
  const FixedScratch fixed{static_cast<char *>(ptr), kStaticScratchBytes};
  return run_engine<AccJerkF64>(TUPAN_K_ACC_JERK, TUPAN_MODE_FP64, ni, ip, nj,
                                jp, op, {}, as_stream(stream), &fixed);
}

int tupan_snap_crackle_dev(int mode, size_t ni, const double *const *ip,
                           size_t nj, const double *const *jp,
                           double *const *op, void *stream) {
  return dispatch<SnapCrackleF64, SnapCrackleF32c>(
      TUPAN_K_SNAP_CRACKLE, mode, ni, ip, nj, jp, op, as_stream(stream));
}
int tupan_snap_crackle_host(int mode, size_t ni, const double *const *ip,
                            size_t nj, const double *const *jp,
                            double *const *op) {
  return dispatch_host<SnapCrackleF64, SnapCrackleF32c>(
      TUPAN_K_SNAP_CRACKLE, mode, 14, 6, ni, ip, nj, jp, op);
}

int tupan_tstep_dev(int mode, size_t ni, const double *const *ip,
                    size_t nj, const double *const *jp, double eta,
                    double *const *op, void *stream) {
  if (mode == TUPAN_MODE_FP64_IEEE || mode == TUPAN_MODE_FP32_IEEE)
    return run_ieee_tstep(mode == TUPAN_MODE_FP32_IEEE, ni, ip, nj, jp, eta,
                          op, as_stream(stream));
  if (mode != TUPAN_MODE_FP64 && mode != TUPAN_MODE_FP32C)
    return set_error(TUPAN_ERR_ARG, "unknown mode %d", mode);
  TstepF64::Params prm{eta};
  return run_engine<TstepF64>(TUPAN_K_TSTEP, TUPAN_MODE_FP64, ni, ip, nj, jp,
                              op, prm, as_stream(stream));
}
int tupan_tstep_host(int mode, size_t ni, const double *const *ip, size_t nj,
                     const double *const *jp, double eta, double *const *op) {
  return host_call(ni, ip, nj, jp, 8, op, 2,
                   [&](const double *const *di, const double *const *dj,
                       double *const *dout, hipStream_t st) {
                     return tupan_tstep_dev(mode, ni, di, nj, dj, eta, dout,
                                            st);
                   });
}

}  // extern "C"
