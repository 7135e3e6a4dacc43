// adjoint.hip -- reverse-mode derivatives of Acc and Phi on the pair engine.
//
// No counterpart in the reference.  For a self-gravitating set and a loss
// L with upstream gradients g_k = dL/da_k (Acc) or h_k = dL/dphi_k (Phi),
// the gradients with respect to every position and mass are again O(N^2)
// pair sums (r = x_k - x_j, s^2 = r^2 + e2_k + e2_j, y = 1/s, pairs with
// r^2 = 0 masked as in the forward kernels):
//
//   Acc  dL/dx_k = sum_j y^3 (u - 3 (r.u) y^2 r),  u = m_k g_j - m_j g_k
//        dL/dm_k = sum_j y^3 (g_j . r)
//   Phi  dL/dx_k = sum_j (h_k m_j + h_j m_k) y^3 r
//        dL/dm_k = -sum_j h_j y
//
// The first line is the softened tidal tensor M(r) = I/s^3 - 3 r r^T/s^5
// applied to u: a_k depends on x_k through every r_kj (-m_j M g_k) and each
// a_i depends on x_k through r_ik (+m_k M g_i); M is even in r, so both
// terms share one pair evaluation.  The FP32C variants follow the forward
// FP32C scheme (FP64-exact separations rounded once, FP32 pair arithmetic,
// FP64 tile folds, FP64 near field).
#include "f32c.hpp"
#include "launch.hpp"

namespace tupan {

// ---- field slots of the adjoint inputs: F5, then the upstream gradient
enum AdjField { GX = VX, GY = VY, GZ = VZ, H = VX };

struct AccAdjF64 : SumOut<4> {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 8, kMaxIpl = 2, kPrefIpl = 1;
  using Params = NoParams;
  struct JRec { double m, x, y, z, e2, gx, gy, gz; };
  using I = JRec;
  using A64 = F64Acc<4>;
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j],  s.f[RX][j], s.f[RY][j], s.f[RZ][j],
            s.f[E2][j], s.f[GX][j], s.f[GY][j], s.f[GZ][j]};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return pack(s, i, 0.0);
  }
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    double r2 = rx * rx + ry * ry + rz * rz;
    double y = rsq64(r2 + (q.e2 + r.e2));
    y = r2 > 0.0 ? y : 0.0;
    double y2 = y * y;
    double y3 = y * y2;
    double ux = q.m * r.gx - r.m * q.gx;
    double uy = q.m * r.gy - r.m * q.gy;
    double uz = q.m * r.gz - r.m * q.gz;
    double t = 3.0 * (rx * ux + ry * uy + rz * uz) * y2;
    a.v[0] += y3 * (ux - t * rx);
    a.v[1] += y3 * (uy - t * ry);
    a.v[2] += y3 * (uz - t * rz);
    a.v[3] += y3 * (r.gx * rx + r.gy * ry + r.gz * rz);
  }
};

struct AccAdjF32c : F32cBase<AccAdjF32c, AccAdjF64, 4> {
  static constexpr int kFields = 8, kMaxIpl = 2, kPrefIpl = 1;
  struct JRec {
    double x, y, z;
    float m, e2, near2, gx, gy, gz;
  };
  struct I {
    double x, y, z;
    float m, e2, gx, gy, gz;
    int i;
  };
  __device__ static JRec pack(const Soa &s, int j, const PackCtx &c) {
    return {s.f[RX][j],        s.f[RY][j],        s.f[RZ][j],
            (float)s.f[M][j],  (float)s.f[E2][j],
            near_threshold(s.f[M][j], s.f[E2][j], c),
            (float)s.f[GX][j], (float)s.f[GY][j], (float)s.f[GZ][j]};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return {s.f[RX][i],        s.f[RY][i],        s.f[RZ][i],
            (float)s.f[M][i],  (float)s.f[E2][i], (float)s.f[GX][i],
            (float)s.f[GY][i], (float)s.f[GZ][i], i};
  }
  __device__ static void body32(const I &q, const JRec &r, float rx, float ry,
                                float rz, float s2, bool far, A32 &a) {
    float y = rsq32(s2);
    y = far ? y : 0.0f;
    float y2 = y * y;
    float y3 = y * y2;
    float ux = q.m * r.gx - r.m * q.gx;
    float uy = q.m * r.gy - r.m * q.gy;
    float uz = q.m * r.gz - r.m * q.gz;
    float t = 3.0f * (rx * ux + ry * uy + rz * uz) * y2;
    a.v[0] += y3 * (ux - t * rx);
    a.v[1] += y3 * (uy - t * ry);
    a.v[2] += y3 * (uz - t * rz);
    a.v[3] += y3 * (r.gx * rx + r.gy * ry + r.gz * rz);
  }
};

struct PhiAdjF64 : SumOut<4> {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 6, kMaxIpl = 2, kPrefIpl = 1;
  using Params = NoParams;
  struct JRec { double m, x, y, z, e2, h; };
  using I = JRec;
  using A64 = F64Acc<4>;
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j], s.f[RX][j], s.f[RY][j], s.f[RZ][j], s.f[E2][j],
            s.f[H][j]};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return pack(s, i, 0.0);
  }
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    double r2 = rx * rx + ry * ry + rz * rz;
    double y = rsq64(r2 + (q.e2 + r.e2));
    y = r2 > 0.0 ? y : 0.0;
    double w = (q.h * r.m + r.h * q.m) * (y * y * y);
    a.v[0] += w * rx;
    a.v[1] += w * ry;
    a.v[2] += w * rz;
    a.v[3] -= r.h * y;
  }
};

struct PhiAdjF32c : F32cBase<PhiAdjF32c, PhiAdjF64, 4> {
  static constexpr int kFields = 6, kMaxIpl = 2, kPrefIpl = 1;
  struct JRec {
    double x, y, z;
    float m, e2, near2, h;
  };
  struct I {
    double x, y, z;
    float m, e2, h;
    int i;
  };
  __device__ static JRec pack(const Soa &s, int j, const PackCtx &c) {
    return {s.f[RX][j],       s.f[RY][j],        s.f[RZ][j],
            (float)s.f[M][j], (float)s.f[E2][j],
            near_threshold(s.f[M][j], s.f[E2][j], c), (float)s.f[H][j]};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return {s.f[RX][i],       s.f[RY][i],        s.f[RZ][i],
            (float)s.f[M][i], (float)s.f[E2][i], (float)s.f[H][i], i};
  }
  __device__ static void body32(const I &q, const JRec &r, float rx, float ry,
                                float rz, float s2, bool far, A32 &a) {
    float y = rsq32(s2);
    y = far ? y : 0.0f;
    float w = (q.h * r.m + r.h * q.m) * (y * y * y);
    a.v[0] += w * rx;
    a.v[1] += w * ry;
    a.v[2] += w * rz;
    a.v[3] -= r.h * y;
  }
};

}  // namespace tupan

using namespace tupan;

template <class P64, class P32c>
static int adj_dispatch(int kid, int mode, size_t ni, const double *const *ip,
                        size_t nj, const double *const *jp, double *const *op,
                        void *stream) {
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (mode == TUPAN_MODE_FP64)
    return run_engine<P64>(kid, mode, ni, ip, nj, jp, op, {}, st);
  if (mode == TUPAN_MODE_FP32C)
    return run_engine<P32c>(kid, mode, ni, ip, nj, jp, op, {}, st);
  return set_error(TUPAN_ERR_ARG, "unknown mode %d", mode);
}

extern "C" {

int tupan_acc_adjoint_dev(int mode, size_t ni, const double *const *ip,
                          size_t nj, const double *const *jp,
                          double *const *op, void *stream) {
  return adj_dispatch<AccAdjF64, AccAdjF32c>(TUPAN_K_ACC_ADJOINT, mode, ni,
                                             ip, nj, jp, op, stream);
}

int tupan_phi_adjoint_dev(int mode, size_t ni, const double *const *ip,
                          size_t nj, const double *const *jp,
                          double *const *op, void *stream) {
  return adj_dispatch<PhiAdjF64, PhiAdjF32c>(TUPAN_K_PHI_ADJOINT, mode, ni,
                                             ip, nj, jp, op, stream);
}

}  // extern "C"
