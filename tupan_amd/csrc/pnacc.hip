// pnacc.hip -- post-Newtonian pair acceleration (FP64) on the pair engine.
//
// Pair acceleration a_pn = pn.a * r + pn.b * v with pn.a, pn.b the sums of
// the 1PN, 2PN, 2.5PN, 3PN and 3.5PN coefficient polynomials of the
// harmonic-gauge two-body equations of motion, gated by the PN order and
// scaled by m_j/s^3 and m_j/s^2 (tupan/lib/src/pn_terms.h:126-567,
// tupan/lib/src/pnacc_kernel_common.h:8-68).  The unit vector n = r/s uses
// the softened 1/s, as the reference does (pn_terms.h:575-577).  The
// polynomials keep the reference's grouping and term order: the terms have
// alternating signs and this keeps the FP64 results within a few ulp of the
// reference C.
#include "ieee_engine.hpp"
#include "launch.hpp"

namespace tupan {

// rational constant n/d in FP64
#define Q(n, d) (((double)(n)) / (d))

constexpr double kPi2 = 9.869604401089358;  // common.h:21

struct PnParams {
  unsigned order;
  double inv1, inv2, inv3, inv4, inv5, inv6, inv7;
};

struct PnAB {
  double a, b;
};

namespace pnfast {
#include "pn_terms.inc"
}  // namespace pnfast

// the same terms, one IEEE operation at a time (TUPAN_MODE_FP64_IEEE)
#pragma clang fp contract(off)
namespace pnexact {
#include "pn_terms.inc"
}  // namespace pnexact
#pragma clang fp contract(fast)



struct PnAccF64 : SumOut<3> {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 8, kMaxIpl = 1, kPrefIpl = 1;
  using Params = PnParams;
  struct JRec { double m, x, y, z, e2, vx, vy, vz; };
  using I = JRec;
  using A64 = F64Acc<3>;
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j],  s.f[RX][j], s.f[RY][j], s.f[RZ][j],
            s.f[E2][j], s.f[VX][j], s.f[VY][j], s.f[VZ][j]};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return pack(s, i, 0.0);
  }
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &c) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    double vx = q.vx - r.vx, vy = q.vy - r.vy, vz = q.vz - r.vz;
    double r2 = rx * rx + ry * ry + rz * rz;
    double v2 = vx * vx + vy * vy + vz * vz;
    double inv_r = rsq64(r2 + (q.e2 + r.e2));
    inv_r = r2 > 0.0 ? inv_r : 0.0;
    double inv_r2 = inv_r * inv_r;
    double inv_r3 = inv_r2 * inv_r;
    PnAB pn = pnfast::pn_pair(q.m, r.m, rx, ry, rz, vx, vy, vz, v2, q.vx, q.vy, q.vz,
                      r.vx, r.vy, r.vz, inv_r, inv_r2, inv_r3, c);
    a.v[0] += pn.a * rx + pn.b * vx;
    a.v[1] += pn.a * ry + pn.b * vy;
    a.v[2] += pn.a * rz + pn.b * vz;
  }
};

// The same pair acceleration in the reference's arithmetic
// (TUPAN_MODE_FP64_IEEE, pnacc_kernel_common.h:8-68): 1/s^2 by a correctly
// rounded division, 1/s = sqrt(1/s^2), 1/s^3 = 1/s^2 * 1/s
// (smoothing.h:508-520), the uncontracted terms, and one accumulator per i
// in ascending j order.
struct PnAccIeee : ieee::IeeeBase<3> {
  static constexpr int kFields = 8;
  using Params = PnParams;
  using JRec = PnAccF64::JRec;
  using I = JRec;
  __device__ static JRec pack(const Soa &s, int j, double c) {
    return PnAccF64::pack(s, j, c);
  }
  __device__ static I load_i(const Soa &s, int i) {
    return PnAccF64::pack(s, i, 0.0);
  }
  __device__ static void terms(const I &q, const JRec &r, double *t,
                               const Params &c) {
#pragma clang fp contract(off)
    double rx = q.x - r.x;
    double ry = q.y - r.y;
    double rz = q.z - r.z;
    double e2 = q.e2 + r.e2;
    double vx = q.vx - r.vx;
    double vy = q.vy - r.vy;
    double vz = q.vz - r.vz;
    double r2 = rx * rx + ry * ry + rz * rz;
    double v2 = vx * vx + vy * vy + vz * vz;
    double inv_r2 = ieee::ieee_inv_r2(r2, e2);
    double inv_r1 = sqrt(inv_r2);
    double inv_r3 = inv_r2 * inv_r1;
    PnAB pn = pnexact::pn_pair(q.m, r.m, rx, ry, rz, vx, vy, vz, v2, q.vx,
                               q.vy, q.vz, r.vx, r.vy, r.vz, inv_r1, inv_r2,
                               inv_r3, c);
    t[0] = pn.a * rx + pn.b * vx;
    t[1] = pn.a * ry + pn.b * vy;
    t[2] = pn.a * rz + pn.b * vz;
  }
  __device__ static void fold(double *acc, const double *t) {
#pragma clang fp contract(off)
    for (int k = 0; k < 3; ++k) acc[k] = acc[k] + t[k];
  }
};

#undef Q

}  // namespace tupan

using namespace tupan;

static int pn_params(unsigned order, const double *inv, PnParams *p) {
  if (!inv) return set_error(TUPAN_ERR_ARG, "NULL inverse-c array");
  *p = PnParams{order, inv[0], inv[1], inv[2], inv[3], inv[4], inv[5], inv[6]};
  return TUPAN_OK;
}

extern "C" {

int tupan_pnacc_dev(int mode, size_t ni, const double *const *ip,
                    size_t nj, const double *const *jp, unsigned order,
                    const double *inv, double *const *op, void *stream) {
  PnParams prm;
  int rc = pn_params(order, inv, &prm);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (mode == TUPAN_MODE_FP64_IEEE)
    return ieee::run_ieee_engine<PnAccIeee, 1>(ni, ip, nj, jp, op, st, prm);
  if (mode != TUPAN_MODE_FP64 && mode != TUPAN_MODE_FP32C)
    return set_error(TUPAN_ERR_ARG, "unknown mode %d", mode);
  return run_engine<PnAccF64>(TUPAN_K_PNACC, TUPAN_MODE_FP64, ni, ip, nj, jp,
                              op, prm, st);
}

int tupan_pnacc_host(int mode, size_t ni, const double *const *ip, size_t nj,
                     const double *const *jp, unsigned order,
                     const double *inv, double *const *op) {
  return host_call(ni, ip, nj, jp, 8, op, 3,
                   [&](const double *const *di, const double *const *dj,
                       double *const *dout, hipStream_t st) {
                     return tupan_pnacc_dev(mode, ni, di, nj, dj, order, inv,
                                            dout, st);
                   });
}

}  // extern "C"
