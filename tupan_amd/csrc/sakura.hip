// sakura.hip -- Sakura pairwise two-body drift (FP64) on the pair engine.
//
// For every pair the relative orbit (r0, v0) = (x_i - x_j, v_i - v_j) with
// m = m_i + m_j is evolved by pair_evolve (twobody.hpp) and the i-particle
// receives dr_i += m_j/m (r1 - r0), dv_i += m_j/m (v1 - v0)
// (tupan/lib/src/sakura_kernel_common.h:194-242).  Lanes whose pairs need
// Kepler retries run while the others idle: the cost is data-dependent by
// nature.  A pair that exhausts its solve budget sets bit 0 of the caller's
// status word.
#include "launch.hpp"
#include "twobody.hpp"

namespace tupan {

struct SakuraParams {
  double dt;
  int flag;
  int budget;
  int *status;
  double rc2;  // only pairs with r^2 < rc2 evolve (infinity: all pairs)
  const uint32_t *mask;  // filled by the launcher (near_mask_kernel)
  int mask_tiles;
};

struct SakuraF64 : SumOut<6> {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 8, kMaxIpl = 1, kPrefIpl = 1;
  using Params = SakuraParams;
  struct JRec { double m, x, y, z, e2, vx, vy, vz; };
  using I = JRec;
  using A64 = F64Acc<6>;
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j],  s.f[RX][j], s.f[RY][j], s.f[RZ][j],
            s.f[E2][j], s.f[VX][j], s.f[VY][j], s.f[VZ][j]};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return pack(s, i, 0.0);
  }
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  // Cut-off runs (finite rc2) test every pair before evolving any.
  static constexpr bool kSparse = true;
  __host__ __device__ static bool sparse(const Params &prm) {
    return prm.rc2 < INFINITY;
  }
  // the near mask's cut-off on r^2: rc2 widened by 2^-30 (r^2 computed
  // with another rounding differs from pair()'s by a few ulps)
  __device__ static double mask_cut(const Params &prm) {
    return prm.rc2 * (1.0 + 0x1p-30);
  }
  // particle j paired with itself is the zero state; it adds nothing when
  // its weight and dt are finite (zero_state below)
  __device__ static bool self_noop(const Soa &s, int j, const Params &prm) {
    const double m = s.f[M][j] + s.f[M][j];
    return m > 0 && m < INFINITY && fabs(prm.dt) < INFINITY;
  }
  // The pairs pair() evolves: r^2 < rc2, except the zero state (r = v = 0:
  // a particle paired with itself, or an exact duplicate) of a pair whose
  // weight m_j / m is finite, with a finite dt.  pair_evolve leaves the
  // zero state at zero, so such a pair adds exactly +-0 to accumulators
  // that start at +0, and skipping it changes no bit.  (A massless pair
  // adds NaN, as in the reference, and is kept.)
  __device__ static bool zero_state(double r2, double dvx, double dvy,
                                    double dvz, double m, double dt) {
    return !(r2 > 0) && dvx == 0 && dvy == 0 && dvz == 0 && m > 0 &&
           m < INFINITY && fabs(dt) < INFINITY;
  }
  __device__ static bool near(const I &q, const JRec &r, const Params &prm) {
    double dx = q.x - r.x, dy = q.y - r.y, dz = q.z - r.z;
    double r2 = dx * dx + dy * dy + dz * dz;
    return r2 < prm.rc2 && !zero_state(r2, q.vx - r.vx, q.vy - r.vy,
                                       q.vz - r.vz, q.m + r.m, prm.dt);
  }
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &prm) {
    double st[6] = {q.x - r.x,   q.y - r.y,   q.z - r.z,
                    q.vx - r.vx, q.vy - r.vy, q.vz - r.vz};
    const double r2 = st[0] * st[0] + st[1] * st[1] + st[2] * st[2];
    if (prm.rc2 < INFINITY && !(r2 < prm.rc2)) return;
    if (zero_state(r2, st[3], st[4], st[5], q.m + r.m, prm.dt)) return;
    const double r0[6] = {st[0], st[1], st[2], st[3], st[4], st[5]};
    double m = q.m + r.m;
    double e2 = q.e2 + r.e2;
    if (!pair_evolve(prm.dt, prm.flag, m, e2, st, prm.budget) && prm.status)
      atomicOr(prm.status, 1);
    double w = r.m / m;
#pragma unroll
    for (int k = 0; k < 6; ++k) a.v[k] += w * (st[k] - r0[k]);
  }
};

// Sakura restricted to the pairs whose classification positions (fields
// 8-10: cx, cy, cz) are closer than rcut: the close pairs of the hybrid
// scheme of AccFarF64, split by the same positions (fixed for the step,
// so that both Sakura passes of a drift and both kicks agree on every
// pair).  The zero state (a particle paired with itself) is
// recognised from the current state, as in SakuraF64.
struct SakuraNearF64 : SakuraF64 {
  static constexpr int kFields = 11;
  static constexpr int kClassX = 8;  // near_mask_kernel tests fields 8-10
  struct JRec { double m, x, y, z, e2, vx, vy, vz, cx, cy, cz; };
  using I = JRec;
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[M][j],  s.f[RX][j], s.f[RY][j], s.f[RZ][j],
            s.f[E2][j], s.f[VX][j], s.f[VY][j], s.f[VZ][j],
            s.f[8][j],  s.f[9][j],  s.f[10][j]};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return pack(s, i, 0.0);
  }
  __device__ static double class_r2(const I &q, const JRec &r) {
    double cx = q.cx - r.cx, cy = q.cy - r.cy, cz = q.cz - r.cz;
    return cx * cx + cy * cy + cz * cz;
  }
  __device__ static bool near(const I &q, const JRec &r, const Params &prm) {
    double dx = q.x - r.x, dy = q.y - r.y, dz = q.z - r.z;
    double r2 = dx * dx + dy * dy + dz * dz;
    return class_r2(q, r) < prm.rc2 &&
           !zero_state(r2, q.vx - r.vx, q.vy - r.vy, q.vz - r.vz, q.m + r.m,
                       prm.dt);
  }
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &prm) {
    if (!(class_r2(q, r) < prm.rc2)) return;
    double st[6] = {q.x - r.x,   q.y - r.y,   q.z - r.z,
                    q.vx - r.vx, q.vy - r.vy, q.vz - r.vz};
    const double r2 = st[0] * st[0] + st[1] * st[1] + st[2] * st[2];
    if (zero_state(r2, st[3], st[4], st[5], q.m + r.m, prm.dt)) return;
    const double r0[6] = {st[0], st[1], st[2], st[3], st[4], st[5]};
    double m = q.m + r.m;
    double e2 = q.e2 + r.e2;
    if (!pair_evolve(prm.dt, prm.flag, m, e2, st, prm.budget) && prm.status)
      atomicOr(prm.status, 1);
    double w = r.m / m;
#pragma unroll
    for (int k = 0; k < 6; ++k) a.v[k] += w * (st[k] - r0[k]);
  }
};

}  // namespace tupan

using namespace tupan;

extern "C" {

int tupan_sakura_dev(size_t ni, const double *const *ip, size_t nj,
                     const double *const *jp, double dt, int flag,
                     double *const *op, int *d_status, void *stream) {
  SakuraParams prm{dt, flag, sakura_budget(), d_status, INFINITY};
  return run_engine<SakuraF64>(TUPAN_K_SAKURA, TUPAN_MODE_FP64, ni, ip, nj,
                               jp, op, prm,
                               reinterpret_cast<hipStream_t>(stream));
}

int tupan_sakura_near_dev(size_t ni, const double *const *ip, size_t nj,
                          const double *const *jp, double dt, int flag,
                          double rcut, double *const *op, int *d_status,
                          void *stream) {
  if (!(rcut >= 0.0))
    return set_error(TUPAN_ERR_ARG, "rcut must be >= 0 (got %g)", rcut);
  SakuraParams prm{dt, flag, sakura_budget(), d_status, rcut * rcut};
  return run_engine<SakuraNearF64>(TUPAN_K_SAKURA, TUPAN_MODE_FP64, ni, ip,
                                   nj, jp, op, prm,
                                   reinterpret_cast<hipStream_t>(stream));
}

int tupan_sakura_host(size_t ni, const double *const *ip, size_t nj,
                      const double *const *jp, double dt, int flag,
                      double *const *op) {
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1)
    return set_error(TUPAN_ERR_HIP, "no HIP device available");
  int *status = nullptr;
  int rc = device_alloc((void **)&status, sizeof(int), nullptr);
  if (rc) return rc;
  hipError_t e = hipMemsetAsync(status, 0, sizeof(int), nullptr);
  if (e != hipSuccess) {
    device_free(status, nullptr);
    return hip_fail(e, "hipMemsetAsync");
  }
  rc = host_call(ni, ip, nj, jp, 8, op, 6,
                 [&](const double *const *di, const double *const *dj,
                     double *const *dout, hipStream_t st) {
                   SakuraParams prm{dt, flag, sakura_budget(), status,
                                    INFINITY};
                   return run_engine<SakuraF64>(TUPAN_K_SAKURA,
                                                TUPAN_MODE_FP64, ni, di, nj,
                                                dj, dout, prm, st);
                 });
  int hstat = 0;
  e = hipMemcpy(&hstat, status, sizeof(int), hipMemcpyDeviceToHost);
  device_free(status, nullptr);
  (void)hipStreamSynchronize(nullptr);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "status copy");
  if (hstat)
    return set_error(TUPAN_ERR_NOT_CONVERGED,
                     "a Kepler solve exhausted its budget of %d solves",
                     sakura_budget());
  return TUPAN_OK;
}

}  // extern "C"
