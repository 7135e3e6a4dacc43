// nreg.hip -- the NREG pair sums and the two-body Kepler kernel, FP64.
//
//   NREG X  tupan/lib/src/nreg_kernels_common.h:7-58 (driver
//           nreg_kernels.c:5-61): sums at the drifted separation
//           r_ij + v_ij dt of m_j r, the softened acceleration and
//           m_j / s; the last is finished as u_i = m_i * sum.
//   NREG V  nreg_kernels_common.h:61-99 (driver nreg_kernels.c:64-115):
//           sums at the kicked relative velocity w = v_ij + a_ij dt of
//           m_j w and m_j |w|^2, the last finished as k_i = m_i * sum.
//   Kepler  kepler_solver_kernel_common.h:8-79 (driver
//           kepler_solver_kernel.c:5-49): particles 0 and 1 advanced by dt
//           along their softened Kepler orbit about the freely drifting
//           centre of mass (twobody.hpp's kepler_drift).
#include "ieee_engine.hpp"
#include "launch.hpp"
#include "twobody.hpp"

namespace tupan {

struct NregParams {
  double dt;
  const double *im;   // i-side masses, for the finishing factor
  const double *dtp;  // when set, dt is read from here at kernel start
};

// dt from device memory when the caller passed it there
struct NregDt {
  __device__ static NregParams prepare(NregParams p) {
    if (p.dtp) p.dt = *p.dtp;
    return p;
  }
};

struct NregXF64 : NregDt {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 8, kMaxIpl = 2, kPrefIpl = 1, kOut = 7;
  using Params = NregParams;
  struct JRec { double m, x, y, z, e2, vx, vy, vz; };
  using I = JRec;
  using A64 = F64Acc<7>;
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
                              const Params &prm) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    rx += (q.vx - r.vx) * prm.dt;
    ry += (q.vy - r.vy) * prm.dt;
    rz += (q.vz - r.vz) * prm.dt;
    double r2 = rx * rx + ry * ry + rz * rz;
    double y = rsq64(r2 + (q.e2 + r.e2));
    y = r2 > 0.0 ? y : 0.0;
    double k = r.m * y * (y * y);
    a.v[0] += r.m * rx;
    a.v[1] += r.m * ry;
    a.v[2] += r.m * rz;
    a.v[3] -= k * rx;
    a.v[4] -= k * ry;
    a.v[5] -= k * rz;
    a.v[6] += r.m * y;
  }
  __device__ static void combine(double *v, const double *w) {
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] += w[k];
  }
  __device__ static void finish(const double *v, Out out, int i,
                                const Params &prm) {
#pragma unroll
    for (int k = 0; k < 6; ++k) out.f[k][i] = v[k];
    out.f[6][i] = prm.im[i] * v[6];
  }
};

// Fields (reference order, libtupan.h:181-197): m, vx, vy, vz, ax, ay, az.
struct NregVF64 : NregDt {
  static constexpr bool kTwoLevel = false;
  static constexpr int kFields = 7, kMaxIpl = 2, kPrefIpl = 1, kOut = 4;
  using Params = NregParams;
  struct JRec { double m, vx, vy, vz, ax, ay, az, pad; };
  using I = JRec;
  using A64 = F64Acc<4>;
  __device__ static JRec pack(const Soa &s, int j, double) {
    return {s.f[0][j], s.f[1][j], s.f[2][j], s.f[3][j],
            s.f[4][j], s.f[5][j], s.f[6][j], 0.0};
  }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return pack(s, i, 0.0);
  }
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  __device__ static void pair(const I &q, const JRec &r, A64 &a,
                              const Params &prm) {
    double wx = (q.vx - r.vx) + (q.ax - r.ax) * prm.dt;
    double wy = (q.vy - r.vy) + (q.ay - r.ay) * prm.dt;
    double wz = (q.vz - r.vz) + (q.az - r.az) * prm.dt;
    double w2 = wx * wx + wy * wy + wz * wz;
    a.v[0] += r.m * wx;
    a.v[1] += r.m * wy;
    a.v[2] += r.m * wz;
    a.v[3] += r.m * w2;
  }
  __device__ static void combine(double *v, const double *w) {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] += w[k];
  }
  __device__ static void finish(const double *v, Out out, int i,
                                const Params &prm) {
#pragma unroll
    for (int k = 0; k < 3; ++k) out.f[k][i] = v[k];
    out.f[3][i] = prm.im[i] * v[3];
  }
};

// The same sums in the reference's arithmetic (TUPAN_MODE_FP64_IEEE): its
// operations in its order without contraction, correctly rounded division
// and sqrt, one accumulator per i in ascending j order.
struct NregXIeee : ieee::IeeeBase<7>, NregDt {
  static constexpr int kFields = 8;
  using Params = NregParams;
  using JRec = NregXF64::JRec;
  using I = JRec;
  __device__ static JRec pack(const Soa &s, int j, double c) {
    return NregXF64::pack(s, j, c);
  }
  __device__ static I load_i(const Soa &s, int i) {
    return NregXF64::pack(s, i, 0.0);
  }
  // nreg_kernels_common.h:7-58
  __device__ static void terms(const I &q, const JRec &r, double *t,
                               const Params &p) {
#pragma clang fp contract(off)
    double rx = q.x - r.x;
    double ry = q.y - r.y;
    double rz = q.z - r.z;
    double e2 = q.e2 + r.e2;
    double vx = q.vx - r.vx;
    double vy = q.vy - r.vy;
    double vz = q.vz - r.vz;
    rx += vx * p.dt;
    ry += vy * p.dt;
    rz += vz * p.dt;
    double r2 = rx * rx + ry * ry + rz * rz;
    double inv_r2 = ieee::ieee_inv_r2(r2, e2);
    double inv_r1 = sqrt(inv_r2);
    double inv_r3 = inv_r2 * inv_r1;
    inv_r3 *= r.m;
    t[0] = r.m * rx;
    t[1] = r.m * ry;
    t[2] = r.m * rz;
    t[3] = inv_r3 * rx;
    t[4] = inv_r3 * ry;
    t[5] = inv_r3 * rz;
    t[6] = r.m * inv_r1;
  }
  __device__ static void fold(double *acc, const double *t) {
#pragma clang fp contract(off)
    for (int c = 0; c < 3; ++c) acc[c] = acc[c] + t[c];
    for (int c = 3; c < 6; ++c) acc[c] = acc[c] - t[c];
    acc[6] = acc[6] + t[6];
  }
  __device__ static void finish(const double *acc, Out out, int i,
                                const Params &p) {
#pragma clang fp contract(off)
    for (int c = 0; c < 6; ++c) out.f[c][i] = acc[c];
    out.f[6][i] = p.im[i] * acc[6];
  }
};

struct NregVIeee : ieee::IeeeBase<4>, NregDt {
  static constexpr int kFields = 7;
  using Params = NregParams;
  using JRec = NregVF64::JRec;
  using I = JRec;
  __device__ static JRec pack(const Soa &s, int j, double c) {
    return NregVF64::pack(s, j, c);
  }
  __device__ static I load_i(const Soa &s, int i) {
    return NregVF64::pack(s, i, 0.0);
  }
  // nreg_kernels_common.h:61-99
  __device__ static void terms(const I &q, const JRec &r, double *t,
                               const Params &p) {
#pragma clang fp contract(off)
    double vx = q.vx - r.vx;
    double vy = q.vy - r.vy;
    double vz = q.vz - r.vz;
    double ax = q.ax - r.ax;
    double ay = q.ay - r.ay;
    double az = q.az - r.az;
    vx += ax * p.dt;
    vy += ay * p.dt;
    vz += az * p.dt;
    double v2 = vx * vx + vy * vy + vz * vz;
    t[0] = r.m * vx;
    t[1] = r.m * vy;
    t[2] = r.m * vz;
    t[3] = r.m * v2;
  }
  __device__ static void fold(double *acc, const double *t) {
#pragma clang fp contract(off)
    for (int c = 0; c < 4; ++c) acc[c] = acc[c] + t[c];
  }
  __device__ static void finish(const double *acc, Out out, int i,
                                const Params &p) {
#pragma clang fp contract(off)
    for (int c = 0; c < 3; ++c) out.f[c][i] = acc[c];
    out.f[3][i] = p.im[i] * acc[3];
  }
};

// One thread: reads particles 0 and 1 completely before writing, so the
// outputs may alias the inputs (the reference updates ips in place,
// tupan/lib/extensions.py:642-646).
__global__ void kepler_kernel(Soa s, double dt, double *r1x, double *r1y,
                              double *r1z, double *v1x, double *v1y,
                              double *v1z, int budget, int *status) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const double m0 = s.f[M][0], m1 = s.f[M][1];
  const double x0[3] = {s.f[RX][0], s.f[RY][0], s.f[RZ][0]};
  const double x1[3] = {s.f[RX][1], s.f[RY][1], s.f[RZ][1]};
  const double u0[3] = {s.f[VX][0], s.f[VY][0], s.f[VZ][0]};
  const double u1[3] = {s.f[VX][1], s.f[VY][1], s.f[VZ][1]};
  const double e2 = s.f[E2][0] + s.f[E2][1];
  const double m = m0 + m1;
  const double imu = m0 / m, jmu = m1 / m;
  double st[6], rcm[3], vcm[3];
  for (int k = 0; k < 3; ++k) {
    st[k] = x0[k] - x1[k];
    st[3 + k] = u0[k] - u1[k];
    rcm[k] = imu * x0[k] + jmu * x1[k];
    vcm[k] = imu * u0[k] + jmu * u1[k];
    rcm[k] += vcm[k] * dt;
  }
  int left = budget;
  if (!kepler_drift(dt, m, e2, st, left) && status) atomicOr(status, 1);
  double *out_r[3] = {r1x, r1y, r1z}, *out_v[3] = {v1x, v1y, v1z};
  for (int k = 0; k < 3; ++k) {
    out_r[k][0] = rcm[k] + jmu * st[k];
    out_v[k][0] = vcm[k] + jmu * st[3 + k];
    out_r[k][1] = rcm[k] - imu * st[k];
    out_v[k][1] = vcm[k] - imu * st[3 + k];
  }
}

// One step of NREG's scalar recurrences (nreg.py:22-97) on the scalar
// block (TUPAN_NREG_* slots), in Python's double arithmetic: one operation
// per statement, no contraction, correctly rounded division.
__global__ void nreg_scalar_kernel(double *sc, int op, double h) {
#pragma clang fp contract(off)
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  switch (op) {
    case TUPAN_NREG_OP_BEGIN:
      sc[TUPAN_NREG_T0] = sc[TUPAN_NREG_T];
      break;
    case TUPAN_NREG_OP_H:
      sc[TUPAN_NREG_HSUB] = h;
      break;
    case TUPAN_NREG_OP_HS:
      sc[TUPAN_NREG_HSUB] = 0.5 * (h * sc[TUPAN_NREG_S]);
      break;
    case TUPAN_NREG_OP_XDT:
      sc[TUPAN_NREG_DT] = 0.5 * (sc[TUPAN_NREG_HSUB] / sc[TUPAN_NREG_W]);
      break;
    case TUPAN_NREG_OP_VDT:
      sc[TUPAN_NREG_DT] = sc[TUPAN_NREG_HSUB] / sc[TUPAN_NREG_U];
      break;
    case TUPAN_NREG_OP_XEND:
      sc[TUPAN_NREG_U] = 0.5 * sc[TUPAN_NREG_SUM];
      sc[TUPAN_NREG_T] = sc[TUPAN_NREG_T] + sc[TUPAN_NREG_DT];
      break;
    case TUPAN_NREG_OP_VEND: {
      const double kin = 0.25 * sc[TUPAN_NREG_SUM] / sc[TUPAN_NREG_MTOT];
      sc[TUPAN_NREG_W] = kin - sc[TUPAN_NREG_E0];
      break;
    }
    case TUPAN_NREG_OP_S:
      sc[TUPAN_NREG_S] =
          1 / (2 / sc[TUPAN_NREG_W] - 1 / sc[TUPAN_NREG_S]);
      break;
    default:
      break;
  }
}

}  // namespace tupan

using namespace tupan;

extern "C" {

int tupan_nreg_scalars_dev(double *sc, int op, double h, void *stream) {
  if (!sc) return set_error(TUPAN_ERR_ARG, "NULL scalar block");
  if (op < TUPAN_NREG_OP_BEGIN || op > TUPAN_NREG_OP_S)
    return set_error(TUPAN_ERR_ARG, "unknown NREG scalar op %d", op);
  hipLaunchKernelGGL(nreg_scalar_kernel, dim3(1), dim3(64), 0,
                     reinterpret_cast<hipStream_t>(stream), sc, op, h);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? TUPAN_OK : hip_fail(e, "nreg scalar launch");
}

static int nreg_mode_ok(int mode) {
  if (mode == TUPAN_MODE_FP64 || mode == TUPAN_MODE_FP32C ||
      mode == TUPAN_MODE_FP64_IEEE)
    return TUPAN_OK;
  return set_error(TUPAN_ERR_ARG, "unknown mode %d", mode);
}

static int nreg_x(int mode, size_t ni, const double *const *ip, size_t nj,
                  const double *const *jp, double dt, const double *dtp,
                  double *const *op, void *stream) {
  if (int rc = nreg_mode_ok(mode)) return rc;
  if (ni && (!ip || !ip[M]))
    return set_error(TUPAN_ERR_ARG, "NULL i mass field");
  NregParams prm{dt, ni ? ip[M] : nullptr, dtp};
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (mode == TUPAN_MODE_FP64_IEEE)
    return ieee::run_ieee_engine<NregXIeee, 4>(ni, ip, nj, jp, op, st, prm);
  return run_engine<NregXF64>(TUPAN_K_NREG_X, TUPAN_MODE_FP64, ni, ip, nj,
                              jp, op, prm, st);
}

static int nreg_v(int mode, size_t ni, const double *const *ip, size_t nj,
                  const double *const *jp, double dt, const double *dtp,
                  double *const *op, void *stream) {
  if (int rc = nreg_mode_ok(mode)) return rc;
  if (ni && (!ip || !ip[0]))
    return set_error(TUPAN_ERR_ARG, "NULL i mass field");
  NregParams prm{dt, ni ? ip[0] : nullptr, dtp};
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (mode == TUPAN_MODE_FP64_IEEE)
    return ieee::run_ieee_engine<NregVIeee, 4>(ni, ip, nj, jp, op, st, prm);
  return run_engine<NregVF64>(TUPAN_K_NREG_V, TUPAN_MODE_FP64, ni, ip, nj,
                              jp, op, prm, st);
}

int tupan_nreg_x_dev(int mode, size_t ni, const double *const *ip, size_t nj,
                     const double *const *jp, double dt, double *const *op,
                     void *stream) {
  return nreg_x(mode, ni, ip, nj, jp, dt, nullptr, op, stream);
}

int tupan_nreg_x_devdt(int mode, size_t ni, const double *const *ip,
                       size_t nj, const double *const *jp, const double *d_dt,
                       double *const *op, void *stream) {
  if (!d_dt) return set_error(TUPAN_ERR_ARG, "NULL device dt");
  return nreg_x(mode, ni, ip, nj, jp, 0.0, d_dt, op, stream);
}

int tupan_nreg_x_host(int mode, size_t ni, const double *const *ip,
                      size_t nj, const double *const *jp, double dt,
                      double *const *op) {
  return host_call(ni, ip, nj, jp, 8, op, 7,
                   [&](const double *const *di, const double *const *dj,
                       double *const *dout, hipStream_t st) {
                     return tupan_nreg_x_dev(mode, ni, di, nj, dj, dt, dout,
                                             st);
                   });
}

int tupan_nreg_v_dev(int mode, size_t ni, const double *const *ip, size_t nj,
                     const double *const *jp, double dt, double *const *op,
                     void *stream) {
  return nreg_v(mode, ni, ip, nj, jp, dt, nullptr, op, stream);
}

int tupan_nreg_v_devdt(int mode, size_t ni, const double *const *ip,
                       size_t nj, const double *const *jp, const double *d_dt,
                       double *const *op, void *stream) {
  if (!d_dt) return set_error(TUPAN_ERR_ARG, "NULL device dt");
  return nreg_v(mode, ni, ip, nj, jp, 0.0, d_dt, op, stream);
}

int tupan_nreg_v_host(int mode, size_t ni, const double *const *ip,
                      size_t nj, const double *const *jp, double dt,
                      double *const *op) {
  return host_call(ni, ip, nj, jp, 7, op, 4,
                   [&](const double *const *di, const double *const *dj,
                       double *const *dout, hipStream_t st) {
                     return tupan_nreg_v_dev(mode, ni, di, nj, dj, dt, dout,
                                             st);
                   });
}

int tupan_kepler_dev(const double *const *ip, double dt, double *const *op,
                     int *d_status, void *stream) {
  if (!ip || !op) return set_error(TUPAN_ERR_ARG, "NULL field array");
  Soa s{};
  for (int k = 0; k < 8; ++k) {
    if (!ip[k]) return set_error(TUPAN_ERR_ARG, "NULL input field %d", k);
    s.f[k] = ip[k];
  }
  for (int k = 0; k < 6; ++k)
    if (!op[k]) return set_error(TUPAN_ERR_ARG, "NULL output field %d", k);
  hipLaunchKernelGGL(kepler_kernel, dim3(1), dim3(64), 0,
                     reinterpret_cast<hipStream_t>(stream), s, dt, op[0],
                     op[1], op[2], op[3], op[4], op[5], sakura_budget(),
                     d_status);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? TUPAN_OK : hip_fail(e, "kepler launch");
}

int tupan_kepler_host(const double *const *ip, double dt, double *const *op) {
  // two particles: stage them in one small device buffer
  double h[8 * 2], *d = nullptr;
  for (int k = 0; k < 8; ++k) {
    if (!ip || !ip[k]) return set_error(TUPAN_ERR_ARG, "NULL input field");
    h[2 * k] = ip[k][0];
    h[2 * k + 1] = ip[k][1];
  }
  if (!op) return set_error(TUPAN_ERR_ARG, "NULL output field array");
  int rc = device_alloc((void **)&d, sizeof(double) * (16 + 12) + 64, nullptr);
  if (rc) return rc;
  int *status = reinterpret_cast<int *>(d + 28);
  hipError_t e = hipMemcpy(d, h, sizeof h, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(status, 0, sizeof(int));
  if (e != hipSuccess) {
    device_free(d, nullptr);
    return hip_fail(e, "kepler staging");
  }
  const double *dip[8];
  double *dop[6];
  for (int k = 0; k < 8; ++k) dip[k] = d + 2 * k;
  for (int k = 0; k < 6; ++k) dop[k] = d + 16 + 2 * k;
  rc = tupan_kepler_dev(dip, dt, dop, status, nullptr);
  double o[12];
  int hs = 0;
  if (!rc) {
    e = hipMemcpy(o, d + 16, sizeof o, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = hipMemcpy(&hs, status, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "kepler copy-back");
  }
  device_free(d, nullptr);
  (void)hipStreamSynchronize(nullptr);
  if (rc) return rc;
  for (int k = 0; k < 6; ++k) {
    if (!op[k]) return set_error(TUPAN_ERR_ARG, "NULL output field %d", k);
    op[k][0] = o[2 * k];
    op[k][1] = o[2 * k + 1];
  }
  if (hs)
    return set_error(TUPAN_ERR_NOT_CONVERGED,
                     "the Kepler solve exhausted its budget of %d solves",
                     sakura_budget());
  return TUPAN_OK;
}

}  // extern "C"
