// integrate.hip -- per-particle update steps of Tupan's integrators, FP64.
//
// The reference performs these updates with numpy expressions on SoA
// arrays; every kernel here evaluates the same expression in the same order,
// one IEEE operation at a time (no contraction into FMA: `fp contract(off)`;
// FP64 division stays correctly rounded), so that a device-resident driver
// produces the numpy driver's bits given the same pair-kernel outputs.
//
//   hermite_predict  tupan/integrator/hermite.py:25-41 (H2), 75-90 (H4),
//                    127-159 (H6), 203-241 (H8)
//   hermite_correct  hermite.py:43-57 (H2), 93-119 (H4), 162-195 (H6),
//                    244-284 (H8)
//   drift / kick     tupan/integrator/sia.py:65-87 (drift_n, kick_n),
//                    sakura.py:26-28, 45-47 (half drifts)
//   kick_pn          sia.py:142-156 with tupan/particles/body.py:471-525
//                    (pn_kick_ke, pn_kick_lmom, pn_kick_amom) and
//                    body.py:481-492 (pn_drift_com_r)
//   apply_delta      sakura.py:31-36, 39-44
//   divide           nreg.py:28-30, 54-56 (r = mr / M, v = mv / M)
#include "hermite.hpp"
#include "launch.hpp"

namespace tupan {
namespace {

constexpr int kThreads = 256;

struct P18 {
  double *f[18];
};

inline int grid_for(size_t n) {
  size_t g = (n + kThreads - 1) / kThreads;
  return (int)(g < 65535 * 8 ? g : 65535 * 8);
}

// x = {rx,ry,rz, vx,vy,vz, ax,ay,az, jx,jy,jz, sx,sy,sz, cx,cy,cz}
// On entry x holds the step's initial state and its derivatives; x0
// receives that snapshot (the reference's ps0 = ps.copy()), then x's
// positions and velocities are advanced by the Taylor predictor.
template <int ORDER>
__global__ void __launch_bounds__(kThreads)
    hermite_predict_kernel(int n, P18 x, P18 x0, double tau) {
  constexpr int nd = 6 + 3 * (ORDER / 2);  // r, v + derivatives used
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
    double s[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) s[k] = 0.0;
#pragma unroll
    for (int k = 0; k < nd; ++k) {
      s[k] = x.f[k][i];
      x0.f[k][i] = s[k];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double r1, v1;
      hermite_predict_axis<ORDER>(s[d], s[3 + d], s[6 + d], s[9 + d],
                                  s[12 + d], s[15 + d], tau, r1, v1);
      x.f[d][i] = r1;
      x.f[3 + d][i] = v1;
    }
  }
}

// x: predicted r, v (overwritten with the corrected ones) and the
// derivatives evaluated at the predicted state; x0: the step's snapshot.
template <int ORDER>
__global__ void __launch_bounds__(kThreads)
    hermite_correct_kernel(int n, P18 x, P18 x0, double tau) {
  constexpr int nd = 6 + 3 * (ORDER / 2);
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
    double d0[18], d1[18];
#pragma unroll
    for (int k = 6; k < 18; ++k) {
      d0[k] = k < nd ? x0.f[k][i] : 0.0;
      d1[k] = k < nd ? x.f[k][i] : 0.0;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double r1, v1;
      hermite_correct_axis<ORDER>(x0.f[d][i], x0.f[3 + d][i], d0[6 + d],
                                  d1[6 + d], d0[9 + d], d1[9 + d],
                                  d0[12 + d], d1[12 + d], d0[15 + d],
                                  d1[15 + d], tau, r1, v1);
      x.f[3 + d][i] = v1;
      x.f[d][i] = r1;
    }
  }
}

// r += v * tau (half = 0) or r += v * tau / 2 (half = 1); with PN
// bookkeeping also pn_mr += pn_mv * tau.
__global__ void __launch_bounds__(kThreads)
    drift_kernel(int n, P18 p, double tau, int half, int pn) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double dr = p.f[3 + d][i] * tau;
      if (half) dr = dr / 2;
      p.f[d][i] = p.f[d][i] + dr;
      if (pn) p.f[6 + d][i] = p.f[6 + d][i] + p.f[9 + d][i] * tau;
    }
  }
}

// v += a * tau
__global__ void __launch_bounds__(kThreads)
    kick_kernel(int n, P18 p, double tau) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
#pragma unroll
    for (int d = 0; d < 3; ++d) p.f[d][i] = p.f[d][i] + p.f[3 + d][i] * tau;
  }
}

// v += (a * tau + w) / 2
__global__ void __launch_bounds__(kThreads)
    kick_pn_half_kernel(int n, P18 p, double tau) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
      p.f[d][i] = p.f[d][i] + (p.f[3 + d][i] * tau + p.f[6 + d][i]) / 2;
  }
}

// In  p.f[0..9]  = m, rx, ry, rz, vx, vy, vz, pnax, pnay, pnaz
// Out p.f[10..17]= wx, wy, wz, pn_ke, pn_mvx..z (in/out);
// q.f[0..2]      = pn_amx..z (in/out)
__global__ void __launch_bounds__(kThreads)
    pn_flux_kernel(int n, P18 p, P18 q, double tau) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
    const double m = p.f[0][i];
    const double rx = p.f[1][i], ry = p.f[2][i], rz = p.f[3][i];
    const double vx = p.f[4][i], vy = p.f[5][i], vz = p.f[6][i];
    const double ax = p.f[7][i], ay = p.f[8][i], az = p.f[9][i];
    const double fx = m * ax, fy = m * ay, fz = m * az;
    // pn_kick_ke
    p.f[13][i] = p.f[13][i] - (vx * fx + vy * fy + vz * fz) * tau;
    // pn_kick_lmom
    p.f[14][i] = p.f[14][i] - fx * tau;
    p.f[15][i] = p.f[15][i] - fy * tau;
    p.f[16][i] = p.f[16][i] - fz * tau;
    // pn_kick_amom
    q.f[0][i] = q.f[0][i] - (ry * fz - rz * fy) * tau;
    q.f[1][i] = q.f[1][i] - (rz * fx - rx * fz) * tau;
    q.f[2][i] = q.f[2][i] - (rx * fy - ry * fx) * tau;
    // w = 2 * pna * tau - w
    p.f[10][i] = 2 * ax * tau - p.f[10][i];
    p.f[11][i] = 2 * ay * tau - p.f[11][i];
    p.f[12][i] = 2 * az * tau - p.f[12][i];
  }
}

// pn_flux_kernel followed by the second half of kick_pn on the same
// particle: the flux sees the velocity before the half kick, the half kick
// the new w, as when the two kernels run one after the other.
// p as pn_flux_kernel; q.f[0..2] = pn_amx..z (in/out), q.f[3..5] = ax..az
__global__ void __launch_bounds__(kThreads)
    pn_flux_kick_kernel(int n, P18 p, P18 q, double tau) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
    const double m = p.f[0][i];
    const double rx = p.f[1][i], ry = p.f[2][i], rz = p.f[3][i];
    const double vx = p.f[4][i], vy = p.f[5][i], vz = p.f[6][i];
    const double ax = p.f[7][i], ay = p.f[8][i], az = p.f[9][i];
    const double fx = m * ax, fy = m * ay, fz = m * az;
    p.f[13][i] = p.f[13][i] - (vx * fx + vy * fy + vz * fz) * tau;
    p.f[14][i] = p.f[14][i] - fx * tau;
    p.f[15][i] = p.f[15][i] - fy * tau;
    p.f[16][i] = p.f[16][i] - fz * tau;
    q.f[0][i] = q.f[0][i] - (ry * fz - rz * fy) * tau;
    q.f[1][i] = q.f[1][i] - (rz * fx - rx * fz) * tau;
    q.f[2][i] = q.f[2][i] - (rx * fy - ry * fx) * tau;
    const double wx = 2 * ax * tau - p.f[10][i];
    const double wy = 2 * ay * tau - p.f[11][i];
    const double wz = 2 * az * tau - p.f[12][i];
    p.f[10][i] = wx;
    p.f[11][i] = wy;
    p.f[12][i] = wz;
    p.f[4][i] = vx + (q.f[3][i] * tau + wx) / 2;
    p.f[5][i] = vy + (q.f[4][i] * tau + wy) / 2;
    p.f[6][i] = vz + (q.f[5][i] * tau + wz) / 2;
  }
}

// the end-of-step bookkeeping of a shared step: tstep = tau, time += tau,
// nstep += 1
__global__ void __launch_bounds__(kThreads)
    advance_clock_kernel(int n, double *tstep, double *time,
                         long long *nstep, double tau) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
    tstep[i] = tau;
    time[i] = time[i] + tau;
    nstep[i] = nstep[i] + 1;
  }
}

// NREG's end of step from the scalar block: tstep = the physical time
// advanced (T - T0), time += tau, nstep += 1
__global__ void __launch_bounds__(kThreads)
    nreg_step_end_kernel(int n, double *tstep, double *time,
                         long long *nstep, const double *sc, double tau) {
#pragma clang fp contract(off)
  const double dt = sc[TUPAN_NREG_T] - sc[TUPAN_NREG_T0];
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
    tstep[i] = dt;
    time[i] = time[i] + tau;
    nstep[i] = nstep[i] + 1;
  }
}

// x += d over six fields (r, v)
__global__ void __launch_bounds__(kThreads)
    apply_delta_kernel(int n, P18 p) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
#pragma unroll
    for (int k = 0; k < 6; ++k) p.f[k][i] = p.f[k][i] + p.f[6 + k][i];
  }
}

// dst[k][i] = src[k][i] / s for k < nk (correctly rounded division)
__global__ void __launch_bounds__(kThreads)
    divide_kernel(int n, int nk, P18 p, double s, const double *sp) {
#pragma clang fp contract(off)
  if (sp) s = *sp;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
    for (int k = 0; k < nk; ++k) p.f[9 + k][i] = p.f[k][i] / s;
  }
}

int check_fields(const void *const *f, int k, const char *what) {
  if (!f) return set_error(TUPAN_ERR_ARG, "NULL %s field array", what);
  for (int i = 0; i < k; ++i)
    if (!f[i]) return set_error(TUPAN_ERR_ARG, "NULL %s field %d", what, i);
  return TUPAN_OK;
}

int check_n(size_t n) {
  if (n > (size_t)INT_MAX)
    return set_error(TUPAN_ERR_ARG, "n too large (%zu)", n);
  return TUPAN_OK;
}

int finish(const char *what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? TUPAN_OK : hip_fail(e, what);
}

template <class K>
int launch_order(int order, K &&k) {
  switch (order) {
    case 2: k(std::integral_constant<int, 2>{}); break;
    case 4: k(std::integral_constant<int, 4>{}); break;
    case 6: k(std::integral_constant<int, 6>{}); break;
    case 8: k(std::integral_constant<int, 8>{}); break;
    default:
      return set_error(TUPAN_ERR_ARG, "Hermite order must be 2, 4, 6 or 8 "
                                      "(got %d)", order);
  }
  return TUPAN_OK;
}

}  // namespace
}  // namespace tupan

using namespace tupan;

extern "C" {

int tupan_hermite_predict_dev(int order, size_t n, double *const *x,
                              double *const *x0, double tau, void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  const int nf = 6 + 3 * (order / 2);
  if (order == 2 || order == 4 || order == 6 || order == 8) {
    if ((rc = check_fields((const void *const *)x, nf, "state"))) return rc;
    if ((rc = check_fields((const void *const *)x0, nf, "snapshot")))
      return rc;
  }
  if (n == 0) return launch_order(order, [](auto) {});
  P18 a{}, b{};
  for (int k = 0; k < nf && k < 18; ++k) {
    a.f[k] = x[k];
    b.f[k] = x0[k];
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  rc = launch_order(order, [&](auto o) {
    hipLaunchKernelGGL((hermite_predict_kernel<decltype(o)::value>),
                       dim3(grid_for(n)), dim3(kThreads), 0, st, (int)n, a, b,
                       tau);
  });
  return rc ? rc : finish("hermite_predict");
}

int tupan_hermite_correct_dev(int order, size_t n, double *const *x,
                              const double *const *x0, double tau,
                              void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  const int nf = 6 + 3 * (order / 2);
  if (order == 2 || order == 4 || order == 6 || order == 8) {
    if ((rc = check_fields((const void *const *)x, nf, "state"))) return rc;
    if ((rc = check_fields((const void *const *)x0, nf, "snapshot")))
      return rc;
  }
  if (n == 0) return launch_order(order, [](auto) {});
  P18 a{}, b{};
  for (int k = 0; k < nf && k < 18; ++k) {
    a.f[k] = x[k];
    b.f[k] = const_cast<double *>(x0[k]);
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  rc = launch_order(order, [&](auto o) {
    hipLaunchKernelGGL((hermite_correct_kernel<decltype(o)::value>),
                       dim3(grid_for(n)), dim3(kThreads), 0, st, (int)n, a, b,
                       tau);
  });
  return rc ? rc : finish("hermite_correct");
}

int tupan_drift_dev(size_t n, double *const *r, const double *const *v,
                    double tau, int half, double *const *pn_mr,
                    const double *const *pn_mv, void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if ((rc = check_fields((const void *const *)r, 3, "position"))) return rc;
  if ((rc = check_fields((const void *const *)v, 3, "velocity"))) return rc;
  const int pn = pn_mr != nullptr;
  if (pn) {
    if ((rc = check_fields((const void *const *)pn_mr, 3, "pn_mr"))) return rc;
    if ((rc = check_fields((const void *const *)pn_mv, 3, "pn_mv"))) return rc;
  }
  if (n == 0) return TUPAN_OK;
  P18 p{};
  for (int k = 0; k < 3; ++k) {
    p.f[k] = r[k];
    p.f[3 + k] = const_cast<double *>(v[k]);
    if (pn) {
      p.f[6 + k] = pn_mr[k];
      p.f[9 + k] = const_cast<double *>(pn_mv[k]);
    }
  }
  hipLaunchKernelGGL(drift_kernel, dim3(grid_for(n)), dim3(kThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), (int)n, p, tau,
                     half ? 1 : 0, pn);
  return finish("drift");
}

int tupan_kick_dev(size_t n, double *const *v, const double *const *a,
                   double tau, void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if ((rc = check_fields((const void *const *)v, 3, "velocity"))) return rc;
  if ((rc = check_fields((const void *const *)a, 3, "acceleration")))
    return rc;
  if (n == 0) return TUPAN_OK;
  P18 p{};
  for (int k = 0; k < 3; ++k) {
    p.f[k] = v[k];
    p.f[3 + k] = const_cast<double *>(a[k]);
  }
  hipLaunchKernelGGL(kick_kernel, dim3(grid_for(n)), dim3(kThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), (int)n, p, tau);
  return finish("kick");
}

int tupan_kick_pn_half_dev(size_t n, double *const *v, const double *const *a,
                           const double *const *w, double tau, void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if ((rc = check_fields((const void *const *)v, 3, "velocity"))) return rc;
  if ((rc = check_fields((const void *const *)a, 3, "acceleration")))
    return rc;
  if ((rc = check_fields((const void *const *)w, 3, "w"))) return rc;
  if (n == 0) return TUPAN_OK;
  P18 p{};
  for (int k = 0; k < 3; ++k) {
    p.f[k] = v[k];
    p.f[3 + k] = const_cast<double *>(a[k]);
    p.f[6 + k] = const_cast<double *>(w[k]);
  }
  hipLaunchKernelGGL(kick_pn_half_kernel, dim3(grid_for(n)), dim3(kThreads),
                     0, reinterpret_cast<hipStream_t>(stream), (int)n, p, tau);
  return finish("kick_pn_half");
}

int tupan_pn_flux_dev(size_t n, const double *const *in, double *const *acc,
                      double tau, void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if ((rc = check_fields((const void *const *)in, 10, "input"))) return rc;
  if ((rc = check_fields((const void *const *)acc, 10, "accumulator")))
    return rc;
  if (n == 0) return TUPAN_OK;
  P18 p{}, q{};
  for (int k = 0; k < 10; ++k) p.f[k] = const_cast<double *>(in[k]);
  for (int k = 0; k < 7; ++k) p.f[10 + k] = acc[k];
  for (int k = 0; k < 3; ++k) q.f[k] = acc[7 + k];
  hipLaunchKernelGGL(pn_flux_kernel, dim3(grid_for(n)), dim3(kThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), (int)n, p, q, tau);
  return finish("pn_flux");
}

int tupan_pn_flux_kick_dev(size_t n, double *const *in, double *const *acc,
                           const double *const *a, double tau, void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if ((rc = check_fields((const void *const *)in, 10, "input"))) return rc;
  if ((rc = check_fields((const void *const *)acc, 10, "accumulator")))
    return rc;
  if ((rc = check_fields((const void *const *)a, 3, "acceleration")))
    return rc;
  if (n == 0) return TUPAN_OK;
  P18 p{}, q{};
  for (int k = 0; k < 10; ++k) p.f[k] = in[k];
  for (int k = 0; k < 7; ++k) p.f[10 + k] = acc[k];
  for (int k = 0; k < 3; ++k) {
    q.f[k] = acc[7 + k];
    q.f[3 + k] = const_cast<double *>(a[k]);
  }
  hipLaunchKernelGGL(pn_flux_kick_kernel, dim3(grid_for(n)), dim3(kThreads),
                     0, reinterpret_cast<hipStream_t>(stream), (int)n, p, q,
                     tau);
  return finish("pn_flux_kick");
}

int tupan_advance_clock_dev(size_t n, double *tstep, double *time,
                            long long *nstep, double tau, void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if (!tstep || !time || !nstep)
    return set_error(TUPAN_ERR_ARG, "NULL clock field");
  if (n == 0) return TUPAN_OK;
  hipLaunchKernelGGL(advance_clock_kernel, dim3(grid_for(n)), dim3(kThreads),
                     0, reinterpret_cast<hipStream_t>(stream), (int)n, tstep,
                     time, nstep, tau);
  return finish("advance_clock");
}

int tupan_nreg_step_end_dev(size_t n, double *tstep, double *time,
                            long long *nstep, const double *sc, double tau,
                            void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if (!tstep || !time || !nstep || !sc)
    return set_error(TUPAN_ERR_ARG, "NULL clock field or scalar block");
  if (n == 0) return TUPAN_OK;
  hipLaunchKernelGGL(nreg_step_end_kernel, dim3(grid_for(n)), dim3(kThreads),
                     0, reinterpret_cast<hipStream_t>(stream), (int)n, tstep,
                     time, nstep, sc, tau);
  return finish("nreg_step_end");
}

int tupan_apply_delta_dev(size_t n, double *const *x, const double *const *d,
                          void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if ((rc = check_fields((const void *const *)x, 6, "state"))) return rc;
  if ((rc = check_fields((const void *const *)d, 6, "delta"))) return rc;
  if (n == 0) return TUPAN_OK;
  P18 p{};
  for (int k = 0; k < 6; ++k) {
    p.f[k] = x[k];
    p.f[6 + k] = const_cast<double *>(d[k]);
  }
  hipLaunchKernelGGL(apply_delta_kernel, dim3(grid_for(n)), dim3(kThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), (int)n, p);
  return finish("apply_delta");
}

static int divide(size_t n, int k, const double *const *src,
                  double *const *dst, double s, const double *sp,
                  void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if (k < 1 || k > 9) return set_error(TUPAN_ERR_ARG, "k must be in 1..9");
  if ((rc = check_fields((const void *const *)src, k, "source"))) return rc;
  if ((rc = check_fields((const void *const *)dst, k, "destination")))
    return rc;
  if (n == 0) return TUPAN_OK;
  P18 p{};
  for (int c = 0; c < k; ++c) {
    p.f[c] = const_cast<double *>(src[c]);
    p.f[9 + c] = dst[c];
  }
  hipLaunchKernelGGL(divide_kernel, dim3(grid_for(n)), dim3(kThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), (int)n, k, p, s,
                     sp);
  return finish("divide");
}

// out = r + (v * tau) / 2: drift_kernel's half drift, written elsewhere
__global__ __launch_bounds__(kThreads) void half_drift_to_kernel(int n, P18 p,
                                                                double tau) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double dr = p.f[3 + d][i] * tau;
      dr = dr / 2;
      p.f[6 + d][i] = p.f[d][i] + dr;
    }
  }
}

int tupan_half_drift_to_dev(size_t n, const double *const *r,
                            const double *const *v, double tau,
                            double *const *out, void *stream) {
  int rc;
  if ((rc = check_n(n))) return rc;
  if ((rc = check_fields((const void *const *)r, 3, "position"))) return rc;
  if ((rc = check_fields((const void *const *)v, 3, "velocity"))) return rc;
  if ((rc = check_fields((const void *const *)out, 3, "output"))) return rc;
  if (n == 0) return TUPAN_OK;
  P18 p{};
  for (int d = 0; d < 3; ++d) {
    p.f[d] = const_cast<double *>(r[d]);
    p.f[3 + d] = const_cast<double *>(v[d]);
    p.f[6 + d] = out[d];
  }
  hipLaunchKernelGGL(half_drift_to_kernel, dim3(grid_for(n)), dim3(kThreads),
                     0, reinterpret_cast<hipStream_t>(stream), (int)n, p, tau);
  return finish("half_drift_to");
}

int tupan_divide_dev(size_t n, int k, const double *const *src,
                     double *const *dst, double s, void *stream) {
  return divide(n, k, src, dst, s, nullptr, stream);
}

int tupan_divide_by_dev(size_t n, int k, const double *const *src,
                        double *const *dst, const double *d_s, void *stream) {
  if (!d_s) return set_error(TUPAN_ERR_ARG, "NULL device divisor");
  return divide(n, k, src, dst, 0.0, d_s, stream);
}

}  // extern "C"
