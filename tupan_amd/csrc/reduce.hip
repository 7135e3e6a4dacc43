// reduce.hip -- deterministic sums over particles, FP64.
//
//   tupan_sum_dev          sum of one field (the integrators' ps.x.sum(),
//                          e.g. tupan/integrator/nreg.py:26, 31, 57)
//   tupan_diagnostics_dev  the twelve sums behind the energy/momentum report
//                          (tupan/simulation.py:96-129), one pass over the
//                          particles with the per-particle expressions of
//                          tupan/particles/body.py
//
// Two launches: B <= 256 blocks accumulate grid-strided per-thread sums and
// reduce them through LDS in a fixed tree; one block adds the B partials
// the same way.  B depends on n only, so a sum is bit-reproducible run to
// run (it is not numpy's pairwise order: results agree with the host to
// round-off, not bit for bit).
#include "launch.hpp"

namespace tupan {
namespace {

constexpr int kRThreads = 256;
constexpr int kMaxPartials = 256;

struct Fields {
  const double *f[18];
};

struct SumFn {
  static constexpr int K = 1;
  __device__ static void add(const Fields &p, int i, double *a) {
    a[0] += p.f[0][i];
  }
};

// Per-particle report quantities, body.py: m (:79-83); ke = 0.5 m v^2
// [+ pn_ke] (:261-275); m phi (:294-300, halved on the host); m r [+ pn_mr]
// (:86-108); p = m v [+ pn_mv] (:66-76, :180-200); r x p [+ pn_am]
// (:219-242).  Fields: 0 m, 1-3 r, 4-6 v, 7 phi; with PN 8 pn_ke, 9-11
// pn_mr, 12-14 pn_mv, 15-17 pn_am.
template <bool PN>
struct DiagFn {
  static constexpr int K = 12;
  __device__ static void add(const Fields &p, int i, double *a) {
#pragma clang fp contract(off)
    const double m = p.f[0][i];
    const double rx = p.f[1][i], ry = p.f[2][i], rz = p.f[3][i];
    const double vx = p.f[4][i], vy = p.f[5][i], vz = p.f[6][i];
    double ke = 0.5 * m * (vx * vx + vy * vy + vz * vz);
    const double px = m * vx, py = m * vy, pz = m * vz;
    double mrx = m * rx, mry = m * ry, mrz = m * rz;
    double lmx = px, lmy = py, lmz = pz;
    double amx = (ry * pz) - (rz * py);
    double amy = (rz * px) - (rx * pz);
    double amz = (rx * py) - (ry * px);
    if (PN) {
      ke += p.f[8][i];
      mrx += p.f[9][i];
      mry += p.f[10][i];
      mrz += p.f[11][i];
      lmx += p.f[12][i];
      lmy += p.f[13][i];
      lmz += p.f[14][i];
      amx += p.f[15][i];
      amy += p.f[16][i];
      amz += p.f[17][i];
    }
    a[0] += m;
    a[1] += ke;
    a[2] += m * p.f[7][i];
    a[3] += mrx;
    a[4] += mry;
    a[5] += mrz;
    a[6] += lmx;
    a[7] += lmy;
    a[8] += lmz;
    a[9] += amx;
    a[10] += amy;
    a[11] += amz;
  }
};

template <int K>
__device__ void block_tree(double (*sh)[kRThreads]) {
  for (int s = kRThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k < K; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    }
    __syncthreads();
  }
}

template <class Fn>
__global__ __launch_bounds__(kRThreads) void reduce_partial(Fields p, int n,
                                                            double *part) {
  constexpr int K = Fn::K;
  __shared__ double sh[K][kRThreads];
  double a[K];
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = 0.0;
  for (int i = blockIdx.x * kRThreads + threadIdx.x; i < n;
       i += gridDim.x * kRThreads)
    Fn::add(p, i, a);
#pragma unroll
  for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = a[k];
  __syncthreads();
  block_tree<K>(sh);
  if ((int)threadIdx.x < K)
    part[blockIdx.x * K + threadIdx.x] = sh[threadIdx.x][0];
}

template <int K>
__global__ __launch_bounds__(kRThreads) void reduce_final(
    const double *__restrict__ part, int nb, double *__restrict__ out) {
  __shared__ double sh[K][kRThreads];
#pragma unroll
  for (int k = 0; k < K; ++k)
    sh[k][threadIdx.x] =
        (int)threadIdx.x < nb ? part[threadIdx.x * K + k] : 0.0;
  __syncthreads();
  block_tree<K>(sh);
  if ((int)threadIdx.x < K) out[threadIdx.x] = sh[threadIdx.x][0];
}

template <class Fn>
int run_reduce(const Fields &p, size_t n, double *out, hipStream_t st) {
  constexpr int K = Fn::K;
  if (n > (size_t)INT_MAX)
    return set_error(TUPAN_ERR_ARG, "n too large (%zu)", n);
  if (!out) return set_error(TUPAN_ERR_ARG, "NULL output");
  const size_t nb0 = (n + kRThreads - 1) / kRThreads;
  const int nb = nb0 < 1 ? 1 : (nb0 > kMaxPartials ? kMaxPartials : (int)nb0);
  double *part = nullptr;
  int rc = device_alloc((void **)&part, sizeof(double) * K * nb, st);
  if (rc) return rc;
  hipLaunchKernelGGL((reduce_partial<Fn>), dim3(nb), dim3(kRThreads), 0, st,
                     p, (int)n, part);
  hipLaunchKernelGGL((reduce_final<K>), dim3(1), dim3(kRThreads), 0, st,
                     (const double *)part, nb, out);
  hipError_t e = hipGetLastError();
  device_free(part, st);
  if (e != hipSuccess) return hip_fail(e, "reduction launch");
  return TUPAN_OK;
}

}  // namespace
}  // namespace tupan

using namespace tupan;

extern "C" {

int tupan_sum_dev(size_t n, const double *x, double *out, void *stream) {
  if (n && !x) return set_error(TUPAN_ERR_ARG, "NULL field");
  Fields p{};
  p.f[0] = x;
  return run_reduce<SumFn>(p, n, out, reinterpret_cast<hipStream_t>(stream));
}

int tupan_diagnostics_dev(size_t n, const double *const *f, int pn,
                          double *out, void *stream) {
  if (!f) return set_error(TUPAN_ERR_ARG, "NULL field array");
  const int nf = pn ? 18 : 8;
  Fields p{};
  for (int k = 0; k < nf; ++k) {
    if (n && !f[k]) return set_error(TUPAN_ERR_ARG, "NULL field %d", k);
    p.f[k] = f[k];
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  return pn ? run_reduce<DiagFn<true>>(p, n, out, st)
            : run_reduce<DiagFn<false>>(p, n, out, st);
}

}  // extern "C"
