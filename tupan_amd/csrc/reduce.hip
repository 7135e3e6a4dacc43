// reduce.hip -- deterministic sums over particles, FP64.
//
//   tupan_sum_dev          sum of one field (the integrators' ps.x.sum(),
//                          e.g. tupan/integrator/nreg.py:26, 31, 57)
//   tupan_diagnostics_dev  the twelve sums behind the energy/momentum report
//                          (tupan/simulation.py:96-129), one pass over the
//                          particles with the per-particle expressions of
//                          tupan/particles/body.py
//   tupan_sum_pairwise_dev the sum in numpy's order for a contiguous float64
//                          array (what the reference's ps.x.sum() computes):
//                          0 + the pairwise sums of consecutive 8192-element
//                          chunks, each split in halves rounded to multiples
//                          of 8 down to <= 128 elements, which are summed
//                          with 8 interleaved accumulators
//   tupan_energy_terms_dev the per-particle 0.5 m v^2 and m phi of body.py
//                          (:261-275, :294-300), for sums in numpy's order
//   tupan_report_terms_dev the report's twelve per-particle quantities
//                          (those of tupan_diagnostics_dev), stored for sums
//                          in numpy's order
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
  int rc = scratch_alloc(kSlotReduce, (void **)&part, sizeof(double) * K * nb,
                         st);
  if (rc) return rc;
  hipLaunchKernelGGL((reduce_partial<Fn>), dim3(nb), dim3(kRThreads), 0, st,
                     p, (int)n, part);
  hipLaunchKernelGGL((reduce_final<K>), dim3(1), dim3(kRThreads), 0, st,
                     (const double *)part, nb, out);
  hipError_t e = hipGetLastError();
  scratch_free(kSlotReduce, part, st);
  if (e != hipSuccess) return hip_fail(e, "reduction launch");
  return TUPAN_OK;
}

// ---- numpy's summation order -------------------------------------------
constexpr int kNpChunk = 8192;  // numpy's reduction buffer (elements)
constexpr int kNpLeaf = 128;    // pairwise-sum block size

// a block of n <= 128 elements (n < 8: one running sum from 0)
__device__ double np_leaf(const double *a, int n) {
#pragma clang fp contract(off)
  if (n < 8) {
    double res = 0.0;
    for (int i = 0; i < n; ++i) res = res + a[i];
    return res;
  }
  double r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = a[j];
  int i = 8;
  const int lim = n - (n % 8);
  for (; i < lim; i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = r[j] + a[i + j];
  }
  double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res = res + a[i];
  return res;
}

// halves rounded down to a multiple of 8; n <= 8192 needs 7 levels
template <int D>
__device__ __noinline__ double np_pairwise(const double *a, int n) {
  if (n <= kNpLeaf || D == 0) return np_leaf(a, n);
  int n2 = n / 2;
  n2 -= n2 % 8;
  return np_pairwise<(D > 0 ? D - 1 : 0)>(a, n2) +
         np_pairwise<(D > 0 ? D - 1 : 0)>(a + n2, n - n2);
}

// One wave per chunk, staged in LDS.  A full chunk splits into 64 leaves of
// 128 that the recursion combines as a perfect binary tree: one leaf per
// lane, then the same tree level by level.  A partial (last) chunk follows
// the recursion on one lane.
__global__ __launch_bounds__(64) void np_chunk_kernel(const double *x, int n,
                                                     double *part) {
  __shared__ double buf[kNpChunk];
  const int lane = threadIdx.x;
  const int base = blockIdx.x * kNpChunk;
  const int m = min(kNpChunk, n - base);
  for (int i = lane; i < m; i += 64) buf[i] = x[base + i];
  __syncthreads();
  if (m == kNpChunk) {
    const double v = np_leaf(buf + kNpLeaf * lane, kNpLeaf);
    __syncthreads();
    buf[lane] = v;
    __syncthreads();
    for (int s = 1; s < 64; s *= 2) {
      if (lane % (2 * s) == 0) buf[lane] = buf[lane] + buf[lane + s];
      __syncthreads();
    }
    if (lane == 0) part[blockIdx.x] = buf[0];
  } else if (lane == 0) {
    part[blockIdx.x] = np_pairwise<10>(buf, m);
  }
}

__global__ void np_final_kernel(const double *part, int nchunks, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double r = 0.0;
  for (int c = 0; c < nchunks; ++c) r = r + part[c];
  out[0] = r;
}

// The report's per-particle quantities (DiagFn's, stored instead of
// summed) for totals in numpy's order.
template <bool PN>
__global__ __launch_bounds__(256) void report_terms_kernel(int n, Fields p,
                                                          Fields out) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n;
       i += gridDim.x * 256) {
    double a[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) a[k] = 0.0;
    DiagFn<PN>::add(p, i, a);
#pragma unroll
    for (int k = 0; k < 12; ++k) const_cast<double *>(out.f[k])[i] = a[k];
  }
}

__global__ __launch_bounds__(256) void energy_terms_kernel(
    int n, const double *m, const double *vx, const double *vy,
    const double *vz, const double *phi, double *ke, double *mphi) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n;
       i += gridDim.x * 256) {
    ke[i] = 0.5 * m[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    mphi[i] = m[i] * phi[i];
  }
}

// Lane-parallel scan of up to 8 arrays for NaN/inf; a block that finds one
// ORs the flag into the status word from one lane (vector atomic).
struct FinitePtrs {
  const double *a[8];
};
__global__ __launch_bounds__(256) void check_finite_kernel(
    int n, int na, FinitePtrs p, int *status) {
  bool bad = false;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    for (int k = 0; k < na; ++k) bad = bad || !isfinite(p.a[k][i]);
  if (__syncthreads_or(bad) && threadIdx.x == 0)
    atomicOr(status, TUPAN_STATUS_NONFINITE);
}

}  // namespace
}  // namespace tupan

using namespace tupan;

extern "C" {

int tupan_check_finite_dev(size_t n, int narrays, const double *const *arrays,
                           int *d_status, void *stream) {
  if (narrays < 0 || narrays > 8)
    return set_error(TUPAN_ERR_ARG, "narrays must be in 0..8 (got %d)",
                     narrays);
  if (!d_status) return set_error(TUPAN_ERR_ARG, "NULL status");
  if (n > (size_t)INT_MAX) return set_error(TUPAN_ERR_ARG, "n too large");
  if (n == 0 || narrays == 0) return TUPAN_OK;
  if (!arrays) return set_error(TUPAN_ERR_ARG, "NULL array list");
  FinitePtrs p{};
  for (int k = 0; k < narrays; ++k) {
    if (!arrays[k]) return set_error(TUPAN_ERR_ARG, "NULL array %d", k);
    p.a[k] = arrays[k];
  }
  const int blocks = (int)std::min<size_t>(256, (n + 255) / 256);
  hipLaunchKernelGGL(check_finite_kernel, dim3(blocks), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), (int)n, narrays, p,
                     d_status);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? TUPAN_OK : hip_fail(e, "check_finite_kernel");
}

int tupan_sum_dev(size_t n, const double *x, double *out, void *stream) {
  if (n && !x) return set_error(TUPAN_ERR_ARG, "NULL field");
  Fields p{};
  p.f[0] = x;
  return run_reduce<SumFn>(p, n, out, reinterpret_cast<hipStream_t>(stream));
}

int tupan_sum_pairwise_dev(size_t n, const double *x, double *out,
                           void *stream) {
  if (n > (size_t)INT_MAX)
    return set_error(TUPAN_ERR_ARG, "n too large (%zu)", n);
  if ((n && !x) || !out) return set_error(TUPAN_ERR_ARG, "NULL pointer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int nchunks = (int)((n + kNpChunk - 1) / kNpChunk);
  double *part = nullptr;
  if (nchunks > 0) {
    int rc = scratch_alloc(kSlotPairwise, (void **)&part,
                           sizeof(double) * nchunks, st);
    if (rc) return rc;
    hipLaunchKernelGGL(np_chunk_kernel, dim3(nchunks), dim3(64), 0, st, x,
                       (int)n, part);
  }
  hipLaunchKernelGGL(np_final_kernel, dim3(1), dim3(1), 0, st,
                     (const double *)part, nchunks, out);
  hipError_t e = hipGetLastError();
  scratch_free(kSlotPairwise, part, st);
  if (e != hipSuccess) return hip_fail(e, "pairwise sum launch");
  return TUPAN_OK;
}

int tupan_energy_terms_dev(size_t n, const double *const *f, double *ke,
                           double *mphi, void *stream) {
  if (n > (size_t)INT_MAX)
    return set_error(TUPAN_ERR_ARG, "n too large (%zu)", n);
  if (!f || !ke || !mphi) return set_error(TUPAN_ERR_ARG, "NULL pointer");
  for (int k = 0; k < 5; ++k)
    if (n && !f[k]) return set_error(TUPAN_ERR_ARG, "NULL field %d", k);
  if (n == 0) return TUPAN_OK;
  const size_t g0 = (n + 255) / 256;
  const int g = g0 > 65535 * 8 ? 65535 * 8 : (int)g0;
  hipLaunchKernelGGL(energy_terms_kernel, dim3(g), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), (int)n, f[0],
                     f[1], f[2], f[3], f[4], ke, mphi);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? TUPAN_OK : hip_fail(e, "energy terms launch");
}

int tupan_report_terms_dev(size_t n, const double *const *f, int pn,
                           double *const *out, void *stream) {
  if (n > (size_t)INT_MAX)
    return set_error(TUPAN_ERR_ARG, "n too large (%zu)", n);
  if (!f || !out) return set_error(TUPAN_ERR_ARG, "NULL field array");
  const int nf = pn ? 18 : 8;
  Fields p{}, o{};
  for (int k = 0; k < nf; ++k) {
    if (n && !f[k]) return set_error(TUPAN_ERR_ARG, "NULL field %d", k);
    p.f[k] = f[k];
  }
  for (int k = 0; k < 12; ++k) {
    if (n && !out[k]) return set_error(TUPAN_ERR_ARG, "NULL output %d", k);
    o.f[k] = out[k];
  }
  if (n == 0) return TUPAN_OK;
  const size_t g0 = (n + 255) / 256;
  const int g = g0 > 65535 * 8 ? 65535 * 8 : (int)g0;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (pn)
    hipLaunchKernelGGL(report_terms_kernel<true>, dim3(g), dim3(256), 0, st,
                       (int)n, p, o);
  else
    hipLaunchKernelGGL(report_terms_kernel<false>, dim3(g), dim3(256), 0, st,
                       (int)n, p, o);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? TUPAN_OK : hip_fail(e, "report terms launch");
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
