// ieee_engine.hpp -- the engine of the reference-arithmetic modes
// (TUPAN_MODE_FP64_IEEE, and TUPAN_MODE_FP32_IEEE: the reference's FP32
// build), shared by ieee.hip, pnacc.hip and nreg.hip.
//
// The reference sums every j into one accumulator per i in ascending
// order, which leaves one independent chain per (i, output): at N = 65536
// the chip holds one wave per SIMD.  The engine therefore feeds that wave
// without waits: the block's four waves share j tiles of 64 packed records
// that are double-buffered through LDS (the next tile's global loads are in
// flight while the current one is summed), and each wave computes the pair
// terms of U consecutive j's independently -- their division and
// square-root chains overlap -- before folding them into the accumulators
// in j order.
#pragma once
#include "launch.hpp"

namespace tupan {
namespace ieee {

constexpr int kIeeeTile = 64;  // j records per LDS tile

// A policy supplies Real (the arithmetic type: double, or float for the
// FP32 build), JRec (a multiple of 16 bytes, read from LDS as 16-byte
// chunks), I, pack, load_i and terms(q, r, t, prm): the kTerms per-pair
// terms t; IeeeBase's fold and finish are the reference's sum -= t and
// out = sum (a policy with another accumulation overrides them).  Inputs
// and outputs stay FP64 arrays: the FP32 policies round each input to
// float when packing (exact for values that came from float arrays) and
// widen their float results exactly.
template <int N, class T = double>
struct IeeeBase {
  using Real = T;
  static constexpr bool kTwoLevel = false;
  static constexpr int kOut = N, kTerms = N, kAcc = N;
  struct Params {};
  __device__ static void init(T *acc) {
#pragma unroll
    for (int c = 0; c < N; ++c) acc[c] = 0;
  }
  __device__ static void fold(T *acc, const T *t) {
#pragma clang fp contract(off)
#pragma unroll
    for (int c = 0; c < N; ++c) acc[c] = acc[c] - t[c];
  }
  template <class Prm>
  __device__ static void finish(const T *acc, Out out, int i, const Prm &) {
#pragma unroll
    for (int c = 0; c < N; ++c) out.f[c][i] = (double)acc[c];
  }
};

// smoothing.h:382-390 / :405-414 -- 1/(r^2+e2), zero unless r^2 > 0; for
// float a correctly rounded float division (as the reference's FP32 build)
template <class T>
__device__ __forceinline__ T ieee_inv_r2(T r2, T e2) {
#pragma clang fp contract(off)
  T inv = 1 / (r2 + e2);
  return r2 > 0 ? inv : (T)0;
}

// One wave per 64 i-particles (one per lane); blocks of 4 waves share the
// j tiles.  Every thread of the block takes part in every tile copy and
// barrier; waves past ni skip only the sums.
template <class Pol, int U>
__global__ __launch_bounds__(256) void ieee_engine(
    Soa is, int ni, const typename Pol::JRec *__restrict__ jr, int nj,
    Out out, typename Pol::Params prm) {
  using JRec = typename Pol::JRec;
  static_assert(sizeof(JRec) % 16 == 0, "records are copied in 16 B chunks");
  constexpr int kRecChunks = (int)(sizeof(JRec) / 16);
  constexpr int kPer = (kIeeeTile * kRecChunks + 255) / 256;
  __shared__ __align__(16) JRec tile[2][kIeeeTile];

  prepare_params<Pol>(prm);
  const int lane = threadIdx.x & 63;
  const int i0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
  const bool active = i0 < ni;  // wave-uniform
  const int i = min(i0 + lane, ni - 1);
  using R = typename Pol::Real;
  const typename Pol::I q = Pol::load_i(is, i);
  R acc[Pol::kAcc];
  Pol::init(acc);

  const int ntiles = (nj + kIeeeTile - 1) / kIeeeTile;
  double2 buf[kPer];
  auto fetch = [&](int t) {
    const int nch = min(kIeeeTile, nj - t * kIeeeTile) * kRecChunks;
    const double2 *src =
        reinterpret_cast<const double2 *>(jr + (size_t)t * kIeeeTile);
#pragma unroll
    for (int p = 0; p < kPer; ++p) {
      const int c = (int)threadIdx.x + p * 256;
      if (c < nch) buf[p] = src[c];
    }
  };
  auto store = [&](int b, int t) {
    const int nch = min(kIeeeTile, nj - t * kIeeeTile) * kRecChunks;
    double2 *dst = reinterpret_cast<double2 *>(&tile[b][0]);
#pragma unroll
    for (int p = 0; p < kPer; ++p) {
      const int c = (int)threadIdx.x + p * 256;
      if (c < nch) dst[c] = buf[p];
    }
  };

  if (ntiles > 0) {
    fetch(0);
    store(0, 0);
  }
  __syncthreads();
  for (int t = 0; t < ntiles; ++t) {
    if (t + 1 < ntiles) fetch(t + 1);
    if (active) {
      const JRec *L = tile[t & 1];
      const int nrec = min(kIeeeTile, nj - t * kIeeeTile);
      int k = 0;
      for (; k + U <= nrec; k += U) {
        R tt[U][Pol::kTerms];
#pragma unroll
        for (int u = 0; u < U; ++u) Pol::terms(q, L[k + u], tt[u], prm);
#pragma unroll
        for (int u = 0; u < U; ++u) Pol::fold(acc, tt[u]);
      }
      for (; k < nrec; ++k) {
        R t1[Pol::kTerms];
        Pol::terms(q, L[k], t1, prm);
        Pol::fold(acc, t1);
      }
    }
    if (t + 1 < ntiles) store((t + 1) & 1, t + 1);
    __syncthreads();
  }
  if (active && i0 + lane < ni) Pol::finish(acc, out, i0 + lane, prm);
}

template <class Pol, int U>
int run_ieee_engine(size_t ni, const double *const *ip, size_t nj,
                    const double *const *jp, double *const *op,
                    hipStream_t st, typename Pol::Params prm = {}) {
  if (ni == 0) return TUPAN_OK;
  if (!ip || !jp || !op) return set_error(TUPAN_ERR_ARG, "NULL field array");
  if (ni > (size_t)INT_MAX / 8 || nj > (size_t)INT_MAX / 8)
    return set_error(TUPAN_ERR_ARG, "ni/nj too large (%zu, %zu)", ni, nj);
  Soa is{}, js{};
  for (int k = 0; k < Pol::kFields; ++k) {
    if (!ip[k] || (nj && !jp[k]))
      return set_error(TUPAN_ERR_ARG, "NULL input field %d", k);
    is.f[k] = ip[k];
    js.f[k] = jp[k];
  }
  Out out{};
  for (int k = 0; k < Pol::kOut; ++k) {
    if (!op[k]) return set_error(TUPAN_ERR_ARG, "NULL output field %d", k);
    out.f[k] = op[k];
  }
  const int n_i = (int)ni, n_j = (int)nj;
  typename Pol::JRec *jr = nullptr;
  if (n_j > 0) {
    const size_t bytes =
        ((sizeof(typename Pol::JRec) * (size_t)n_j + 255) / 256) * 256;
    const int rc = scratch_alloc(kSlotIeee, (void **)&jr, bytes, st);
    if (rc) return rc;
    hipLaunchKernelGGL((pack_kernel<Pol>), dim3((n_j + 255) / 256),
                       dim3(256), 0, st, js, n_j, jr, (const double *)nullptr);
  }
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  const bool timed = timing_begin(&ev0, &ev1, st);
  if (timed) (void)hipEventRecord(ev0, st);
  hipLaunchKernelGGL((ieee_engine<Pol, U>), dim3((n_i + 255) / 256),
                     dim3(256), 0, st, is, n_i,
                     (const typename Pol::JRec *)jr, n_j, out, prm);
  if (timed) (void)hipEventRecord(ev1, st);
  hipError_t e = hipGetLastError();
  scratch_free(kSlotIeee, jr, st);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return TUPAN_OK;
}

}  // namespace ieee
}  // namespace tupan
