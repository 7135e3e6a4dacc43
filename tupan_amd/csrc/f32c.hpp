// f32c.hpp -- common shape of the FP32C ("FP32-compensated") policies.
//
//   * pair separations are formed in FP64 and rounded once to FP32
//     (dsub32); velocity differences are formed from FP32 offsets to a
//     per-tile FP64 origin (policies with kVelOrigins): the i-side offset is
//     rounded once per tile, so its rounding is shared by 32 pairs only and
//     averages out over the tiles instead of adding up over all j;
//   * the rest of the pair arithmetic is FP32 with v_rsq_f32, whose mean
//     bias the single-system policies fold into their stored masses and
//     velocity offsets (kRsqBias), or refine by a Newton step (rsq32n);
//   * sums run in FP32 over a tile of kTile j's and are folded into FP64
//     accumulators after every tile;
//   * near field: a pair whose softened s^2 is below the j record's
//     threshold contributes nothing in FP32 and is re-evaluated by the FP64
//     policy from the caller's FP64 arrays.
#pragma once
#include "engine.hpp"

namespace tupan {

// v_rsq_f32 on MI355X is within 1 ulp but biased low: over every float in
// [1, 4) (which fixes every argument: rsq(4x) = rsq(x)/2), 9.8% of results
// are one ulp below the correctly rounded value and 1.3% one ulp above,
// a mean relative error of -6.98e-9 (log-uniform arguments; -7.39e-9
// uniform) and an RMS of 2.6e-8 (2.4e-8 when correctly rounded).  That
// bias enters y^3 of every pair the same way, and it is what limited the
// FP32 accelerations to ~1e-6.  The single-system policies remove it for
// free: the stored masses carry (1 + 3 b) (k = m y^3) and the velocity
// offsets (1 + 2 b) (alpha = 3 (r.v) y^2), rounded once to FP32 with them.
constexpr double kRsqBias = 6.98e-9;
constexpr double kMassY3 = 1.0 + 3.0 * kRsqBias;  // masses multiplying y^3
constexpr double kMassY1 = 1.0 + kRsqBias;        // masses multiplying y
constexpr double kVelY2 = 1.0 + 2.0 * kRsqBias;   // velocity offsets

// 1/sqrt(x) in FP32 to about 1/2 ulp: v_rsq_f32 refined by one Newton step
// (SnapCrackle, whose pair terms take y^3..y^9 through the recurrences),
//   e = 1 - x y^2  (an FMA: the residual keeps its full precision),
//   y' = y + (y / 2) e.
// x = 0 gives NaN (0 * inf): callers select 0 for masked pairs afterwards.
__device__ __forceinline__ float rsq32n(float x) {
  const float y = rsq32(x);
  const float e = __builtin_fmaf(-(x * y), y, 1.0f);
  return __builtin_fmaf(0.5f * y, e, y);
}

// Threshold on the FP32 s^2 = r^2 + e2_i + e2_j below which a pair is near:
// (near2_j + e2_j + e2max) widened by 2^-20, so that every pair with r^2 = 0
// (s^2 = e2_i + e2_j rounded) is near whatever the rounding, and a pair
// with r^2 < near2_j always is.  near2_j = max(m_j * factor, FLT_MIN).
__device__ __forceinline__ float near_threshold(double m, double e2j,
                                                const PackCtx &c) {
  const double n2 = fmax(m * c.factor, 1.1754943508222875e-38);
  return (float)((n2 + e2j + c.e2max) * (1.0 + 0x1p-20));
}

// Common shape of the FP32C policies: P64 is the FP64 policy used for the
// near field; the derived policy supplies JRec (with a float `near2`
// threshold), I (with float `e2` and int `i`), pack, load_i and body32.
template <class Derived, class P64, int N>
struct F32cBase : SumOut<N> {
  static constexpr bool kTwoLevel = true;
  static constexpr int kVelOrigins = 0;
  // factor of the stored per-tile velocity origins (midpoints times it)
  static constexpr double kOrgScale = 1.0;
  using Exact = P64;  // the FP64 policy (near pairs, recomputed rows)
  using Sorted = Derived;  // the policy of the sorted engine's calls
  // minimum waves per SIMD the sorted engine is compiled for (8: Phi and
  // Acc 6-7% faster than at 7, AccJerk the same, no spill)
  static constexpr int kWavesPerEu = 8;
  using Params = NoParams;
  using A64 = F64Acc<N>;
  using A32 = F32Acc<N>;
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void zero(A32 &a) { zero_acc(a); }
  __device__ static void fold(const A32 &a, A64 &b) { fold_acc(a, b); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }
  template <class I>
  __device__ static void tile_begin(I &, const double *, int) {}
  // FP64 fields of I that the sorted engine keeps in LDS during its tile
  // loop (none by default)
  static constexpr int kI64 = 0;
  template <class I>
  __device__ static void save64(const I &, double *) {}
  template <class I>
  __device__ static void load64(I &, const double *) {}

  // Error estimate of the FP32 tile sums (used by the sorted engine): per
  // output group -- a 3-vector, or the one scalar output -- the sum over
  // tiles of |tile partial|^2, accumulated in FP32 after every tile.  The
  // FP32 error of a row is a random walk whose steps scale with the tile
  // partials, so sqrt(stat) / |result| predicts it.
  static constexpr int kGroupLen = N >= 3 ? 3 : 1;
  static constexpr int kGroups = N / kGroupLen;
  static constexpr int kPart = N + kGroups;  // partials: outputs, then stats
  __device__ static void stat(const A32 &a, float *g) {
#pragma unroll
    for (int q = 0; q < kGroups; ++q)
#pragma unroll
      for (int k = 0; k < kGroupLen; ++k)
        g[q] = __builtin_fmaf(a.v[q * kGroupLen + k], a.v[q * kGroupLen + k],
                              g[q]);
  }

  // Separation, softened s^2 and the near-field test.  s^2 uses explicit
  // FMAs so that pair32 and near_pair take bit-identical decisions; a NaN
  // separation is near too (the FP64 path propagates it).
  template <class I, class JRec>
  __device__ static bool near_test(const I &q, const JRec &r, float &rx,
                                   float &ry, float &rz, float &s2) {
    rx = dsub32(q.x, r.x);
    ry = dsub32(q.y, r.y);
    rz = dsub32(q.z, r.z);
    s2 = __builtin_fmaf(
        rz, rz, __builtin_fmaf(ry, ry, __builtin_fmaf(rx, rx, q.e2 + r.e2)));
    return !(s2 >= r.near2);
  }
  // FP32 contribution of a far pair; near pairs contribute zero here.
  // Returns the wave's mask of far lanes.
  template <class I, class JRec>
  __device__ static uint64_t pair32(const I &q, const JRec &r, A32 &a) {
    float rx, ry, rz, s2;
    const bool far = !near_test(q, r, rx, ry, rz, s2);
    Derived::body32(q, r, rx, ry, rz, s2, far, a);
    return __builtin_amdgcn_ballot_w64(far);
  }
  // FP32 contribution of a pair known to be far (a tile whose bounding box
  // is beyond every near-field threshold of its j's): no near test.
  template <class I, class JRec>
  __device__ static void pair32_far(const I &q, const JRec &r, A32 &a) {
    const float rx = dsub32(q.x, r.x), ry = dsub32(q.y, r.y),
                rz = dsub32(q.z, r.z);
    const float s2 = __builtin_fmaf(
        rz, rz, __builtin_fmaf(ry, ry, __builtin_fmaf(rx, rx, q.e2 + r.e2)));
    Derived::body32(q, r, rx, ry, rz, s2, true, a);
  }
  // Far tiles from FarRec records (policies with kFarRec, sorted engine):
  // the positions are FP32 offsets to the j tile's FP64 position origin
  // (the centre of its bounding box), the j offset rounded once when
  // packed and the i offset once per tile (tile_begin_far), so a pair forms
  // its separation with three FP32 subtractions in place of dsub32's three
  // FP64 subtractions and conversions.  On far tiles both offsets are about
  // as large as the separation, so the two roundings add an error of the
  // order of dsub32's one.
  static constexpr bool kFarRec = false;
  // far records per scalar load in full far tiles (tile_loop)
  static constexpr int kFarGroup = 1;
  // kFar16 policies (a pair body that reads only the position and the
  // mass of j): 16-byte far records {x, y, z, m} for the far tiles whose
  // j's share one e2, stored after the FarRec array, kFar16Group per
  // scalar load
  static constexpr bool kFar16 = false;
  static constexpr int kFar16Group = 4;
  struct alignas(16) FarRec16 {
    float x, y, z, m;
  };
  struct FarRec {
    float x, y, z, e2;
  };
  template <class I>
  __device__ static void tile_begin_far(I &q, const double *po) {
    q.px = (float)(q.x - po[0]);
    q.py = (float)(q.y - po[1]);
    q.pz = (float)(q.z - po[2]);
  }
  template <class I, class R>
  __device__ static void pair32_farrec(const I &q, const R &r, A32 &a) {
    pair32_farrec_u(q, r, q.e2 + r.e2, a);
  }
  // the same with the softening sum e2_i + e2_j given: a tile whose j's
  // share one e2 forms it once per tile (the same float sum, so the same
  // bits), one VALU instruction less per pair
  template <class I, class R>
  __device__ static void pair32_farrec_u(const I &q, const R &r, float e2ij,
                                         A32 &a) {
    const float rx = q.px - r.x, ry = q.py - r.y, rz = q.pz - r.z;
    const float s2 = __builtin_fmaf(
        rz, rz, __builtin_fmaf(ry, ry, __builtin_fmaf(rx, rx, e2ij)));
    Derived::body32(q, r, rx, ry, rz, s2, true, a);
  }
  // FP64 contribution of a near pair, from the caller's FP64 arrays.
  template <class I, class JRec>
  __device__ static void near_pair(const I &q, const JRec &r, A64 &a64,
                                   const Soa &is, const Soa &js, int j) {
    float rx, ry, rz, s2;
    if (near_test(q, r, rx, ry, rz, s2)) {
      typename P64::Params p64{};
      P64::pair(Derived::exact_i(q, is), P64::pack(js, j, 0.0), a64, p64);
    }
  }
  // the same with the FP64 policy's record of j given (the sorted engine
  // reads it with scalar loads from a packed array)
  template <class I, class JRec>
  __device__ static void near_pair64(const I &q, const JRec &r, A64 &a64,
                                     const Soa &is,
                                     const typename P64::JRec &r64) {
    float rx, ry, rz, s2;
    if (near_test(q, r, rx, ry, rz, s2)) {
      typename P64::Params p64{};
      P64::pair(Derived::exact_i(q, is), r64, a64, p64);
    }
  }
  // the FP64 policy's i record of q (policies holding q's FP64 fields
  // build it from them)
  template <class I>
  __device__ static typename P64::I exact_i(const I &q, const Soa &is) {
    typename P64::Params p64{};
    return P64::load_i(is, q.i, p64);
  }
};

}  // namespace tupan
