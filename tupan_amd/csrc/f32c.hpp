// f32c.hpp -- common shape of the FP32C ("FP32-compensated") policies.
//
//   * pair separations (and velocity differences) are formed in FP64 and
//     rounded once to FP32 (dsub32);
//   * the rest of the pair arithmetic is FP32 with v_rsq_f32;
//   * sums run in FP32 over a tile of kTile j's and are folded into FP64
//     accumulators after every tile;
//   * near field: a pair with r^2 < near2_j contributes nothing in FP32 and
//     is re-evaluated by the FP64 policy from the caller's FP64 arrays.
#pragma once
#include "engine.hpp"

namespace tupan {

// Common shape of the FP32C policies: P64 is the FP64 policy used for the
// near field; the derived policy supplies JRec, I, pack, load_i and body32.
template <class Derived, class P64, int N>
struct F32cBase : SumOut<N> {
  static constexpr bool kTwoLevel = true;
  using Params = NoParams;
  using A64 = F64Acc<N>;
  using A32 = F32Acc<N>;
  __device__ static void zero(A64 &a) { zero_acc(a); }
  __device__ static void zero(A32 &a) { zero_acc(a); }
  __device__ static void fold(const A32 &a, A64 &b) { fold_acc(a, b); }
  __device__ static void values(const A64 &a, double *v) { values_acc(a, v); }

  // Separation and the near-field test.  r^2 uses explicit FMAs so that
  // pair32 and near_pair take bit-identical decisions.  near2 >= FLT_MIN,
  // so coincident particles (r^2 = 0, masked by the reference) always take
  // the FP64 path, which applies the mask (as does a NaN separation).
  template <class I, class JRec>
  __device__ static bool near_test(const I &q, const JRec &r, float &rx,
                                   float &ry, float &rz, float &r2) {
    rx = dsub32(q.x, r.x);
    ry = dsub32(q.y, r.y);
    rz = dsub32(q.z, r.z);
    r2 = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
    return !(r2 >= r.near2);
  }
  // FP32 contribution of a far pair; near pairs contribute zero here.
  // Returns the wave's mask of far lanes.
  template <class I, class JRec>
  __device__ static uint64_t pair32(const I &q, const JRec &r, A32 &a) {
    float rx, ry, rz, r2;
    const bool far = !near_test(q, r, rx, ry, rz, r2);
    Derived::body32(q, r, rx, ry, rz, r2, far, a);
    return __builtin_amdgcn_ballot_w64(far);
  }
  // FP64 contribution of a near pair, from the caller's FP64 arrays.
  template <class I, class JRec>
  __device__ static void near_pair(const I &q, const JRec &r, A64 &a64,
                                   const Soa &is, const Soa &js, int j) {
    float rx, ry, rz, r2;
    if (near_test(q, r, rx, ry, rz, r2)) {
      typename P64::Params p64{};
      P64::pair(P64::load_i(is, q.i, p64), P64::pack(js, j, 0.0), a64, p64);
    }
  }
};

__device__ __forceinline__ float near2_of(double m, double sc) {
  return fmaxf((float)(m * sc), 1.17549435e-38f);  // >= FLT_MIN
}

}  // namespace tupan
