// hermite.hpp -- one axis of the Hermite predictor / corrector, FP64, in
// the reference's numpy operation order (tupan/integrator/hermite.py:25-41,
// 75-90, 127-159, 203-241 predict; 43-57, 93-119, 162-195, 244-284
// correct), one IEEE operation at a time (no FMA contraction).
#pragma once
#include <hip/hip_runtime.h>

namespace tupan {

// r, v advanced from the step's initial state (r, v) and derivatives a, j,
// s(nap), c(rackle); only the ones the order uses are read.
template <int ORDER>
__device__ __forceinline__ void hermite_predict_axis(double r, double v,
                                                     double a, double j,
                                                     double sn, double c,
                                                     double tau, double &r1,
                                                     double &v1) {
#pragma clang fp contract(off)
  double dr, dv;
  if (ORDER == 2) {
    dr = (a * tau / 2 + v) * tau;
    dv = a * tau;
  } else if (ORDER == 4) {
    dr = ((j * tau / 3 + a) * tau / 2 + v) * tau;
    dv = (j * tau / 2 + a) * tau;
  } else if (ORDER == 6) {
    dr = (((sn * tau / 4 + j) * tau / 3 + a) * tau / 2 + v) * tau;
    dv = ((sn * tau / 3 + j) * tau / 2 + a) * tau;
  } else {
    dr = ((((c * tau / 5 + sn) * tau / 4 + j) * tau / 3 + a) * tau / 2 + v) *
         tau;
    dv = (((c * tau / 4 + sn) * tau / 3 + j) * tau / 2 + a) * tau;
  }
  r1 = r + dr;
  v1 = v + dv;
}

// Corrected r1, v1 from the snapshot (suffix 0) and the derivatives at the
// predicted state (suffix 1).
template <int ORDER>
__device__ __forceinline__ void hermite_correct_axis(
    double r0, double v0, double a0, double a1, double j0, double j1,
    double s0, double s1, double c0, double c1, double tau, double &r1,
    double &v1) {
#pragma clang fp contract(off)
  if (ORDER == 2) {
    v1 = (a0 + a1) * tau / 2 + v0;
    r1 = (v0 + v1) * tau / 2 + r0;
  } else if (ORDER == 4) {
    v1 = ((j0 - j1) * tau / 6 + (a0 + a1)) * tau / 2 + v0;
    r1 = ((a0 - a1) * tau / 6 + (v0 + v1)) * tau / 2 + r0;
  } else if (ORDER == 6) {
    v1 = (((s0 + s1) * tau / 12 + (j0 - j1)) * tau / 5 + (a0 + a1)) * tau /
             2 +
         v0;
    r1 = (((j0 + j1) * tau / 12 + (a0 - a1)) * tau / 5 + (v0 + v1)) * tau /
             2 +
         r0;
  } else {
    v1 = ((((c0 - c1) * tau / 20 + (s0 + s1)) * tau / 3 + 3 * (j0 - j1)) *
              tau / 14 +
          (a0 + a1)) * tau / 2 +
         v0;
    r1 = ((((s0 - s1) * tau / 20 + (j0 + j1)) * tau / 3 + 3 * (a0 - a1)) *
              tau / 14 +
          (v0 + v1)) * tau / 2 +
         r0;
  }
}

}  // namespace tupan
