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

/* 1PN, pn_terms.h:126-152 */
__device__ __forceinline__ PnAB pn_1pn(double im, double jm, double inv_r, double iv2, double jv2,
                           double ivjv, double niv, double njv, double njv2,
                           const PnParams &c)
{
    PnAB t;
    t.a = - iv2
          - 2 * jv2
          + 4 * ivjv
          + Q(3, 2) * njv2
          + inv_r * (+ 5 * im
                     + 4 * jm);
    t.b = + 4 * niv
          - 3 * njv;
    t.a *= c.inv2;
    t.b *= c.inv2;
    return t;
}

/* 2PN, pn_terms.h:155-213 */
__device__ __forceinline__ PnAB pn_2pn(double im, double jm, double im2, double jm2, double imjm,
                           double inv_r, double inv_r2, double iv2, double jv2,
                           double jv4, double ivjv, double ivjvivjv, double nv,
                           double niv, double njv, double niv2, double njv2,
                           double nivnjv, const PnParams &c)
{
    PnAB t;
    t.a = - 2 * (+ jv4
                 + ivjvivjv)
          + 4 * jv2 * ivjv
          + njv2 * (+ Q(3, 2) * iv2
                    + Q(9, 2) * jv2
                    - 6 * ivjv
                    - Q(15, 8) * njv2)
          - inv_r2 * (+ Q(57, 4) * im2
                      + 9 * jm2
                      + Q(69, 2) * imjm)
          + inv_r * (+ im * (- Q(15, 4) * iv2
                             + Q(5, 4) * jv2
                             - Q(5, 2) * ivjv
                             + Q(39, 2) * niv2
                             - 39 * nivnjv
                             + Q(17, 2) * njv2)
                     + jm * (+ 4 * jv2
                             - 8 * ivjv
                             + 2 * niv2
                             - 4 * nivnjv
                             - 6 * njv2));
    t.b = + inv_r * (+ im * (- Q(63, 4) * niv
                             + Q(55, 4) * njv)
                     - jm * 2 * (+ niv
                                 + njv))
          + iv2 * njv
          - ivjv * nv * 4
          + jv2 * (+ 4 * niv
                   - 5 * njv)
          + njv2 * (- 6 * niv
                    + Q(9, 2) * njv);
    t.a *= c.inv4;
    t.b *= c.inv4;
    return t;
}

/* 2.5PN (radiation reaction), pn_terms.h:216-237 */
__device__ __forceinline__ PnAB pn_25pn(double im, double jm, double inv_r, double v2, double nv,
                            const PnParams &c)
{
    PnAB t;
    t.a = + nv * (+ inv_r * (- Q(24, 5) * im
                             + Q(208, 15) * jm)
                  + Q(12, 5) * v2);
    t.b = - v2
          + inv_r * (+ Q(8, 5) * im
                     - Q(32, 5) * jm);
    double scale = (im * inv_r) * c.inv5;
    t.a *= scale;
    t.b *= scale;
    return t;
}

/* 3PN, pn_terms.h:240-372 */
__device__ __forceinline__ PnAB pn_3pn(double im, double jm, double im2, double jm2, double imjm,
                           double inv_r, double inv_r2, double v2, double iv2,
                           double jv2, double jv4, double ivjv, double ivjvivjv,
                           double nv, double nvnv, double niv, double njv, double niv2,
                           double njv2, double nivnjv, const PnParams &c)
{
    PnAB t;
    t.a = + njv2 * (+ 3 * ivjvivjv
                    + Q(3, 2) * iv2 * jv2
                    - 12 * ivjv * jv2
                    + Q(15, 2) * jv4
                    + njv2 * (+ Q(15, 2) * (+ ivjv
                                            - jv2
                                            - Q(1, 4) * iv2)
                              + Q(35, 16) * njv2))
          + 2 * jv2 * (- ivjvivjv
                       + jv2 * (+ 2 * ivjv
                                - jv2))
          + im * inv_r * (+ niv * (+ njv * (+ 244 * ivjv
                                            - Q(205, 2) * iv2
                                            - Q(283, 2) * jv2
                                            + Q(383, 2) * njv2)
                                   + niv * (+ Q(229, 4) * (+ iv2
                                                           + jv2
                                                           - 2 * ivjv)
                                            - Q(723, 4) * njv2
                                            + niv * (+ Q(171, 2) * (+ njv
                                                                    - Q(1, 4) * niv))))
                          + njv2 * (+ Q(191, 4) * iv2
                                    + Q(259, 4) * jv2
                                    - Q(225, 2) * ivjv
                                    - Q(455, 8) * njv2)
                          + ivjv * (+ Q(91, 2) * iv2
                                    + 43 * jv2
                                    - Q(177, 4) * ivjv)
                          - Q(91, 8) * iv2 * (+ iv2
                                              + 2 * jv2)
                          - Q(81, 8) * jv4)
          + jm * inv_r * 4 * (+ jv4
                              + njv * (+ niv * (+ ivjv
                                                - jv2)
                                       + njv * (+ 3 * (+ ivjv
                                                       - jv2)
                                                - Q(3, 2) * niv2
                                                + njv * (+ 3 * niv
                                                         + Q(3, 2) * njv)))
                              + ivjv * (+ ivjv
                                        - 2 * jv2))
          + jm2 * inv_r2 * (- niv2
                            + 2 * nivnjv
                            + Q(43, 2) * njv2
                            + 18 * ivjv
                            - 9 * jv2)
          + imjm * inv_r2 * (+ Q(415, 8) * niv2
                             - Q(375, 4) * nivnjv
                             + Q(1113, 8) * njv2
                             + 18 * iv2
                             + kPi2 * (+ Q(123, 64) * v2
                                        - Q(615, 64) * nvnv)
                             + 33 * (+ ivjv
                                     - Q(1, 2) * jv2))
          + im2 * inv_r2 * (- Q(2069, 8) * niv2
                            + 543 * nivnjv
                            - Q(939, 4) * njv2
                            + Q(471, 8) * iv2
                            + Q(357, 8) * (+ jv2
                                           - 2 * ivjv))
          + inv_r * inv_r2 * (+ 16 * jm * jm2
                              + im2 * jm * (+ Q(547, 3)
                                            - Q(41, 16) * kPi2)
                              - Q(13, 12) * im * im2
                              + im * jm2 * (+ Q(545, 3)
                                            - Q(41, 16) * kPi2));

    t.b = + njv * (+ jv2 * (+ iv2
                            + 8 * ivjv
                            - 7 * jv2)
                   - 2 * ivjvivjv
                   + njv * (+ 6 * niv * (+ ivjv
                                         - 2 * jv2)
                            + njv * (+ 6 * (+ 2 * jv2
                                            - ivjv
                                            - Q(1, 4) * iv2)
                                     + njv * (+ Q(15, 2) * (+ niv
                                                            - Q(3, 4) * njv)))))
          + 4 * niv * (+ jv4
                       - ivjv * jv2)
          + jm * inv_r * (+ njv * (+ 4 * (+ ivjv
                                          - jv2
                                          - Q(1, 2) * niv2)
                                   + njv * (+ 2 * (+ 4 * niv
                                                   + njv)))
                          + 2 * niv * (+ ivjv
                                       - jv2))
          + im * inv_r * (+ niv * (+ Q(207, 8) * iv2
                                   + Q(81, 8) * jv2
                                   - 36 * ivjv
                                   - Q(269, 4) * njv2
                                   + niv * (+ Q(565, 4) * njv
                                            - Q(243, 4) * niv))
                          + njv * (+ Q(83, 8) * jv2
                                   + Q(27, 4) * ivjv
                                   - Q(137, 8) * iv2
                                   - Q(95, 12) * njv2))
          + inv_r2 * (+ jm2 * (+ 4 * niv
                               + 5 * njv)
                      + im2 * (+ Q(311, 4) * niv
                               - Q(357, 4) * njv)
                      + imjm * (+ Q(479, 8) * njv
                                - Q(307, 8) * niv
                                + Q(123, 32) * kPi2 * nv));
    t.a *= c.inv6;
    t.b *= c.inv6;
    return t;
}

/* 3.5PN, pn_terms.h:375-470 */
__device__ __forceinline__ PnAB pn_35pn(double im, double jm, double im2, double jm2, double imjm,
                            double inv_r, double inv_r2, double v2, double iv2,
                            double jv2, double iv4, double jv4, double ivjv, double nv,
                            double nvnv, double niv, double njv, double niv2,
                            double njv2, double nivnjv, const PnParams &c)
{
    PnAB t;
    t.a = + im2 * inv_r2 * (+ Q(3992, 105) * niv
                            - Q(4328, 105) * njv)
          + imjm * inv_r * inv_r2 * (- Q(13576, 105) * niv
                                     + Q(2872, 21) * njv)
          + jm2 * inv_r * inv_r2 * (- Q(3172, 21) * nv)
          + im * inv_r * (+ niv * (+ 48 * niv2
                                   - Q(4888, 105) * iv2
                                   + Q(2056, 21) * ivjv
                                   - Q(1028, 21) * jv2)
                          + nivnjv * (- Q(696, 5) * niv
                                      + Q(744, 5) * njv)
                          + njv * (- Q(288, 5) * njv2
                                   + Q(5056, 105) * iv2
                                   - Q(2224, 21) * ivjv
                                   + Q(5812, 105) * jv2))
          + jm * inv_r * (+ niv * (- Q(582, 5) * niv2
                                   - Q(2864, 35) * ivjv
                                   + Q(1432, 35) * jv2)
                          + nivnjv * (+ Q(1746, 5) * niv
                                      - Q(1954, 5) * njv)
                          + Q(3568, 105) * nv * iv2
                          + njv * (+ 158 * njv2
                                   - Q(5752, 105) * jv2
                                   + Q(10048, 105) * ivjv))
          + (+ nv * (- 56 * nvnv * nvnv
                     - Q(246, 35) * iv4)
             + niv * (+ v2 * (+ 60 * niv2
                              - 180 * nivnjv
                              + 174 * njv2)
                      + ivjv * (+ Q(1068, 35) * (+ iv2
                                                 - ivjv)
                                + Q(984, 35) * jv2)
                      - Q(534, 35) * iv2 * jv2
                      - Q(204, 35) * jv4)
             + njv * (- 54 * njv2 * v2
                      + ivjv * (- Q(984, 35) * iv2
                                + Q(180, 7) * ivjv
                                - Q(732, 35) * jv2)
                      + Q(90, 7) * iv2 * jv2
                      + Q(24, 7) * jv4));

    t.b = - im2 * inv_r2 * Q(184, 21)
          + imjm * inv_r2 * Q(6224, 105)
          + jm2 * inv_r2 * Q(6388, 105)
          + im * inv_r * (+ Q(52, 15) * niv2
                          - Q(56, 15) * nivnjv
                          - Q(44, 15) * njv2
                          - Q(132, 35) * iv2
                          + Q(152, 35) * ivjv
                          - Q(48, 35) * jv2)
          + jm * inv_r * (+ Q(454, 15) * niv2
                          - Q(372, 5) * nivnjv
                          + Q(854, 15) * njv2
                          - Q(152, 21) * iv2
                          + Q(2864, 105) * ivjv
                          - Q(1768, 105) * jv2)
          + (+ 60 * nvnv * nvnv
             + v2 * (- Q(348, 5) * niv2
                     + Q(684, 5) * nivnjv
                     - 66 * njv2)
             + Q(334, 35) * iv4
             + ivjv * (- Q(1336, 35) * iv2
                       + Q(1308, 35) * ivjv
                       - Q(1252, 35) * jv2)
             + Q(654, 35) * iv2 * jv2
             + Q(292, 35) * jv4);
    double scale = (im * inv_r) * c.inv7;
    t.a *= scale;
    t.b *= scale;
    return t;
}

/* Order gating and summation, pn_terms.h:473-567: order>1 adds 1PN, >3 adds
 * 2PN, >4 adds 2.5PN, >5 adds 3PN, >6 adds 3.5PN (the odd slots are
 * unimplemented and contribute zero). */
__device__ __forceinline__ PnAB pn_pair(
    double im, double jm, double rx, double ry, double rz,
                            double vx, double vy, double vz, double v2,
                            double ivx, double ivy, double ivz,
                            double jvx, double jvy, double jvz,
                            double inv_r, double inv_r2, double inv_r3,
                            const PnParams &c)
{
    PnAB p1 = {0, 0}, p2 = {0, 0}, p3 = {0, 0}, p4 = {0, 0};
    PnAB p5 = {0, 0}, p6 = {0, 0}, p7 = {0, 0};
    double nx = rx * inv_r;
    double ny = ry * inv_r;
    double nz = rz * inv_r;
    double iv2 = ivx * ivx + ivy * ivy + ivz * ivz;
    double jv2 = jvx * jvx + jvy * jvy + jvz * jvz;

    if (c.order > 1) {
        double niv = nx * ivx + ny * ivy + nz * ivz;
        double njv = nx * jvx + ny * jvy + nz * jvz;
        double njv2 = njv * njv;
        double ivjv = ivx * jvx + ivy * jvy + ivz * jvz;
        p2 = pn_1pn(im, jm, inv_r, iv2, jv2, ivjv, niv, njv, njv2, c);
        if (c.order > 3) {
            double im2 = im * im;
            double jm2 = jm * jm;
            double imjm = im * jm;
            double jv4 = jv2 * jv2;
            double ivjvivjv = ivjv * ivjv;
            double nv = nx * vx + ny * vy + nz * vz;
            double niv_ = nx * ivx + ny * ivy + nz * ivz;
            double njv_ = nx * jvx + ny * jvy + nz * jvz;
            double niv2 = niv_ * niv_;
            double njv2_ = njv_ * njv_;
            double nivnjv = niv_ * njv_;
            p4 = pn_2pn(im, jm, im2, jm2, imjm, inv_r, inv_r2, iv2, jv2, jv4,
                        ivjv, ivjvivjv, nv, niv_, njv_, niv2, njv2_, nivnjv,
                        c);
            if (c.order > 4) {
                p5 = pn_25pn(im, jm, inv_r, v2, nv, c);
                if (c.order > 5) {
                    double nvnv = nv * nv;
                    p6 = pn_3pn(im, jm, im2, jm2, imjm, inv_r, inv_r2, v2,
                                iv2, jv2, jv4, ivjv, ivjvivjv, nv, nvnv,
                                niv_, njv_, niv2, njv2_, nivnjv, c);
                    if (c.order > 6) {
                        double iv4 = iv2 * iv2;
                        p7 = pn_35pn(im, jm, im2, jm2, imjm, inv_r, inv_r2,
                                     v2, iv2, jv2, iv4, jv4, ivjv, nv, nvnv,
                                     niv_, njv_, niv2, njv2_, nivnjv, c);
                    }
                }
            }
        }
    }
    PnAB pn = {0, 0};
    pn.a += ((((((p7.a + p6.a) + p5.a) + p4.a) + p3.a) + p2.a) + p1.a);
    pn.b += ((((((p7.b + p6.b) + p5.b) + p4.b) + p3.b) + p2.b) + p1.b);
    pn.a *= jm * inv_r3;
    pn.b *= jm * inv_r2;
    return pn;
}


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
    PnAB pn = pn_pair(q.m, r.m, rx, ry, rz, vx, vy, vz, v2, q.vx, q.vy, q.vz,
                      r.vx, r.vy, r.vz, inv_r, inv_r2, inv_r3, c);
    a.v[0] += pn.a * rx + pn.b * vx;
    a.v[1] += pn.a * ry + pn.b * vy;
    a.v[2] += pn.a * rz + pn.b * vz;
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

int tupan_pnacc_dev(size_t ni, const double *const *ip, size_t nj,
                    const double *const *jp, unsigned order,
                    const double *inv, double *const *op, void *stream) {
  PnParams prm;
  int rc = pn_params(order, inv, &prm);
  if (rc) return rc;
  return run_engine<PnAccF64>(TUPAN_K_PNACC, TUPAN_MODE_FP64, ni, ip, nj, jp,
                              op, prm, reinterpret_cast<hipStream_t>(stream));
}

int tupan_pnacc_host(size_t ni, const double *const *ip, size_t nj,
                     const double *const *jp, unsigned order,
                     const double *inv, double *const *op) {
  PnParams prm;
  int rc = pn_params(order, inv, &prm);
  if (rc) return rc;
  return host_call(ni, ip, nj, jp, 8, op, 3,
                   [&](const double *const *di, const double *const *dj,
                       double *const *dout, hipStream_t st) {
                     return run_engine<PnAccF64>(TUPAN_K_PNACC,
                                                 TUPAN_MODE_FP64, ni, di, nj,
                                                 dj, dout, prm, st);
                   });
}

}  // extern "C"
