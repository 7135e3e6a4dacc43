/*
 * oracle_kernels.c -- CPU restatement of Tupan's direct-summation kernel
 * family.  TEST INFRASTRUCTURE ONLY: this file is the checker that the HIP
 * kernels are compared against.  It is never linked into, loaded by or called
 * from the product library (tupan_amd/), which fails loudly without its HIP
 * build.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.
 *
 * Built twice from this one source (see oracle/Makefile):
 *   liboracle_f64.so : real = double, sint = long, uint = unsigned long
 *   liboracle_f32.so : real = float,  sint = int,  uint = unsigned int
 * matching the two precisions of the reference library
 * (tupan/lib/src/c_common.h:7-43, tupan/lib/cffi_backend.py:31-87).
 *
 * Arithmetic contract: every floating-point operation is performed in the
 * same order as the reference C kernels so that, compiled with
 * `-O3 -std=c99 -ffp-contract=off` (the reference's ISO-C99 flags,
 * tupan/lib/cffi_backend.py:63-77), results are bit-identical to the
 * compiled reference on the same inputs.  Exported symbols take the
 * libtupan.h argument lists (tupan/lib/src/libtupan.h:2-229) under an
 * `oracle_` prefix.
 *
 * Pinning (tests/test_oracle.py): bitwise against the golden fixtures
 * produced by the compiled reference C (tests/golden/), plus closed forms
 * (two-body force, jerk as d/dt of acc by finite differences, Kepler period
 * return and energy conservation, PN 1PN two-body EIH formula).
 */
#include <math.h>
#include <stddef.h>

#ifdef ORACLE_FP64
typedef double real;
typedef long sint;
typedef unsigned long uint_t;
#define R_SQRT sqrt
#define R_FABS fabs
#define R_COS cos
#define R_SIN sin
#define R_COSH cosh
#define R_SINH sinh
#define R_LOG log
#define R_FMAX fmax
/* universal_kepler_solver.h:7-9 : 2^-42 */
#define KEPLER_TOL ((real)(2.2737367544323205948e-13))
#else
typedef float real;
typedef int sint;
typedef unsigned int uint_t;
#define R_SQRT sqrtf
#define R_FABS fabsf
#define R_COS cosf
#define R_SIN sinf
#define R_COSH coshf
#define R_SINH sinhf
#define R_LOG logf
#define R_FMAX fmaxf
/* universal_kepler_solver.h:9-10 : 2^-16 */
#define KEPLER_TOL ((real)(1.52587890625e-5))
#endif

#define KEPLER_MAXITER 64    /* universal_kepler_solver.h:12 */
#define LAGUERRE_ORDER 5     /* universal_kepler_solver.h:225 */
#define R_PI ((real)(3.141592653589793))   /* common.h:20 */
#define R_PI2 ((real)(9.869604401089358))  /* common.h:21 */

#define EXPORT __attribute__((visibility("default")))
#define RP const real *restrict
#define WP real *restrict

/* rational constant n/d evaluated in the working precision, as the
 * reference's pn_terms.h:7-123 macros do */
#define Q(n, d) (((real)(n)) / (d))

static inline sint sign_of(real x) { return (x > 0) - (x < 0); }

/* Plummer-softened 1/(r^2+e^2), zeroed unless the *unsoftened* r^2 > 0.
 * smoothing.h:382-390 (inv_r2), used by :405-414 (inv_r3) etc. */
static inline real soft_inv_r2(real r2, real e2)
{
    real inv = 1 / (r2 + e2);
    return (r2 > 0) ? inv : (real)0;
}

/* ------------------------------------------------------------------ Phi */
/* phi_kernel_common.h:7-31, driver phi_kernel.c:5-33 */
EXPORT void oracle_phi_kernel(
    const uint_t ni, RP im, RP irx, RP iry, RP irz, RP ie2,
    const uint_t nj, RP jm, RP jrx, RP jry, RP jrz, RP je2,
    WP iphi)
{
    (void)im;
    for (uint_t i = 0; i < ni; ++i) {
        real acc = 0;
        for (uint_t j = 0; j < nj; ++j) {
            real rx = irx[i] - jrx[j];
            real ry = iry[i] - jry[j];
            real rz = irz[i] - jrz[j];
            real e2 = ie2[i] + je2[j];
            real r2 = rx * rx + ry * ry + rz * rz;
            real inv_r1 = R_SQRT(soft_inv_r2(r2, e2));
            acc -= jm[j] * inv_r1;
        }
        iphi[i] = acc;
    }
}

/* ------------------------------------------------------------------ Acc */
/* acc_kernel_common.h:7-37, driver acc_kernel.c:5-39 */
EXPORT void oracle_acc_kernel(
    const uint_t ni, RP im, RP irx, RP iry, RP irz, RP ie2,
    const uint_t nj, RP jm, RP jrx, RP jry, RP jrz, RP je2,
    WP iax, WP iay, WP iaz)
{
    (void)im;
    for (uint_t i = 0; i < ni; ++i) {
        real ax = 0, ay = 0, az = 0;
        for (uint_t j = 0; j < nj; ++j) {
            real rx = irx[i] - jrx[j];
            real ry = iry[i] - jry[j];
            real rz = irz[i] - jrz[j];
            real e2 = ie2[i] + je2[j];
            real r2 = rx * rx + ry * ry + rz * rz;
            real inv_r2 = soft_inv_r2(r2, e2);
            real k = inv_r2 * R_SQRT(inv_r2);   /* smoothing.h:413 */
            k *= jm[j];
            ax -= k * rx;
            ay -= k * ry;
            az -= k * rz;
        }
        iax[i] = ax; iay[i] = ay; iaz[i] = az;
    }
}

/* -------------------------------------------------------------- AccJerk */
/* acc_jerk_kernel_common.h:7-55, driver acc_jerk_kernel.c:5-59 */
EXPORT void oracle_acc_jerk_kernel(
    const uint_t ni, RP im, RP irx, RP iry, RP irz, RP ie2,
    RP ivx, RP ivy, RP ivz,
    const uint_t nj, RP jm, RP jrx, RP jry, RP jrz, RP je2,
    RP jvx, RP jvy, RP jvz,
    WP iax, WP iay, WP iaz, WP ijx, WP ijy, WP ijz)
{
    (void)im;
    for (uint_t i = 0; i < ni; ++i) {
        real ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0;
        for (uint_t j = 0; j < nj; ++j) {
            real rx = irx[i] - jrx[j];
            real ry = iry[i] - jry[j];
            real rz = irz[i] - jrz[j];
            real e2 = ie2[i] + je2[j];
            real vx = ivx[i] - jvx[j];
            real vy = ivy[i] - jvy[j];
            real vz = ivz[i] - jvz[j];
            real r2 = rx * rx + ry * ry + rz * rz;
            real rv = rx * vx + ry * vy + rz * vz;
            real inv_r2 = soft_inv_r2(r2, e2);
            real k = inv_r2 * R_SQRT(inv_r2);
            real alpha = 3 * rv * inv_r2;
            vx -= alpha * rx;
            vy -= alpha * ry;
            vz -= alpha * rz;
            k *= jm[j];
            ax -= k * rx;
            ay -= k * ry;
            az -= k * rz;
            jx -= k * vx;
            jy -= k * vy;
            jz -= k * vz;
        }
        iax[i] = ax; iay[i] = ay; iaz[i] = az;
        ijx[i] = jx; ijy[i] = jy; ijz[i] = jz;
    }
}

/* ---------------------------------------------------------- SnapCrackle */
/* snap_crackle_kernel_common.h:7-94, driver snap_crackle_kernel.c:5-81 */
EXPORT void oracle_snap_crackle_kernel(
    const uint_t ni, RP im, RP irx, RP iry, RP irz, RP ie2,
    RP ivx, RP ivy, RP ivz, RP iax, RP iay, RP iaz, RP ijx, RP ijy, RP ijz,
    const uint_t nj, RP jm, RP jrx, RP jry, RP jrz, RP je2,
    RP jvx, RP jvy, RP jvz, RP jax, RP jay, RP jaz, RP jjx, RP jjy, RP jjz,
    WP isx, WP isy, WP isz, WP icx, WP icy, WP icz)
{
    (void)im;
    for (uint_t i = 0; i < ni; ++i) {
        real sx = 0, sy = 0, sz = 0, cx = 0, cy = 0, cz = 0;
        for (uint_t j = 0; j < nj; ++j) {
            real rx = irx[i] - jrx[j];
            real ry = iry[i] - jry[j];
            real rz = irz[i] - jrz[j];
            real e2 = ie2[i] + je2[j];
            real vx = ivx[i] - jvx[j];
            real vy = ivy[i] - jvy[j];
            real vz = ivz[i] - jvz[j];
            real Ax = iax[i] - jax[j];
            real Ay = iay[i] - jay[j];
            real Az = iaz[i] - jaz[j];
            real Jx = ijx[i] - jjx[j];
            real Jy = ijy[i] - jjy[j];
            real Jz = ijz[i] - jjz[j];
            real r2 = rx * rx + ry * ry + rz * rz;
            real rv = rx * vx + ry * vy + rz * vz;
            real v2 = vx * vx + vy * vy + vz * vz;
            real rJ = rx * Jx + ry * Jy + rz * Jz;
            real rA = rx * Ax + ry * Ay + rz * Az;
            real vA = vx * Ax + vy * Ay + vz * Az;
            real inv_r2 = soft_inv_r2(r2, e2);
            real k = inv_r2 * R_SQRT(inv_r2);

            /* Taylor recurrences of d^n/dt^n (1/r^3), :63-84 */
            real alpha = rv * inv_r2;
            real alpha2 = alpha * alpha;
            real beta = 3 * ((v2 + rA) * inv_r2 + alpha2);
            real gamma = (3 * vA + rJ) * inv_r2 + alpha * (beta - 4 * alpha2);
            alpha *= 3;
            gamma *= 3;
            vx -= alpha * rx;
            vy -= alpha * ry;
            vz -= alpha * rz;
            real two_alpha = 2 * alpha;
            Ax -= two_alpha * vx + beta * rx;
            Ay -= two_alpha * vy + beta * ry;
            Az -= two_alpha * vz + beta * rz;
            alpha *= 3;
            beta *= 3;
            Jx -= alpha * Ax + beta * vx + gamma * rx;
            Jy -= alpha * Ay + beta * vy + gamma * ry;
            Jz -= alpha * Az + beta * vz + gamma * rz;

            k *= jm[j];
            sx -= k * Ax;
            sy -= k * Ay;
            sz -= k * Az;
            cx -= k * Jx;
            cy -= k * Jy;
            cz -= k * Jz;
        }
        isx[i] = sx; isy[i] = sy; isz[i] = sz;
        icx[i] = cx; icy[i] = cy; icz[i] = cz;
    }
}

/* ---------------------------------------------------------------- Tstep */
/* tstep_kernel_common.h:7-56, driver + finalisation tstep_kernel.c:5-48 */
EXPORT void oracle_tstep_kernel(
    const uint_t ni, RP im, RP irx, RP iry, RP irz, RP ie2,
    RP ivx, RP ivy, RP ivz,
    const uint_t nj, RP jm, RP jrx, RP jry, RP jrz, RP je2,
    RP jvx, RP jvy, RP jvz,
    const real eta, WP idt_a, WP idt_b)
{
    for (uint_t i = 0; i < ni; ++i) {
        real w2_sum = 0, w2_max = 0;
        for (uint_t j = 0; j < nj; ++j) {
            real rx = irx[i] - jrx[j];
            real ry = iry[i] - jry[j];
            real rz = irz[i] - jrz[j];
            real e2 = ie2[i] + je2[j];
            real vx = ivx[i] - jvx[j];
            real vy = ivy[i] - jvy[j];
            real vz = ivz[i] - jvz[j];
            real m = im[i] + jm[j];
            real r2 = rx * rx + ry * ry + rz * rz;
            real rv = rx * vx + ry * vy + rz * vz;
            real v2 = vx * vx + vy * vy + vz * vz;
            int pair = (r2 > 0);
            real inv_r2 = soft_inv_r2(r2, e2);
            real inv_r1 = R_SQRT(inv_r2);
            /* :41-55, with a = 2 so that b = (1 + a/2) / s^2 */
            real a = 2;
            real b = (1 + a / 2) * inv_r2;
            real w2 = inv_r2;
            real gam = inv_r2;
            real phi = m * inv_r1;
            w2 *= (v2 + a * phi);
            gam *= (w2 + b * phi);
            gam *= (eta / R_SQRT(w2));
            w2 -= gam * rv;
            /* masked pairs may carry 0*inf = NaN here; discard them */
            w2 = pair ? w2 : (real)0;
            w2_sum += w2;
            w2_max = R_FMAX(w2, w2_max);
        }
        idt_a[i] = eta / R_SQRT(1 + w2_sum);
        idt_b[i] = eta / R_SQRT(1 + w2_max);
    }
}

/* ---------------------------------------------------------------- PNAcc */
/* Post-Newtonian pair terms, pn_terms.h:126-567.  The polynomial grouping
 * and term order below follow the reference term for term: this is required
 * for bit-identical FP64 results (floating-point addition does not
 * associate).  Each block returns the (a, b) coefficients of the pair
 * acceleration a*r + b*v before the mj/s^3, mj/s^2 scaling. */
typedef struct { real a, b; } pn_ab;

typedef struct {
    uint_t order;
    real inv1, inv2, inv3, inv4, inv5, inv6, inv7;
} pn_clight;

/* 1PN, pn_terms.h:126-152 */
static inline pn_ab pn_1pn(real im, real jm, real inv_r, real iv2, real jv2,
                           real ivjv, real niv, real njv, real njv2,
                           const pn_clight *c)
{
    pn_ab t;
    t.a = - iv2
          - 2 * jv2
          + 4 * ivjv
          + Q(3, 2) * njv2
          + inv_r * (+ 5 * im
                     + 4 * jm);
    t.b = + 4 * niv
          - 3 * njv;
    t.a *= c->inv2;
    t.b *= c->inv2;
    return t;
}

/* 2PN, pn_terms.h:155-213 */
static inline pn_ab pn_2pn(real im, real jm, real im2, real jm2, real imjm,
                           real inv_r, real inv_r2, real iv2, real jv2,
                           real jv4, real ivjv, real ivjvivjv, real nv,
                           real niv, real njv, real niv2, real njv2,
                           real nivnjv, const pn_clight *c)
{
    pn_ab t;
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
    t.a *= c->inv4;
    t.b *= c->inv4;
    return t;
}

/* 2.5PN (radiation reaction), pn_terms.h:216-237 */
static inline pn_ab pn_25pn(real im, real jm, real inv_r, real v2, real nv,
                            const pn_clight *c)
{
    pn_ab t;
    t.a = + nv * (+ inv_r * (- Q(24, 5) * im
                             + Q(208, 15) * jm)
                  + Q(12, 5) * v2);
    t.b = - v2
          + inv_r * (+ Q(8, 5) * im
                     - Q(32, 5) * jm);
    real scale = (im * inv_r) * c->inv5;
    t.a *= scale;
    t.b *= scale;
    return t;
}

/* 3PN, pn_terms.h:240-372 */
static inline pn_ab pn_3pn(real im, real jm, real im2, real jm2, real imjm,
                           real inv_r, real inv_r2, real v2, real iv2,
                           real jv2, real jv4, real ivjv, real ivjvivjv,
                           real nv, real nvnv, real niv, real njv, real niv2,
                           real njv2, real nivnjv, const pn_clight *c)
{
    pn_ab t;
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
                             + R_PI2 * (+ Q(123, 64) * v2
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
                                            - Q(41, 16) * R_PI2)
                              - Q(13, 12) * im * im2
                              + im * jm2 * (+ Q(545, 3)
                                            - Q(41, 16) * R_PI2));

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
                                + Q(123, 32) * R_PI2 * nv));
    t.a *= c->inv6;
    t.b *= c->inv6;
    return t;
}

/* 3.5PN, pn_terms.h:375-470 */
static inline pn_ab pn_35pn(real im, real jm, real im2, real jm2, real imjm,
                            real inv_r, real inv_r2, real v2, real iv2,
                            real jv2, real iv4, real jv4, real ivjv, real nv,
                            real nvnv, real niv, real njv, real niv2,
                            real njv2, real nivnjv, const pn_clight *c)
{
    pn_ab t;
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
    real scale = (im * inv_r) * c->inv7;
    t.a *= scale;
    t.b *= scale;
    return t;
}

/* Order gating and summation, pn_terms.h:473-567: order>1 adds 1PN, >3 adds
 * 2PN, >4 adds 2.5PN, >5 adds 3PN, >6 adds 3.5PN (the odd slots are
 * unimplemented and contribute zero). */
static inline pn_ab pn_pair(real im, real jm, real rx, real ry, real rz,
                            real vx, real vy, real vz, real v2,
                            real ivx, real ivy, real ivz,
                            real jvx, real jvy, real jvz,
                            real inv_r, real inv_r2, real inv_r3,
                            const pn_clight *c)
{
    pn_ab p1 = {0, 0}, p2 = {0, 0}, p3 = {0, 0}, p4 = {0, 0};
    pn_ab p5 = {0, 0}, p6 = {0, 0}, p7 = {0, 0};
    real nx = rx * inv_r;
    real ny = ry * inv_r;
    real nz = rz * inv_r;
    real iv2 = ivx * ivx + ivy * ivy + ivz * ivz;
    real jv2 = jvx * jvx + jvy * jvy + jvz * jvz;

    if (c->order > 1) {
        real niv = nx * ivx + ny * ivy + nz * ivz;
        real njv = nx * jvx + ny * jvy + nz * jvz;
        real njv2 = njv * njv;
        real ivjv = ivx * jvx + ivy * jvy + ivz * jvz;
        p2 = pn_1pn(im, jm, inv_r, iv2, jv2, ivjv, niv, njv, njv2, c);
        if (c->order > 3) {
            real im2 = im * im;
            real jm2 = jm * jm;
            real imjm = im * jm;
            real jv4 = jv2 * jv2;
            real ivjvivjv = ivjv * ivjv;
            real nv = nx * vx + ny * vy + nz * vz;
            real niv_ = nx * ivx + ny * ivy + nz * ivz;
            real njv_ = nx * jvx + ny * jvy + nz * jvz;
            real niv2 = niv_ * niv_;
            real njv2_ = njv_ * njv_;
            real nivnjv = niv_ * njv_;
            p4 = pn_2pn(im, jm, im2, jm2, imjm, inv_r, inv_r2, iv2, jv2, jv4,
                        ivjv, ivjvivjv, nv, niv_, njv_, niv2, njv2_, nivnjv,
                        c);
            if (c->order > 4) {
                p5 = pn_25pn(im, jm, inv_r, v2, nv, c);
                if (c->order > 5) {
                    real nvnv = nv * nv;
                    p6 = pn_3pn(im, jm, im2, jm2, imjm, inv_r, inv_r2, v2,
                                iv2, jv2, jv4, ivjv, ivjvivjv, nv, nvnv,
                                niv_, njv_, niv2, njv2_, nivnjv, c);
                    if (c->order > 6) {
                        real iv4 = iv2 * iv2;
                        p7 = pn_35pn(im, jm, im2, jm2, imjm, inv_r, inv_r2,
                                     v2, iv2, jv2, iv4, jv4, ivjv, nv, nvnv,
                                     niv_, njv_, niv2, njv2_, nivnjv, c);
                    }
                }
            }
        }
    }
    pn_ab pn = {0, 0};
    pn.a += ((((((p7.a + p6.a) + p5.a) + p4.a) + p3.a) + p2.a) + p1.a);
    pn.b += ((((((p7.b + p6.b) + p5.b) + p4.b) + p3.b) + p2.b) + p1.b);
    pn.a *= jm * inv_r3;
    pn.b *= jm * inv_r2;
    return pn;
}

/* pnacc_kernel_common.h:8-68, driver pnacc_kernel.c:5-59 */
EXPORT void oracle_pnacc_kernel(
    const uint_t ni, RP im, RP irx, RP iry, RP irz, RP ie2,
    RP ivx, RP ivy, RP ivz,
    const uint_t nj, RP jm, RP jrx, RP jry, RP jrz, RP je2,
    RP jvx, RP jvy, RP jvz,
    uint_t order, const real inv1, const real inv2, const real inv3,
    const real inv4, const real inv5, const real inv6, const real inv7,
    WP ipnax, WP ipnay, WP ipnaz)
{
    pn_clight c = {order, inv1, inv2, inv3, inv4, inv5, inv6, inv7};
    for (uint_t i = 0; i < ni; ++i) {
        real ax = 0, ay = 0, az = 0;
        for (uint_t j = 0; j < nj; ++j) {
            real rx = irx[i] - jrx[j];
            real ry = iry[i] - jry[j];
            real rz = irz[i] - jrz[j];
            real e2 = ie2[i] + je2[j];
            real vx = ivx[i] - jvx[j];
            real vy = ivy[i] - jvy[j];
            real vz = ivz[i] - jvz[j];
            real r2 = rx * rx + ry * ry + rz * rz;
            real v2 = vx * vx + vy * vy + vz * vz;
            /* smoothing.h:508-520 (inv_r1r2r3) */
            real inv_r2 = soft_inv_r2(r2, e2);
            real inv_r1 = R_SQRT(inv_r2);
            real inv_r3 = inv_r2 * inv_r1;
            pn_ab pn = pn_pair(im[i], jm[j], rx, ry, rz, vx, vy, vz, v2,
                               ivx[i], ivy[i], ivz[i], jvx[j], jvy[j], jvz[j],
                               inv_r1, inv_r2, inv_r3, &c);
            ax += pn.a * rx + pn.b * vx;
            ay += pn.a * ry + pn.b * vy;
            az += pn.a * rz + pn.b * vz;
        }
        ipnax[i] = ax; ipnay[i] = ay; ipnaz[i] = az;
    }
}

/* --------------------------------------------------------------- Sakura */
/* Stumpff functions c0..c3, universal_kepler_solver.h:17-79 */
static inline real stumpff_c0(real z)
{
    if (z < 0) return R_COS(R_SQRT(-z));
    if (z > 0) return R_COSH(R_SQRT(z));
    return (real)1;
}
static inline real stumpff_c1(real z)
{
    if (z < 0) { real s = R_SQRT(-z); return R_SIN(s) / s; }
    if (z > 0) { real s = R_SQRT(z); return R_SINH(s) / s; }
    return (real)1;
}
static inline real stumpff_c2(real z)
{
    if (z < 0) { real s = R_SQRT(-z); return (R_COS(s) - 1) / z; }
    if (z > 0) { real s = R_SQRT(z); return (R_COSH(s) - 1) / z; }
    return 1 / ((real)2);
}
static inline real stumpff_c3(real z)
{
    if (z < 0) { real s = R_SQRT(-z); return (R_SIN(s) / s - 1) / z; }
    if (z > 0) { real s = R_SQRT(z); return (R_SINH(s) / s - 1) / z; }
    return 1 / ((real)6);
}

/* S_k(s) = s^k c_k(alpha s^2), :82-120 */
static inline real uv_S0(real s, real alpha) { real s2 = s * s; return stumpff_c0(alpha * s2); }
static inline real uv_S1(real s, real alpha) { real s2 = s * s; return s * stumpff_c1(alpha * s2); }
static inline real uv_S2(real s, real alpha) { real s2 = s * s; return s2 * stumpff_c2(alpha * s2); }
static inline real uv_S3(real s, real alpha)
{
    real s2 = s * s;
    real s3 = s * s2;
    return s3 * stumpff_c3(alpha * s2);
}

/* Kepler's equation in universal variables and its s-derivatives,
 * :166-198.  k = {dt, r0, r0.v0, m, alpha}. */
typedef struct { real dt, r0, rv, m, alpha; } kepler_eq;

static inline real kep_f(real s, const kepler_eq *k)
{
    real val = k->r0 * s + k->rv * uv_S2(s, k->alpha)
               + (k->m + k->alpha * k->r0) * uv_S3(s, k->alpha);
    return val - k->dt;
}
static inline real kep_df(real s, const kepler_eq *k)
{
    return k->r0 + k->rv * uv_S1(s, k->alpha)
           + (k->m + k->alpha * k->r0) * uv_S2(s, k->alpha);
}
static inline real kep_ddf(real s, const kepler_eq *k)
{
    return k->rv * uv_S0(s, k->alpha)
           + (k->m + k->alpha * k->r0) * uv_S1(s, k->alpha);
}

/* Laguerre-Conway root finder (order 5), :225-254.
 * Returns 0, or -1 (zero denominator), -2 (iteration cap), -3 (sign flip). */
static inline sint laguerre_solve(real s0, real *s, const kepler_eq *k)
{
    sint it = 0;
    real delta;
    *s = s0;
    do {
        real fv = kep_f(*s, k);
        real dfv = kep_df(*s, k);
        real ddfv = kep_ddf(*s, k);
        real a = dfv;
        real a2 = a * a;
        real b = a2 - fv * ddfv;
        real g = LAGUERRE_ORDER * fv;
        real h = a + sign_of(a) * R_SQRT(R_FABS((LAGUERRE_ORDER - 1)
                                                * (LAGUERRE_ORDER * b - a2)));
        if (h == 0) return -1;
        delta = -g / h;
        *s += delta;
        it += 1;
        if (it > KEPLER_MAXITER) return -2;
    } while (R_FABS(delta) > KEPLER_TOL);
    if (sign_of(*s) != sign_of(s0)) return -3;
    return 0;
}

/* Lagrange f, g and their derivatives applied to (r, v), :312-357 */
static inline void kepler_fg_update(real dt, real s, real r0, real m, real e2,
                                    real alpha, real st[6])
{
    real r0x = st[0], r0y = st[1], r0z = st[2];
    real v0x = st[3], v0y = st[4], v0z = st[5];
    real lf = 1 - m * uv_S2(s, alpha) / r0;
    real lg = dt - m * uv_S3(s, alpha);
    real r1x = r0x * lf + v0x * lg;
    real r1y = r0y * lf + v0y * lg;
    real r1z = r0z * lf + v0z * lg;
    real r1sqr = r1x * r1x + r1y * r1y + r1z * r1z;
    real r1 = R_SQRT(r1sqr + e2);            /* softened, :341 */
    real ldf = -m * uv_S1(s, alpha) / (r0 * r1);
    real ldg = (1 + lg * ldf) / lf;
    st[0] = r1x; st[1] = r1y; st[2] = r1z;
    st[3] = r0x * ldf + v0x * ldg;
    st[4] = r0y * ldf + v0y * ldg;
    st[5] = r0z * ldf + v0z * ldg;
}

/* One universal-variable Kepler step of the softened two-body problem,
 * :360-463.  st holds (r, v) in and out; unchanged on failure. */
static inline sint kepler_step_once(real dt0, real m, real e2, real st[6])
{
    real rx = st[0], ry = st[1], rz = st[2];
    real vx = st[3], vy = st[4], vz = st[5];
    real r2 = rx * rx + ry * ry + rz * rz;
    if (!(r2 > 0)) return 0;
    r2 += e2;
    real r = R_SQRT(r2);
    real v2 = vx * vx + vy * vy + vz * vz;
    real rv = rx * vx + ry * vy + rz * vz;
    real beta = 2 - (e2 / r2);
    real alpha = v2 - beta * m / r;

    real dt = dt0;
    if (alpha < 0) {
        /* bound orbit: reduce dt modulo the period, :406-412 */
        real a = m / R_FABS(alpha);
        real T = 2 * R_PI * a * R_SQRT(a / m);
        real ratio = dt0 / T;
        dt = (ratio - (sint)(ratio)) * T;
    }
    real s0 = dt / r;
    if (alpha > 0) {
        /* hyperbolic first guess, :416-425 */
        real salpha = R_SQRT(alpha);
        real ss = R_FABS(2 * alpha * dt / (rv + (m + alpha * r) / salpha));
        if (ss > 1) s0 = sign_of(dt) * R_LOG(ss) / salpha;
    }
    kepler_eq k = {dt, r, rv, m, alpha};
    real s;
    sint err = laguerre_solve(s0, &s, &k);
    if (err == 0) kepler_fg_update(dt, s, r, m, e2, alpha, st);
    return err;
}

/* Retry with n = 1, 2, 4, ... equal sub-steps until every sub-step
 * converges, :466-507 */
static inline void kepler_substeps(real dt, real m, real e2,
                                   const real in[6], real out[6])
{
    sint n = 1;
    for (;;) {
        real st[6] = {in[0], in[1], in[2], in[3], in[4], in[5]};
        sint i, failed = 0;
        for (i = 0; i < n; ++i) {
            if (kepler_step_once(dt / n, m, e2, st) != 0) { failed = 1; break; }
        }
        if (!failed) {
            for (int q = 0; q < 6; ++q) out[q] = st[q];
            return;
        }
        n *= 2;
    }
}

/* Softened Kepler drift with an energy check that doubles the number of
 * sub-steps until the softened two-body energy is conserved, :510-614 */
static inline void kepler_drift(real dt, real m, real e2, const real in[6],
                                real out[6])
{
    real st[6];
    kepler_substeps(dt, m, e2, in, st);
    if (e2 == 0) {
        for (int q = 0; q < 6; ++q) out[q] = st[q];
        return;
    }
    real r2 = in[0] * in[0] + in[1] * in[1] + in[2] * in[2];
    if (!(r2 > 0)) {
        for (int q = 0; q < 6; ++q) out[q] = in[q];
        return;
    }
    r2 += e2;
    real r = R_SQRT(r2);
    real v2 = in[3] * in[3] + in[4] * in[4] + in[5] * in[5];
    real u0 = 2 * m / r;
    real e0 = v2 - u0;

    r2 = st[0] * st[0] + st[1] * st[1] + st[2] * st[2];
    r2 += e2;
    r = R_SQRT(r2);
    v2 = st[3] * st[3] + st[4] * st[4] + st[5] * st[5];
    real u1 = 2 * m / r;
    real e1 = v2 - u1;

    sint n = 1;
    real tol = 64 * KEPLER_TOL;
    real err = R_FABS(e1 - e0);
    if (2 * err < tol * (u1 + u0)) {
        for (int q = 0; q < 6; ++q) out[q] = st[q];
        return;
    }
    for (;;) {
        n *= 2;
        for (int q = 0; q < 6; ++q) st[q] = in[q];
        sint i, redo = 0;
        for (i = 0; i < n; ++i) {
            real nxt[6];
            kepler_substeps(dt / n, m, e2, st, nxt);
            for (int q = 0; q < 6; ++q) st[q] = nxt[q];
            r2 = st[0] * st[0] + st[1] * st[1] + st[2] * st[2];
            r2 += e2;
            r = R_SQRT(r2);
            v2 = st[3] * st[3] + st[4] * st[4] + st[5] * st[5];
            u1 = 2 * m / r;
            e1 = v2 - u1;
            err = R_FABS(e1 - e0);
            if (2 * err > tol * (u1 + u0)) { redo = 1; break; }
        }
        if (!redo) break;
    }
    for (int q = 0; q < 6; ++q) out[q] = st[q];
}

/* Softened kick-drift-kick for wide pairs, sakura_kernel_common.h:24-91 */
static inline void pair_leapfrog(real dt, real m, real e2, real st[6])
{
    real dt_2 = dt / 2;
    real rx = st[0], ry = st[1], rz = st[2];
    real vx = st[3], vy = st[4], vz = st[5];
    rx += vx * dt_2;
    ry += vy * dt_2;
    rz += vz * dt_2;
    real r2 = rx * rx + ry * ry + rz * rz;
    real inv_r2 = soft_inv_r2(r2, e2);
    real inv_r3 = inv_r2 * R_SQRT(inv_r2);
    real m_r3 = m * inv_r3;
    real ax = -m_r3 * rx;
    real ay = -m_r3 * ry;
    real az = -m_r3 * rz;
    vx += ax * dt;
    vy += ay * dt;
    vz += az * dt;
    rx += vx * dt_2;
    ry += vy * dt_2;
    rz += vz * dt_2;
    st[0] = rx; st[1] = ry; st[2] = rz;
    st[3] = vx; st[4] = vy; st[5] = vz;
}

/* leapfrog when r > 64 m / v^2, Kepler otherwise, :94-123 */
static inline void pair_twobody(real dt, real m, real e2, real st[6])
{
    real r2 = st[0] * st[0] + st[1] * st[1] + st[2] * st[2];
    real v2 = st[3] * st[3] + st[4] * st[4] + st[5] * st[5];
    real R = 64 * (m / v2);
    if (r2 > R * R) {
        pair_leapfrog(dt, m, e2, st);
    } else {
        real out[6];
        kepler_drift(dt, m, e2, st, out);
        for (int q = 0; q < 6; ++q) st[q] = out[q];
    }
}

/* flag semantics, :126-191 */
static inline void pair_evolve(real dt, sint flag, real m, real e2, real st[6])
{
    if (flag == -1) {
        st[0] -= st[3] * dt;
        st[1] -= st[4] * dt;
        st[2] -= st[5] * dt;
        pair_twobody(dt, m, e2, st);
    }
    if (flag == 1) {
        pair_twobody(dt, m, e2, st);
        st[0] -= st[3] * dt;
        st[1] -= st[4] * dt;
        st[2] -= st[5] * dt;
    }
    if (flag == -2) {
        st[0] -= st[3] * dt / 2;
        st[1] -= st[4] * dt / 2;
        st[2] -= st[5] * dt / 2;
        pair_twobody(dt, m, e2, st);
        st[0] -= st[3] * dt / 2;
        st[1] -= st[4] * dt / 2;
        st[2] -= st[5] * dt / 2;
    }
    if (flag == 2) {
        pair_twobody(dt / 2, m, e2, st);
        st[0] -= st[3] * dt;
        st[1] -= st[4] * dt;
        st[2] -= st[5] * dt;
        pair_twobody(dt / 2, m, e2, st);
    }
}

/* sakura_kernel_common.h:194-242, driver sakura_kernel.c:5-62 */
EXPORT void oracle_sakura_kernel(
    const uint_t ni, RP im, RP irx, RP iry, RP irz, RP ie2,
    RP ivx, RP ivy, RP ivz,
    const uint_t nj, RP jm, RP jrx, RP jry, RP jrz, RP je2,
    RP jvx, RP jvy, RP jvz,
    const real dt, const sint flag,
    WP idrx, WP idry, WP idrz, WP idvx, WP idvy, WP idvz)
{
    for (uint_t i = 0; i < ni; ++i) {
        real d[6] = {0, 0, 0, 0, 0, 0};
        for (uint_t j = 0; j < nj; ++j) {
            real r0x = irx[i] - jrx[j];
            real r0y = iry[i] - jry[j];
            real r0z = irz[i] - jrz[j];
            real e2 = ie2[i] + je2[j];
            real v0x = ivx[i] - jvx[j];
            real v0y = ivy[i] - jvy[j];
            real v0z = ivz[i] - jvz[j];
            real m = im[i] + jm[j];
            real st[6] = {r0x, r0y, r0z, v0x, v0y, v0z};
            pair_evolve(dt, flag, m, e2, st);
            real jmu = jm[j] / m;
            d[0] += jmu * (st[0] - r0x);
            d[1] += jmu * (st[1] - r0y);
            d[2] += jmu * (st[2] - r0z);
            d[3] += jmu * (st[3] - v0x);
            d[4] += jmu * (st[4] - v0y);
            d[5] += jmu * (st[5] - v0z);
        }
        idrx[i] = d[0]; idry[i] = d[1]; idrz[i] = d[2];
        idvx[i] = d[3]; idvy[i] = d[4]; idvz[i] = d[5];
    }
}

/* One relative-orbit Sakura evolution for a single pair, exported so tests
 * can check the two-body solver against closed forms (period return,
 * energy and angular-momentum conservation) without the pair sum. */
EXPORT void oracle_twobody_evolve(real dt, sint flag, real m, real e2,
                                  real *state6)
{
    pair_evolve(dt, flag, m, e2, state6);
}

/* ------------------------------------------------------------- NREG X */
/* nreg_kernels_common.h:7-58, driver nreg_kernels.c:5-61: pair sums at the
 * drifted separation r + v dt; u is finished as m_i * sum_j m_j / s. */
EXPORT void oracle_nreg_Xkernel(
    const uint_t ni, RP im, RP irx, RP iry, RP irz, RP ie2,
    RP ivx, RP ivy, RP ivz,
    const uint_t nj, RP jm, RP jrx, RP jry, RP jrz, RP je2,
    RP jvx, RP jvy, RP jvz,
    const real dt,
    WP idrx, WP idry, WP idrz, WP iax, WP iay, WP iaz, WP iu)
{
    for (uint_t i = 0; i < ni; ++i) {
        real drx = 0, dry = 0, drz = 0, ax = 0, ay = 0, az = 0, u = 0;
        for (uint_t j = 0; j < nj; ++j) {
            real rx = irx[i] - jrx[j];
            real ry = iry[i] - jry[j];
            real rz = irz[i] - jrz[j];
            real e2 = ie2[i] + je2[j];
            real vx = ivx[i] - jvx[j];
            real vy = ivy[i] - jvy[j];
            real vz = ivz[i] - jvz[j];
            rx += vx * dt;
            ry += vy * dt;
            rz += vz * dt;
            real r2 = rx * rx + ry * ry + rz * rz;
            real inv_r2 = soft_inv_r2(r2, e2);
            real inv_r1 = R_SQRT(inv_r2);
            real inv_r3 = inv_r2 * inv_r1;
            inv_r3 *= jm[j];
            drx += jm[j] * rx;
            dry += jm[j] * ry;
            drz += jm[j] * rz;
            ax -= inv_r3 * rx;
            ay -= inv_r3 * ry;
            az -= inv_r3 * rz;
            u += jm[j] * inv_r1;
        }
        idrx[i] = drx; idry[i] = dry; idrz[i] = drz;
        iax[i] = ax; iay[i] = ay; iaz[i] = az;
        iu[i] = im[i] * u;
    }
}

/* ------------------------------------------------------------- NREG V */
/* nreg_kernels_common.h:61-99, driver nreg_kernels.c:64-115: pair sums at
 * the kicked relative velocity v + a dt; k is finished as m_i * sum. */
EXPORT void oracle_nreg_Vkernel(
    const uint_t ni, RP im, RP ivx, RP ivy, RP ivz, RP iax, RP iay, RP iaz,
    const uint_t nj, RP jm, RP jvx, RP jvy, RP jvz, RP jax, RP jay, RP jaz,
    const real dt,
    WP idvx, WP idvy, WP idvz, WP ik)
{
    for (uint_t i = 0; i < ni; ++i) {
        real dvx = 0, dvy = 0, dvz = 0, k = 0;
        for (uint_t j = 0; j < nj; ++j) {
            real vx = ivx[i] - jvx[j];
            real vy = ivy[i] - jvy[j];
            real vz = ivz[i] - jvz[j];
            real ax = iax[i] - jax[j];
            real ay = iay[i] - jay[j];
            real az = iaz[i] - jaz[j];
            vx += ax * dt;
            vy += ay * dt;
            vz += az * dt;
            real v2 = vx * vx + vy * vy + vz * vz;
            dvx += jm[j] * vx;
            dvy += jm[j] * vy;
            dvz += jm[j] * vz;
            k += jm[j] * v2;
        }
        idvx[i] = dvx; idvy[i] = dvy; idvz[i] = dvz;
        ik[i] = im[i] * k;
    }
}

/* ------------------------------------------------------------- Kepler */
/* kepler_solver_kernel_common.h:8-79, driver kepler_solver_kernel.c:5-49:
 * particles 0 and 1 advanced by dt along their (softened) Kepler orbit
 * about the centre of mass, which drifts freely. */
EXPORT void oracle_kepler_solver_kernel(
    RP m, RP rx, RP ry, RP rz, RP e2_, RP vx, RP vy, RP vz,
    const real dt,
    WP r1x, WP r1y, WP r1z, WP v1x, WP v1y, WP v1z)
{
    real r0x = rx[0] - rx[1];
    real r0y = ry[0] - ry[1];
    real r0z = rz[0] - rz[1];
    real e2 = e2_[0] + e2_[1];
    real v0x = vx[0] - vx[1];
    real v0y = vy[0] - vy[1];
    real v0z = vz[0] - vz[1];
    real mt = m[0] + m[1];
    real imu = m[0] / mt;
    real jmu = m[1] / mt;
    real rcmx = imu * rx[0] + jmu * rx[1];
    real rcmy = imu * ry[0] + jmu * ry[1];
    real rcmz = imu * rz[0] + jmu * rz[1];
    real vcmx = imu * vx[0] + jmu * vx[1];
    real vcmy = imu * vy[0] + jmu * vy[1];
    real vcmz = imu * vz[0] + jmu * vz[1];
    rcmx += vcmx * dt;
    rcmy += vcmy * dt;
    rcmz += vcmz * dt;
    real in[6] = {r0x, r0y, r0z, v0x, v0y, v0z};
    real o[6];
    kepler_drift(dt, mt, e2, in, o);
    r1x[0] = rcmx + jmu * o[0];
    r1y[0] = rcmy + jmu * o[1];
    r1z[0] = rcmz + jmu * o[2];
    v1x[0] = vcmx + jmu * o[3];
    v1y[0] = vcmy + jmu * o[4];
    v1z[0] = vcmz + jmu * o[5];
    r1x[1] = rcmx - imu * o[0];
    r1y[1] = rcmy - imu * o[1];
    r1z[1] = rcmz - imu * o[2];
    v1x[1] = vcmx - imu * o[3];
    v1y[1] = vcmy - imu * o[4];
    v1z[1] = vcmz - imu * o[5];
}
