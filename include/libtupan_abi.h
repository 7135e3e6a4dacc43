/*
 * libtupan_abi.h -- the drop-in libraries libtupan_f64.so and
 * libtupan_f32.so.  They export exactly the kernel symbols of the
 * reference's tupan/lib/src/libtupan.h:2-246 with the same argument lists,
 * for REAL = double, INT = long, UINT = unsigned long (libtupan_f64.so,
 * compile users with -DTUPAN_ABI_F64) or REAL = float, INT = int,
 * UINT = unsigned int (libtupan_f32.so), as the reference builds its two
 * precisions (tupan/lib/cffi_backend.py:31-87, c_common.h:7-43).
 *
 * Each call is synchronous and runs on the MI355X through libtupan_hip.so
 * in the library's mode (tupan_abi_set_mode below): libtupan_f64.so follows
 * the process default (tupan_set_default_mode; FP64 unless set);
 * libtupan_f32.so computes Phi, Acc, AccJerk, SnapCrackle and Tstep in the
 * reference FP32 build's arithmetic (TUPAN_MODE_FP32_IEEE: its results bit
 * for bit) and the other kernels in FP64, rounded to float.
 * The reference functions report nothing; on failure these write NaN to
 * every output and tupan_abi_last_status() returns the TUPAN_ERR_* code of
 * the last call: 0 on success, TUPAN_ERR_NOT_CONVERGED when a Sakura or
 * Kepler solve hit its cap (results written), TUPAN_ERR_NONFINITE when the
 * results hold NaN or inf (results written).  All ten libtupan.h kernels are
 * provided, nreg_Xkernel, nreg_Vkernel and kepler_solver_kernel
 * (libtupan.h:152-201, 231-246) included.
 */
#ifndef LIBTUPAN_ABI_H
#define LIBTUPAN_ABI_H

#ifdef TUPAN_ABI_F64
typedef double REAL;
typedef long INT;
typedef unsigned long UINT;
#else
typedef float REAL;
typedef int INT;
typedef unsigned int UINT;
#endif

#ifdef __cplusplus
extern "C" {
#endif

int tupan_abi_last_status(void);
/* No counterpart in the reference: the library's arithmetic mode, a
 * TUPAN_MODE_* of tupan_hip.h or -1 (follow tupan_set_default_mode);
 * kernels without TUPAN_MODE_FP32_IEEE run FP64 in that mode.  Returns 0,
 * or TUPAN_ERR_ARG for an unknown mode. */
int tupan_abi_set_mode(int mode);
int tupan_abi_get_mode(void);

/* libtupan.h:2-15 */
void phi_kernel(const UINT ni, const REAL *_im, const REAL *_irx,
                const REAL *_iry, const REAL *_irz, const REAL *_ie2,
                const UINT nj, const REAL *_jm, const REAL *_jrx,
                const REAL *_jry, const REAL *_jrz, const REAL *_je2,
                REAL *_iphi);

/* libtupan.h:17-32 */
void acc_kernel(const UINT ni, const REAL *_im, const REAL *_irx,
                const REAL *_iry, const REAL *_irz, const REAL *_ie2,
                const UINT nj, const REAL *_jm, const REAL *_jrx,
                const REAL *_jry, const REAL *_jrz, const REAL *_je2,
                REAL *_iax, REAL *_iay, REAL *_iaz);

/* libtupan.h:34-58 */
void acc_jerk_kernel(const UINT ni, const REAL *_im, const REAL *_irx,
                     const REAL *_iry, const REAL *_irz, const REAL *_ie2,
                     const REAL *_ivx, const REAL *_ivy, const REAL *_ivz,
                     const UINT nj, const REAL *_jm, const REAL *_jrx,
                     const REAL *_jry, const REAL *_jrz, const REAL *_je2,
                     const REAL *_jvx, const REAL *_jvy, const REAL *_jvz,
                     REAL *_iax, REAL *_iay, REAL *_iaz, REAL *_ijx,
                     REAL *_ijy, REAL *_ijz);

/* libtupan.h:60-96 */
void snap_crackle_kernel(
    const UINT ni, const REAL *_im, const REAL *_irx, const REAL *_iry,
    const REAL *_irz, const REAL *_ie2, const REAL *_ivx, const REAL *_ivy,
    const REAL *_ivz, const REAL *_iax, const REAL *_iay, const REAL *_iaz,
    const REAL *_ijx, const REAL *_ijy, const REAL *_ijz, const UINT nj,
    const REAL *_jm, const REAL *_jrx, const REAL *_jry, const REAL *_jrz,
    const REAL *_je2, const REAL *_jvx, const REAL *_jvy, const REAL *_jvz,
    const REAL *_jax, const REAL *_jay, const REAL *_jaz, const REAL *_jjx,
    const REAL *_jjy, const REAL *_jjz, REAL *_isx, REAL *_isy, REAL *_isz,
    REAL *_icx, REAL *_icy, REAL *_icz);

/* libtupan.h:98-119 */
void tstep_kernel(const UINT ni, const REAL *_im, const REAL *_irx,
                  const REAL *_iry, const REAL *_irz, const REAL *_ie2,
                  const REAL *_ivx, const REAL *_ivy, const REAL *_ivz,
                  const UINT nj, const REAL *_jm, const REAL *_jrx,
                  const REAL *_jry, const REAL *_jrz, const REAL *_je2,
                  const REAL *_jvx, const REAL *_jvy, const REAL *_jvz,
                  const REAL eta, REAL *_idt_a, REAL *_idt_b);

/* libtupan.h:121-150 */
void pnacc_kernel(const UINT ni, const REAL *_im, const REAL *_irx,
                  const REAL *_iry, const REAL *_irz, const REAL *_ie2,
                  const REAL *_ivx, const REAL *_ivy, const REAL *_ivz,
                  const UINT nj, const REAL *_jm, const REAL *_jrx,
                  const REAL *_jry, const REAL *_jrz, const REAL *_je2,
                  const REAL *_jvx, const REAL *_jvy, const REAL *_jvz,
                  UINT order, const REAL inv1, const REAL inv2,
                  const REAL inv3, const REAL inv4, const REAL inv5,
                  const REAL inv6, const REAL inv7, REAL *_ipnax,
                  REAL *_ipnay, REAL *_ipnaz);

/* libtupan.h:203-229 */
void sakura_kernel(const UINT ni, const REAL *_im, const REAL *_irx,
                   const REAL *_iry, const REAL *_irz, const REAL *_ie2,
                   const REAL *_ivx, const REAL *_ivy, const REAL *_ivz,
                   const UINT nj, const REAL *_jm, const REAL *_jrx,
                   const REAL *_jry, const REAL *_jrz, const REAL *_je2,
                   const REAL *_jvx, const REAL *_jvy, const REAL *_jvz,
                   const REAL dt, const INT flag, REAL *_idrx, REAL *_idry,
                   REAL *_idrz, REAL *_idvx, REAL *_idvy, REAL *_idvz);

/* libtupan.h:152-179 */
void nreg_Xkernel(const UINT ni, const REAL *_im, const REAL *_irx,
                  const REAL *_iry, const REAL *_irz, const REAL *_ie2,
                  const REAL *_ivx, const REAL *_ivy, const REAL *_ivz,
                  const UINT nj, const REAL *_jm, const REAL *_jrx,
                  const REAL *_jry, const REAL *_jrz, const REAL *_je2,
                  const REAL *_jvx, const REAL *_jvy, const REAL *_jvz,
                  const REAL dt, REAL *_idrx, REAL *_idry, REAL *_idrz,
                  REAL *_iax, REAL *_iay, REAL *_iaz, REAL *_iu);

/* libtupan.h:181-201 */
void nreg_Vkernel(const UINT ni, const REAL *_im, const REAL *_ivx,
                  const REAL *_ivy, const REAL *_ivz, const REAL *_iax,
                  const REAL *_iay, const REAL *_iaz, const UINT nj,
                  const REAL *_jm, const REAL *_jvx, const REAL *_jvy,
                  const REAL *_jvz, const REAL *_jax, const REAL *_jay,
                  const REAL *_jaz, const REAL dt, REAL *_idvx,
                  REAL *_idvy, REAL *_idvz, REAL *_ik);

/* libtupan.h:231-246 -- particles 0 and 1; outputs may alias inputs */
void kepler_solver_kernel(const REAL *_im, const REAL *_irx,
                          const REAL *_iry, const REAL *_irz,
                          const REAL *_ie2, const REAL *_ivx,
                          const REAL *_ivy, const REAL *_ivz, const REAL dt,
                          REAL *_ir1x, REAL *_ir1y, REAL *_ir1z,
                          REAL *_iv1x, REAL *_iv1y, REAL *_iv1z);

#ifdef __cplusplus
}
#endif
#endif /* LIBTUPAN_ABI_H */
