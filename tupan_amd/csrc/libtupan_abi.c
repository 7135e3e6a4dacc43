/*
 * libtupan_abi.c -- drop-in replacement for the reference's kernel
 * libraries.  Built twice (tupan_amd/build.py):
 *
 *   libtupan_f64.so  (-DTUPAN_ABI_F64): REAL = double, INT = long,
 *                    UINT = unsigned long
 *   libtupan_f32.so                   : REAL = float, INT = int,
 *                    UINT = unsigned int
 *
 * exactly as the reference builds its FP64 and FP32 libraries with the same
 * symbol names (tupan/lib/cffi_backend.py:31-87, tupan/lib/src/c_common.h:
 * 7-43).  Each exported function has the signature of its declaration in
 * tupan/lib/src/libtupan.h and forwards to the synchronous host-pointer
 * entry points of libtupan_hip.so in the library's mode (tupan_abi_set_mode:
 * the FP64 library follows tupan_set_default_mode, the FP32 library computes
 * in the reference FP32 build's arithmetic where that mode exists).  FP32
 * arrays are widened to FP64 for the device (exactly) and the results
 * narrowed back (exactly, for FP32 results).
 *
 * The reference functions return void and report nothing.  On a device or
 * argument failure these shims write NaN to every output, print the error
 * once to stderr and record it for tupan_abi_last_status(); a call whose
 * results hold NaN or inf records TUPAN_ERR_NONFINITE there (results kept).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "libtupan_abi.h"
#include "tupan_hip.h"

#define EXPORT __attribute__((visibility("default")))

static int g_status = 0;

EXPORT int tupan_abi_last_status(void) { return g_status; }

/* The library's arithmetic mode: -1 follows the process default
 * (tupan_set_default_mode).  The FP32 library starts in
 * TUPAN_MODE_FP32_IEEE, the reference's FP32 build bit for bit, for the
 * kernels that have that mode (Phi, Acc, AccJerk, SnapCrackle, Tstep); the
 * others run FP64 there. */
#ifdef TUPAN_ABI_F64
static int g_mode = -1;
#else
static int g_mode = TUPAN_MODE_FP32_IEEE;
#endif

EXPORT int tupan_abi_set_mode(int mode)
{
    if (mode < -1 || mode >= TUPAN_MODE_COUNT) return TUPAN_ERR_ARG;
    g_mode = mode;
    return TUPAN_OK;
}

EXPORT int tupan_abi_get_mode(void) { return g_mode; }

static int kernel_mode(int has_fp32_ieee)
{
    int mode = g_mode < 0 ? tupan_get_default_mode() : g_mode;
    if (mode == TUPAN_MODE_FP32_IEEE && !has_fp32_ieee) mode = TUPAN_MODE_FP64;
    return mode;
}

static void report(int rc, const char *who, size_t ni, REAL **outs, int nout)
{
    g_status = rc;
    if (rc == TUPAN_OK) {
        /* results written: note NaN/inf among them (the reference would
         * return them silently) */
        for (int k = 0; k < nout && g_status == TUPAN_OK; ++k)
            for (size_t i = 0; i < ni; ++i)
                if (!isfinite(outs[k][i])) {
                    g_status = TUPAN_ERR_NONFINITE;
                    break;
                }
        return;
    }
    static int printed = 0;
    if (!printed) {
        fprintf(stderr, "%s: %s\n", who, tupan_last_error());
        printed = 1;
    }
    if (rc == TUPAN_ERR_NOT_CONVERGED) return; /* results are valid */
    for (int k = 0; k < nout; ++k)
        for (size_t i = 0; i < ni; ++i) outs[k][i] = (REAL)NAN;
}

#ifdef TUPAN_ABI_F64
/* FP64: pass the caller's arrays straight through. */
#define WIDEN(dst, src, n) ((dst) = (const double *)(src), 0)
#define UNWIDEN(dst, src, n) ((void)0)
#define ALLOC_OUT(dst, src, n) ((dst) = (double *)(src), 0)
#define FREE_W(p) ((void)0)
#define FREE_O(p) ((void)0)
#else
static int widen(const double **dst, const float *src, size_t n)
{
    double *d = (double *)malloc((n ? n : 1) * sizeof(double));
    if (!d) return 1;
    for (size_t i = 0; i < n; ++i) d[i] = (double)src[i];
    *dst = d;
    return 0;
}
static int alloc_out(double **dst, size_t n)
{
    *dst = (double *)malloc((n ? n : 1) * sizeof(double));
    return *dst == NULL;
}
#define WIDEN(dst, src, n) widen(&(dst), (src), (n))
#define ALLOC_OUT(dst, src, n) alloc_out(&(dst), (n))
#define UNWIDEN(dst, src, n) \
    do { for (size_t q_ = 0; q_ < (n); ++q_) (dst)[q_] = (float)(src)[q_]; } while (0)
#define FREE_W(p) free((void *)(p))
#define FREE_O(p) free((void *)(p))
#endif

typedef int (*mode_fn)(int, size_t, const double *const *, size_t,
                       const double *const *, double *const *);

/* Generic body: nf input fields per side, nout outputs. */
static void forward(const char *who, int kind, mode_fn fn, size_t ni,
                    const REAL *const *ih, size_t nj, const REAL *const *jh,
                    int nf, REAL **oh, int nout, double eta, unsigned order,
                    const double *inv, double dt, int flag)
{
    const double *iw[14] = {0}, *jw[14] = {0};
    double *ow[8] = {0};
    int rc = TUPAN_OK, bad = 0;
    if (ni == 0) return;
    for (int k = 0; k < nf && !bad; ++k) bad |= WIDEN(iw[k], ih[k], ni);
    for (int k = 0; k < nf && !bad; ++k) {
#ifdef TUPAN_ABI_F64
        bad |= WIDEN(jw[k], jh[k], nj);
#else
        if (jh[k] == ih[k] && nj == ni) jw[k] = iw[k];
        else bad |= WIDEN(jw[k], jh[k], nj);
#endif
    }
    for (int k = 0; k < nout && !bad; ++k) bad |= ALLOC_OUT(ow[k], oh[k], ni);
    if (bad) {
        rc = TUPAN_ERR_ARG;
    } else {
        /* kinds 0 (Phi, Acc, AccJerk, SnapCrackle) and 1 (Tstep) have the
         * FP32 IEEE mode */
        int mode = kernel_mode(kind <= 1);
        switch (kind) {
        case 0: rc = fn(mode, ni, iw, nj, jw, ow); break;
        case 1: rc = tupan_tstep_host(mode, ni, iw, nj, jw, eta, ow); break;
        case 2: rc = tupan_pnacc_host(mode, ni, iw, nj, jw, order, inv, ow); break;
        case 3: rc = tupan_sakura_host(ni, iw, nj, jw, dt, flag, ow); break;
        case 4: rc = tupan_nreg_x_host(mode, ni, iw, nj, jw, dt, ow); break;
        case 5: rc = tupan_nreg_v_host(mode, ni, iw, nj, jw, dt, ow); break;
        }
    }
    for (int k = 0; k < nout; ++k) {
        if (ow[k]) UNWIDEN(oh[k], ow[k], ni);
        FREE_O(ow[k]);
    }
#ifndef TUPAN_ABI_F64
    for (int k = 0; k < nf; ++k) {
        if (jw[k] && jw[k] != iw[k]) FREE_W(jw[k]);
        if (iw[k]) FREE_W(iw[k]);
    }
#endif
    report(rc, who, ni, oh, nout);
}

#define R const REAL *restrict

/* libtupan.h:2-15 */
EXPORT void phi_kernel(const UINT ni, R im, R irx, R iry, R irz, R ie2,
                       const UINT nj, R jm, R jrx, R jry, R jrz, R je2,
                       REAL *restrict iphi)
{
    const REAL *i5[5] = {im, irx, iry, irz, ie2};
    const REAL *j5[5] = {jm, jrx, jry, jrz, je2};
    REAL *o[1] = {iphi};
    forward("phi_kernel", 0, tupan_phi_host, ni, i5, nj, j5, 5, o, 1, 0, 0,
            NULL, 0, 0);
}

/* libtupan.h:17-32 */
EXPORT void acc_kernel(const UINT ni, R im, R irx, R iry, R irz, R ie2,
                       const UINT nj, R jm, R jrx, R jry, R jrz, R je2,
                       REAL *restrict iax, REAL *restrict iay,
                       REAL *restrict iaz)
{
    const REAL *i5[5] = {im, irx, iry, irz, ie2};
    const REAL *j5[5] = {jm, jrx, jry, jrz, je2};
    REAL *o[3] = {iax, iay, iaz};
    forward("acc_kernel", 0, tupan_acc_host, ni, i5, nj, j5, 5, o, 3, 0, 0,
            NULL, 0, 0);
}

/* libtupan.h:34-58 */
EXPORT void acc_jerk_kernel(const UINT ni, R im, R irx, R iry, R irz, R ie2,
                            R ivx, R ivy, R ivz, const UINT nj, R jm, R jrx,
                            R jry, R jrz, R je2, R jvx, R jvy, R jvz,
                            REAL *restrict iax, REAL *restrict iay,
                            REAL *restrict iaz, REAL *restrict ijx,
                            REAL *restrict ijy, REAL *restrict ijz)
{
    const REAL *i8[8] = {im, irx, iry, irz, ie2, ivx, ivy, ivz};
    const REAL *j8[8] = {jm, jrx, jry, jrz, je2, jvx, jvy, jvz};
    REAL *o[6] = {iax, iay, iaz, ijx, ijy, ijz};
    forward("acc_jerk_kernel", 0, tupan_acc_jerk_host, ni, i8, nj, j8, 8, o,
            6, 0, 0, NULL, 0, 0);
}

/* libtupan.h:60-96 */
EXPORT void snap_crackle_kernel(
    const UINT ni, R im, R irx, R iry, R irz, R ie2, R ivx, R ivy, R ivz,
    R iax, R iay, R iaz, R ijx, R ijy, R ijz, const UINT nj, R jm, R jrx,
    R jry, R jrz, R je2, R jvx, R jvy, R jvz, R jax, R jay, R jaz, R jjx,
    R jjy, R jjz, REAL *restrict isx, REAL *restrict isy, REAL *restrict isz,
    REAL *restrict icx, REAL *restrict icy, REAL *restrict icz)
{
    const REAL *i14[14] = {im, irx, iry, irz, ie2, ivx, ivy, ivz,
                           iax, iay, iaz, ijx, ijy, ijz};
    const REAL *j14[14] = {jm, jrx, jry, jrz, je2, jvx, jvy, jvz,
                           jax, jay, jaz, jjx, jjy, jjz};
    REAL *o[6] = {isx, isy, isz, icx, icy, icz};
    forward("snap_crackle_kernel", 0, tupan_snap_crackle_host, ni, i14, nj,
            j14, 14, o, 6, 0, 0, NULL, 0, 0);
}

/* libtupan.h:98-119 */
EXPORT void tstep_kernel(const UINT ni, R im, R irx, R iry, R irz, R ie2,
                         R ivx, R ivy, R ivz, const UINT nj, R jm, R jrx,
                         R jry, R jrz, R je2, R jvx, R jvy, R jvz,
                         const REAL eta, REAL *restrict idt_a,
                         REAL *restrict idt_b)
{
    const REAL *i8[8] = {im, irx, iry, irz, ie2, ivx, ivy, ivz};
    const REAL *j8[8] = {jm, jrx, jry, jrz, je2, jvx, jvy, jvz};
    REAL *o[2] = {idt_a, idt_b};
    forward("tstep_kernel", 1, NULL, ni, i8, nj, j8, 8, o, 2, (double)eta, 0,
            NULL, 0, 0);
}

/* libtupan.h:121-150 */
EXPORT void pnacc_kernel(const UINT ni, R im, R irx, R iry, R irz, R ie2,
                         R ivx, R ivy, R ivz, const UINT nj, R jm, R jrx,
                         R jry, R jrz, R je2, R jvx, R jvy, R jvz, UINT order,
                         const REAL inv1, const REAL inv2, const REAL inv3,
                         const REAL inv4, const REAL inv5, const REAL inv6,
                         const REAL inv7, REAL *restrict ipnax,
                         REAL *restrict ipnay, REAL *restrict ipnaz)
{
    const REAL *i8[8] = {im, irx, iry, irz, ie2, ivx, ivy, ivz};
    const REAL *j8[8] = {jm, jrx, jry, jrz, je2, jvx, jvy, jvz};
    REAL *o[3] = {ipnax, ipnay, ipnaz};
    double inv[7] = {inv1, inv2, inv3, inv4, inv5, inv6, inv7};
    forward("pnacc_kernel", 2, NULL, ni, i8, nj, j8, 8, o, 3, 0,
            (unsigned)order, inv, 0, 0);
}

/* libtupan.h:203-229 */
EXPORT void sakura_kernel(const UINT ni, R im, R irx, R iry, R irz, R ie2,
                          R ivx, R ivy, R ivz, const UINT nj, R jm, R jrx,
                          R jry, R jrz, R je2, R jvx, R jvy, R jvz,
                          const REAL dt, const INT flag, REAL *restrict idrx,
                          REAL *restrict idry, REAL *restrict idrz,
                          REAL *restrict idvx, REAL *restrict idvy,
                          REAL *restrict idvz)
{
    const REAL *i8[8] = {im, irx, iry, irz, ie2, ivx, ivy, ivz};
    const REAL *j8[8] = {jm, jrx, jry, jrz, je2, jvx, jvy, jvz};
    REAL *o[6] = {idrx, idry, idrz, idvx, idvy, idvz};
    forward("sakura_kernel", 3, NULL, ni, i8, nj, j8, 8, o, 6, 0, 0, NULL,
            (double)dt, (int)flag);
}

/* libtupan.h:152-179 */
EXPORT void nreg_Xkernel(const UINT ni, R im, R irx, R iry, R irz, R ie2,
                         R ivx, R ivy, R ivz, const UINT nj, R jm, R jrx,
                         R jry, R jrz, R je2, R jvx, R jvy, R jvz,
                         const REAL dt, REAL *restrict idrx,
                         REAL *restrict idry, REAL *restrict idrz,
                         REAL *restrict iax, REAL *restrict iay,
                         REAL *restrict iaz, REAL *restrict iu)
{
    const REAL *i8[8] = {im, irx, iry, irz, ie2, ivx, ivy, ivz};
    const REAL *j8[8] = {jm, jrx, jry, jrz, je2, jvx, jvy, jvz};
    REAL *o[7] = {idrx, idry, idrz, iax, iay, iaz, iu};
    forward("nreg_Xkernel", 4, NULL, ni, i8, nj, j8, 8, o, 7, 0, 0, NULL,
            (double)dt, 0);
}

/* libtupan.h:181-201 */
EXPORT void nreg_Vkernel(const UINT ni, R im, R ivx, R ivy, R ivz, R iax,
                         R iay, R iaz, const UINT nj, R jm, R jvx, R jvy,
                         R jvz, R jax, R jay, R jaz, const REAL dt,
                         REAL *restrict idvx, REAL *restrict idvy,
                         REAL *restrict idvz, REAL *restrict ik)
{
    const REAL *i7[7] = {im, ivx, ivy, ivz, iax, iay, iaz};
    const REAL *j7[7] = {jm, jvx, jvy, jvz, jax, jay, jaz};
    REAL *o[4] = {idvx, idvy, idvz, ik};
    forward("nreg_Vkernel", 5, NULL, ni, i7, nj, j7, 7, o, 4, 0, 0, NULL,
            (double)dt, 0);
}

/* libtupan.h:231-246: particles 0 and 1, outputs may alias the inputs
 * (the reference's Kepler extension updates them in place) -- all inputs
 * are read before any output is written. */
EXPORT void kepler_solver_kernel(const REAL *im, const REAL *irx,
                                 const REAL *iry, const REAL *irz,
                                 const REAL *ie2, const REAL *ivx,
                                 const REAL *ivy, const REAL *ivz,
                                 const REAL dt, REAL *ir1x, REAL *ir1y,
                                 REAL *ir1z, REAL *iv1x, REAL *iv1y,
                                 REAL *iv1z)
{
    const REAL *in[8] = {im, irx, iry, irz, ie2, ivx, ivy, ivz};
    REAL *out[6] = {ir1x, ir1y, ir1z, iv1x, iv1y, iv1z};
    double buf[8][2], res[6][2];
    const double *ip[8];
    double *op[6];
    for (int k = 0; k < 8; ++k) {
        buf[k][0] = (double)in[k][0];
        buf[k][1] = (double)in[k][1];
        ip[k] = buf[k];
    }
    for (int k = 0; k < 6; ++k) op[k] = res[k];
    int rc = tupan_kepler_host(ip, (double)dt, op);
    g_status = rc;
    if (rc != TUPAN_OK) {
        static int printed = 0;
        if (!printed) {
            fprintf(stderr, "kepler_solver_kernel: %s\n", tupan_last_error());
            printed = 1;
        }
        if (rc != TUPAN_ERR_NOT_CONVERGED) {
            for (int k = 0; k < 6; ++k) out[k][0] = out[k][1] = (REAL)NAN;
            return;
        }
    }
    for (int k = 0; k < 6; ++k) {
        out[k][0] = (REAL)res[k][0];
        out[k][1] = (REAL)res[k][1];
    }
}
