/*
 * tupan_hip.h -- C ABI of libtupan_hip.so, the MI355X (gfx950) build of
 * Tupan's O(N^2) direct-summation kernel family.
 *
 * Plain C: pointers, sizes, ints and doubles only.  Two flavours per kernel:
 *
 *   tupan_<kernel>_dev(...)   device pointers, asynchronous on `stream`
 *                             (a hipStream_t passed as void*; NULL = the
 *                             default stream).  No copies.
 *   tupan_<kernel>_host(...)  host pointers, synchronous: copies inputs to
 *                             HBM, runs, copies outputs back.  This is the
 *                             call the reference's C backend makes
 *                             (tupan/lib/cffi_backend.py:125-126).
 *
 * Inputs are passed as arrays of field pointers, `ip[k]` / `jp[k]`, in the
 * reference's argument order (tupan/lib/src/libtupan.h), all SoA FP64 of
 * length ni (i side) or nj (j side):
 *
 *   F5  = m, rx, ry, rz, e2
 *   F8  = F5, vx, vy, vz
 *   F14 = F8, ax, ay, az, jx, jy, jz
 *
 * Outputs `op[k]` likewise, in libtupan.h order.  `e2` holds half the pair
 * softening (a pair uses ie2 + je2).  Outputs are overwritten (not
 * accumulated) for i < ni; i and j arrays may alias (ips is jps); ni or
 * nj = 0 is valid (nj = 0 writes the empty sums).
 *
 * mode: TUPAN_MODE_FP64      -- FP64 arithmetic (oracle-grade)
 *       TUPAN_MODE_FP32C     -- FP32 pair arithmetic on FP64-exact position
 *                               differences with FP64 block accumulation
 *       TUPAN_MODE_FP64_IEEE -- the reference C's arithmetic bit for bit:
 *                               its operations in its order, no FMA
 *                               contraction, correctly rounded division and
 *                               sqrt, j summed in ascending order (one j
 *                               split); results equal the reference's bits
 *       TUPAN_MODE_FP32_IEEE -- the same for the reference's FP32 build:
 *                               each input rounded to float (exact for
 *                               values that came from float arrays), float
 *                               arithmetic in the reference's order, the
 *                               float results widened exactly
 *       (Phi, Acc, AccJerk and SnapCrackle have all four; Tstep has all but
 *       FP32C, where it runs FP64; PNAcc and NREG X/V have FP64 and
 *       FP64_IEEE, run FP64 for FP32C and refuse FP32_IEEE; Sakura and
 *       Kepler always run FP64 and take no mode).
 *
 * Every function returns 0 on success or a TUPAN_ERR_* code; the text of the
 * last error on the calling thread is tupan_last_error().
 */
#ifndef TUPAN_HIP_H
#define TUPAN_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TUPAN_OK 0
#define TUPAN_ERR_HIP 1           /* HIP runtime failure (no device, OOM) */
#define TUPAN_ERR_ARG 2           /* invalid argument (mode, NULL, size) */
#define TUPAN_ERR_NOT_CONVERGED 3 /* Sakura: a Kepler solve hit its cap */
#define TUPAN_ERR_NONFINITE 4     /* drop-in status: an output is NaN/inf */

/* Bits of a device status word (d_status arguments) */
#define TUPAN_STATUS_NOT_CONVERGED 1  /* a Kepler solve hit its cap */
#define TUPAN_STATUS_NONFINITE 2      /* tupan_check_finite_dev found NaN/inf */

#define TUPAN_MODE_FP64 0
#define TUPAN_MODE_FP32C 1
#define TUPAN_MODE_FP64_IEEE 2
#define TUPAN_MODE_FP32_IEEE 3
#define TUPAN_MODE_COUNT 4

/* kernel ids for tupan_set_launch */
#define TUPAN_K_PHI 0
#define TUPAN_K_ACC 1
#define TUPAN_K_ACC_JERK 2
#define TUPAN_K_SNAP_CRACKLE 3
#define TUPAN_K_TSTEP 4
#define TUPAN_K_PNACC 5
#define TUPAN_K_SAKURA 6
#define TUPAN_K_ACC_ADJOINT 7
#define TUPAN_K_PHI_ADJOINT 8
#define TUPAN_K_NREG_X 9
#define TUPAN_K_NREG_V 10
#define TUPAN_K_COUNT 11

/* ---- runtime ------------------------------------------------------- */
const char *tupan_version(void);
const char *tupan_last_error(void);
int tupan_device_count(int *count);
/* default mode of the libtupan.h drop-in libraries (initially FP64) */
int tupan_set_default_mode(int mode);
int tupan_get_default_mode(void);
/* Launch shape of the pair engine (autotuner hook): ipl = i-particles per
 * lane (1, 2 or 4; 0 = automatic), splits = number of j-range splits
 * (0 = automatic).  TUPAN_MODE_FP64_IEEE runs its own engine (one
 * i-particle per lane, one ascending j sweep): it ignores ipl and accepts
 * splits 0 or 1 only. */
int tupan_set_launch(int kernel_id, int mode, int ipl, int splits);
/* Live timing of the pair-engine kernel: while on, every engine launch is
 * bracketed by HIP events on its stream.  tupan_take_kernel_times waits for
 * the recorded launches, returns their summed duration (ms) and count, and
 * clears the record. */
int tupan_set_timing(int on);
/* FP32C near field of the single-system kernels: pairs with
 * r^2 < (m_j/<m>) (c nj^(-1/3))^2 are evaluated in FP64 (a threshold on
 * the pair force, c mean interparticle spacings of a unit-size system for
 * a mean-mass j).  c = 0 leaves only the r^2 = 0 pairs; default 5.  (The
 * ensemble's FP32C passes use their own factor, 0.5.) */
int tupan_set_near_field(double c);
/* FP32C calls with nj > 4096 run over Morton-sorted copies of the
 * particles: tiles whose bounding box lies beyond the near field skip the
 * per-pair near test, and rows whose estimated FP32 error exceeds 0.7 of
 * the largest measured error-to-estimate ratio (7.1e-7) are recomputed in
 * FP64.  on = 0 selects the unsorted engine for them (default 1). */
int tupan_set_sorted(int on);
/* Sorted FP32C calls recompute in FP64 the rows whose error estimate
 * sqrt(sum over tiles of |FP32 tile partial|^2) exceeds tau |result| in
 * some output group; default 0.7 (about 0.5% of the rows of a Plummer
 * model, every other row within ~6e-7 of FP64).  tau = 0 recomputes every
 * row with any FP32 tile. */
int tupan_set_flag_tau(double tau);
/* Rows of the last sorted FP32C call (on the calling process's last stream
 * used for one) that were flagged and recomputed in FP64; synchronises
 * that stream.  0 before any such call. */
int tupan_last_flagged(int *count);
/* Sakura: Kepler solves a pair may spend (default 2^20) before it is
 * reported as not converged. */
int tupan_set_sakura_budget(int solves);
int tupan_take_kernel_times(double *total_ms, int *count);
/* The asynchronous entry points keep their scratch (packed j records,
 * partial sums) in per-(device, stream) buffers that grow to the largest
 * call and are reused, so steady-state calls allocate nothing and can be
 * captured into a HIP graph once the same call has run outside the
 * capture.  A graph shares these buffers with later calls on its capture
 * stream: do not replay it concurrently with them.  tupan_release_scratch
 * synchronises the devices and frees every buffer (graphs captured before
 * must not be replayed afterwards). */
int tupan_release_scratch(void);

/* ---- Phi: replaces phi_kernel, tupan/lib/src/libtupan.h:2-15 ---------
 * ip, jp: F5.  op: phi. */
int tupan_phi_dev(int mode, size_t ni, const double *const *ip, size_t nj,
                  const double *const *jp, double *const *op, void *stream);
int tupan_phi_host(int mode, size_t ni, const double *const *ip, size_t nj,
                   const double *const *jp, double *const *op);

/* ---- Acc: replaces acc_kernel, libtupan.h:17-32 ----------------------
 * ip, jp: F5.  op: ax, ay, az. */
int tupan_acc_dev(int mode, size_t ni, const double *const *ip, size_t nj,
                  const double *const *jp, double *const *op, void *stream);
int tupan_acc_host(int mode, size_t ni, const double *const *ip, size_t nj,
                   const double *const *jp, double *const *op);

/* ---- AccJerk: replaces acc_jerk_kernel, libtupan.h:34-58 -------------
 * ip, jp: F8.  op: ax, ay, az, jx, jy, jz. */
int tupan_acc_jerk_dev(int mode, size_t ni, const double *const *ip,
                       size_t nj, const double *const *jp, double *const *op,
                       void *stream);
int tupan_acc_jerk_host(int mode, size_t ni, const double *const *ip,
                        size_t nj, const double *const *jp,
                        double *const *op);

/* ---- AccJerk without allocation -------------------------------------
 * No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  tupan_acc_jerk_dev in TUPAN_MODE_FP64 (same fields, outputs
 * and launch shape; the same bits whenever the j records and block-row
 * partials fit) whose scratch is a file-scope 32 MiB device array, one per
 * device, held until the process ends: the call allocates and frees
 * nothing.  The j split shrinks until the records (64 B per j, none for
 * nj <= 4096) and partials (48 B per i per block row) fit; a call whose
 * records alone exceed 32 MiB (nj > 524288) is refused with TUPAN_ERR_ARG.
 * Every call on a device shares that array: calls that overlap on one
 * device (different streams or host threads) race and must be serialised
 * by the caller.  Nothing in the array needs clearing between calls. */
int tupan_acc_jerk_static_dev(size_t ni, const double *const *ip,
                              size_t nj, const double *const *jp,
                              double *const *op, void *stream);

/* ---- SnapCrackle: replaces snap_crackle_kernel, libtupan.h:60-96 -----
 * ip, jp: F14.  op: sx, sy, sz, cx, cy, cz. */
int tupan_snap_crackle_dev(int mode, size_t ni, const double *const *ip,
                           size_t nj, const double *const *jp,
                           double *const *op, void *stream);
int tupan_snap_crackle_host(int mode, size_t ni, const double *const *ip,
                            size_t nj, const double *const *jp,
                            double *const *op);

/* ---- Tstep: replaces tstep_kernel, libtupan.h:98-119 -----------------
 * ip, jp: F8.  op: dt_a (-> tstep), dt_b (-> tstepij).  mode: FP64 (FP32C
 * has no Tstep variant and runs FP64) or FP64_IEEE. */
int tupan_tstep_dev(int mode, size_t ni, const double *const *ip, size_t nj,
                    const double *const *jp, double eta, double *const *op,
                    void *stream);
int tupan_tstep_host(int mode, size_t ni, const double *const *ip, size_t nj,
                     const double *const *jp, double eta, double *const *op);

/* ---- PNAcc: replaces pnacc_kernel, libtupan.h:121-150 ----------------
 * ip, jp: F8.  inv[0..6] = 1/c^1 .. 1/c^7.  op: pnax, pnay, pnaz.
 * mode: FP64 (FP32C runs FP64) or FP64_IEEE. */
int tupan_pnacc_dev(int mode, size_t ni, const double *const *ip, size_t nj,
                    const double *const *jp, unsigned order,
                    const double *inv, double *const *op, void *stream);
int tupan_pnacc_host(int mode, size_t ni, const double *const *ip, size_t nj,
                     const double *const *jp, unsigned order,
                     const double *inv, double *const *op);

/* ---- Sakura: replaces sakura_kernel, libtupan.h:203-229 --------------
 * ip, jp: F8.  op: drx, dry, drz, dvx, dvy, dvz.  d_status (device int,
 * may be NULL) is OR-ed with 1 when a pair's Kepler sub-stepping hit the
 * device cap (the reference would loop without bound); the host variant
 * then returns TUPAN_ERR_NOT_CONVERGED with the results still written. */
int tupan_sakura_dev(size_t ni, const double *const *ip, size_t nj,
                     const double *const *jp, double dt, int flag,
                     double *const *op, int *d_status, void *stream);
int tupan_sakura_host(size_t ni, const double *const *ip, size_t nj,
                      const double *const *jp, double dt, int flag,
                      double *const *op);

/* ---- NREG X: replaces nreg_Xkernel, libtupan.h:152-179 ---------------
 * ip, jp: F8.  op: mrx, mry, mrz, ax, ay, az, u -- sums over j at the drifted
 * separation r_ij + v_ij dt of m_j r, the softened acceleration, and
 * u_i = m_i sum_j m_j / s.  mode: FP64 (FP32C runs FP64) or FP64_IEEE. */
int tupan_nreg_x_dev(int mode, size_t ni, const double *const *ip, size_t nj,
                     const double *const *jp, double dt, double *const *op,
                     void *stream);
int tupan_nreg_x_host(int mode, size_t ni, const double *const *ip,
                      size_t nj, const double *const *jp, double dt,
                      double *const *op);

/* ---- NREG V: replaces nreg_Vkernel, libtupan.h:181-201 ---------------
 * ip, jp: m, vx, vy, vz, ax, ay, az.  op: mvx, mvy, mvz, mk -- sums over j
 * of m_j w and k_i = m_i sum_j m_j |w|^2 with w = v_ij + a_ij dt.  mode as
 * for NREG X. */
int tupan_nreg_v_dev(int mode, size_t ni, const double *const *ip, size_t nj,
                     const double *const *jp, double dt, double *const *op,
                     void *stream);
int tupan_nreg_v_host(int mode, size_t ni, const double *const *ip,
                      size_t nj, const double *const *jp, double dt,
                      double *const *op);

/* ---- NREG with its scalars in device memory (no reference counterpart:
 * the reference keeps E0, W, U, S on the host, nreg.py:22-163).  A step
 * runs without host round trips, so it can be captured into a HIP graph.
 * The scalar block holds TUPAN_NREG_SCALARS doubles at these slots: */
#define TUPAN_NREG_E0 0   /* initial energy */
#define TUPAN_NREG_W 1    /* W, U, S of nreg.py */
#define TUPAN_NREG_U 2
#define TUPAN_NREG_S 3
#define TUPAN_NREG_MTOT 4 /* total mass */
#define TUPAN_NREG_DT 5   /* the current sub-step length */
#define TUPAN_NREG_T 6    /* physical time */
#define TUPAN_NREG_T0 7   /* physical time at the start of the step */
#define TUPAN_NREG_SUM 8  /* a sum written by tupan_sum_pairwise_dev */
#define TUPAN_NREG_HSUB 9 /* the regularised length of the anreg step */
#define TUPAN_NREG_SCALARS 10
/* tupan_nreg_x_dev / tupan_nreg_v_dev with dt read from device memory
 * (d_dt, e.g. sc + TUPAN_NREG_DT) when the kernels start. */
int tupan_nreg_x_devdt(int mode, size_t ni, const double *const *ip,
                       size_t nj, const double *const *jp, const double *d_dt,
                       double *const *op, void *stream);
int tupan_nreg_v_devdt(int mode, size_t ni, const double *const *ip,
                       size_t nj, const double *const *jp, const double *d_dt,
                       double *const *op, void *stream);
/* One scalar recurrence of nreg.py on the block sc (one thread, double
 * arithmetic without contraction, as Python evaluates it): */
#define TUPAN_NREG_OP_BEGIN 0 /* T0 = T                        (:154, :158) */
#define TUPAN_NREG_OP_H 1     /* HSUB = h       (anreg: h = tau/2, :155)   */
#define TUPAN_NREG_OP_HS 2    /* HSUB = 0.5*(h*S)              (:92, :94)  */
#define TUPAN_NREG_OP_XDT 3   /* DT = 0.5*(HSUB/W)             (:81, :83)  */
#define TUPAN_NREG_OP_VDT 4   /* DT = HSUB/U                   (:82)       */
#define TUPAN_NREG_OP_XEND 5  /* U = 0.5*SUM; T = T + DT       (:31-33)    */
#define TUPAN_NREG_OP_VEND 6  /* W = 0.25*SUM/MTOT - E0        (:57-58)    */
#define TUPAN_NREG_OP_S 7     /* S = 1/(2/W - 1/S)             (:93)       */
int tupan_nreg_scalars_dev(double *sc, int op, double h, void *stream);
/* end of an NREG step (nreg.py:162-166): tstep[i] = T - T0, time[i] += tau,
 * nstep[i] += 1. */
int tupan_nreg_step_end_dev(size_t n, double *tstep, double *time,
                            long long *nstep, const double *sc, double tau,
                            void *stream);

/* ---- Kepler: replaces kepler_solver_kernel, libtupan.h:231-246 -------
 * ip: F8 arrays of which particles 0 and 1 are used; op: rx, ry, rz, vx,
 * vy, vz of the two particles after dt (may alias ip's fields, as the
 * reference updates in place).  FP64; d_status as for Sakura. */
int tupan_kepler_dev(const double *const *ip, double dt, double *const *op,
                     int *d_status, void *stream);
int tupan_kepler_host(const double *const *ip, double dt, double *const *op);

/* ---- pair-selected variants for hybrid schemes -------------------------
 * No counterpart in the reference (which applies Sakura to every pair).
 * For a splitting whose close pairs are drifted by the Sakura two-body
 * Kepler solve and whose other pairs are kicked, the pairs are split by
 * their separation at classification positions c = (cx, cy, cz), extra
 * fields after the usual ones: a scheme passes positions fixed for the
 * whole step (tupan_amd's sakura_pn: the free-drift midpoint r + v dt/2
 * predicted at the step's start), so that every kick and drift of the
 * step splits the pairs the same way (passing the positions themselves
 * splits by the current separation):
 *   tupan_acc_far_dev:    Acc (FP64) over the pairs with |c_i - c_j| >=
 *                          rcut only; ip, jp = {m, rx, ry, rz, e2, cx, cy,
 *                          cz};
 *   tupan_sakura_near_dev: Sakura over the pairs with |c_i - c_j| < rcut
 *                          only (the other pairs contribute nothing); ip,
 *                          jp = {m, rx, ry, rz, e2, vx, vy, vz, cx, cy, cz}.
 * Otherwise arguments as tupan_acc_dev / tupan_sakura_dev; rcut >= 0. */
int tupan_acc_far_dev(size_t ni, const double *const *ip, size_t nj,
                      const double *const *jp, double rcut, double *const *op,
                      void *stream);
/* tupan_acc_far_dev of a set on itself (ip = {m, rx, ry, rz, e2, cx, cy,
 * cz}, n each) followed by tupan_kick_pn_half_dev(v, acc, w, tau) -- kick_pn's first
 * half kick, tupan/integrator/sia.py:142-156 -- in the same launches: acc
 * and v end bit-identical to the two calls.  No reference counterpart. */
int tupan_acc_far_kick_pn_half_dev(size_t n, const double *const *ip,
                                   double rcut, double *const *acc,
                                   double *const *v, const double *const *w,
                                   double tau, void *stream);
int tupan_sakura_near_dev(size_t ni, const double *const *ip, size_t nj,
                          const double *const *jp, double dt, int flag,
                          double rcut, double *const *op, int *d_status,
                          void *stream);

/* ---- adjoints of Acc and Phi ------------------------------------------
 * No counterpart in the reference (it has no differentiation).  For a
 * self-gravitating set (i = j = all particles) with a = acc(ps, ps) or
 * phi = phi(ps, ps) and a loss L, given the upstream gradient
 * g_k = dL/da_k (Acc) or h_k = dL/dphi_k (Phi) these return the full
 * gradients with respect to every position and mass:
 *
 *   Acc:  dL/dx_k = sum_j M(r_kj) (m_k g_j - m_j g_k),
 *         M(r) = I/s^3 - 3 r r^T/s^5 (the softened tidal tensor, negated),
 *         dL/dm_k = sum_j g_j . r_kj / s_kj^3
 *   Phi:  dL/dx_k = sum_j (h_k m_j + h_j m_k) r_kj / s_kj^3,
 *         dL/dm_k = -sum_j h_j / s_kj
 *
 * with r_kj = x_k - x_j, s^2 = r^2 + e2_k + e2_j and the reference's r^2 > 0
 * mask.  Both are pair sums on the same engine and modes as Acc/Phi.
 * ip, jp: Acc {m, rx, ry, rz, e2, gx, gy, gz}; Phi {m, rx, ry, rz, e2, h}.
 * op: dL/drx, dL/dry, dL/drz, dL/dm. */
int tupan_acc_adjoint_dev(int mode, size_t ni, const double *const *ip,
                          size_t nj, const double *const *jp,
                          double *const *op, void *stream);
int tupan_phi_adjoint_dev(int mode, size_t ni, const double *const *ip,
                          size_t nj, const double *const *jp,
                          double *const *op, void *stream);

/* ---- ragged ensemble of independent clusters --------------------------
 * No single counterpart in the reference: each cluster follows the
 * reference's adaptive Hermite-4 ("ahermite4", tupan/integrator/
 * hermite.py:340-410 with the block steps of tupan/integrator/
 * __init__.py:48-78) on its own clock, and the ensemble batches the pair
 * sums of all clusters into one persistent kernel per phase.  One ensemble
 * step (Tstep and AccJerk at the start state in one pair pass, block-step
 * choice, predict, 2 x (AccJerk + correct), clock) is captured once into a
 * HIP graph and replayed.  FP64 arithmetic by default;
 * tupan_ensemble_set_mode(TUPAN_MODE_FP32C) evaluates the three pair
 * passes (Tstep with AccJerk at the start state, the two corrector AccJerk
 * sums) in the FP32C arithmetic of the single-system kernels (FP32 tile
 * sums folded into FP64, FP32 offsets to per-tile FP64 origins, near pairs
 * in FP64 with a near-field factor per cluster); the updates and the
 * block-step choice stay FP64.  Device memory is
 * allocated at create and at the first step after set_launch or set_mode;
 * steps allocate nothing. */
typedef struct tupan_ensemble tupan_ensemble;
/* sizes[c] = particles of cluster c; clusters are stored back to back. */
int tupan_ensemble_create(tupan_ensemble **out, int nclusters,
                          const int *sizes, double eta, double t_end);
int tupan_ensemble_destroy(tupan_ensemble *e);
/* Autotuner hook: i-particles per lane (1, 2), j-chunk length of a work
 * item, resident waves per SIMD of the persistent kernels. */
int tupan_ensemble_set_launch(tupan_ensemble *e, int ipl, int jchunk,
                              int waves_per_simd);
/* mode = TUPAN_MODE_FP64 (default) or TUPAN_MODE_FP32C (the pair passes;
 * one i-particle per lane there, jchunk a multiple of 32). */
int tupan_ensemble_set_mode(tupan_ensemble *e, int mode);
/* Storage order of each cluster's particles on the device, applied at the
 * next set_state (get_state returns input order either way):
 * TUPAN_ENS_ORDER_MORTON (default) sorts them by the Morton code of their
 * positions, so that neighbours share waves and j tiles;
 * TUPAN_ENS_ORDER_INPUT keeps the input order. */
#define TUPAN_ENS_ORDER_INPUT 0
#define TUPAN_ENS_ORDER_MORTON 1
int tupan_ensemble_set_order(tupan_ensemble *e, int order);
/* fields = host {m, rx, ry, rz, e2, vx, vy, vz}, each of the total length;
 * t0 = per-cluster start times (NULL = 0).  Synchronous. */
int tupan_ensemble_set_state(tupan_ensemble *e, const double *const *fields,
                             const double *t0);
/* out = host {rx, ry, rz, vx, vy, vz} (entries may be NULL); tc, nsteps =
 * per-cluster clocks and step counts (may be NULL).  Synchronous. */
int tupan_ensemble_get_state(tupan_ensemble *e, double *const *out,
                             double *tc, long long *nsteps);
/* Queue `steps` ensemble steps (graph = 1: replay the captured graph;
 * 0: launch the same kernels one by one).  Finished clusters idle. */
int tupan_ensemble_step(tupan_ensemble *e, int steps, int graph);
/* Wait for queued steps; unfinished clusters, pair interactions evaluated
 * since set_state, error flag (1 = a cluster met a zero or non-finite time
 * step and was stopped). */
int tupan_ensemble_poll(tupan_ensemble *e, int *unfinished, double *pairs,
                        int *error);

/* ---- integrator updates (device pointers, FP64, asynchronous) --------
 * The reference performs these with numpy on SoA arrays; each entry point
 * evaluates the same expression in the same order, one IEEE operation at a
 * time (no FMA contraction), so a device-resident driver reproduces the
 * numpy driver's bits given the same pair-kernel outputs.
 *
 * Hermite predictor, replaces H2/H4/H6/H8.epredict
 * (tupan/integrator/hermite.py:25-41, 75-90, 127-159, 203-241).
 * x = {rx,ry,rz, vx,vy,vz, ax,ay,az, jx,jy,jz, sx,sy,sz, cx,cy,cz}; the
 * first 6 + 3*order/2 are used.  On entry x holds the step's state and its
 * derivatives; they are copied to x0 (the reference's ps0 = ps.copy()) and
 * x's r, v are advanced.  order: 2, 4, 6 or 8. */
int tupan_hermite_predict_dev(int order, size_t n, double *const *x,
                              double *const *x0, double tau, void *stream);
/* Hermite corrector, replaces H*.ecorrect (hermite.py:43-57, 93-119,
 * 162-195, 244-284): x holds the predicted r, v (overwritten) and the
 * derivatives evaluated there; x0 the snapshot written by the predictor. */
int tupan_hermite_correct_dev(int order, size_t n, double *const *x,
                              const double *const *x0, double tau,
                              void *stream);
/* drift_n / drift_pn (tupan/integrator/sia.py:65-103) and the Sakura half
 * drifts (tupan/integrator/sakura.py:26-28, 45-47):
 * r += v*tau (half = 0) or r += v*tau/2 (half = 1); when pn_mr is not
 * NULL also pn_mr += pn_mv*tau (body.py:481-492). */
int tupan_drift_dev(size_t n, double *const *r, const double *const *v,
                    double tau, int half, double *const *pn_mr,
                    const double *const *pn_mv, void *stream);
/* kick_n (sia.py:79-87): v += a*tau. */
int tupan_kick_dev(size_t n, double *const *v, const double *const *a,
                   double tau, void *stream);
/* half of kick_pn (sia.py:142-144, 154-156): v += (a*tau + w)/2. */
int tupan_kick_pn_half_dev(size_t n, double *const *v, const double *const *a,
                           const double *const *w, double tau, void *stream);
/* middle of kick_pn (sia.py:146-152 with body.py:471-525):
 * in  = {m, rx,ry,rz, vx,vy,vz, pnax,pnay,pnaz};
 * acc = {wx,wy,wz, pn_ke, pn_mvx,pn_mvy,pn_mvz, pn_amx,pn_amy,pn_amz},
 * updated: pn_ke -= (v . m pna) tau, pn_mv -= m pna tau,
 * pn_am -= (r x m pna) tau, w = 2 pna tau - w. */
int tupan_pn_flux_dev(size_t n, const double *const *in, double *const *acc,
                      double tau, void *stream);
/* tupan_pn_flux_dev followed by the second tupan_kick_pn_half_dev, fused
 * (no reference counterpart; the same bits as the two calls):
 * in = {m, rx,ry,rz, vx,vy,vz, pnax,pnay,pnaz} with v updated in place,
 * acc as tupan_pn_flux_dev, a = {ax, ay, az} (the Newtonian part). */
int tupan_pn_flux_kick_dev(size_t n, double *const *in, double *const *acc,
                           const double *const *a, double tau, void *stream);
/* end of a shared step (integrator/__init__.py and the drivers' clock
 * updates): tstep[i] = tau, time[i] += tau, nstep[i] += 1 (no reference
 * counterpart as one call). */
int tupan_advance_clock_dev(size_t n, double *tstep, double *time,
                            long long *nstep, double tau, void *stream);
/* sakura_step's application of the pair deltas (sakura.py:31-36, 39-44):
 * x = {rx..vz} += d = {drx..dvz}. */
int tupan_apply_delta_dev(size_t n, double *const *x, const double *const *d,
                          void *stream);
/* NREG's recentring (nreg.py:28-30, 54-56): dst[c] = src[c] / s for
 * c < k (1 <= k <= 9), correctly rounded; dst may alias src. */
int tupan_divide_dev(size_t n, int k, const double *const *src,
                     double *const *dst, double s, void *stream);
/* the same with the divisor read from device memory (no reference
 * counterpart: NREG with device-resident scalars). */
int tupan_divide_by_dev(size_t n, int k, const double *const *src,
                        double *const *dst, const double *d_s, void *stream);
/* out = r + (v * tau) / 2, with tupan_drift_dev's operations for its half
 * drift, r and v unchanged (no reference counterpart: the classification
 * positions of the hybrid scheme, the free-drift midpoint of the step). */
int tupan_half_drift_to_dev(size_t n, const double *const *r,
                            const double *const *v, double tau,
                            double *const *out, void *stream);

/* ---- reductions ---------------------------------------------------------
 * No counterpart in the reference (it sums with numpy on the host).
 * Deterministic (fixed tree for a given n: bit-reproducible run to run),
 * written to a device array `out`.
 *   tupan_sum_dev:         out[0] = sum_i x[i] (ps.x.sum(), e.g.
 *                          nreg.py:26, 31, 57).
 *   tupan_diagnostics_dev: the sums behind the report of
 *                          simulation.py:96-129, one pass with body.py's
 *                          per-particle expressions.  f: m, rx, ry, rz, vx,
 *                          vy, vz, phi, and with pn != 0 also pn_ke,
 *                          pn_mrx..z, pn_mvx..z, pn_amx..z.  out[12]: M,
 *                          KE (with pn_ke), sum m phi (PE = half of it),
 *                          sum m r (+pn_mr), sum m v (+pn_mv), sum r x m v
 *                          (+pn_am). */
int tupan_sum_dev(size_t n, const double *x, double *out, void *stream);
/* out[0] = sum_i x[i] in numpy's order for a contiguous float64 array
 * (0 + the pairwise sums of consecutive 8192-element chunks): the bits of
 * the reference drivers' ps.x.sum(). */
int tupan_sum_pairwise_dev(size_t n, const double *x, double *out,
                           void *stream);
/* Per-particle ke = 0.5 m (vx^2 + vy^2 + vz^2) and mphi = m phi (body.py
 * :261-275, :294-300) from f = {m, vx, vy, vz, phi}, evaluated as numpy
 * does, for totals summed in numpy's order. */
int tupan_energy_terms_dev(size_t n, const double *const *f, double *ke,
                           double *mphi, void *stream);
/* The twelve per-particle quantities whose totals tupan_diagnostics_dev
 * returns (same f and pn), written to out[0..11] for totals summed in
 * numpy's order (tupan_sum_pairwise_dev): the reference report's bits. */
int tupan_report_terms_dev(size_t n, const double *const *f, int pn,
                           double *const *out, void *stream);
int tupan_diagnostics_dev(size_t n, const double *const *f, int pn,
                          double *out, void *stream);

/* ---- non-finite outputs -------------------------------------------------
 * No counterpart in the reference (its kernels report nothing).  OR-s
 * TUPAN_STATUS_NONFINITE into the device int *d_status when any of
 * arrays[0..narrays)[0..n) is NaN or infinite; leaves it alone otherwise.
 * Asynchronous on the stream, allocates nothing, records no memset. */
int tupan_check_finite_dev(size_t n, int narrays, const double *const *arrays,
                           int *d_status, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TUPAN_HIP_H */
