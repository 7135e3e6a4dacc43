// ensemble.hip -- a ragged ensemble of independent clusters integrated with
// adaptive Hermite-4 in one device-resident loop.
//
// Every cluster c (n_c particles, stored contiguously from off_c) follows
// exactly the reference's adaptive Hermite-4 step ("ahermite4",
// tupan/integrator/hermite.py:340-410 with the block steps of
// tupan/integrator/__init__.py:48-78) on its own clock t_c:
//
//   tstep, a0, j0 <- Tstep(c, c, eta), AccJerk(c, c)   (one pair pass)
//   tau    <- block step from min |tstep| and min(t_end - t_c, eta)
//   snapshot; predict
//   twice: a1, j1 <- AccJerk(c, c); correct
//   t_c += tau
//
// Clusters do not interact, so they need no common time line: one "step"
// of the ensemble advances every unfinished cluster by one of its own steps.
// The step is a fixed sequence of kernels that read each cluster's state
// (active flag, tau) from device memory, so it is captured once into a HIP
// graph and replayed; the host only polls how many clusters are unfinished.
//
// Pair sums: a persistent kernel walks a static list of work items
// (cluster, i-wave, j-chunk), sorted by cost and dealt round-robin to the
// resident waves, so that 64-particle and 4096-particle clusters share the
// chip evenly.  (Taking items from an atomic counter instead was measured
// 15% slower: the waves then wait for the counter before every item.)  Each item writes FP64 partial sums; a finishing kernel adds a
// cluster's chunks in chunk order (bit-reproducible).  The pair arithmetic
// is the single-evaluation FP64 policies' (gravity.hip); the Tstep and
// AccJerk sums at the start state share one pass (TSAJ).  In FP32C mode
// the three pair passes run in FP32C arithmetic (Tsaj32, AJ32E below).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "accjerk.hpp"
#include "hermite.hpp"
#include "launch.hpp"

namespace tupan {

// j record of the AccJerk and Tstep pair terms (the reference's argument
// order, tupan/lib/src/libtupan.h:34-58)
struct AccJerkF64Rec { double m, x, y, z, e2, vx, vy, vz; };

namespace {

constexpr int kThreads = 256;

struct ClusterInfo {
  int off, n, nchunk;
  long long base;  // offset of the cluster's partials (doubles)
};

struct Item {
  int c, i0, j0, j1, chunk;
};

// ---------------------------------------------------------------- pairs
// FP64 AccJerk and Tstep pair terms, as in gravity.hip's AccJerkF64 and
// TstepF64 (acc_jerk_kernel_common.h:7-55, tstep_kernel_common.h:7-56).
struct AJ {
  static constexpr int kOut = 6;
  struct A { double v[6]; };
  struct Params {};
  __device__ static void zero(A &a) {
#pragma unroll
    for (int k = 0; k < 6; ++k) a.v[k] = 0.0;
  }
  __device__ static void pair(const AccJerkF64Rec &q, const AccJerkF64Rec &r,
                              A &a, const Params &) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    double vx = q.vx - r.vx, vy = q.vy - r.vy, vz = q.vz - r.vz;
    double r2 = rx * rx + ry * ry + rz * rz;
    double rv = rx * vx + ry * vy + rz * vz;
    double y = rsq64(r2 + (q.e2 + r.e2));
    y = r2 > 0.0 ? y : 0.0;
    double y2 = y * y;
    double k = r.m * y * y2;
    double alpha = 3.0 * rv * y2;
    vx -= alpha * rx;
    vy -= alpha * ry;
    vz -= alpha * rz;
    a.v[0] -= k * rx;
    a.v[1] -= k * ry;
    a.v[2] -= k * rz;
    a.v[3] -= k * vx;
    a.v[4] -= k * vy;
    a.v[5] -= k * vz;
  }
  __device__ static void values(const A &a, double *v) {
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = a.v[k];
  }
  __device__ static void combine(double *v, const double *w) {
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] += w[k];
  }
  __device__ static void finish(const double *v, double *const *out, int i,
                                const Params &) {
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k][i] = v[k];
  }
};

// Tstep and AccJerk of one pair together (the two evaluations at the start
// of a step share the state): the Tstep terms first (gravity.hip's TstepF64,
// tstep_kernel_common.h:7-56), then AJ's, each with the same operations as
// its own kernel; the shared 1/sqrt(r^2 + e2) is the same value in both.
// Outputs: tstep, tstepij, ax..jz.
struct TSAJ {
  static constexpr int kOut = 8;
  struct A { double v[8]; };
  struct Params { double eta; };
  __device__ static void zero(A &a) {
#pragma unroll
    for (int k = 0; k < 8; ++k) a.v[k] = 0.0;
  }
  __device__ static void pair(const AccJerkF64Rec &q, const AccJerkF64Rec &r,
                              A &a, const Params &prm) {
    double rx = q.x - r.x, ry = q.y - r.y, rz = q.z - r.z;
    double vx = q.vx - r.vx, vy = q.vy - r.vy, vz = q.vz - r.vz;
    double r2 = rx * rx + ry * ry + rz * rz;
    double rv = rx * vx + ry * vy + rz * vz;
    const bool real_pair = r2 > 0.0;
    double y = rsq64(r2 + (q.e2 + r.e2));
    y = real_pair ? y : 0.0;
    double y2 = y * y;
    // Tstep
    {
      double v2 = vx * vx + vy * vy + vz * vz;
      double phi = (q.m + r.m) * y;
      double w2 = y2 * (v2 + 2.0 * phi);
      double gam = y2 * (w2 + 2.0 * y2 * phi);
      gam *= prm.eta * rsq64(w2);
      w2 -= gam * rv;
      w2 = real_pair ? w2 : 0.0;
      a.v[0] += w2;
      a.v[1] = fmax(w2, a.v[1]);
    }
    // AccJerk (AJ::pair)
    double k = r.m * y * y2;
    double alpha = 3.0 * rv * y2;
    vx -= alpha * rx;
    vy -= alpha * ry;
    vz -= alpha * rz;
    a.v[2] -= k * rx;
    a.v[3] -= k * ry;
    a.v[4] -= k * rz;
    a.v[5] -= k * vx;
    a.v[6] -= k * vy;
    a.v[7] -= k * vz;
  }
  __device__ static void values(const A &a, double *v) {
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = a.v[k];
  }
  __device__ static void combine(double *v, const double *w) {
    v[0] += w[0];
    v[1] = fmax(w[1], v[1]);
#pragma unroll
    for (int k = 2; k < 8; ++k) v[k] += w[k];
  }
  __device__ static void finish(const double *v, double *const *out, int i,
                                const Params &prm) {
    out[0][i] = prm.eta / sqrt(1.0 + v[0]);
    out[1][i] = prm.eta / sqrt(1.0 + v[1]);
#pragma unroll
    for (int k = 2; k < 8; ++k) out[k][i] = v[k];
  }
};

// Largest number of outputs of a pair policy (partials per chunk).
constexpr int kMaxOut = 8;

struct Outs {
  double *f[kMaxOut];
};

// Persistent pair kernel.  Wave w handles items w, w + W, w + 2W, ... of
// the cost-sorted list.  Item and j indices are made wave-uniform
// (readfirstlane) so that item descriptors and j records are scalar loads.
template <class Pol, int P>
__global__ __launch_bounds__(kThreads) void ens_pairs(
    const AccJerkF64Rec *__restrict__ rec, const Item *__restrict__ items,
    int nitems, const ClusterInfo *__restrict__ cl,
    const int *__restrict__ active, double *__restrict__ part,
    typename Pol::Params prm) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(
      (int)(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)));
  const int nwaves = gridDim.x * (kThreads / 64);
  for (int it = wave; it < nitems; it += nwaves) {
    const Item w = items[it];
    if (!active[w.c]) continue;
    const ClusterInfo ci = cl[w.c];
    typename Pol::A acc[P];
    AccJerkF64Rec q[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      int li = w.i0 + p * 64 + lane;
      li = li < ci.n ? li : ci.n - 1;
      q[p] = rec[ci.off + li];
      Pol::zero(acc[p]);
    }
    const int jb = ci.off + w.j0, je = ci.off + w.j1;
    AccJerkF64Rec r = rec[jb];
#pragma unroll 2
    for (int j = jb; j < je; ++j) {
      const AccJerkF64Rec rn = rec[min(j + 1, je - 1)];
#pragma unroll
      for (int p = 0; p < P; ++p) Pol::pair(q[p], r, acc[p], prm);
      r = rn;
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int li = w.i0 + p * 64 + lane;
      if (li >= ci.n) continue;
      double v[Pol::kOut];
      Pol::values(acc[p], v);
#pragma unroll
      for (int k = 0; k < Pol::kOut; ++k)
        part[ci.base + ((long long)w.chunk * Pol::kOut + k) * ci.n + li] =
            v[k];
    }
  }
}

// Adds a cluster's chunk partials in chunk order and finishes.
template <class Pol>
__global__ __launch_bounds__(kThreads) void ens_finish(
    int ntot, const int *__restrict__ cid, const ClusterInfo *__restrict__ cl,
    const int *__restrict__ active, const double *__restrict__ part, Outs out,
    typename Pol::Params prm) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < ntot;
       i += gridDim.x * kThreads) {
    const int c = cid[i];
    if (!active[c]) continue;
    const ClusterInfo ci = cl[c];
    const int li = i - ci.off;
    double v[Pol::kOut];
#pragma unroll
    for (int k = 0; k < Pol::kOut; ++k) v[k] = part[ci.base + (long long)k * ci.n + li];
    for (int s = 1; s < ci.nchunk; ++s) {
      double w[Pol::kOut];
#pragma unroll
      for (int k = 0; k < Pol::kOut; ++k)
        w[k] = part[ci.base + ((long long)s * Pol::kOut + k) * ci.n + li];
      Pol::combine(v, w);
    }
    Pol::finish(v, out.f, i, prm);
  }
}

// Field indices of the ensemble state.
enum EF {
  FM = 0, FX, FY, FZ, FE2, FVX, FVY, FVZ,  // base (reference order)
  FAX, FAY, FAZ, FJX, FJY, FJZ,            // derivatives at the state
  S_X, S_Y, S_Z, S_VX, S_VY, S_VZ, S_AX, S_AY, S_AZ, S_JX, S_JY, S_JZ,
  FTS, FTSIJ,                              // tstep, tstepij
  NFIELD
};

struct Fields {
  double *f[NFIELD];
};

__global__ __launch_bounds__(kThreads) void ens_pack(int ntot, Fields s,
                                                     const int *cid,
                                                     const int *active,
                                                     AccJerkF64Rec *rec) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < ntot;
       i += gridDim.x * kThreads) {
    if (!active[cid[i]]) continue;
    rec[i] = {s.f[FM][i],  s.f[FX][i],  s.f[FY][i],  s.f[FZ][i],
              s.f[FE2][i], s.f[FVX][i], s.f[FVY][i], s.f[FVZ][i]};
  }
}

// ---------------------------------------------------------------- FP32C
// The pair passes in the FP32C arithmetic of the single-system kernels
// (f32c.hpp, accjerk.hpp) without the Newton refinement of 1/sqrt and
// without the debiasing factors: FP32 tile sums folded into FP64,
// velocity offsets to per-tile FP64 origins, and near pairs (below the j
// record's threshold) re-evaluated in FP64 from the slab.  As in the
// sorted single-system engine, each (wave, tile) pair of boxes is tested
// first: a tile whose box is beyond every near threshold of its j's from
// the box of the wave's 64 i's runs the untested loop on far records, the
// others the tested loop.  Both take the positions as FP32 offsets to the
// tile's box centre (the j offset rounded once when packed, the i offset
// once per tile); in a near tile a pair's offsets can be larger than its
// separation, which costs the pair sums accuracy (about 1e-6 of a pair
// term at the near-field radius of a whole-cluster tile) that the
// ensemble's states do not see (below).  Tiles are numbered within each cluster
// (tbase: the cluster's first tile), so a work item's j range -- a
// multiple of the tile length from the cluster's start -- covers whole
// tiles.  The near-field threshold uses each cluster's own factor and
// largest e2 (ClusterCtx, set with the state).
using AJ32 = AccJerkF32cTile;
// Near-field factor of the ensemble's FP32C passes (the single-system
// kernels' rule with c = 0.5: radius 0.5 N_c^(-1/3) for equal masses).
// The ensemble needs its states, not each pair sum, close to FP64's: on
// the 512 Plummer clusters (N_c 64-4096) every cluster stays within 6e-9
// of FP64 after 6 steps.  (Measured with FP64 separations in the tested
// loop: 3e-10, 6e-10 and 1.2e-9 at c = 2, 0.5 and 0.25, and c = 0.5 12%
// faster than c = 2.)
constexpr double kEnsNearC = 0.5;

struct ClusterCtx {
  double factor, e2max;
  int tbase, pad;
};

// Per tile of the active clusters, one thread per tile: the velocity
// origin (midpoint of the tile's velocity range, tile_origin_kernel's
// rule), the bounding box of the positions with the largest near-field
// threshold of the tile's j's, and the box centre (the far records'
// position origin), with porg[4 g + 3] = the records' e2 when every j of
// the tile has the same one (-1 otherwise, NaN included).
__global__ __launch_bounds__(kThreads) void ens_tiles(
    int ntiles, const int *__restrict__ tcid, const ClusterInfo *__restrict__ cl,
    const ClusterCtx *__restrict__ cx, const int *__restrict__ active,
    Fields s, double *__restrict__ orig, TileBox *__restrict__ boxes,
    double *__restrict__ porg) {
  for (int g = blockIdx.x * kThreads + threadIdx.x; g < ntiles;
       g += gridDim.x * kThreads) {
    const int c = tcid[g];
    if (!active[c]) continue;
    const ClusterInfo ci = cl[c];
    const ClusterCtx x = cx[c];
    const PackCtx pc{x.factor, x.e2max, nullptr};
    const int j0 = ci.off + (g - x.tbase) * kTile;
    const int je = min(ci.off + ci.n, j0 + kTile);
    TileBox bx;
    for (int k = 0; k < 3; ++k) {
      const double *v = s.f[FVX + k];
      const double *r = s.f[FX + k];
      double lo = v[j0], hi = lo, plo = r[j0], phi = plo;
      for (int j = j0 + 1; j < je; ++j) {
        lo = fmin(lo, v[j]);
        hi = fmax(hi, v[j]);
        plo = fmin(plo, r[j]);
        phi = fmax(phi, r[j]);
      }
      orig[4 * g + k] = 0.5 * (lo + hi);
      bx.lo[k] = plo;
      bx.hi[k] = phi;
      porg[4 * g + k] = 0.5 * (plo + phi);
    }
    double n2 = 0.0;
    const float e20 = (float)s.f[FE2][j0];
    bool same = e20 == e20;
    for (int j = j0; j < je; ++j) {
      n2 = fmax(n2, (double)near_threshold(s.f[FM][j], s.f[FE2][j], pc));
      same = same && (float)s.f[FE2][j] == e20;
    }
    bx.near2 = n2;
    bx.pad = 0.0;
    boxes[g] = bx;
    orig[4 * g + 3] = 0.0;
    porg[4 * g + 3] = same ? (double)e20 : -1.0;
  }
}

// j record of the tested (near-tile) loop: AJ32's far record plus the
// near-field threshold.
struct alignas(16) EnsNearRec {
  float x, y, z, dvx, dvy, dvz, m, e2, near2, pad[3];
};

// FP32C j records of the tested and untested loops, with the cluster's
// tile origins and near-field context.
__global__ __launch_bounds__(kThreads) void ens_pack32(
    int ntot, Fields s, const int *__restrict__ cid,
    const ClusterInfo *__restrict__ cl, const ClusterCtx *__restrict__ cx,
    const int *__restrict__ active, const double *__restrict__ orig,
    const double *__restrict__ porg, EnsNearRec *__restrict__ rec,
    AJ32::FarRec *__restrict__ frec) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < ntot;
       i += gridDim.x * kThreads) {
    const int c = cid[i];
    if (!active[c]) continue;
    const ClusterCtx x = cx[c];
    const int t = x.tbase + (i - cl[c].off) / kTile;
    const double *o = orig + 4 * t;
    const PackCtx pc{x.factor, x.e2max, nullptr};
    const AJ32::JRec r = {s.f[FX][i],
                          s.f[FY][i],
                          s.f[FZ][i],
                          dsub32(s.f[FVX][i], o[0]),
                          dsub32(s.f[FVY][i], o[1]),
                          dsub32(s.f[FVZ][i], o[2]),
                          (float)s.f[FM][i],
                          (float)s.f[FE2][i],
                          near_threshold(s.f[FM][i], s.f[FE2][i], pc)};
    const AJ32::FarRec f = AJ32::far_rec(r, porg + 4 * t);
    frec[i] = f;
    rec[i] = {f.x, f.y, f.z, f.dvx, f.dvy, f.dvz, f.m, f.e2, r.near2,
              {0.0f, 0.0f, 0.0f}};
  }
}

// Near test on the position offsets: rx.. = i offset - j offset, s^2 in
// the FMA order of F32cBase::near_test, near when s^2 < near2 (or NaN).
template <class I>
__device__ __forceinline__ bool ens_near_test(const I &q, const EnsNearRec &r,
                                              float &rx, float &ry, float &rz,
                                              float &s2) {
  rx = q.px - r.x;
  ry = q.py - r.y;
  rz = q.pz - r.z;
  s2 = __builtin_fmaf(
      rz, rz, __builtin_fmaf(ry, ry, __builtin_fmaf(rx, rx, q.e2 + r.e2)));
  return !(s2 >= r.near2);
}

// AccJerk (the corrector passes) in the tile kernel's terms.
struct AJ32E {
  using P = AJ32;
  using JRec = EnsNearRec;
  using FarRec = P::FarRec;
  using I = P::I;
  using A32 = P::A32;
  using A64 = P::A64;
  using Params = AJ::Params;
  static constexpr int kOut = 6;
  static constexpr int kWavesPerEu = 8;  // ens_tiles32's occupancy
  // the FP64 i fields ens_tiles32 keeps in LDS
  static constexpr int kI64 = P::kI64;
  __device__ static void save64(const I &q, double *f) { P::save64(q, f); }
  __device__ static void load64(I &q, const double *f) { P::load64(q, f); }
  __device__ static I load_i(const Soa &s, int i, const Params &) {
    return P::load_i(s, i, NoParams{});
  }
  __device__ static void zero(A32 &a) { P::zero(a); }
  __device__ static void zero(A64 &a) { P::zero(a); }
  __device__ static void tile_begin(I &q, const double *orig, int t) {
    P::tile_begin(q, orig, t);
  }
  __device__ static void tile_begin_far(I &q, const double *po) {
    P::tile_begin_far(q, po);
  }
  __device__ static uint64_t pair32(const I &q, const JRec &r, A32 &a) {
    float rx, ry, rz, s2;
    const bool far = !ens_near_test(q, r, rx, ry, rz, s2);
    P::body32(q, r, rx, ry, rz, s2, far, a);
    return __builtin_amdgcn_ballot_w64(far);
  }
  __device__ static void pair32_farrec(const I &q, const FarRec &r, A32 &a) {
    P::pair32_farrec(q, r, a);
  }
  __device__ static void pair32_farrec_u(const I &q, const FarRec &r,
                                         float e2ij, A32 &a) {
    P::pair32_farrec_u(q, r, e2ij, a);
  }
  __device__ static void fold(const A32 &a, A64 &b) { P::fold(a, b); }
  __device__ static void values(const A64 &a, double *v) { P::values(a, v); }
  __device__ static void near_pair(const I &q, const JRec &r, A64 &a,
                                   const Soa &s, int j, const Params &) {
    float rx, ry, rz, s2;
    if (ens_near_test(q, r, rx, ry, rz, s2)) {
      AccJerkF64::Params p64{};
      AccJerkF64::pair(P::exact_i(q, s), AccJerkF64::pack(s, j, 0.0), a, p64);
    }
  }
};

// Tstep and AccJerk at the step's start state (TSAJ) in FP32C: AJ32's
// records and AccJerk terms, and the Tstep terms of TSAJ::pair
// (tstep_kernel_common.h:7-56) in FP32 on the same separation, velocity
// offsets and 1/s.  The tstep sum folds into FP64 like the others; its
// pair maximum is taken in FP32 and folded with fmax (exact).  Near pairs
// are TSAJ::pair in FP64.  A near or self pair gives y = 0 and then
// 0 * inf in gam; w2 is selected to 0 after, as in TSAJ::pair.
struct Tsaj32 {
  using P = AJ32;
  using JRec = EnsNearRec;
  using FarRec = P::FarRec;
  struct I : P::I {
    float m, eta;
  };
  using Params = TSAJ::Params;
  static constexpr int kOut = 8;
  static constexpr int kWavesPerEu = 7;  // ens_tiles32's occupancy
  // the FP64 i fields ens_tiles32 keeps in LDS
  static constexpr int kI64 = P::kI64;
  __device__ static void save64(const I &q, double *f) { P::save64(q, f); }
  __device__ static void load64(I &q, const double *f) { P::load64(q, f); }
  struct A32 {
    float v[8];
  };
  using A64 = TSAJ::A;
  __device__ static I load_i(const Soa &s, int i, const Params &prm) {
    I q;
    static_cast<P::I &>(q) = P::load_i(s, i, NoParams{});
    q.m = (float)s.f[M][i];
    q.eta = (float)prm.eta;
    return q;
  }
  __device__ static void zero(A32 &a) {
#pragma unroll
    for (int k = 0; k < 8; ++k) a.v[k] = 0.0f;
  }
  __device__ static void zero(A64 &a) { TSAJ::zero(a); }
  __device__ static void tile_begin(I &q, const double *orig, int t) {
    P::tile_begin(q, orig, t);
  }
  __device__ static void tile_begin_far(I &q, const double *po) {
    P::tile_begin_far(q, po);
  }
  template <class R>  // JRec or FarRec
  __device__ static void body32(const I &q, const R &r, float rx, float ry,
                                float rz, float s2, bool far, A32 &a) {
    float vx = q.dvx - r.dvx, vy = q.dvy - r.dvy, vz = q.dvz - r.dvz;
    const float rv = rx * vx + ry * vy + rz * vz;
    float y = rsq32(s2);
    y = far ? y : 0.0f;
    const float y2 = y * y;
    {
      const float v2 = vx * vx + vy * vy + vz * vz;
      const float phi = (q.m + r.m) * y;
      float w2 = y2 * (v2 + 2.0f * phi);
      float gam = y2 * (w2 + 2.0f * y2 * phi);
      gam *= q.eta * rsq32(w2);
      w2 -= gam * rv;
      w2 = far ? w2 : 0.0f;
      a.v[0] += w2;
      a.v[1] = fmaxf(w2, a.v[1]);
    }
    const float k = r.m * y * y2;
    const float alpha = 3.0f * rv * y2;
    vx -= alpha * rx;
    vy -= alpha * ry;
    vz -= alpha * rz;
    a.v[2] -= k * rx;
    a.v[3] -= k * ry;
    a.v[4] -= k * rz;
    a.v[5] -= k * vx;
    a.v[6] -= k * vy;
    a.v[7] -= k * vz;
  }
  __device__ static uint64_t pair32(const I &q, const JRec &r, A32 &a) {
    float rx, ry, rz, s2;
    const bool far = !ens_near_test(q, r, rx, ry, rz, s2);
    body32(q, r, rx, ry, rz, s2, far, a);
    return __builtin_amdgcn_ballot_w64(far);
  }
  __device__ static void pair32_farrec(const I &q, const FarRec &r, A32 &a) {
    pair32_farrec_u(q, r, q.e2 + r.e2, a);
  }
  __device__ static void pair32_farrec_u(const I &q, const FarRec &r,
                                         float e2ij, A32 &a) {
    const float rx = q.px - r.x, ry = q.py - r.y, rz = q.pz - r.z;
    const float s2 = __builtin_fmaf(
        rz, rz, __builtin_fmaf(ry, ry, __builtin_fmaf(rx, rx, e2ij)));
    body32(q, r, rx, ry, rz, s2, true, a);
  }
  __device__ static void fold(const A32 &a, A64 &b) {
    b.v[0] += (double)a.v[0];
    b.v[1] = fmax(b.v[1], (double)a.v[1]);
#pragma unroll
    for (int k = 2; k < 8; ++k) b.v[k] += (double)a.v[k];
  }
  __device__ static void values(const A64 &a, double *v) { TSAJ::values(a, v); }
  __device__ static void near_pair(const I &q, const JRec &r, A64 &a,
                                   const Soa &s, int j, const Params &prm) {
    float rx, ry, rz, s2;
    if (ens_near_test(q, r, rx, ry, rz, s2)) {
      const int i = q.i;
      const AccJerkF64Rec qi = {s.f[M][i], q.x,  q.y,  q.z,
                                s.f[E2][i], q.vx, q.vy, q.vz};
      const AccJerkF64Rec rj = {s.f[M][j],  s.f[RX][j], s.f[RY][j],
                                s.f[RZ][j], s.f[E2][j], s.f[VX][j],
                                s.f[VY][j], s.f[VZ][j]};
      TSAJ::pair(qi, rj, a, prm);
    }
  }
};

// Persistent FP32C pair kernel over the work items (ens_pairs' scheme, one
// i-particle per lane), writing the same FP64 chunk partials.  Per 64 tiles
// of an item, lane l tests tile l's box against the wave's box (far: the
// squared gap reaches the tile's largest near threshold, so that every
// pair has s^2 >= r^2 >= near2_j and r^2 > 0).
template <class Pol>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(Pol::kWavesPerEu, 8))) void ens_tiles32(
    const typename Pol::JRec *__restrict__ rec,
    const typename Pol::FarRec *__restrict__ frec,
    const TileBox *__restrict__ boxes, const double *__restrict__ orig,
    const double *__restrict__ porg, const Item *__restrict__ items,
    int nitems, const ClusterInfo *__restrict__ cl,
    const ClusterCtx *__restrict__ cx, const int *__restrict__ active,
    Fields s, double *__restrict__ part, typename Pol::Params prm) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(
      (int)(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)));
  const int nwaves = gridDim.x * (kThreads / 64);
  // each wave's i's FP64 fields during its tile loops: read at tile starts
  // and in the tested loop only, they would otherwise cost the far loop
  // its last registers and spill
  __shared__ double i64[kThreads / 64][Pol::kI64][64];
  const int wslot = threadIdx.x >> 6;
  Soa soa{};
  for (int k = 0; k < 8; ++k) soa.f[k] = s.f[k];
  for (int it = wave; it < nitems; it += nwaves) {
    const Item w = items[it];
    if (!active[w.c]) continue;
    const ClusterInfo ci = cl[w.c];
    const int tbase = cx[w.c].tbase;
    int li = w.i0 + lane;
    li = li < ci.n ? li : ci.n - 1;
    typename Pol::I q = Pol::load_i(soa, ci.off + li, prm);
    {
      double f[Pol::kI64];
      Pol::save64(q, f);
#pragma unroll
      for (int k = 0; k < Pol::kI64; ++k) i64[wslot][k][lane] = f[k];
    }
    double wlo[3], whi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double x = soa.f[RX + k][ci.off + li];
      wlo[k] = wave_min(x);
      whi[k] = wave_max(x);
    }
    typename Pol::A64 acc;
    Pol::zero(acc);
    uint64_t farbits = 0;
    int tt = 0;
    for (int j0 = w.j0; j0 < w.j1; j0 += kTile, ++tt) {
      const int t = tbase + j0 / kTile;
      if ((tt & 63) == 0) {
        bool far = false;
        if (j0 + lane * kTile < w.j1) {
          const TileBox b = boxes[t + lane];
          double d2 = 0.0;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double gap =
                fmax(fmax(b.lo[k] - whi[k], wlo[k] - b.hi[k]), 0.0);
            d2 = __builtin_fma(gap, gap, d2);
          }
          far = d2 >= b.near2;
        }
        farbits = __builtin_amdgcn_ballot_w64(far);
      }
      const int nk = min(kTile, w.j1 - j0);
      {
        // the barrier keeps the compiler from hoisting the loads out of
        // the tile loop (which would keep the fields in registers again)
        __asm__ volatile("" ::: "memory");
        double f[Pol::kI64];
#pragma unroll
        for (int k = 0; k < Pol::kI64; ++k) f[k] = i64[wslot][k][lane];
        Pol::load64(q, f);
      }
      Pol::tile_begin(q, orig, t);
      Pol::tile_begin_far(q, porg + 4 * t);
      typename Pol::A32 a;
      Pol::zero(a);
      // near pairs of a tested tile, re-evaluated in FP64 after the one
      // fold of both kinds of tile
      uint32_t near = 0u;
      const typename Pol::JRec *jr = rec + ci.off + j0;
      if ((farbits >> (tt & 63)) & 1) {
        const typename Pol::FarRec *fr = frec + ci.off + j0;
        // a tile whose j's share one e2 forms e2_i + e2_j once
        const double e2u = porg[4 * t + 3];
        if (e2u >= 0.0) {
          const float e2ij = q.e2 + (float)e2u;
          tile_loop<1>(fr, nk, a, [&](const typename Pol::FarRec &r) {
            Pol::pair32_farrec_u(q, r, e2ij, a);
          });
        } else {
          tile_loop<1>(fr, nk, a, [&](const typename Pol::FarRec &r) {
            Pol::pair32_farrec(q, r, a);
          });
        }
      } else {
        // the record of j+1 is fetched while j is computed; whole tiles
        // unrolled so that the fetch is in flight over a full pair
        typename Pol::JRec r = jr[0];
        if (nk == kTile) {
#pragma unroll 2
          for (int k = 0; k < kTile; ++k) {
            const typename Pol::JRec rn = jr[min(k + 1, kTile - 1)];
            near |= near_bit(Pol::pair32(q, r, a)) << k;
            r = rn;
          }
        } else {
          for (int k = 0; k < nk; ++k) {
            const typename Pol::JRec rn = jr[min(k + 1, nk - 1)];
            near |= near_bit(Pol::pair32(q, r, a)) << k;
            r = rn;
          }
        }
      }
      Pol::fold(a, acc);
      while (near) {
        const int k = __builtin_ctz(near);
        near &= near - 1u;
        Pol::near_pair(q, jr[k], acc, soa, ci.off + j0 + k, prm);
      }
    }
    if (w.i0 + lane < ci.n) {
      double v[Pol::kOut];
      Pol::values(acc, v);
#pragma unroll
      for (int k = 0; k < Pol::kOut; ++k)
        part[ci.base + ((long long)w.chunk * Pol::kOut + k) * ci.n + w.i0 +
             lane] = v[k];
    }
  }
}

// Block step of each active cluster from min |tstep| over its particles:
// get_base_tstep then get_min_block_tstep (integrator/__init__.py:48-78).
// One block per cluster; the minimum is exact, so the reduction order does
// not matter.
__global__ __launch_bounds__(kThreads) void ens_choose_dt(
    const ClusterInfo *__restrict__ cl, const double *__restrict__ tstep,
    const double *__restrict__ tc, const int *__restrict__ active,
    double *__restrict__ dt, double eta, double t_end,
    int *__restrict__ err) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  const ClusterInfo ci = cl[c];
  __shared__ double red[kThreads];
  double mn = INFINITY;
  for (int i = threadIdx.x; i < ci.n; i += kThreads)
    mn = fmin(mn, fabs(tstep[ci.off + i]));
  red[threadIdx.x] = mn;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double t = tc[c];
  // get_base_tstep
  double tau = fmin(fabs(t_end) - fabs(t), fabs(eta));
  tau = fmax(tau, fabs(t_end) * (2 * 2.220446049250313e-16));
  tau = copysign(tau, eta);
  // get_min_block_tstep
  const double min_ts = red[0];
  if (!(min_ts > 0.0) || !(min_ts < INFINITY)) {
    // a vanishing or non-finite time step: stop this cluster
    err[0] = 1;
    dt[c] = 0.0;
    return;
  }
  const int power = (int)(log2(min_ts) - 1.0);
  double min_bts = ldexp(1.0, power);
  const double t_next = t + min_bts;
  for (int k = 0; k < 2200 && fmod(t_next, min_bts) != 0.0; ++k)
    min_bts /= 2;
  min_bts = copysign(min_bts, tau);
  if (fabs(min_bts) > fabs(tau)) min_bts = tau;
  dt[c] = min_bts;
}

__global__ __launch_bounds__(kThreads) void ens_predict(
    int ntot, Fields s, const int *__restrict__ cid,
    const int *__restrict__ active, const double *__restrict__ dt) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < ntot;
       i += gridDim.x * kThreads) {
    const int c = cid[i];
    if (!active[c]) continue;
    const double tau = dt[c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double r = s.f[FX + d][i], v = s.f[FVX + d][i];
      const double a = s.f[FAX + d][i], j = s.f[FJX + d][i];
      s.f[S_X + d][i] = r;
      s.f[S_VX + d][i] = v;
      s.f[S_AX + d][i] = a;
      s.f[S_JX + d][i] = j;
      double r1, v1;
      hermite_predict_axis<4>(r, v, a, j, 0.0, 0.0, tau, r1, v1);
      s.f[FX + d][i] = r1;
      s.f[FVX + d][i] = v1;
    }
  }
}

__global__ __launch_bounds__(kThreads) void ens_correct(
    int ntot, Fields s, const int *__restrict__ cid,
    const int *__restrict__ active, const double *__restrict__ dt) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < ntot;
       i += gridDim.x * kThreads) {
    const int c = cid[i];
    if (!active[c]) continue;
    const double tau = dt[c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double r1, v1;
      hermite_correct_axis<4>(s.f[S_X + d][i], s.f[S_VX + d][i],
                              s.f[S_AX + d][i], s.f[FAX + d][i],
                              s.f[S_JX + d][i], s.f[FJX + d][i], 0.0, 0.0,
                              0.0, 0.0, tau, r1, v1);
      s.f[FVX + d][i] = v1;
      s.f[FX + d][i] = r1;
    }
  }
}

// Advances the clocks of the clusters that stepped, counts pair work and
// refreshes the active flags and the count of unfinished clusters.  One
// block; fixed reduction order.
__global__ __launch_bounds__(kThreads) void ens_clock(
    int ncl, const ClusterInfo *__restrict__ cl, double *__restrict__ tc,
    const double *__restrict__ dt, int *__restrict__ active,
    long long *__restrict__ nsteps, double *__restrict__ pairs,
    double t_end, int *__restrict__ count) {
  __shared__ int red[kThreads];
  __shared__ double pw[kThreads];
  int unfinished = 0;
  double work = 0.0;
  for (int c = threadIdx.x; c < ncl; c += kThreads) {
    if (active[c]) {
      if (dt[c] != 0.0) {
        tc[c] += dt[c];
        nsteps[c] += 1;
        const double n = (double)cl[c].n;
        work += 4.0 * n * n;  // Tstep + 3 AccJerk evaluations
      }
      const int a = (fabs(tc[c]) < t_end) && dt[c] != 0.0;
      active[c] = a;
      unfinished += a;
    }
  }
  red[threadIdx.x] = unfinished;
  pw[threadIdx.x] = work;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x] += red[threadIdx.x + w];
      pw[threadIdx.x] += pw[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    count[0] = red[0];
    pairs[0] += pw[0];
  }
}

__global__ void ens_init_flags(int ncl, const double *tc, double t_end,
                               int *active, int *count, int *err,
                               double *pairs) {
  int unfinished = 0;
  for (int c = 0; c < ncl; ++c) {
    active[c] = fabs(tc[c]) < t_end;
    unfinished += active[c];
  }
  count[0] = unfinished;
  err[0] = 0;
  pairs[0] = 0.0;
}

// Spreads the low 21 bits of x to every third bit.
inline uint64_t spread3(uint64_t x) {
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}

// perm[off .. off+n) = the cluster's input indices, sorted by the Morton
// code of their positions (21 bits per axis over the cluster's bounding
// box; ties and non-finite coordinates keep input order) when `morton`,
// else in input order.  Nearby particles then share waves and j tiles, so
// the FP32C near-field re-evaluations concentrate in a few tiles.
void morton_order(const double *x, const double *y, const double *z, int off,
                  int n, bool morton, int *perm) {
  for (int k = 0; k < n; ++k) perm[off + k] = off + k;
  if (!morton || n < 2) return;
  const double *a[3] = {x, y, z};
  double lo[3], sc[3];
  for (int d = 0; d < 3; ++d) {
    double mn = INFINITY, mx = -INFINITY;
    for (int k = off; k < off + n; ++k)
      if (std::isfinite(a[d][k])) {
        mn = std::fmin(mn, a[d][k]);
        mx = std::fmax(mx, a[d][k]);
      }
    lo[d] = mn;
    sc[d] = (mx > mn) ? 2097151.0 / (mx - mn) : 0.0;
  }
  std::vector<uint64_t> code(n);
  for (int k = 0; k < n; ++k) {
    uint64_t c = 0;
    for (int d = 0; d < 3; ++d) {
      const double v = a[d][off + k];
      const double q = std::isfinite(v) ? (v - lo[d]) * sc[d] : 0.0;
      const uint64_t b = q > 0.0 ? (uint64_t)std::fmin(q, 2097151.0) : 0;
      c |= spread3(b) << d;
    }
    code[k] = c;
  }
  std::stable_sort(perm + off, perm + off + n,
                   [&](int p, int q) { return code[p - off] < code[q - off]; });
}

inline int blocks_for(long n) {
  long g = (n + kThreads - 1) / kThreads;
  return (int)(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

}  // namespace
}  // namespace tupan

using namespace tupan;

struct tupan_ensemble {
  int ncl = 0, ntot = 0;
  double eta = 0, t_end = 0;
  std::vector<int> n, off, tbase;
  // storage order: slot k of the device fields holds input particle perm[k]
  // (each cluster's particles in Morton order of their positions at
  // set_state, or in input order)
  int order = TUPAN_ENS_ORDER_MORTON;
  std::vector<int> perm;
  hipStream_t st = nullptr;
  // device
  double *fields[NFIELD] = {};
  double *slab = nullptr;
  AccJerkF64Rec *rec = nullptr;
  ClusterInfo *cl = nullptr;
  int *cid = nullptr;
  Item *items = nullptr;
  int nitems = 0;
  double *part = nullptr;
  double *tc = nullptr, *dt = nullptr, *pairs = nullptr;
  int *active = nullptr, *count = nullptr, *err = nullptr;
  long long *nsteps = nullptr;
  // FP32C AccJerk (mode TUPAN_MODE_FP32C)
  int mode = TUPAN_MODE_FP64;
  int ntiles = 0;
  EnsNearRec *rec32 = nullptr;
  AJ32::FarRec *frec32 = nullptr;
  TileBox *boxes = nullptr;
  double *orig = nullptr, *porg = nullptr;
  ClusterCtx *cctx = nullptr;
  int *tcid = nullptr;
  // launch shape (autotuner)
  int ipl = 1, jchunk = 512, waves_per_simd = 8, blocks = 0;
  bool built = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void ens_release_graph(tupan_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

void ens_free_plan(tupan_ensemble *e) {
  ens_release_graph(e);
  if (e->items) (void)hipFree(e->items);
  if (e->part) (void)hipFree(e->part);
  if (e->cl) (void)hipFree(e->cl);
  e->items = nullptr;
  e->part = nullptr;
  e->cl = nullptr;
  e->built = false;
}

// Work items and partial-sum layout for the current launch shape.
int ens_plan(tupan_ensemble *e) {
  ens_free_plan(e);
  // the FP32C pair kernel takes one i-particle per lane
  const int P = e->mode == TUPAN_MODE_FP32C ? 1 : e->ipl;
  std::vector<ClusterInfo> cl(e->ncl);
  std::vector<Item> items;
  long long base = 0;
  for (int c = 0; c < e->ncl; ++c) {
    const int nc = e->n[c];
    const int nchunk = std::max(1, (nc + e->jchunk - 1) / e->jchunk);
    cl[c] = {e->off[c], nc, nchunk, base};
    base += (long long)nchunk * kMaxOut * nc;
    for (int i0 = 0; i0 < nc; i0 += 64 * P)
      for (int s = 0; s < nchunk; ++s) {
        const int j0 = s * e->jchunk;
        items.push_back({c, i0, j0, std::min(nc, j0 + e->jchunk), s});
      }
  }
  // longest first, ties by cluster then position (deterministic)
  std::stable_sort(items.begin(), items.end(), [](const Item &a,
                                                  const Item &b) {
    return (a.j1 - a.j0) > (b.j1 - b.j0);
  });
  e->nitems = (int)items.size();
  hipError_t er;
  if ((er = hipMalloc(&e->cl, sizeof(ClusterInfo) * e->ncl)) != hipSuccess ||
      (er = hipMalloc(&e->items, sizeof(Item) * std::max(1, e->nitems))) !=
          hipSuccess ||
      (er = hipMalloc(&e->part, sizeof(double) * std::max(1LL, base))) !=
          hipSuccess)
    return hip_fail(er, "ensemble plan allocation");
  if ((er = hipMemcpy(e->cl, cl.data(), sizeof(ClusterInfo) * e->ncl,
                      hipMemcpyHostToDevice)) != hipSuccess ||
      (er = hipMemcpy(e->items, items.data(), sizeof(Item) * e->nitems,
                      hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(er, "ensemble plan upload");
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  const long waves = (long)cus * 4 * e->waves_per_simd;
  e->blocks = (int)std::max(1L, std::min((long)(e->nitems + 3) / 4,
                                         waves / (kThreads / 64)));
  e->built = true;
  return TUPAN_OK;
}

template <class Pol>
void launch_pairs(tupan_ensemble *e, typename Pol::Params prm) {
  // the FP32C plan holds items of 64 i's (one per lane)
  if (e->ipl == 2 && e->mode != TUPAN_MODE_FP32C)
    hipLaunchKernelGGL((ens_pairs<Pol, 2>), dim3(e->blocks), dim3(kThreads),
                       0, e->st, e->rec, e->items, e->nitems, e->cl,
                       e->active, e->part, prm);
  else
    hipLaunchKernelGGL((ens_pairs<Pol, 1>), dim3(e->blocks), dim3(kThreads),
                       0, e->st, e->rec, e->items, e->nitems, e->cl,
                       e->active, e->part, prm);
}

// The kernels of one ensemble step, in stream order.
void ens_enqueue_step(tupan_ensemble *e) {
  Fields f;
  for (int k = 0; k < NFIELD; ++k) f.f[k] = e->fields[k];
  const int gb = blocks_for(e->ntot);
  Outs tsaj{{e->fields[FTS], e->fields[FTSIJ], e->fields[FAX],
             e->fields[FAY], e->fields[FAZ], e->fields[FJX], e->fields[FJY],
             e->fields[FJZ]}};
  Outs aj{{e->fields[FAX], e->fields[FAY], e->fields[FAZ], e->fields[FJX],
           e->fields[FJY], e->fields[FJZ]}};
  TSAJ::Params tp{e->eta};
  AJ::Params ap{};
  const bool f32c = e->mode == TUPAN_MODE_FP32C;
  // FP32C: tile boxes and origins, then the j records, of the current state
  auto tiles32 = [&]() {
    hipLaunchKernelGGL(ens_tiles, dim3(blocks_for(e->ntiles)),
                       dim3(kThreads), 0, e->st, e->ntiles, e->tcid, e->cl,
                       e->cctx, e->active, f, e->orig, e->boxes, e->porg);
    hipLaunchKernelGGL(ens_pack32, dim3(gb), dim3(kThreads), 0, e->st,
                       e->ntot, f, e->cid, e->cl, e->cctx, e->active,
                       e->orig, e->porg, e->rec32, e->frec32);
  };
  // Tstep and the derivatives at the step's start state: one pass
  if (f32c) {
    tiles32();
    hipLaunchKernelGGL((ens_tiles32<Tsaj32>), dim3(e->blocks),
                       dim3(kThreads), 0, e->st, e->rec32, e->frec32,
                       e->boxes, e->orig, e->porg, e->items, e->nitems,
                       e->cl, e->cctx, e->active, f, e->part, tp);
  } else {
    hipLaunchKernelGGL(ens_pack, dim3(gb), dim3(kThreads), 0, e->st, e->ntot,
                       f, e->cid, e->active, e->rec);
    launch_pairs<TSAJ>(e, tp);
  }
  hipLaunchKernelGGL((ens_finish<TSAJ>), dim3(gb), dim3(kThreads), 0, e->st,
                     e->ntot, e->cid, e->cl, e->active, e->part, tsaj, tp);
  hipLaunchKernelGGL(ens_choose_dt, dim3(e->ncl), dim3(kThreads), 0, e->st,
                     e->cl, e->fields[FTS], e->tc, e->active, e->dt, e->eta,
                     e->t_end, e->err);
  hipLaunchKernelGGL(ens_predict, dim3(gb), dim3(kThreads), 0, e->st,
                     e->ntot, f, e->cid, e->active, e->dt);
  for (int pass = 1; pass < 3; ++pass) {
    if (f32c) {
      tiles32();
      hipLaunchKernelGGL((ens_tiles32<AJ32E>), dim3(e->blocks),
                         dim3(kThreads), 0, e->st, e->rec32, e->frec32,
                         e->boxes, e->orig, e->porg, e->items, e->nitems,
                         e->cl, e->cctx, e->active, f, e->part, ap);
    } else {
      hipLaunchKernelGGL(ens_pack, dim3(gb), dim3(kThreads), 0, e->st,
                         e->ntot, f, e->cid, e->active, e->rec);
      launch_pairs<AJ>(e, ap);
    }
    hipLaunchKernelGGL((ens_finish<AJ>), dim3(gb), dim3(kThreads), 0, e->st,
                       e->ntot, e->cid, e->cl, e->active, e->part, aj, ap);
    hipLaunchKernelGGL(ens_correct, dim3(gb), dim3(kThreads), 0, e->st,
                       e->ntot, f, e->cid, e->active, e->dt);
  }
  hipLaunchKernelGGL(ens_clock, dim3(1), dim3(kThreads), 0, e->st, e->ncl,
                     e->cl, e->tc, e->dt, e->active, e->nsteps, e->pairs,
                     e->t_end, e->count);
}

int ens_capture(tupan_ensemble *e) {
  ens_release_graph(e);
  hipError_t er = hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal);
  if (er != hipSuccess) return hip_fail(er, "hipStreamBeginCapture");
  ens_enqueue_step(e);
  er = hipStreamEndCapture(e->st, &e->graph);
  if (er != hipSuccess) return hip_fail(er, "hipStreamEndCapture");
  er = hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0);
  if (er != hipSuccess) return hip_fail(er, "hipGraphInstantiate");
  return TUPAN_OK;
}

}  // namespace

extern "C" {

int tupan_ensemble_create(tupan_ensemble **out, int ncl, const int *sizes,
                          double eta, double t_end) {
  if (!out || !sizes || ncl <= 0)
    return set_error(TUPAN_ERR_ARG, "ensemble: bad arguments");
  if (!(eta != 0.0) || !(t_end > 0.0))
    return set_error(TUPAN_ERR_ARG, "ensemble: eta must be non-zero and "
                                    "t_end positive");
  auto *e = new tupan_ensemble();
  e->ncl = ncl;
  e->eta = eta;
  e->t_end = t_end;
  e->n.assign(sizes, sizes + ncl);
  e->off.resize(ncl);
  long long tot = 0;
  for (int c = 0; c < ncl; ++c) {
    if (sizes[c] <= 0 || sizes[c] > (1 << 24)) {
      delete e;
      return set_error(TUPAN_ERR_ARG, "ensemble: cluster %d has size %d", c,
                       sizes[c]);
    }
    e->off[c] = (int)tot;
    tot += sizes[c];
  }
  if (tot > (1LL << 30)) {
    delete e;
    return set_error(TUPAN_ERR_ARG, "ensemble: too many particles");
  }
  e->ntot = (int)tot;
  hipError_t er = hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking);
  if (er != hipSuccess) {
    delete e;
    return hip_fail(er, "hipStreamCreate");
  }
  std::vector<int> cid(e->ntot);
  for (int c = 0; c < ncl; ++c)
    std::fill(cid.begin() + e->off[c], cid.begin() + e->off[c] + e->n[c], c);
  // tiles numbered within each cluster, clusters back to back
  std::vector<int> tcid;
  e->tbase.resize(ncl);
  for (int c = 0; c < ncl; ++c) {
    e->tbase[c] = (int)tcid.size();
    tcid.insert(tcid.end(), (e->n[c] + kTile - 1) / kTile, c);
  }
  e->ntiles = (int)tcid.size();
  const size_t nt = (size_t)e->ntot;
  if ((er = hipMalloc(&e->slab, sizeof(double) * nt * NFIELD)) !=
          hipSuccess ||
      (er = hipMalloc(&e->rec, sizeof(AccJerkF64Rec) * nt)) != hipSuccess ||
      (er = hipMalloc(&e->cid, sizeof(int) * nt)) != hipSuccess ||
      (er = hipMalloc(&e->tc, sizeof(double) * ncl)) != hipSuccess ||
      (er = hipMalloc(&e->dt, sizeof(double) * ncl)) != hipSuccess ||
      (er = hipMalloc(&e->active, sizeof(int) * ncl)) != hipSuccess ||
      (er = hipMalloc(&e->nsteps, sizeof(long long) * ncl)) != hipSuccess ||
      (er = hipMalloc(&e->pairs, sizeof(double))) != hipSuccess ||
      (er = hipMalloc(&e->count, sizeof(int))) != hipSuccess ||
      (er = hipMalloc(&e->err, sizeof(int))) != hipSuccess ||
      (er = hipMalloc(&e->rec32, sizeof(EnsNearRec) * nt)) !=
          hipSuccess ||
      (er = hipMalloc(&e->frec32, sizeof(AJ32::FarRec) * nt)) !=
          hipSuccess ||
      (er = hipMalloc(&e->orig, sizeof(double) * 4 * e->ntiles)) !=
          hipSuccess ||
      (er = hipMalloc(&e->porg, sizeof(double) * 4 * e->ntiles)) !=
          hipSuccess ||
      (er = hipMalloc(&e->boxes, sizeof(TileBox) * e->ntiles)) !=
          hipSuccess ||
      (er = hipMalloc(&e->cctx, sizeof(ClusterCtx) * ncl)) != hipSuccess ||
      (er = hipMalloc(&e->tcid, sizeof(int) * e->ntiles)) != hipSuccess) {
    int rc = hip_fail(er, "ensemble allocation");
    tupan_ensemble_destroy(e);
    return rc;
  }
  for (int k = 0; k < NFIELD; ++k) e->fields[k] = e->slab + (size_t)k * nt;
  (void)hipMemsetAsync(e->slab, 0, sizeof(double) * nt * NFIELD, e->st);
  (void)hipMemsetAsync(e->tc, 0, sizeof(double) * ncl, e->st);
  (void)hipMemsetAsync(e->dt, 0, sizeof(double) * ncl, e->st);
  (void)hipMemsetAsync(e->nsteps, 0, sizeof(long long) * ncl, e->st);
  er = hipMemcpyAsync(e->cid, cid.data(), sizeof(int) * nt,
                      hipMemcpyHostToDevice, e->st);
  if (er == hipSuccess)
    er = hipMemcpyAsync(e->tcid, tcid.data(), sizeof(int) * e->ntiles,
                        hipMemcpyHostToDevice, e->st);
  if (er == hipSuccess) er = hipStreamSynchronize(e->st);
  if (er != hipSuccess) {
    int rc = hip_fail(er, "ensemble init");
    tupan_ensemble_destroy(e);
    return rc;
  }
  *out = e;
  return TUPAN_OK;
}

int tupan_ensemble_destroy(tupan_ensemble *e) {
  if (!e) return TUPAN_OK;
  if (e->st) (void)hipStreamSynchronize(e->st);
  ens_free_plan(e);
  void *ptrs[] = {e->slab,   e->rec,    e->cid,   e->tc,
                  e->dt,     e->active, e->nsteps, e->pairs,
                  e->count,  e->err,    e->rec32, e->orig,
                  e->frec32, e->porg,   e->boxes,
                  e->cctx,   e->tcid};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (e->st) (void)hipStreamDestroy(e->st);
  delete e;
  return TUPAN_OK;
}

int tupan_ensemble_set_launch(tupan_ensemble *e, int ipl, int jchunk,
                              int waves_per_simd) {
  if (!e) return set_error(TUPAN_ERR_ARG, "NULL ensemble");
  if ((ipl != 1 && ipl != 2) || jchunk < 16 || waves_per_simd < 1 ||
      waves_per_simd > 32)
    return set_error(TUPAN_ERR_ARG, "ensemble launch: ipl 1|2, jchunk >= 16,"
                                    " 1 <= waves_per_simd <= 32");
  if (e->mode == TUPAN_MODE_FP32C && jchunk % kTile != 0)
    return set_error(TUPAN_ERR_ARG, "FP32C needs a j chunk that is a "
                                    "multiple of %d", kTile);
  (void)hipStreamSynchronize(e->st);
  e->ipl = ipl;
  e->jchunk = jchunk;
  e->waves_per_simd = waves_per_simd;
  ens_free_plan(e);
  return TUPAN_OK;
}

int tupan_ensemble_set_mode(tupan_ensemble *e, int mode) {
  if (!e) return set_error(TUPAN_ERR_ARG, "NULL ensemble");
  if (mode != TUPAN_MODE_FP64 && mode != TUPAN_MODE_FP32C)
    return set_error(TUPAN_ERR_ARG, "ensemble mode: FP64 or FP32C");
  if (mode == TUPAN_MODE_FP32C && e->jchunk % kTile != 0)
    return set_error(TUPAN_ERR_ARG, "FP32C needs a j chunk that is a "
                                    "multiple of %d", kTile);
  (void)hipStreamSynchronize(e->st);
  e->mode = mode;
  ens_free_plan(e);
  return TUPAN_OK;
}

// fields: m, rx, ry, rz, e2, vx, vy, vz (host, ntot each); t0: per-cluster
// start times (may be NULL = 0).  Resets the step counters.
int tupan_ensemble_set_order(tupan_ensemble *e, int order) {
  if (!e) return set_error(TUPAN_ERR_ARG, "NULL ensemble");
  if (order != TUPAN_ENS_ORDER_INPUT && order != TUPAN_ENS_ORDER_MORTON)
    return set_error(TUPAN_ERR_ARG, "ensemble order: INPUT or MORTON");
  e->order = order;
  return TUPAN_OK;
}

int tupan_ensemble_set_state(tupan_ensemble *e, const double *const *fields,
                             const double *t0) {
  if (!e || !fields) return set_error(TUPAN_ERR_ARG, "NULL argument");
  const size_t nt = (size_t)e->ntot;
  for (int k = 0; k < 8; ++k)
    if (!fields[k]) return set_error(TUPAN_ERR_ARG, "NULL field %d", k);
  e->perm.resize(nt);
  for (int c = 0; c < e->ncl; ++c)
    morton_order(fields[FX], fields[FY], fields[FZ], e->off[c], e->n[c],
                 e->order == TUPAN_ENS_ORDER_MORTON, e->perm.data());
  {
    std::vector<double> stage(8 * nt);
    for (int k = 0; k < 8; ++k)
      for (size_t i = 0; i < nt; ++i)
        stage[k * nt + i] = fields[k][e->perm[i]];
    for (int k = 0; k < 8; ++k) {
      hipError_t er = hipMemcpyAsync(e->fields[k], stage.data() + k * nt,
                                     sizeof(double) * nt,
                                     hipMemcpyHostToDevice, e->st);
      if (er != hipSuccess) return hip_fail(er, "ensemble upload");
    }
    // stage must outlive the copies
    hipError_t er = hipStreamSynchronize(e->st);
    if (er != hipSuccess) return hip_fail(er, "ensemble upload");
  }
  // FP32C near-field context of each cluster (the single-system rule,
  // ctx_final_kernel: factor = c^2 n^(1/3) / sum m, largest e2)
  std::vector<ClusterCtx> cx(e->ncl);
  const double cnear = kEnsNearC;
  for (int c = 0; c < e->ncl; ++c) {
    double msum = 0.0, e2max = 0.0;
    for (int i = e->off[c]; i < e->off[c] + e->n[c]; ++i) {
      msum += fields[FM][i];
      e2max = std::fmax(e2max, fields[FE2][i]);
    }
    const double f = cnear * cnear * std::cbrt((double)e->n[c]) / msum;
    cx[c].factor = (cnear > 0.0 && msum > 0.0 && f < 1e300) ? f : 0.0;
    cx[c].e2max = e2max;
    cx[c].tbase = e->tbase[c];
    cx[c].pad = 0;
  }
  {
    hipError_t er = hipMemcpyAsync(e->cctx, cx.data(),
                                   sizeof(ClusterCtx) * e->ncl,
                                   hipMemcpyHostToDevice, e->st);
    if (er != hipSuccess) return hip_fail(er, "ensemble upload");
    // cx must outlive the copy
    er = hipStreamSynchronize(e->st);
    if (er != hipSuccess) return hip_fail(er, "ensemble upload");
  }
  std::vector<double> t(e->ncl, 0.0);
  if (t0) t.assign(t0, t0 + e->ncl);
  hipError_t er = hipMemcpyAsync(e->tc, t.data(), sizeof(double) * e->ncl,
                                 hipMemcpyHostToDevice, e->st);
  if (er == hipSuccess)
    er = hipMemsetAsync(e->nsteps, 0, sizeof(long long) * e->ncl, e->st);
  if (er != hipSuccess) return hip_fail(er, "ensemble upload");
  hipLaunchKernelGGL(ens_init_flags, dim3(1), dim3(1), 0, e->st, e->ncl,
                     e->tc, e->t_end, e->active, e->count, e->err, e->pairs);
  er = hipStreamSynchronize(e->st);
  return er == hipSuccess ? TUPAN_OK : hip_fail(er, "ensemble upload");
}

// out: rx, ry, rz, vx, vy, vz (host, ntot each); tc, nsteps per cluster (may
// be NULL).
int tupan_ensemble_get_state(tupan_ensemble *e, double *const *out,
                             double *tc, long long *nsteps) {
  if (!e) return set_error(TUPAN_ERR_ARG, "NULL ensemble");
  const size_t nt = (size_t)e->ntot;
  if (out && e->perm.size() != nt)
    return set_error(TUPAN_ERR_ARG, "ensemble: no state set");
  hipError_t er = hipSuccess;
  std::vector<double> stage(out ? nt : 0);
  for (int k = 0; out && k < 6 && er == hipSuccess; ++k)
    if (out[k]) {
      er = hipMemcpyAsync(stage.data(),
                          e->fields[k < 3 ? FX + k : FVX + (k - 3)],
                          sizeof(double) * nt, hipMemcpyDeviceToHost, e->st);
      if (er == hipSuccess) er = hipStreamSynchronize(e->st);
      if (er == hipSuccess)
        for (size_t i = 0; i < nt; ++i) out[k][e->perm[i]] = stage[i];
    }
  if (er == hipSuccess && tc)
    er = hipMemcpyAsync(tc, e->tc, sizeof(double) * e->ncl,
                        hipMemcpyDeviceToHost, e->st);
  if (er == hipSuccess && nsteps)
    er = hipMemcpyAsync(nsteps, e->nsteps, sizeof(long long) * e->ncl,
                        hipMemcpyDeviceToHost, e->st);
  if (er == hipSuccess) er = hipStreamSynchronize(e->st);
  return er == hipSuccess ? TUPAN_OK : hip_fail(er, "ensemble download");
}

// Replays the captured step `steps` times (asynchronously).  graph = 0 runs
// the same kernels launched one by one instead (for comparison).
int tupan_ensemble_step(tupan_ensemble *e, int steps, int graph) {
  if (!e) return set_error(TUPAN_ERR_ARG, "NULL ensemble");
  int rc;
  if (!e->built && (rc = ens_plan(e))) return rc;
  if (graph && !e->exec && (rc = ens_capture(e))) return rc;
  for (int s = 0; s < steps; ++s) {
    if (graph) {
      hipError_t er = hipGraphLaunch(e->exec, e->st);
      if (er != hipSuccess) return hip_fail(er, "hipGraphLaunch");
    } else {
      ens_enqueue_step(e);
    }
  }
  hipError_t er = hipGetLastError();
  return er == hipSuccess ? TUPAN_OK : hip_fail(er, "ensemble step");
}

// Waits for the queued steps; returns the number of unfinished clusters,
// the pair interactions evaluated since set_state, and the error flag
// (1: a cluster met a zero or non-finite time step and was stopped).
int tupan_ensemble_poll(tupan_ensemble *e, int *unfinished, double *pairs,
                        int *error) {
  if (!e) return set_error(TUPAN_ERR_ARG, "NULL ensemble");
  int c = 0, ef = 0;
  double p = 0.0;
  hipError_t er = hipMemcpyAsync(&c, e->count, sizeof(int),
                                 hipMemcpyDeviceToHost, e->st);
  if (er == hipSuccess)
    er = hipMemcpyAsync(&p, e->pairs, sizeof(double), hipMemcpyDeviceToHost,
                        e->st);
  if (er == hipSuccess)
    er = hipMemcpyAsync(&ef, e->err, sizeof(int), hipMemcpyDeviceToHost,
                        e->st);
  if (er == hipSuccess) er = hipStreamSynchronize(e->st);
  if (er != hipSuccess) return hip_fail(er, "ensemble poll");
  if (unfinished) *unfinished = c;
  if (pairs) *pairs = p;
  if (error) *error = ef;
  return TUPAN_OK;
}

}  // extern "C"
