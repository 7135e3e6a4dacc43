// launch.hpp -- host-side launch of the pair engine and the shared runtime
// helpers (errors, launch-shape choice, host<->device staging).
#pragma once
#include <hip/hip_runtime.h>
#include <climits>
#include <cstddef>
#include <functional>

#include "../../include/tupan_hip.h"
#include "engine.hpp"

namespace tupan {

// ---- runtime.hip ----
int set_error(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
struct LaunchShape {
  int ipl;     // i-particles per lane
  int splits;  // j-range splits
};
LaunchShape choose_shape(int kernel_id, int mode, size_t ni, size_t nj,
                         int pref_ipl, int max_ipl, int min_chunk,
                         int waves_per_simd = 64, int max_splits = 0);
int device_alloc(void **p, size_t bytes, hipStream_t st);
void device_free(void *p, hipStream_t st);
// Reusable per-(device, stream, slot) device buffers for the asynchronous
// entry points (no allocation in steady state; safe to capture).  Each
// buffer in simultaneous use within one call needs its own slot.
enum ScratchSlot { kSlotRecords = 0, kSlotPartials, kSlotIeee, kSlotReduce,
                   kSlotPairwise, kSlotMask, kSlotSorted };

// Largest near mask (sparse policies) the launcher builds; above it the
// engine tests the pairs itself.
constexpr size_t kMaxMaskBytes = size_t(64) << 20;
int scratch_alloc(int slot, void **p, size_t bytes, hipStream_t st);
void scratch_free(int slot, void *p, hipStream_t st);
// Optional live timing of the pair-engine launches (tupan_set_timing):
// returns the event pair to record around the launch, or nulls when off.
bool timing_begin(hipEvent_t *e0, hipEvent_t *e1, hipStream_t st);

using DevFn = std::function<int(const double *const *, const double *const *,
                                double *const *, hipStream_t)>;
// Stages host arrays in HBM, runs fn on device pointers, copies the nout
// outputs back and synchronises.  Fields shared by i and j (same host
// pointer, same length) are copied once.
int host_call(size_t ni, const double *const *ip, size_t nj,
              const double *const *jp, int nfields, double *const *op,
              int nout, const DevFn &fn);

// ---- ieee.hip: TUPAN_MODE_FP64_IEEE / FP32_IEEE (Phi, Acc, AccJerk,
// SnapCrackle, Tstep) ----
int run_ieee(int kid, bool fp32, size_t ni, const double *const *ip,
             size_t nj, const double *const *jp, double *const *op,
             hipStream_t st);
int run_ieee_tstep(bool fp32, size_t ni, const double *const *ip, size_t nj,
                   const double *const *jp, double eta, double *const *op,
                   hipStream_t st);

// ---- generic engine launch ----
double near_field_factor();
bool sorted_enabled();
double flag_tau();
// the flag word of the last sorted FP32C call (tupan_last_flagged)
void set_last_flags(const int *flags, hipStream_t st);
int sakura_budget();

// FP64 policies read their j records straight from the caller's arrays
// (pair_engine's Direct variant) up to this many j's: below it the packing
// launch costs more than one-record loads save (measured: the Direct pair
// loop is ~13% slower at nj = 16384 and ~40% at 65536 for AccJerk).
constexpr int kDirectMaxJ = 4096;

template <class Pol, int P>
static void launch_p(dim3 grid, hipStream_t st, const Soa &is, const Soa &js,
                     int ni, const typename Pol::JRec *jr, int nj, int jchunk,
                     int nrows, double *part, const Out &out,
                     const typename Pol::Params &prm, const double *aux) {
  if constexpr (!Pol::kTwoLevel) {
    if (!jr) {
      hipLaunchKernelGGL((pair_engine<Pol, P, true>), grid, dim3(kBlock), 0,
                         st, is, js, ni, jr, nj, jchunk, nrows, part, out,
                         prm, aux);
      return;
    }
  }
  hipLaunchKernelGGL((pair_engine<Pol, P>), grid, dim3(kBlock), 0, st, is,
                     js, ni, jr, nj, jchunk, nrows, part, out, prm, aux);
}

template <class Pol, class = void>
struct VelOrigins {
  static constexpr int value = 0;
};
template <class Pol>
struct VelOrigins<Pol, decltype((void)Pol::kVelOrigins)> {
  static constexpr int value = Pol::kVelOrigins;
};
// factor of the stored velocity origins (1 unless the policy says)
template <class Pol, class = void>
struct OrgScale {
  static constexpr double value = 1.0;
};
template <class Pol>
struct OrgScale<Pol, decltype((void)Pol::kOrgScale)> {
  static constexpr double value = Pol::kOrgScale;
};
// doubles per tile of the sorted path's origins: 4 per origin triple
template <class Pol>
struct OrgStride {
  static constexpr int value =
      4 * (VelOrigins<Pol>::value > 0 ? VelOrigins<Pol>::value : 1);
};

}  // namespace tupan
#include "sorted.hpp"
namespace tupan {

// FP32C over Morton-sorted particles (sorted.hpp), one i per lane.
template <class Pol>
int run_sorted(int kernel_id, int mode, int n_i, const Soa &is, int n_j,
               const Soa &js, const Out &out, hipStream_t st) {
  using P64 = typename Pol::Exact;
  bool same = n_i == n_j;
  for (int k = 0; k < Pol::kFields; ++k) same = same && is.f[k] == js.f[k];
  const SortGeom gj = sort_geom(n_j), gi = sort_geom(n_i);
  const int nseg = recompute_groups(n_j);
  const SortedLayout L = sorted_layout<Pol>(n_i, n_j, same, gj, gi, nseg);
  // 128 waves per SIMD over the grid's lifetime, in at most 128 j chunks:
  // the best or within 1% of it for Phi, Acc and AccJerk at N = 16K-64K in
  // a sweep of 16-512 chunks (at N = 32768 the 256 chunks that the wave
  // target alone gives cost Phi 8% and AccJerk 4%: shorter tile loops for
  // the same per-wave set-up and block-row partials)
  const LaunchShape sh =
      choose_shape(kernel_id, mode, (size_t)n_i, (size_t)n_j, 1, 1, kTile,
                   128, 128);
  int S = sh.splits < kSWaves ? kSWaves : sh.splits;
  S = (S + kSWaves - 1) / kSWaves * kSWaves;
  int jchunk = (n_j + S - 1) / S;
  jchunk = (jchunk + kTile - 1) / kTile * kTile;
  const int used = (n_j + jchunk - 1) / jchunk;
  const int SB = used > kSWaves ? (used + kSWaves - 1) / kSWaves : 1;
  const int waves = (n_i + 63) / 64;
  if (SB > 65535)
    return set_error(TUPAN_ERR_ARG, "too many j chunks (%d block rows)", SB);

  char *buf = nullptr;
  double *part = nullptr;
  int rc = scratch_alloc(kSlotSorted, (void **)&buf, L.total, st);
  if (rc) return rc;
  rc = scratch_alloc(kSlotPartials, (void **)&part,
                     sizeof(double) * (size_t)SB * Pol::kPart * (size_t)n_i,
                     st);
  if (rc) {
    scratch_free(kSlotSorted, buf, st);
    return rc;
  }
  auto at = [&](size_t off) { return (void *)(buf + off); };
  double *ctxpart = (double *)at(L.ctxpart);
  double *ctx = (double *)at(L.ctx);
  int *flags = (int *)at(L.flags);
  SoaW sjw{}, siw{};
  Soa sjr{}, sir{};
  for (int k = 0; k < Pol::kFields; ++k) {
    sjw.f[k] = (double *)at(L.sj) + (size_t)k * n_j;
    siw.f[k] = (double *)at(L.si) + (size_t)k * n_i;
    sjr.f[k] = sjw.f[k];
    sir.f[k] = siw.f[k];
  }
  hipLaunchKernelGGL(sctx_partial_kernel, dim3(kCtxBlocks), dim3(256), 0, st,
                     js, n_j, is, n_i, ctxpart, (int *)at(L.zero),
                     (int)L.nzero);
  // one LSD radix sort per set: keys and indices ping-pong 0 -> 1 -> 0 -> 1
  auto sort_set = [&](int set, const Soa &f, int n, int box, SortGeom g,
                      double *ctx_out) -> const int * {
    uint32_t *kb[2] = {(uint32_t *)at(L.keys[set][0]),
                       (uint32_t *)at(L.keys[set][1])};
    int *vb[2] = {(int *)at(L.vals[set][0]), (int *)at(L.vals[set][1])};
    int *hist[kSortPasses];
    for (int q = 0; q < kSortPasses; ++q) hist[q] = (int *)at(L.hist[set][q]);
    hipLaunchKernelGGL(morton_key_kernel, dim3(g.nb), dim3(256), 0, st,
                       f.f[RX], f.f[RY], f.f[RZ], n, (const double *)ctxpart,
                       box, g, kb[0], vb[0], hist[0], ctx_out, n_j,
                       near_field_factor());
    int dest_bits = 0;
    while ((1 << dest_bits) < g.nb) ++dest_bits;
    for (int q = 0; q < kSortPasses; ++q)
      hipLaunchKernelGGL(radix_scatter_kernel, dim3(g.nb), dim3(256), 0, st,
                         (const uint32_t *)kb[q & 1], (const int *)vb[q & 1],
                         kb[(q + 1) & 1], vb[(q + 1) & 1], n, g,
                         (const int *)hist[q], q * kSortDigit,
                         q + 1 < kSortPasses ? hist[q + 1] : nullptr,
                         dest_bits);
    return vb[kSortPasses & 1];
  };
  const int *perm_j = sort_set(0, js, n_j, 2, gj, ctx);
  const int *perm_i = same ? perm_j : sort_set(1, is, n_i, 8, gi, nullptr);
  double *origins = (double *)at(L.origins);
  const TileBox *boxes = (const TileBox *)at(L.boxes);
  auto *recs = (typename Pol::JRec *)at(L.recs);
  double *porg = (double *)at(L.porg);
  auto *frecs = (typename Pol::FarRec *)at(L.frecs);
  auto *recs64 = (typename Pol::Exact::JRec *)at(L.recs64);
  hipLaunchKernelGGL((sorted_pack_kernel<Pol>), dim3((n_j + 255) / 256),
                     dim3(256), 0, st, js, n_j, perm_j, sjw,
                     (const double *)ctx, recs, origins, (TileBox *)boxes,
                     porg, frecs, recs64);
  if (!same)
    hipLaunchKernelGGL((sorted_gather_kernel<Pol::kFields>),
                       dim3((n_i + 255) / 256), dim3(256), 0, st, is, n_i,
                       perm_i, siw);
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  const bool timed = timing_begin(&ev0, &ev1, st);
  if (timed) (void)hipEventRecord(ev0, st);
  hipLaunchKernelGGL((sorted_engine<Pol>), dim3((unsigned)waves, (unsigned)SB),
                     dim3(kSBlock), 0, st, sir, n_i, sjr,
                     (const typename Pol::JRec *)recs, n_j, boxes, jchunk, part,
                     (const double *)origins,
                     (const typename Pol::FarRec *)frecs,
                     (const double *)porg,
                     (const typename Pol::Exact::JRec *)recs64);
  if (timed) (void)hipEventRecord(ev1, st);
  hipLaunchKernelGGL((sorted_finish_kernel<Pol>), dim3((n_i + 255) / 256),
                     dim3(256), 0, st, (const double *)part, SB, n_i, perm_i,
                     out, flag_tau() * flag_tau(), flags);
  double *rpart = (double *)at(L.rpart);
  {
    // at most nseg groups of kWaves chunks per row-wave (recompute_shape)
    const int rblocks = (kRecomputeRows / 64) * nseg;
    hipLaunchKernelGGL((sorted_recompute_kernel<P64>), dim3(rblocks),
                       dim3(kBlock), 0, st, sir, sjr, n_j, (const int *)flags,
                       nseg, rpart);
    hipLaunchKernelGGL((sorted_recompute_finish<P64>), dim3(kRecomputeBlocks),
                       dim3(256), 0, st, sir, sjr, n_j, (const int *)flags,
                       rblocks, nseg, (const double *)rpart, perm_i, out);
    set_last_flags(flags, st);
  }
  const hipError_t e = hipGetLastError();
  scratch_free(kSlotPartials, part, st);
  scratch_free(kSlotSorted, buf, st);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return TUPAN_OK;
}

// FP32C calls with at most this many j's (and kSmallCtxI i's) build their
// context, tile origins and records in one single-block launch
// (small_pack_kernel) instead of four: at these sizes each launch costs more
// than its work.
constexpr int kSmallCtxJ = 4096;
constexpr int kSmallCtxI = 65536;

// ctx_partial_kernel + ctx_final_kernel + tile_origin_kernel + pack_kernel
// in one block of 256 threads: sum of the j masses and largest i-side e2
// (tree-reduced in LDS, a fixed order), the near-field factor, the per-tile
// origins (written to ctx + 2 for the pair kernel and kept in LDS for the
// packing), then the records.  nj <= kSmallCtxJ.
template <class Pol>
__global__ __launch_bounds__(256) void small_pack_kernel(
    Soa is, int ni, Soa js, int nj, double c, typename Pol::JRec *jr,
    double *ctx) {
  __shared__ double ps[256], px[256];
  __shared__ double org[4 * (kSmallCtxJ / kTile)];
  const int t = threadIdx.x;
  double s = 0.0, x = 0.0;
  for (int j = t; j < nj; j += 256) s += js.f[M][j];
  for (int i = t; i < ni; i += 256) x = fmax(x, is.f[E2][i]);
  ps[t] = s;
  px[t] = x;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      ps[t] += ps[t + w];
      px[t] = fmax(px[t], px[t + w]);
    }
    __syncthreads();
  }
  const double msum = ps[0], e2max = px[0];
  const double f = c * c * cbrt((double)nj) / msum;
  const double factor = (c > 0.0 && msum > 0.0 && f < 1e300) ? f : 0.0;
  if (VelOrigins<Pol>::value) {
    const int ntiles = (nj + kTile - 1) / kTile;
    for (int tt = t; tt < ntiles; tt += 256) {
      const int j0 = tt * kTile, je = min(nj, j0 + kTile);
      for (int k = 0; k < 3; ++k) {
        const double *v = js.f[VX + k];
        double lo = v[j0], hi = lo;
        for (int j = j0 + 1; j < je; ++j) {
          lo = fmin(lo, v[j]);
          hi = fmax(hi, v[j]);
        }
        org[4 * tt + k] = 0.5 * (lo + hi) * OrgScale<Pol>::value;
        ctx[2 + 4 * tt + k] = 0.5 * (lo + hi) * OrgScale<Pol>::value;
      }
    }
  }
  if (t == 0) {
    ctx[0] = factor;
    ctx[1] = e2max;
  }
  __syncthreads();
  for (int j = t; j < nj; j += 256)
    jr[j] = Pol::pack(js, j, PackCtx{factor, e2max, org});
}

// A caller-owned device buffer that replaces the scratch cache for one
// call (the no-allocation entry points): plain FP64 policies only; the j
// split is reduced until the records and block-row partials fit.
struct FixedScratch {
  char *base;
  size_t bytes;
};

template <class Pol>
int run_engine(int kernel_id, int mode, size_t ni, const double *const *ip,
               size_t nj, const double *const *jp, double *const *op,
               const typename Pol::Params &prm, hipStream_t st,
               const FixedScratch *fixed = nullptr) {
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
  static_assert(Pol::kOut <= 8, "at most 8 outputs");
  for (int k = 0; k < Pol::kOut; ++k) {
    if (!op[k]) return set_error(TUPAN_ERR_ARG, "NULL output field %d", k);
    out.f[k] = op[k];
  }
  const int n_i = (int)ni, n_j = (int)nj;
  if constexpr (Pol::kTwoLevel) {
    // large FP32C calls: Morton-sorted particles (sorted.hpp)
    if (n_j > kSmallCtxJ && sorted_enabled())
      return run_sorted<typename Pol::Sorted>(kernel_id, mode, n_i, is, n_j,
                                              js, out, st);
  }
  // j chunks are whole tiles for the tiled loops (FP32C, sparse policies);
  // the plain FP64 loop takes chunks of 8 records, which tiny systems use
  constexpr int kChunk =
      (Pol::kTwoLevel || SparsePairs<Pol>::value) ? kTile : 8;
  LaunchShape sh = choose_shape(kernel_id, mode, ni, nj, Pol::kPrefIpl,
                                Pol::kMaxIpl, kChunk);
  const int P = sh.ipl;
  // S j chunks of jchunk records (a multiple of kChunk), dealt kWaves per
  // block: SB block rows (grid: i-waves x SB)
  int S = sh.splits < kWaves ? kWaves : sh.splits;
  S = (S + kWaves - 1) / kWaves * kWaves;
  int jchunk = (n_j + S - 1) / S;
  jchunk = (jchunk + kChunk - 1) / kChunk * kChunk;
  if (jchunk < kChunk) jchunk = kChunk;
  int used = (n_j + jchunk - 1) / jchunk;  // chunks holding records
  int SB = used > kWaves ? (used + kWaves - 1) / kWaves : 1;
  const int waves = (n_i + 64 * P - 1) / (64 * P);  // i-waves
  if (SB > 65535)
    return set_error(TUPAN_ERR_ARG, "too many j chunks (%d block rows)", SB);

  typename Pol::JRec *jr = nullptr;
  double *part = nullptr;
  double *ctx = nullptr;  // FP32C: factor, e2max, per-tile origins
  int rc;
  const size_t ntiles = ((size_t)n_j + kTile - 1) / kTile;
  const size_t rec_bytes =
      ((sizeof(typename Pol::JRec) * (size_t)n_j + 255) / 256) * 256;
  const size_t ctx_bytes =
      Pol::kTwoLevel ? sizeof(double) * (2 + 4 * ntiles + 2 * kCtxBlocks)
                     : 0;
  // FP64 policies at small nj build their records in the pair kernel
  const bool direct = !Pol::kTwoLevel && n_j <= kDirectMaxJ;
  if (fixed) {
    if (Pol::kTwoLevel || SparsePairs<Pol>::value)
      return set_error(TUPAN_ERR_ARG, "fixed scratch: FP64 policies only");
    const size_t rec = direct ? 0 : rec_bytes;
    auto part_bytes = [&](int sb) -> size_t {
      return sb > 1 ? ((sizeof(double) * (size_t)sb * Pol::kOut * n_i + 255) /
                       256) * 256
                    : 0;
    };
    // one block row fewer at a time until records and partials fit
    while (SB > 1 && rec + part_bytes(SB) > fixed->bytes) {
      S = (SB - 1) * kWaves;
      jchunk = (n_j + S - 1) / S;
      jchunk = (jchunk + kChunk - 1) / kChunk * kChunk;
      used = (n_j + jchunk - 1) / jchunk;
      SB = used > kWaves ? (used + kWaves - 1) / kWaves : 1;
    }
    if (rec + part_bytes(SB) > fixed->bytes)
      return set_error(TUPAN_ERR_ARG,
                       "the call needs %zu bytes of scratch; the fixed "
                       "buffer holds %zu", rec + part_bytes(SB),
                       fixed->bytes);
    if (rec) jr = reinterpret_cast<typename Pol::JRec *>(fixed->base);
    if (SB > 1) part = reinterpret_cast<double *>(fixed->base + rec);
  } else if (n_j > 0 && !direct) {
    rc = scratch_alloc(kSlotRecords, (void **)&jr, rec_bytes + ctx_bytes, st);
    if (rc) return rc;
    if (Pol::kTwoLevel)
      ctx = reinterpret_cast<double *>(reinterpret_cast<char *>(jr) +
                                       rec_bytes);
  }
  if (SB > 1 && !fixed) {
    rc = scratch_alloc(kSlotPartials, (void **)&part,
                       sizeof(double) * (size_t)SB * Pol::kOut * (size_t)n_i,
                       st);
    if (rc) {
      scratch_free(kSlotRecords, jr, st);
      return rc;
    }
  }
  const bool small = Pol::kTwoLevel && n_j <= kSmallCtxJ &&
                     n_i <= kSmallCtxI;
  if (n_j > 0 && small) {
    if constexpr (Pol::kTwoLevel)
      hipLaunchKernelGGL((small_pack_kernel<Pol>), dim3(1), dim3(256), 0, st,
                         is, n_i, js, n_j, near_field_factor(), jr, ctx);
  } else if (n_j > 0) {
    if (Pol::kTwoLevel) {
      double *scratch = ctx + 2 + 4 * ntiles;
      hipLaunchKernelGGL(ctx_partial_kernel, dim3(kCtxBlocks), dim3(256), 0,
                         st, js.f[M], n_j, is.f[E2], n_i, scratch);
      hipLaunchKernelGGL(ctx_final_kernel, dim3(1), dim3(64), 0, st, scratch,
                         n_j, near_field_factor(), ctx);
      if (VelOrigins<Pol>::value)
        hipLaunchKernelGGL(tile_origin_kernel,
                           dim3((unsigned)((ntiles + 255) / 256)), dim3(256),
                           0, st, js.f[VX], js.f[VY], js.f[VZ], n_j,
                           OrgScale<Pol>::value, ctx + 2);
    }
    if (!direct)
      hipLaunchKernelGGL((pack_kernel<Pol>), dim3((n_j + 255) / 256),
                         dim3(256), 0, st, js, n_j, jr, (const double *)ctx);
  }
  // sparse policies: the near mask, when it fits (P = 1 only)
  typename Pol::Params prm2 = prm;
  uint32_t *mask = nullptr;
  if constexpr (SparsePairs<Pol>::value) {
    const size_t mtiles = ((size_t)n_j + kTile - 1) / kTile;
    const size_t mbytes = sizeof(uint32_t) * (size_t)waves * mtiles;
    // one wave per (i-block, 64 j's), kWaves per block
    const size_t mrows = ((size_t)n_j + 64 * kWaves - 1) / (64 * kWaves);
    if (P == 1 && n_j > 0 && Pol::sparse(prm) && mbytes <= kMaxMaskBytes &&
        mrows <= 65535) {
      rc = scratch_alloc(kSlotMask, (void **)&mask, mbytes, st);
      if (rc) {
        scratch_free(kSlotPartials, part, st);
        scratch_free(kSlotRecords, jr, st);
        return rc;
      }
      int same = 1;
      for (int k = 0; k < Pol::kFields; ++k) same &= is.f[k] == js.f[k];
      hipLaunchKernelGGL((near_mask_kernel<Pol>),
                         dim3((unsigned)waves, (unsigned)mrows), dim3(kBlock),
                         0, st, is, n_i, js, n_j, (int)mtiles, prm, same,
                         mask);
      prm2.mask = mask;
      prm2.mask_tiles = (int)mtiles;
    }
  }
  dim3 grid((unsigned)waves, (unsigned)SB);
  const double *aux = ctx ? ctx + 2 : nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  const bool timed = !fixed && timing_begin(&ev0, &ev1, st);
  if (timed) (void)hipEventRecord(ev0, st);
  switch (P) {
    case 4:
      launch_p<Pol, (Pol::kMaxIpl >= 4 ? 4 : 1)>(grid, st, is, js, n_i, jr, n_j,
                                                  jchunk, SB, part, out, prm2, aux);
      break;
    case 2:
      launch_p<Pol, (Pol::kMaxIpl >= 2 ? 2 : 1)>(grid, st, is, js, n_i, jr, n_j,
                                                  jchunk, SB, part, out, prm2, aux);
      break;
    default:
      launch_p<Pol, 1>(grid, st, is, js, n_i, jr, n_j, jchunk, SB, part, out,
                       prm2, aux);
  }
  if (timed) (void)hipEventRecord(ev1, st);
  if (SB > 1)
    hipLaunchKernelGGL((finish_kernel<Pol>), dim3((n_i + 255) / 256),
                       dim3(256), 0, st, part, SB, n_i, out, prm);
  hipError_t e = hipGetLastError();
  if (mask) scratch_free(kSlotMask, mask, st);
  if (!fixed) {
    scratch_free(kSlotPartials, part, st);
    scratch_free(kSlotRecords, jr, st);
  }
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return TUPAN_OK;
}

}  // namespace tupan
