// sorted.hpp -- the FP32C pair sums of large calls over Morton-sorted
// particles.
//
// Why: FP32 pair arithmetic meets 1e-6 per particle only when the largest
// pair terms of every particle are summed in FP64 (a near field of several
// mean interparticle spacings) and the rare rows whose sums cancel badly
// are recomputed in FP64.  With the particles in input order the near
// pairs of a wave's 64 i's touch most j's, and each costs an FP64 pair for
// the whole wave; in Morton order a wave's i's and a tile's j's are
// compact, the union of a wave's near j's stays ~1% of nj, and a tile
// whose bounding box is beyond every near-field threshold of its j's needs
// no per-pair near test at all.
//
// One call (ni i's, nj j's; i and j sorted separately unless they are the
// same arrays):
//   sctx_partial_kernel   per-block sums/extrema: sum m_j, max e2_i, the
//                         position boxes; clears the radix histograms of
//                         passes 1..2 and the flagged-row count
//   morton_key_kernel     30-bit Morton key (10 bits per axis over the
//                         set's box), histogram of the first 10-bit digit;
//                         block 0 writes the near-field context
//   radix_scatter_kernel  x3: stable LSD radix passes of 10 bits; each
//                         block sums its digit offsets from the per-block
//                         histograms, ranks its keys with wave ballots and
//                         counts the next digit into the next pass's
//                         histogram -> the permutation (sorted position ->
//                         caller's index)
//   sorted_pack_kernel    sorted FP64 copies of the j fields, per-tile
//                         velocity origins and bounding boxes (with the
//                         tile's largest near-field threshold), j records
//   (sorted_gather_kernel the i fields, when i and j differ)
//   sorted_engine         pair_engine's FP32C loop over sorted particles:
//                         per 64 tiles, one lane per tile tests the tile's
//                         box against the wave's box; far tiles run the
//                         pair loop without the near test, near tiles the
//                         tested loop with FP64 re-evaluation of near pairs;
//                         every tile's FP32 partial also feeds the error
//                         estimate (F32cBase::stat)
//   sorted_finish_kernel  adds the block-row partials in row order, writes
//                         each output to the caller's index and flags the
//                         rows whose estimated FP32 error exceeds
//                         kFlagTau * |result| (tau^2 on squares)
//   sorted_recompute_kernel + sorted_recompute_finish
//                         the flagged rows again in FP64 (the FP64 policy):
//                         64 rows per wave over j chunks (4 per block,
//                         combined in LDS), then the partials of each row
//                         added in a fixed order:
//                         bit-reproducible whatever order the rows were
//                         flagged in
// No memset anywhere (the clears are kernels), so a call can be captured.
#pragma once
#include "f32c.hpp"

namespace tupan {

constexpr int kSortDigit = 10;                 // bits per radix pass
constexpr int kSortBins = 1 << kSortDigit;
constexpr int kSortPasses = 3;                 // 30-bit keys
constexpr int kSortChunk = 1024;               // keys per block (minimum)
constexpr int kSortMaxBlocks = 256;
constexpr int kSctxVals = 14;                  // per-block context values
// Rows whose estimated FP32 error exceeds this fraction of |result| are
// recomputed in FP64 (the default of tupan_set_flag_tau).  The estimate
// sqrt(sum_t |tile partial|^2) bounded the error of every row by 9.2e-7 of
// itself (CPU emulation of these kernels' arithmetic, far records included,
// bit for bit with the GPU's, on the Plummer models N = 16384, 65536 and
// 262144, seed 1), so an unflagged row stays below 0.7 * 9.2e-7 = 6.4e-7
// (4.1e-7 measured); 0.6-0.7% of the rows are flagged at the default near
// field (c = 5).
constexpr double kFlagTau = 0.7;
// The FP64 recomputation runs (kRecomputeRows / 64) * recompute_groups(nj)
// blocks; recompute_shape spreads the flagged rows over them.
constexpr int kRecomputeRows = 1024;
constexpr int kRecomputeMinJ = 64;    // j's per recompute chunk (minimum)
constexpr int kRecomputeGroups = 256; // chunk groups of 4 (one per block)
constexpr int kRecomputeBlocks = 1024;  // sorted_recompute_finish

// Chunk groups of the FP64 recomputation: at most kRecomputeGroups of 4
// chunks of at least kRecomputeMinJ j's each, so that the few row-waves
// still fill the chip.
inline int recompute_groups(int nj) {
  int n = (nj + 4 * kRecomputeMinJ - 1) / (4 * kRecomputeMinJ);
  return n < 1 ? 1 : (n > kRecomputeGroups ? kRecomputeGroups : n);
}

// How the FP64 recomputation of `count` flagged rows uses its `blocks`
// blocks: `rows` rows (a multiple of 64 at most, count beyond 64 * blocks
// is left to sorted_recompute_finish) in row-waves of 64, each row-wave
// over `groups` chunk groups (at most max_groups, so that a chunk keeps at
// least kRecomputeMinJ j's).  Few flagged rows get many short chunks, many
// rows fewer and longer ones; rows * groups <= 64 * blocks always.  Both
// recompute kernels derive it from the same count.
struct RecomputeShape {
  int rows, groups;
};
__host__ __device__ inline RecomputeShape recompute_shape(int count,
                                                          int blocks,
                                                          int max_groups) {
  int rw = (count + 63) / 64;
  if (rw > blocks) rw = blocks;
  int g = rw > 0 ? blocks / rw : 1;
  if (g > max_groups) g = max_groups;
  if (g < 1) g = 1;
  const int rows = count < rw * 64 ? count : rw * 64;
  return RecomputeShape{rows, g};
}

// Geometry of one radix sort: nb blocks of ch keys (ch a multiple of 256).
struct SortGeom {
  int nb, ch;
};
inline SortGeom sort_geom(int n) {
  int nb = (n + kSortChunk - 1) / kSortChunk;
  if (nb > kSortMaxBlocks) nb = kSortMaxBlocks;
  if (nb < 1) nb = 1;
  int ch = (n + nb - 1) / nb;
  ch = (ch + 255) / 256 * 256;
  return SortGeom{nb, ch};
}

struct TileBox {
  double lo[3], hi[3];
  double near2;  // largest near-field threshold of the tile's j records
  double pad;
};

struct SoaW {
  double *f[14];
};

// a value equal in every lane, moved to scalar registers
__device__ __forceinline__ double uniform(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = fmin(v, __shfl_xor(v, m, 64));
  return uniform(v);
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return uniform(v);
}
// min/max over the 32 lanes of a half wave (one tile)
__device__ __forceinline__ double tile_min(double v) {
#pragma unroll
  for (int m = 16; m > 0; m >>= 1) v = fmin(v, __shfl_xor(v, m, 32));
  return v;
}
__device__ __forceinline__ double tile_max(double v) {
#pragma unroll
  for (int m = 16; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m, 32));
  return v;
}

// Per block b: [0] sum m_j, [1] max e2_i, [2..7] j box (lo xyz, hi xyz),
// [8..13] i box.  Grid kCtxBlocks x 256.  Also clears `nzero` ints at
// `zero` (the radix histograms, the flag count).
static __global__ __launch_bounds__(256) void sctx_partial_kernel(
    Soa js, int nj, Soa is, int ni, double *__restrict__ part,
    int *__restrict__ zero, int nzero) {
  __shared__ double red[kSctxVals][256];
  const int t = threadIdx.x;
  double v[kSctxVals];
  v[0] = 0.0;
  v[1] = 0.0;
  for (int k = 0; k < 3; ++k) {
    v[2 + k] = v[8 + k] = INFINITY;
    v[5 + k] = v[11 + k] = -INFINITY;
  }
  const int stride = kCtxBlocks * 256;
  for (int j = blockIdx.x * 256 + t; j < nj; j += stride) {
    v[0] += js.f[M][j];
    for (int k = 0; k < 3; ++k) {
      const double x = js.f[RX + k][j];
      v[2 + k] = fmin(v[2 + k], x);
      v[5 + k] = fmax(v[5 + k], x);
    }
  }
  for (int i = blockIdx.x * 256 + t; i < ni; i += stride) {
    v[1] = fmax(v[1], is.f[E2][i]);
    for (int k = 0; k < 3; ++k) {
      const double x = is.f[RX + k][i];
      v[8 + k] = fmin(v[8 + k], x);
      v[11 + k] = fmax(v[11 + k], x);
    }
  }
  for (int k = 0; k < kSctxVals; ++k) red[k][t] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      red[0][t] += red[0][t + w];
      red[1][t] = fmax(red[1][t], red[1][t + w]);
      for (int k = 0; k < 3; ++k) {
        red[2 + k][t] = fmin(red[2 + k][t], red[2 + k][t + w]);
        red[5 + k][t] = fmax(red[5 + k][t], red[5 + k][t + w]);
        red[8 + k][t] = fmin(red[8 + k][t], red[8 + k][t + w]);
        red[11 + k][t] = fmax(red[11 + k][t], red[11 + k][t + w]);
      }
    }
    __syncthreads();
  }
  if (t < kSctxVals) part[blockIdx.x * kSctxVals + t] = red[t][0];
  for (int q = blockIdx.x * 256 + t; q < nzero; q += stride) zero[q] = 0;
}

// Lanes of `live` whose value of the low `bits` bits of d equals this
// lane's (one ballot per bit).
__device__ __forceinline__ uint64_t match_bits(uint32_t d, int bits,
                                               uint64_t live) {
  uint64_t same = live;
  for (int b = 0; b < bits; ++b) {
    const bool on = (d >> b) & 1u;
    const uint64_t bal = __builtin_amdgcn_ballot_w64(on);
    same &= on ? bal : ~bal;
  }
  return same;
}
__device__ __forceinline__ uint64_t match_digit(int d, uint64_t live) {
  return match_bits((uint32_t)d, kSortDigit, live);
}
// number of lanes of `same` below this lane
__device__ __forceinline__ int lane_rank(uint64_t same) {
  return (int)__builtin_amdgcn_mbcnt_hi(
      (uint32_t)(same >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)same, 0u));
}

// bits 0..9 of v spread to bits 0, 3, 6, ..., 27
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Morton keys of one set and the histogram of their first digit, one block
// per chunk of ch keys: hist[d * nb + b].  box: offset of the set's box in
// the per-block context values.  Block 0 also writes ctx[0] (near-field
// factor, ctx_final_kernel's rule) and ctx[1] (largest e2_i) when ctx is
// given.
static __global__ __launch_bounds__(256) void morton_key_kernel(
    const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, int n, const double *__restrict__ part,
    int box, SortGeom g, uint32_t *__restrict__ keys, int *__restrict__ vals,
    int *__restrict__ hist, double *__restrict__ ctx, int nj, double c) {
  __shared__ int h[kSortBins];
  __shared__ double bx[6][kCtxBlocks];
  const int t = threadIdx.x;
  for (int d = t; d < kSortBins; d += 256) h[d] = 0;
  // the set's box from the kCtxBlocks partials (thread b: partial b)
  if (t < kCtxBlocks)
    for (int k = 0; k < 6; ++k) bx[k][t] = part[t * kSctxVals + box + k];
  __syncthreads();
  for (int w = kCtxBlocks / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int k = 0; k < 3; ++k) {
        bx[k][t] = fmin(bx[k][t], bx[k][t + w]);
        bx[3 + k][t] = fmax(bx[3 + k][t], bx[3 + k][t + w]);
      }
    __syncthreads();
  }
  double lo[3], inv[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = bx[k][0];
    inv[k] = bx[3 + k][0] > lo[k] ? 1024.0 / (bx[3 + k][0] - lo[k]) : 0.0;
  }
  if (ctx && blockIdx.x == 0 && t == 0) {
    double msum = 0.0, e2max = 0.0;
    for (int b = 0; b < kCtxBlocks; ++b) {
      msum += part[b * kSctxVals];
      e2max = fmax(e2max, part[b * kSctxVals + 1]);
    }
    const double f = c * c * cbrt((double)nj) / msum;
    ctx[0] = (c > 0.0 && msum > 0.0 && f < 1e300) ? f : 0.0;
    ctx[1] = e2max;
  }
  const double *f[3] = {x, y, z};
  const int e0 = blockIdx.x * g.ch;
  for (int r = 0; r < g.ch; r += 256) {
    const int e = e0 + r + t;
    const bool valid = e < n;
    uint32_t key = 0;
    if (valid) {
      for (int k = 0; k < 3; ++k) {
        const double q = (f[k][e] - lo[k]) * inv[k];
        // NaN and the low edge -> 0, the high edge -> 1023
        const uint32_t b = q >= 1023.0 ? 1023u : (q > 0.0 ? (uint32_t)q : 0u);
        key |= spread10(b) << k;
      }
      keys[e] = key;
      vals[e] = e;
    }
    const int d = (int)(key & (kSortBins - 1));
    const uint64_t same = match_digit(d, __builtin_amdgcn_ballot_w64(valid));
    if (valid && lane_rank(same) == 0)
      atomicAdd(&h[d], __builtin_popcountll(same));
  }
  __syncthreads();
  for (int d = t; d < kSortBins; d += 256) hist[blockIdx.x * kSortBins + d] = h[d];
}

// One stable LSD pass on digit (key >> shift) & 1023, one block per chunk
// of ch keys.  Digit d of block b starts at (keys of smaller digits) +
// (keys of digit d in blocks before b), both summed here from the
// block-major per-block histograms hist[c * 1024 + d] (coalesced loads);
// within the block, rounds of 256 keys in index order, each wave ranking
// its keys of equal digit with ballots (mbcnt = rank among them) and the
// waves taking their positions in wave order.  hist_next (optional) gets
// the next digit's per-block counts of the output order: lanes of a wave
// with the same (next digit, output block) add their count with one
// atomic.
static __global__ __launch_bounds__(256) void radix_scatter_kernel(
    const uint32_t *__restrict__ kin, const int *__restrict__ vin,
    uint32_t *__restrict__ kout, int *__restrict__ vout, int n, SortGeom g,
    const int *__restrict__ hist, int shift, int *__restrict__ hist_next,
    int dest_bits) {
  __shared__ int base[kSortBins];
  __shared__ int tot[kSortBins];
  __shared__ int tsum[256];
  const int t = threadIdx.x;
  const int b = blockIdx.x;
  {
    int all[4] = {0, 0, 0, 0}, pre[4] = {0, 0, 0, 0};
#pragma unroll 8
    for (int c = 0; c < g.nb; ++c) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int v = hist[c * kSortBins + t + 256 * q];
        all[q] += v;
        pre[q] += c < b ? v : 0;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      tot[t + 256 * q] = all[q];
      base[t + 256 * q] = pre[q];
    }
  }
  __syncthreads();
  // exclusive scan of tot over the digits: 4 consecutive per thread, then
  // the 256 thread sums (Hillis-Steele)
  int run = 0, loc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    loc[q] = run;
    run += tot[4 * t + q];
  }
  tsum[t] = run;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int v = t >= o ? tsum[t - o] : 0;
    __syncthreads();
    tsum[t] += v;
    __syncthreads();
  }
  const int excl = t > 0 ? tsum[t - 1] : 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) base[4 * t + q] += excl + loc[q];
  __syncthreads();

  const int wave = t >> 6;
  const int e0 = b * g.ch;
  // keys of the next round are loaded while this one is placed
  int e = e0 + t;
  uint32_t key = e < n ? kin[e] : 0u;
  int val = e < n ? vin[e] : 0;
  for (int r = 0; r < g.ch; r += 256) {
    const bool valid = e < n;
    const int en = e + 256;
    const bool nvalid = r + 256 < g.ch && en < n;
    const uint32_t nkey = nvalid ? kin[en] : 0u;
    const int nval = nvalid ? vin[en] : 0;
    const int d = (int)((key >> shift) & (kSortBins - 1));
    const uint64_t same = match_digit(d, __builtin_amdgcn_ballot_w64(valid));
    const int rank = lane_rank(same);
    const int cnt = __builtin_popcountll(same);
    int pos = 0;
    for (int w = 0; w < kWaves; ++w) {
      if (wave == w && valid) {
        const int bse = base[d];
        pos = bse + rank;
        // the group's lowest lane advances the digit; its store depends on
        // the value the whole wave has just read
        if (rank == 0) base[d] = bse + cnt;
      }
      __syncthreads();
    }
    if (valid) {
      kout[pos] = key;
      vout[pos] = val;
    }
    if (hist_next) {
      const int dn = (int)((key >> (shift + kSortDigit)) & (kSortBins - 1));
      const int dest = pos / g.ch;
      const uint64_t grp =
          match_bits((uint32_t)dest, dest_bits,
                     match_digit(dn, __builtin_amdgcn_ballot_w64(valid)));
      if (valid && lane_rank(grp) == 0)
        atomicAdd(&hist_next[(size_t)dest * kSortBins + dn],
                  __builtin_popcountll(grp));
    }
    e = en;
    key = nkey;
    val = nval;
  }
}

// Sorted FP64 copies of the j fields, per-tile origins and boxes, and the
// j records, one thread per sorted position p; a tile is the 32 lanes of a
// half wave.  The last tile's spare lanes repeat its last j, which changes
// no minimum or maximum.  kFarRec policies also get the far-tile records,
// positions relative to the centre of the tile's box (porg[4 t + 0..2]),
// and porg[4 t + 3] = the records' e2 when the whole tile shares it (-1
// otherwise, NaN included).
template <class Pol>
__global__ __launch_bounds__(256) void sorted_pack_kernel(
    Soa js, int nj, const int *__restrict__ perm, SoaW sj,
    const double *__restrict__ ctx, typename Pol::JRec *__restrict__ jr,
    double *__restrict__ origins, TileBox *__restrict__ boxes,
    double *__restrict__ porg, typename Pol::FarRec *__restrict__ fr,
    typename Pol::Exact::JRec *__restrict__ jr64) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int o = perm[min(p, nj - 1)];
  if (p < nj) {
    for (int k = 0; k < Pol::kFields; ++k) sj.f[k][p] = js.f[k][o];
    jr64[p] = Pol::Exact::pack(js, o, 0.0);
  }
  // per tile, kOrg doubles: the midpoints of the tile's ranges of the
  // fields VX + 3 g + k (g < VelOrigins, k < 3) at 4 g + k, times the
  // policy's origin scale
  constexpr int kOrg = OrgStride<Pol>::value;
  double org[kOrg];
  for (int k = 0; k < kOrg; ++k) org[k] = 0.0;
  if constexpr (VelOrigins<Pol>::value) {
    for (int g = 0; g < VelOrigins<Pol>::value; ++g)
      for (int k = 0; k < 3; ++k) {
        const double v = js.f[VX + 3 * g + k][o];
        org[4 * g + k] =
            0.5 * (tile_min(v) + tile_max(v)) * OrgScale<Pol>::value;
      }
  }
  const PackCtx pc{ctx[0], ctx[1], nullptr};
  typename Pol::JRec rec;
  if constexpr (VelOrigins<Pol>::value)
    rec = Pol::pack_o(js, o, pc, org);
  else
    rec = Pol::pack(js, o, pc);
  if (p < nj) jr[p] = rec;
  TileBox bx;
  for (int k = 0; k < 3; ++k) {
    const double x = js.f[RX + k][o];
    bx.lo[k] = tile_min(x);
    bx.hi[k] = tile_max(x);
  }
  bx.near2 = tile_max((double)rec.near2);
  bx.pad = 0.0;
  double po[4] = {0.0, 0.0, 0.0, 0.0};
  if constexpr (Pol::kFarRec) {
    for (int k = 0; k < 3; ++k) po[k] = 0.5 * (bx.lo[k] + bx.hi[k]);
    const typename Pol::FarRec f = Pol::far_rec(rec, po);
    if (p < nj) fr[p] = f;
    if constexpr (Pol::kFar16)
      if (p < nj)
        reinterpret_cast<typename Pol::FarRec16 *>(fr + nj)[p] = {f.x, f.y,
                                                                  f.z, f.m};
    const double e2 = (double)rec.e2;
    const bool same = tile_min(e2) == tile_max(e2) &&
                      tile_max(e2 == e2 ? 0.0 : 1.0) == 0.0;
    po[3] = same ? e2 : -1.0;
  }
  if ((threadIdx.x & (kTile - 1)) == 0 && p < nj) {
    const int t = p / kTile;
    for (int k = 0; k < kOrg; ++k) origins[kOrg * t + k] = org[k];
    boxes[t] = bx;
    if constexpr (Pol::kFarRec)
      for (int k = 0; k < 4; ++k) porg[4 * t + k] = po[k];
  }
}

// Sorted FP64 copies of the i fields (i and j differ).
template <int NF>
__global__ __launch_bounds__(256) void sorted_gather_kernel(
    Soa is, int ni, const int *__restrict__ perm, SoaW si) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= ni) return;
  const int o = perm[p];
  for (int k = 0; k < NF; ++k) si.f[k][p] = is.f[k][o];
}

// Marks the words of r as read and rewritten here, and memory as touched:
// emits nothing, but r's pending scalar load completes before this point,
// no load moves across it, and r's uses stay after it.
template <class R>
__device__ __forceinline__ void pin_sgprs(R &r) {
  static_assert(sizeof(R) % 4 == 0, "record of 32-bit words");
  struct W {
    uint32_t w[sizeof(R) / 4];
  };
  W x = __builtin_bit_cast(W, r);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(R) / 4); ++k)
    __asm__ volatile("" : "+s"(x.w[k]) : : "memory");
  r = __builtin_bit_cast(R, x);
}
// The same for the FP32 accumulators: what follows stays after the
// instructions that produced them.
template <class A>
__device__ __forceinline__ void pin_vgprs(A &a) {
#pragma unroll
  for (int k = 0; k < (int)(sizeof(a.v) / sizeof(a.v[0])); ++k)
    __asm__ volatile("" : "+v"(a.v[k]));
}

// f(rec[k]) for k = 0 .. nk-1 (nk <= kTile), the next record fetched while
// the current one is used; a: the FP32 accumulators that f updates.  G > 1
// (kTile / G even): full tiles take their records G at a time (one scalar
// load) in two buffers.  Scalar loads complete out of order, so a wait
// covers every load in flight: each load is issued after the previous G
// pairs and the wait for the buffer about to be used, a whole G-pair step
// ahead of its own use.  That pays for short pair bodies (Phi, Acc: the
// record's latency is no longer exposed every pair) and costs long ones
// (AccJerk: the pins keep the compiler from overlapping consecutive steps).
template <int G, class R, class A, class F>
__device__ __forceinline__ void tile_loop(const R *__restrict__ rec, int nk,
                                          A &a, F &&f) {
  static_assert(G == 1 || (kTile % (2 * G) == 0), "two buffers per tile");
  if (G > 1 && nk == kTile) {
    struct RG {
      R r[G];
    };
    const RG *__restrict__ rr = reinterpret_cast<const RG *>(rec);
    RG c0 = rr[0], c1;
#pragma unroll 1
    for (int k = 0; k < kTile / G; k += 2) {
      pin_vgprs(a);
      pin_sgprs(c0);
      c1 = rr[k + 1];
      pin_sgprs(c0);
#pragma unroll
      for (int g = 0; g < G; ++g) f(c0.r[g]);
      pin_vgprs(a);
      pin_sgprs(c1);
      c0 = rr[min(k + 2, kTile / G - 1)];
      pin_sgprs(c1);
#pragma unroll
      for (int g = 0; g < G; ++g) f(c1.r[g]);
    }
  } else if (nk == kTile) {
    R r = rec[0];
#pragma unroll 2
    for (int k = 0; k < kTile; ++k) {
      const R rn = rec[min(k + 1, kTile - 1)];
      f(r);
      r = rn;
    }
  } else {
    R r = rec[0];
    for (int k = 0; k < nk; ++k) {
      const R rn = rec[min(k + 1, nk - 1)];
      f(r);
      r = rn;
    }
  }
}

// pair_engine's FP32C path (kTwoLevel policies, one i per lane) over
// sorted particles.  Always writes its block-row partials (outputs, then
// the error statistics) for sorted_finish_kernel.
// Waves per block of the sorted engine: a block's waves take the same 64
// i's and consecutive j chunks and combine their sums through LDS, so the
// block rows (and the partial sums sorted_finish_kernel reads) are S / 8.
constexpr int kSWaves = 4;
constexpr int kSBlock = 64 * kSWaves;

// Pol::kWavesPerEu waves per SIMD at least: for AccJerk 8 (63 VGPRs, no
// spill, with its FP64 i fields in LDS), measured faster than the 82
// registers and 5 waves the compiler picks by itself.
template <class Pol>
__global__ __launch_bounds__(kSBlock) __attribute__((amdgpu_waves_per_eu(Pol::kWavesPerEu, 8))) void sorted_engine(
    Soa si, int ni, Soa sj, const typename Pol::JRec *__restrict__ jr, int nj,
    const TileBox *__restrict__ boxes, int jchunk,
    double *__restrict__ part, const double *__restrict__ aux,
    const typename Pol::FarRec *__restrict__ fr,
    const double *__restrict__ porg,
    const typename Pol::Exact::JRec *__restrict__ jr64) {
  constexpr int kPart = Pol::kPart;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int row = blockIdx.y;
  const int i0 = blockIdx.x * 64;
  if (i0 >= ni) return;  // block-uniform
  const int s = row * kSWaves + wave;
  const int i = min(i0 + lane, ni - 1);
  const typename Pol::Params prm{};
  typename Pol::I q = Pol::load_i(si, i, prm);
  // LDS: the i's FP64 fields (Pol::kI64 of them) during the tile loop, the
  // block's partial sums after it.  The FP64 fields are read at tile starts
  // and in the tested loop only; kept in registers they cost the far loop
  // its last registers at 7 waves per SIMD, and the compiler spilled them
  // to scratch (168 MB of writes per launch at N = 65536).
  constexpr int kI64 = Pol::kI64 > 0 ? Pol::kI64 : 1;
  union Lds {
    double i64[kSWaves][kI64][64];
    double red[kSWaves - 1][kPart][64];
  };
  __shared__ Lds lds;
  if constexpr (Pol::kI64 > 0) {
    double f[kI64];
    Pol::save64(q, f);
#pragma unroll
    for (int k = 0; k < kI64; ++k) lds.i64[wave][k][lane] = f[k];
  }
  typename Pol::A64 acc;
  Pol::zero(acc);
  float sig[Pol::kGroups];
#pragma unroll
  for (int g = 0; g < Pol::kGroups; ++g) sig[g] = 0.0f;
  // the wave's box (spare lanes repeat the last i)
  double wlo[3], whi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double x = si.f[RX + k][i];
    wlo[k] = wave_min(x);
    whi[k] = wave_max(x);
  }

  const int jb = s * jchunk;
  const int je = min(nj, jb + jchunk);
  uint64_t farbits = 0;
  int tt = 0;
  for (int j0 = jb; j0 < je; j0 += kTile, ++tt) {
    if ((tt & 63) == 0) {
      // lane l tests tile j0 / kTile + l: far when the squared gap between
      // the boxes reaches the tile's largest near threshold (then every
      // pair has s^2 >= r^2 >= gap^2 >= near2_j, and r^2 > 0)
      const int t = j0 / kTile + lane;
      bool far = false;
      if (j0 + lane * kTile < je) {
        const TileBox b = boxes[t];
        double d2 = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double gap = fmax(fmax(b.lo[k] - whi[k], wlo[k] - b.hi[k]), 0.0);
          d2 = __builtin_fma(gap, gap, d2);
        }
        far = d2 >= b.near2;
      }
      farbits = __builtin_amdgcn_ballot_w64(far);
    }
    const int nk = min(kTile, je - j0);
    typename Pol::A32 a;
    Pol::zero(a);
    if constexpr (Pol::kI64 > 0) {
      // the barrier keeps the compiler from hoisting the loads out of the
      // tile loop (which would keep the fields in registers again)
      __asm__ volatile("" ::: "memory");
      double f[kI64];
#pragma unroll
      for (int k = 0; k < kI64; ++k) f[k] = lds.i64[wave][k][lane];
      Pol::load64(q, f);
    }
    Pol::tile_begin(q, aux, j0 / kTile);
    // the tile's near pairs (tested tiles), re-evaluated in FP64 after the
    // fold; one fold for both kinds of tile (the compiler then keeps the
    // FP64 sums in place instead of copying them back every tile)
    uint32_t near = 0u;
    if ((farbits >> (tt & 63)) & 1) {
      if constexpr (Pol::kFarRec) {
        const double *po = porg + 4 * (j0 / kTile);
        Pol::tile_begin_far(q, po);
        const double e2u = po[3];
        if (e2u >= 0.0) {
          const float e2ij = q.e2 + (float)e2u;
          if constexpr (Pol::kFar16) {
            const auto *f16 =
                reinterpret_cast<const typename Pol::FarRec16 *>(fr + nj);
            tile_loop<Pol::kFar16Group>(
                f16 + j0, nk, a, [&](const typename Pol::FarRec16 &r) {
                  Pol::pair32_farrec_u(q, r, e2ij, a);
                });
          } else {
            tile_loop<Pol::kFarGroup>(fr + j0, nk, a, [&](const typename Pol::FarRec &r) {
              Pol::pair32_farrec_u(q, r, e2ij, a);
            });
          }
        } else {
          tile_loop<Pol::kFarGroup>(fr + j0, nk, a, [&](const typename Pol::FarRec &r) {
            Pol::pair32_farrec(q, r, a);
          });
        }
      } else {
        typename Pol::JRec r = jr[j0];
        if (nk == kTile) {
#pragma unroll 2
          for (int k = 0; k < kTile; ++k) {
            const typename Pol::JRec rn = jr[j0 + min(k + 1, kTile - 1)];
            Pol::pair32_far(q, r, a);
            r = rn;
          }
        } else {
          for (int k = 0; k < nk; ++k) {
            const typename Pol::JRec rn = jr[j0 + min(k + 1, nk - 1)];
            Pol::pair32_far(q, r, a);
            r = rn;
          }
        }
      }
    } else {
      typename Pol::JRec r = jr[j0];
      if (nk == kTile) {
#pragma unroll 2
        for (int k = 0; k < kTile; ++k) {
          const typename Pol::JRec rn = jr[j0 + min(k + 1, kTile - 1)];
          near |= near_bit(Pol::pair32(q, r, a)) << k;
          r = rn;
        }
      } else {
        for (int k = 0; k < nk; ++k) {
          const typename Pol::JRec rn = jr[j0 + min(k + 1, nk - 1)];
          near |= near_bit(Pol::pair32(q, r, a)) << k;
          r = rn;
        }
      }
    }
    Pol::fold(a, acc);
    Pol::stat(a, sig);
    while (near) {
      const int k = __builtin_ctz(near);
      near &= near - 1u;
      Pol::near_pair64(q, jr[j0 + k], acc, si, jr64[j0 + k]);
    }
  }

  double v[kPart];
  Pol::values(acc, v);
#pragma unroll
  for (int g = 0; g < Pol::kGroups; ++g) v[Pol::kOut + g] = (double)sig[g];
  if constexpr (Pol::kI64 > 0) __syncthreads();  // every wave done with i64
  if (wave > 0) {
#pragma unroll
    for (int k = 0; k < kPart; ++k) lds.red[wave - 1][k][lane] = v[k];
  }
  __syncthreads();
  if (wave != 0) return;
  for (int w = 0; w < kSWaves - 1; ++w) {
#pragma unroll
    for (int k = 0; k < kPart; ++k) v[k] += lds.red[w][k][lane];
  }
  if (i0 + lane < ni) {
#pragma unroll
    for (int k = 0; k < kPart; ++k)
      part[((size_t)row * kPart + k) * ni + i0 + lane] = v[k];
  }
}

// Row partials -> outputs at the caller's index; rows whose error estimate
// exceeds tau^2 |result|^2 in some output group go to the flag list
// (flags[0]: count, flags[1..]: sorted row positions).
template <class Pol>
__global__ __launch_bounds__(256) void sorted_finish_kernel(
    const double *__restrict__ part, int S, int ni,
    const int *__restrict__ perm_i, Out out, double tau2,
    int *__restrict__ flags) {
  constexpr int kPart = Pol::kPart;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= ni) return;
  double v[kPart];
#pragma unroll
  for (int k = 0; k < kPart; ++k) v[k] = part[(size_t)k * ni + p];
  for (int r = 1; r < S; ++r) {
#pragma unroll
    for (int k = 0; k < kPart; ++k) v[k] += part[((size_t)r * kPart + k) * ni + p];
  }
  const int o = perm_i[p];
#pragma unroll
  for (int k = 0; k < Pol::kOut; ++k) out.f[k][o] = v[k];
  bool flag = false;
#pragma unroll
  for (int g = 0; g < Pol::kGroups; ++g) {
    double n2 = 0.0;
#pragma unroll
    for (int k = 0; k < Pol::kGroupLen; ++k)
      n2 += v[g * Pol::kGroupLen + k] * v[g * Pol::kGroupLen + k];
    flag = flag || v[Pol::kOut + g] > tau2 * n2;
  }
  if (flag) flags[1 + atomicAdd(flags, 1)] = p;
}

// The flagged rows in FP64 (P64, the FP64 policy of the same kernel), in
// the pair engine's shape: a wave holds 64 flagged rows (one per lane) and
// walks one chunk of j's, so the j data is read once per 64 rows.  The
// chunk passes through LDS 64 records at a time (one coalesced load per
// lane, then wave-uniform LDS reads), so a wave waits for memory once per
// 64 pairs rather than once per pair.  A block's 4 waves take the same 64
// rows and 4 consecutive chunks and add their sums in LDS in wave order.
// With sh = recompute_shape(flag count, gridDim.x, max_groups), block
// b = (row-wave b / sh.groups, chunk group b % sh.groups) writes
// rpart[(row * sh.groups + group) * kOut + k] for the first sh.rows flagged
// rows; blocks beyond those row-waves return at once.
template <class P64>
__global__ __launch_bounds__(kBlock) void sorted_recompute_kernel(
    Soa si, Soa sj, int nj, const int *__restrict__ flags, int max_groups,
    double *__restrict__ rpart) {
  constexpr int K = P64::kOut;
  __shared__ typename P64::JRec jbuf[kWaves][64];
  __shared__ double red[kWaves - 1][K][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const RecomputeShape sh =
      recompute_shape(flags[0], (int)gridDim.x, max_groups);
  const int rw = blockIdx.x / sh.groups, grp = blockIdx.x - rw * sh.groups;
  if (rw * 64 >= sh.rows) return;  // block-uniform, before any barrier
  const int count = sh.rows;
  const int e = rw * 64 + lane;
  const int p = flags[1 + min(e, count - 1)];
  const typename P64::Params prm{};
  const typename P64::I q = P64::load_i(si, p, prm);
  typename P64::A64 a;
  P64::zero(a);
  const int cw = (nj + kWaves * sh.groups - 1) / (kWaves * sh.groups);
  const int jb = (grp * kWaves + w) * cw, je = min(nj, jb + cw);
  for (int j0 = jb; j0 < jb + cw; j0 += 64) {
    if (j0 + lane < je) jbuf[w][lane] = P64::pack(sj, j0 + lane, 0.0);
    __syncthreads();
    const int nk = min(64, je - j0);
#pragma unroll 4
    for (int k = 0; k < nk; ++k) P64::pair(q, jbuf[w][k], a, prm);
    __syncthreads();
  }
  double v[K];
  P64::values(a, v);
  if (w > 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[w - 1][k][lane] = v[k];
  }
  __syncthreads();
  if (w != 0 || e >= count) return;
  for (int o = 0; o < kWaves - 1; ++o) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += red[o][k][lane];
  }
#pragma unroll
  for (int k = 0; k < K; ++k)
    rpart[((size_t)e * sh.groups + grp) * K + k] = v[k];
}

// Outputs of the flagged rows, one block per row (grid-stride): the first
// sh.rows rows (sh as in sorted_recompute_kernel, which ran on
// `blocks` blocks) add their sh.groups partials (thread t the groups t,
// t + 256, ... in order, then a fixed LDS tree); rows beyond, if any, are
// evaluated whole here, the block's threads over the j's.
template <class P64>
__global__ __launch_bounds__(256) void sorted_recompute_finish(
    Soa si, Soa sj, int nj, const int *__restrict__ flags, int blocks,
    int max_groups, const double *__restrict__ rpart,
    const int *__restrict__ perm_i, Out out) {
  constexpr int K = P64::kOut;
  __shared__ double red[K][256];
  const int t = threadIdx.x;
  const int count = flags[0];
  const RecomputeShape sh = recompute_shape(count, blocks, max_groups);
  const typename P64::Params prm{};
  for (int e = blockIdx.x; e < count; e += gridDim.x) {
    const int p = flags[1 + e];
    double v[K];
    if (e < sh.rows) {
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = 0.0;
      for (int c = t; c < sh.groups; c += 256) {
#pragma unroll
        for (int k = 0; k < K; ++k)
          v[k] += rpart[((size_t)e * sh.groups + c) * K + k];
      }
    } else {
      const typename P64::I q = P64::load_i(si, p, prm);
      typename P64::A64 a;
      P64::zero(a);
      for (int j = t; j < nj; j += 256) P64::pair(q, P64::pack(sj, j, 0.0), a, prm);
      P64::values(a, v);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) red[k][t] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) {
#pragma unroll
        for (int k = 0; k < K; ++k) red[k][t] += red[k][t + w];
      }
      __syncthreads();
    }
    if (t == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = red[k][0];
      P64::finish(v, out, perm_i[p], prm);
    }
    __syncthreads();
  }
}

// Byte offsets of the sorted path's scratch, one buffer.
// The histograms of both sets and the flag count lie in one region that
// sctx_partial_kernel clears (zero, nzero ints); the flag list follows the
// count.
struct SortedLayout {
  size_t ctxpart, ctx, keys[2][2], vals[2][2], hist[2][kSortPasses], flags,
      sj, si, origins, boxes, recs, recs64, porg, frecs, rpart, total;
  size_t zero, nzero;
};

template <class Pol>
SortedLayout sorted_layout(int ni, int nj, bool same, SortGeom gj,
                           SortGeom gi, int nseg) {
  SortedLayout L{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += (bytes + 255) / 256 * 256;
    return at;
  };
  const int ntiles = (nj + kTile - 1) / kTile;
  L.ctxpart = take(sizeof(double) * kCtxBlocks * kSctxVals);
  L.ctx = take(sizeof(double) * 2);
  const int nsets = same ? 1 : 2;
  for (int s = 0; s < nsets; ++s) {
    const int n = s == 0 ? nj : ni;
    for (int b = 0; b < 2; ++b) {
      L.keys[s][b] = take(sizeof(uint32_t) * n);
      L.vals[s][b] = take(sizeof(int) * n);
    }
  }
  // histograms and the flag count: contiguous ints (no padding between)
  L.zero = o;
  for (int s = 0; s < nsets; ++s) {
    const SortGeom g = s == 0 ? gj : gi;
    for (int p = 0; p < kSortPasses; ++p) {
      L.hist[s][p] = o;
      o += sizeof(int) * (size_t)kSortBins * g.nb;
    }
  }
  L.flags = o;
  o += sizeof(int);
  L.nzero = (o - L.zero) / sizeof(int);
  o += sizeof(int) * (size_t)ni;  // the flag list
  o = (o + 255) / 256 * 256;
  L.sj = take(sizeof(double) * Pol::kFields * (size_t)nj);
  L.si = same ? L.sj : take(sizeof(double) * Pol::kFields * (size_t)ni);
  L.origins = take(sizeof(double) * OrgStride<Pol>::value * (size_t)ntiles);
  L.boxes = take(sizeof(TileBox) * (size_t)ntiles);
  L.recs = take(sizeof(typename Pol::JRec) * (size_t)nj);
  L.recs64 = take(sizeof(typename Pol::Exact::JRec) * (size_t)nj);
  // far-tile records and the tiles' position origins (kFarRec policies)
  L.porg = take(Pol::kFarRec ? sizeof(double) * 4 * (size_t)ntiles : 0);
  L.frecs = take(Pol::kFarRec ? (sizeof(typename Pol::FarRec) +
                                 (Pol::kFar16 ? sizeof(typename Pol::FarRec16)
                                              : 0)) *
                                    (size_t)nj
                              : 0);
  L.rpart = take(sizeof(double) * (size_t)kRecomputeRows * nseg * Pol::kOut);
  // nseg: chunk groups of the FP64 recomputation
  L.total = o;
  return L;
}

}  // namespace tupan
