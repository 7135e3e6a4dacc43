// engine.hpp -- the O(ni*nj) pair-sum engine for gfx950 (MI355X).
//
// One templated kernel runs every member of the family.  A "policy" class
// supplies the per-kernel pieces:
//
//   JRec             packed j-particle record (built once per call by
//                    pack_kernel from the caller's SoA FP64 arrays)
//   I                the i-particle state a lane keeps in registers
//   A32 / A64        per-lane accumulators: A32 is summed over one j-tile,
//                    then folded into A64 (kTwoLevel); FP64 policies
//                    accumulate straight into A64
//   load_i, pair, fold, combine, finish
//
// and optionally
//
//   kSparse, sparse, near   pairs that are mostly no-ops (Sakura with a
//                           cut-off): tiles are visited through a mask of
//                           the j's some lane needs (near_mask_kernel)
//   prepare                 parameters resolved from device memory when the
//                           kernel starts (NREG's dt in device-resident runs)
//
// Execution model (CDNA4, wave64):
//   * a wave owns 64*P consecutive i-particles, P per lane, in VGPRs;
//   * all 64 lanes walk the same j sequence, so a j record is wave-uniform:
//     it is fetched with scalar loads (s_load) into SGPRs and used directly
//     as a VALU operand -- no LDS round trip, no bank conflicts;
//   * the j range is split into S = 4 * nrows chunks so that the chip
//     holds >= 4 waves per SIMD at any N: a block's 4 waves take the same
//     64*P i-particles and 4 consecutive chunks, and combine their sums
//     through LDS in wave order; with nrows > 1 block rows the row partials
//     land in an FP64 scratch array and a finishing kernel adds them in row
//     order (fixed grouping: bit-reproducible run to run);
//   * blockIdx.x runs over i-waves and blockIdx.y over block rows, so the
//     blocks in flight read the same j chunks (scalar cache and L2 hits);
//     blocks go round-robin to the 8 XCDs, so with a multiple of 8 i-waves
//     each XCD keeps the same i-waves in every row.  (Mapping each XCD to a
//     contiguous i range with rows innermost was measured: it breaks the
//     j-chunk sharing and multiplied the HBM fetches 12x.)
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace tupan {

enum Field { M = 0, RX, RY, RZ, E2, VX, VY, VZ, AX, AY, AZ, JX, JY, JZ };

struct Soa {
  const double *f[14];
};
struct Out {
  double *f[8];
};
// Per-call values the FP32C packers need (see ctx_final_kernel and
// tile_origin_kernel).
struct PackCtx {
  double factor;          // near-field factor: near2_j = m_j * factor
  double e2max;           // largest i-side half softening
  const double *origins;  // per-tile FP64 origins (4 doubles per tile)
};

constexpr int kWaves = 4;       // waves per block
constexpr int kBlock = 64 * kWaves;
constexpr int kTile = 32;       // j per FP32 tile before the FP64 fold

// ---------------------------------------------------------------- math
// 1/sqrt(x) in FP64: hardware estimate (v_rsq_f64, ~2^-23 relative)
// refined by one third-order Newton-Householder step:
//   e = 1 - x y^2,  y <- y (1 + e/2 + 3 e^2/8)  (error ~ e^3).
__device__ __forceinline__ double rsq64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double t = x * y;
  double e = __builtin_fma(-t, y, 1.0);
  double p = __builtin_fma(0.375, e, 0.5);
  return __builtin_fma(y * e, p, y);
}

// 1/sqrt(x) in FP32: v_rsq_f32 (1 ulp).
__device__ __forceinline__ float rsq32(float x) {
  return __builtin_amdgcn_rsqf(x);
}

// FP32 difference of two FP64 coordinates: exact FP64 subtraction
// followed by one rounding to FP32.  This is what the hi/lo position split
// buys -- the pair separation is correct to 1/2 ulp of itself even when
// |x| >> |xi - xj| -- in two VALU ops (v_add_f64 + v_cvt_f32_f64, both
// full rate on CDNA4) instead of three FP32 ops on hi/lo pairs.
__device__ __forceinline__ float dsub32(double a, double b) {
  return static_cast<float>(a - b);
}

// 1 if some active lane is missing from the lane mask `far` (the ballot of
// the lanes whose pair is NOT near).  Two scalar instructions: written out
// because the compiler otherwise round-trips the uniform boolean through a
// VGPR (v_cndmask + v_readfirstlane), and taking the ballot of the "far"
// predicate lets the one v_cmp feed both this mask and the FP32 select.
// No memory access.
__device__ __forceinline__ uint32_t near_bit(uint64_t far) {
  uint32_t b;
  asm volatile("s_cmp_lg_u64 %1, exec\n\ts_cselect_b32 %0, 1, 0"
               : "=s"(b)
               : "s"(far)
               : "scc");
  return b;
}

// Policies whose pairs are mostly no-ops (kSparse: Sakura restricted to
// the pairs closer than a cut-off) provide near(q, r, prm), true for the
// pairs pair() would change, and sparse(prm), true when that test is worth
// a pass of its own.  The engine then tests a tile of kTile j's first and
// calls pair() only for the j's some lane needs, in j order: each lane
// sees the same sequence of contributing pairs, so the sums are the same
// bits as those of the dense loop.
template <class Pol, class = void>
struct SparsePairs {
  static constexpr bool value = false;
};
template <class Pol>
struct SparsePairs<Pol, decltype((void)Pol::kSparse)> {
  static constexpr bool value = Pol::kSparse;
};

// Sparse policies' Params carry `mask` and `mask_tiles`: when the launcher
// has filled a mask (near_mask_kernel), bit k of word
// mask[w * mask_tiles + t] is set when some i of i-block w (64 i's, P = 1)
// may need j = kTile t + k, and the engine reads one word per tile instead
// of testing the tile's pairs itself.  The mask may hold extra bits (it
// is a superset: pair() applies the exact test), never fewer.
//
// near_mask_kernel: lanes own j's (two 32-j tiles per wave, positions in
// VGPRs) and walk the i-block's 64 i's, whose positions are wave-uniform
// scalar loads; bit = r^2 < Pol::mask_cut(prm) (r from the positions in
// fields ClassX<Pol>::value..+2), a cut-off slightly above
// the policy's own so that a different rounding of r^2 cannot drop a pair.
// A separate kernel because the sparse policies' pair() (a Kepler solve)
// needs several times the registers of this test and would run it at a
// fraction of the occupancy.
// `same`: the i and j sets are the same arrays, so j == i is the particle
// paired with itself; where Pol::self_noop says that pair adds exactly
// nothing, it gets no bit.
// Field of the x coordinate the sparse selection tests (Pol::kClassX when
// the policy selects by other positions than its own, else RX).
template <class Pol, class = void>
struct ClassX {
  static constexpr int value = RX;
};
template <class Pol>
struct ClassX<Pol, decltype((void)Pol::kClassX)> {
  static constexpr int value = Pol::kClassX;
};

template <class Pol>
__global__ __launch_bounds__(kBlock) void near_mask_kernel(
    Soa is, int ni, Soa js, int nj, int ntiles, typename Pol::Params prm,
    int same, uint32_t *__restrict__ mask) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int iw = blockIdx.x;
  const int i0 = iw * 64;
  const int j0 = (blockIdx.y * kWaves + wave) * 64;
  if (i0 >= ni || j0 >= nj) return;  // wave-uniform
  const int j = min(j0 + lane, nj - 1);
  constexpr int X = ClassX<Pol>::value;
  const double xj = js.f[X][j], yj = js.f[X + 1][j], zj = js.f[X + 2][j];
  const double cut = Pol::mask_cut(prm);
  // the lane's own index when its self pair is skipped, else -1
  const int self = same && Pol::self_noop(js, j, prm) ? j : -1;
  const int ie = min(ni, i0 + 64);
  bool nb = false;
#pragma unroll 8
  for (int i = i0; i < ie; ++i) {
    const double dx = is.f[X][i] - xj;
    const double dy = is.f[X + 1][i] - yj;
    const double dz = is.f[X + 2][i] - zj;
    nb |= (dx * dx + dy * dy + dz * dz < cut) & (i != self);
  }
  const uint64_t bits = __builtin_amdgcn_ballot_w64(nb && j0 + lane < nj);
  const int t = j0 / kTile;
  if (lane == 0) {
    mask[(size_t)iw * ntiles + t] = (uint32_t)bits;
    if (t + 1 < ntiles) mask[(size_t)iw * ntiles + t + 1] = (uint32_t)(bits >> 32);
  }
}

// Policies whose parameters live partly in device memory (a scalar written
// by an earlier kernel of the same stream) provide prepare(prm), which the
// kernels call once at their start to resolve them into the by-value copy.
template <class Pol, class = void>
struct HasPrepare {
  static constexpr bool value = false;
};
template <class Pol>
struct HasPrepare<Pol, decltype((void)Pol::prepare(
                           *(typename Pol::Params *)nullptr))> {
  static constexpr bool value = true;
};
template <class Pol>
__device__ __forceinline__ void prepare_params(typename Pol::Params &prm) {
  if constexpr (HasPrepare<Pol>::value) prm = Pol::prepare(prm);
}

// ------------------------------------------------------------- engine
// FP32C policies (kTwoLevel) also receive the caller's FP64 arrays and the
// j index: a pair they classify as "near" is re-evaluated from those in FP64.
// Direct (FP64 policies, small nj): no packed records -- each wave-uniform
// j record is built from the caller's SoA arrays with scalar loads where it
// is used, which saves the packing launch when that launch would cost more
// than the pair loop gains from one-record loads.
template <class Pol, int P, bool Direct = false>
__global__ __launch_bounds__(kBlock) void pair_engine(
    Soa is, Soa js, int ni, const typename Pol::JRec *__restrict__ jr, int nj,
    int jchunk, int nrows, double *__restrict__ part, Out out,
    typename Pol::Params prm, const double *__restrict__ aux) {
  static_assert(!(Direct && Pol::kTwoLevel), "FP32C records need packing");
  prepare_params<Pol>(prm);
  auto rec = [&](int j) -> typename Pol::JRec {
    if constexpr (Direct)
      return Pol::pack(js, j, 0.0);
    else
      return jr[j];
  };
  const int lane = threadIdx.x & 63;
  // wave-uniform by construction; readfirstlane lets the compiler keep the
  // j range, and so the j-record loads, scalar
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // blockIdx.x: i-wave (fastest, so the blocks in flight share j chunks);
  // blockIdx.y: block row of kWaves j chunks
  const int row = blockIdx.y;
  const int i0 = blockIdx.x * 64 * P;
  if (i0 >= ni) return;  // block-uniform
  // the block's waves take consecutive j chunks of the same i-particles
  const int s = row * kWaves + wave;

  typename Pol::I ist[P];
  typename Pol::A64 acc[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    int i = i0 + p * 64 + lane;
    i = i < ni ? i : ni - 1;
    ist[p] = Pol::load_i(is, i, prm);
    Pol::zero(acc[p]);
  }

  const int jb = s * jchunk;
  const int je = min(nj, jb + jchunk);
  if constexpr (Pol::kTwoLevel) {
    // FP32 tiles of kTile j's.  pair32 returns the wave's mask of lanes
    // whose pair is in the far field (near pairs contribute zero in FP32);
    // the wave-uniform bit k of near[p] records that some lane of slot p
    // met j0 + k in its near field, and after the tile those (k, p) are
    // re-evaluated in FP64 for exactly those lanes (near_pair).  The hot loop carries no FP64
    // code, so the register budget stays that of the FP32 path.
    // j0 is a multiple of kTile (jchunk is), so tile j0 / kTile is the one
    // the packer's per-tile origins refer to.
    for (int j0 = jb; j0 < je; j0 += kTile) {
      const int nk = min(kTile, je - j0);
      typename Pol::A32 a[P];
      uint32_t near[P];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        Pol::zero(a[p]);
        near[p] = 0u;
        Pol::tile_begin(ist[p], aux, j0 / kTile);
      }
      // the record of j+1 is fetched (s_load) while j is computed
      typename Pol::JRec r = jr[j0];
      if (nk == kTile) {
#pragma unroll 2
        for (int k = 0; k < kTile; ++k) {
          const typename Pol::JRec rn = jr[j0 + min(k + 1, kTile - 1)];
#pragma unroll
          for (int p = 0; p < P; ++p)
            near[p] |= near_bit(Pol::pair32(ist[p], r, a[p])) << k;
          r = rn;
        }
      } else {
        for (int k = 0; k < nk; ++k) {
          const typename Pol::JRec rn = jr[j0 + min(k + 1, nk - 1)];
#pragma unroll
          for (int p = 0; p < P; ++p)
            near[p] |= near_bit(Pol::pair32(ist[p], r, a[p])) << k;
          r = rn;
        }
      }
#pragma unroll
      for (int p = 0; p < P; ++p) {
        Pol::fold(a[p], acc[p]);
        uint32_t bits = near[p];
        while (bits) {
          const int k = __builtin_ctz(bits);
          bits &= bits - 1u;
          Pol::near_pair(ist[p], jr[j0 + k], acc[p], is, js, j0 + k);
        }
      }
    }
  } else if constexpr (SparsePairs<Pol>::value) {
    // One call site of pair() (a Kepler solve: inlined once), fed per tile
    // by a mask of the j's to visit: all of them without a cut-off, else
    // the launcher's near mask or, without one, this wave's own test.
    const bool sparse = Pol::sparse(prm);
    for (int j0 = jb; j0 < je; j0 += kTile) {
      const int nk = min(kTile, je - j0);
      uint32_t need;  // wave-uniform: bit k = some lane needs j0 + k
      if (!sparse) {
        need = nk == kTile ? ~0u : (1u << nk) - 1u;
      } else if (prm.mask) {
        need = prm.mask[(size_t)blockIdx.x * prm.mask_tiles + j0 / kTile];
      } else {
        need = 0u;
        for (int k = 0; k < nk; ++k) {
          const typename Pol::JRec r = rec(j0 + k);
          bool nb = false;
#pragma unroll
          for (int p = 0; p < P; ++p) nb = nb || Pol::near(ist[p], r, prm);
          need |= (__any(nb) ? 1u : 0u) << k;
        }
      }
      if (!need) continue;
      // the record of the next set bit is fetched while this one is used
      int k = __builtin_ctz(need);
      need &= need - 1u;
      typename Pol::JRec r = rec(j0 + k);
      for (;;) {
        const int kn = need ? __builtin_ctz(need) : k;
        const typename Pol::JRec rn = rec(j0 + kn);
#pragma unroll
        for (int p = 0; p < P; ++p) Pol::pair(ist[p], r, acc[p], prm);
        if (!need) break;
        need &= need - 1u;
        r = rn;
      }
    }
  } else if (jb < je) {
    typename Pol::JRec r = rec(jb);
#pragma unroll 2
    for (int j = jb; j < je; ++j) {
      const typename Pol::JRec rn = rec(min(j + 1, je - 1));
#pragma unroll
      for (int p = 0; p < P; ++p) Pol::pair(ist[p], r, acc[p], prm);
      r = rn;
    }
  }

  // Waves 1..3 hand their sums to wave 0 through LDS; wave 0 combines
  // them in wave order (fixed grouping: bit-reproducible) and writes the
  // block's result, or its partial when j is split over several blocks.
  __shared__ double red[kWaves - 1][P][Pol::kOut][64];
  if (wave > 0) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      double v[Pol::kOut];
      Pol::values(acc[p], v);
#pragma unroll
      for (int k = 0; k < Pol::kOut; ++k) red[wave - 1][p][k][lane] = v[k];
    }
  }
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    double v[Pol::kOut];
    Pol::values(acc[p], v);
    for (int w = 0; w < kWaves - 1; ++w) {
      double u[Pol::kOut];
#pragma unroll
      for (int k = 0; k < Pol::kOut; ++k) u[k] = red[w][p][k][lane];
      Pol::combine(v, u);
    }
    const int i = i0 + p * 64 + lane;
    if (i >= ni) continue;
    if (nrows == 1) {
      Pol::finish(v, out, i, prm);
    } else {
#pragma unroll
      for (int k = 0; k < Pol::kOut; ++k)
        part[((size_t)row * Pol::kOut + k) * ni + i] = v[k];
    }
  }
}

// Adds the S block partials in block order and applies the policy's
// finishing step (e.g. Tstep's eta / sqrt(1 + w2)).
template <class Pol>
__global__ __launch_bounds__(256) void finish_kernel(
    const double *__restrict__ part, int S, int ni, Out out,
    typename Pol::Params prm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  double v[Pol::kOut];
#pragma unroll
  for (int k = 0; k < Pol::kOut; ++k) v[k] = part[(size_t)k * ni + i];
  for (int s = 1; s < S; ++s) {
    double w[Pol::kOut];
#pragma unroll
    for (int k = 0; k < Pol::kOut; ++k)
      w[k] = part[((size_t)s * Pol::kOut + k) * ni + i];
    Pol::combine(v, w);
  }
  Pol::finish(v, out, i, prm);
}

// Per-call context of the FP32C packers, in device memory:
//   ctx[0] near-field factor, ctx[1] max i-side e2, ctx[2..] per-tile
//   origins (4 doubles per tile of kTile j's).
// Near field: a pair (i, j) is "near" when r_ij^2 < (m_j / <m>) *
// (c * nj^(-1/3))^2 -- inside c mean interparticle spacings of a unit-size
// system, widened by the square root of the j-mass in units of the mean
// mass, i.e. a fixed threshold on the pair force m_j / r^2 -- so the
// factor is c^2 nj^(1/3) / sum_j m_j (0 when c = 0: no near field beyond
// the r^2 = 0 pairs).
constexpr int kCtxBlocks = 64;

// Partial sums of m_j and maxima of e2_i, one pair per block.
static __global__ __launch_bounds__(256) void ctx_partial_kernel(
    const double *__restrict__ m, int nj, const double *__restrict__ e2,
    int ni, double *__restrict__ part) {
  __shared__ double ps[256], px[256];
  double s = 0.0, x = 0.0;
  for (int j = blockIdx.x * 256 + threadIdx.x; j < nj; j += kCtxBlocks * 256)
    s += m[j];
  for (int i = blockIdx.x * 256 + threadIdx.x; i < ni; i += kCtxBlocks * 256)
    x = fmax(x, e2[i]);
  ps[threadIdx.x] = s;
  px[threadIdx.x] = x;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      ps[threadIdx.x] += ps[threadIdx.x + w];
      px[threadIdx.x] = fmax(px[threadIdx.x], px[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ps[0];
    part[2 * blockIdx.x + 1] = px[0];
  }
}

// Combines the block partials in block order (run-to-run reproducible).
static __global__ void ctx_final_kernel(const double *__restrict__ part,
                                        int nj, double c,
                                        double *__restrict__ ctx) {
  if (threadIdx.x != 0) return;
  double msum = 0.0, e2max = 0.0;
  for (int b = 0; b < kCtxBlocks; ++b) {
    msum += part[2 * b];
    e2max = fmax(e2max, part[2 * b + 1]);
  }
  const double f = c * c * cbrt((double)nj) / msum;
  ctx[0] = (c > 0.0 && msum > 0.0 && f < 1e300) ? f : 0.0;
  ctx[1] = e2max;
}

// Per-tile origin of the fields `f[0..2]` (the midpoint of their range over
// the tile's j's, times `scale`), for policies with kVelOrigins.  One
// thread per tile.
static __global__ __launch_bounds__(256) void tile_origin_kernel(
    const double *__restrict__ f0, const double *__restrict__ f1,
    const double *__restrict__ f2, int nj, double scale,
    double *__restrict__ origins) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t * kTile >= nj) return;
  const double *f[3] = {f0, f1, f2};
  const int je = min(nj, (t + 1) * kTile);
  for (int k = 0; k < 3; ++k) {
    double lo = f[k][t * kTile], hi = lo;
    for (int j = t * kTile + 1; j < je; ++j) {
      lo = fmin(lo, f[k][j]);
      hi = fmax(hi, f[k][j]);
    }
    origins[4 * t + k] = 0.5 * (lo + hi) * scale;
  }
  origins[4 * t + 3] = 0.0;
}

// Packs the caller's SoA j fields into the policy's record; FP32C policies
// read the context written by the kernels above.
template <class Pol>
__global__ __launch_bounds__(256) void pack_kernel(
    Soa js, int nj, typename Pol::JRec *__restrict__ jr,
    const double *__restrict__ ctx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj) return;
  if constexpr (Pol::kTwoLevel)
    jr[j] = Pol::pack(js, j, PackCtx{ctx[0], ctx[1], ctx + 2});
  else
    jr[j] = Pol::pack(js, j, 0.0);
}

// Shared defaults for policies whose outputs are plain sums.
template <int N>
struct SumOut {
  static constexpr int kOut = N;
  __device__ static void combine(double *v, const double *w) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += w[k];
  }
  template <class Params>
  __device__ static void finish(const double *v, Out out, int i,
                                const Params &) {
#pragma unroll
    for (int k = 0; k < N; ++k) out.f[k][i] = v[k];
  }
};

struct NoParams {};


template <int N>
struct F64Acc {
  double v[N];
};
template <int N>
struct F32Acc {
  float v[N];
};

template <int N>
__device__ __forceinline__ void zero_acc(F64Acc<N> &a) {
#pragma unroll
  for (int k = 0; k < N; ++k) a.v[k] = 0.0;
}
template <int N>
__device__ __forceinline__ void zero_acc(F32Acc<N> &a) {
#pragma unroll
  for (int k = 0; k < N; ++k) a.v[k] = 0.0f;
}
template <int N>
__device__ __forceinline__ void fold_acc(const F32Acc<N> &a, F64Acc<N> &b) {
#pragma unroll
  for (int k = 0; k < N; ++k) b.v[k] += static_cast<double>(a.v[k]);
}
template <int N>
__device__ __forceinline__ void values_acc(const F64Acc<N> &a, double *v) {
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = a.v[k];
}

}  // namespace tupan
