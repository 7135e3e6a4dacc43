// runtime.hip -- error reporting, launch-shape selection, device memory and
// host<->HBM staging shared by every kernel of the library.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "launch.hpp"

namespace tupan {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return set_error(TUPAN_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e),
                   (int)e);
}

// ------------------------------------------------------ launch shapes
static std::mutex g_mu;
static LaunchShape g_override[TUPAN_K_COUNT][TUPAN_MODE_COUNT];  // zero = automatic
static int g_default_mode = TUPAN_MODE_FP64;
static double g_near_c = 5.0;          // near-field radius, see engine.hpp
static double g_flag_tau = kFlagTau;   // sorted FP32C: rows recomputed in FP64
static bool g_sorted = true;           // large FP32C calls in Morton order
static int g_sakura_budget = 1 << 20;  // Kepler solves per Sakura pair

double near_field_factor() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_near_c;
}

double flag_tau() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_flag_tau;
}

bool sorted_enabled() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_sorted;
}

static const int *g_last_flags = nullptr;
static hipStream_t g_last_flags_stream = nullptr;
void set_last_flags(const int *flags, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_last_flags = flags;
  g_last_flags_stream = st;
}

int sakura_budget() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_sakura_budget;
}

static int simd_count() {
  static int cached = 0;
  if (cached) return cached;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            dev) == hipSuccess &&
      cus > 0)
    cached = 4 * cus;  // 4 SIMDs per CU on CDNA
  else
    cached = 1024;
  return cached;
}

// Launch shape: i-particles per lane P (the policy's preferred value unless
// overridden) and j splits S.  The pair loop's scalar-load latency is hidden
// by resident waves, so the grid aims at `waves_per_simd` waves per SIMD
// over its lifetime (64 for the pair engine, measured best on MI355X for
// N = 16K..64K; 128 for the sorted FP32C engine, whose near tiles make the
// waves' work uneven), never splitting j below 128 records per wave.
LaunchShape choose_shape(int kid, int mode, size_t ni, size_t nj,
                         int pref_ipl, int max_ipl, int min_chunk,
                         int waves_per_simd, int max_splits) {
  LaunchShape o{0, 0};
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (kid >= 0 && kid < TUPAN_K_COUNT && mode >= 0 &&
        mode < TUPAN_MODE_COUNT)
      o = g_override[kid][mode];
  }
  const long target = (long)waves_per_simd * simd_count();
  // j chunks of at least 128 records, or of min_chunk (a tile, or 8 for the
  // plain FP64 loop) while that keeps the split count at 64 or below: small
  // systems need the extra waves.
  long max_split = nj >= 256 ? (long)(nj / 128) : 1;
  const long fine = std::min<long>((long)(nj / (size_t)min_chunk), 64);
  if (fine > max_split) max_split = fine;
  int P = o.ipl > 0 ? o.ipl : pref_ipl;
  if (P > max_ipl) P = max_ipl;
  if (P != 1 && P != 2 && P != 4) P = 1;
  const long waves_i = (long)((ni + 64 * P - 1) / (64 * P));
  long S = o.splits > 0 ? o.splits : (target + waves_i - 1) / waves_i;
  if (S > max_split && o.splits <= 0) S = max_split;
  if (max_splits > 0 && S > max_splits && o.splits <= 0) S = max_splits;
  if (S < 1) S = 1;
  if (nj > 0 && S > (long)nj) S = (long)nj;
  return LaunchShape{P, (int)S};
}

// ------------------------------------------------------ device memory
// Stream-ordered allocation from the device's default memory pool; the
// pool keeps freed blocks (release threshold raised once per device), so
// steady-state calls do not reach the driver.
int device_alloc(void **p, size_t bytes, hipStream_t st) {
  static bool pool_set[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (dev >= 0 && dev < 64 && !pool_set[dev]) {
      hipMemPool_t pool;
      if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t thr = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold,
                                     &thr);
      }
      pool_set[dev] = true;
    }
  }
  if (bytes == 0) bytes = 8;
  e = hipMallocAsync(p, bytes, st);
  if (e != hipSuccess) return hip_fail(e, "hipMallocAsync");
  return TUPAN_OK;
}

void device_free(void *p, hipStream_t st) {
  if (p) (void)hipFreeAsync(p, st);
}

// ------------------------------------------------------ scratch cache
// Per (device, stream, slot) buffers reused from call to call: calls on one
// stream run in order, so a buffer serves the next call as soon as the
// previous call's kernels are enqueued, and steady-state calls allocate
// nothing (which also lets a caller capture them into a HIP graph).
// Buffers only grow, and a replaced buffer is retired, not freed, until
// tupan_release_scratch: kernels captured into a graph keep valid pointers.
// A buffer in use by another host thread (same stream and slot) is not
// shared: that call allocates its own and frees it stream-ordered.
struct ScratchEntry {
  void *p = nullptr;
  size_t bytes = 0;
  bool busy = false;
};
static std::map<std::tuple<int, hipStream_t, int>, ScratchEntry> g_scratch;
static std::vector<std::pair<int, void *>> g_retired;  // (device, buffer)

static bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) return false;
  return cs != hipStreamCaptureStatusNone;
}

int scratch_alloc(int slot, void **p, size_t bytes, hipStream_t st) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (bytes == 0) bytes = 8;
  const bool cap = capturing(st);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ScratchEntry &s = g_scratch[std::make_tuple(dev, st, slot)];
    if (!s.busy && s.bytes >= bytes) {
      s.busy = true;
      *p = s.p;
      return TUPAN_OK;
    }
    if (!s.busy) {
      if (cap)
        return set_error(TUPAN_ERR_ARG,
                         "scratch buffer too small during stream capture: "
                         "make the same call once before capturing");
      size_t nb = bytes > s.bytes + s.bytes / 2 ? bytes : s.bytes + s.bytes / 2;
      nb = (nb + 255) / 256 * 256;
      void *q = nullptr;
      e = hipMallocAsync(&q, nb, st);
      if (e != hipSuccess) return hip_fail(e, "hipMallocAsync");
      if (s.p) g_retired.emplace_back(dev, s.p);
      s.p = q;
      s.bytes = nb;
      s.busy = true;
      *p = q;
      return TUPAN_OK;
    }
  }
  if (cap)
    return set_error(TUPAN_ERR_ARG,
                     "scratch buffer busy during stream capture");
  return device_alloc(p, bytes, st);
}

void scratch_free(int slot, void *p, hipStream_t st) {
  if (!p) return;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_scratch.find(std::make_tuple(dev, st, slot));
    if (it != g_scratch.end() && it->second.p == p) {
      it->second.busy = false;
      return;
    }
  }
  device_free(p, st);
}

// ------------------------------------------------------ live timing
static bool g_timing = false;
static std::vector<std::pair<hipEvent_t, hipEvent_t>> g_events;

// Launches being captured into a graph are not timed: their events would
// be recorded only when the graph replays.
bool timing_begin(hipEvent_t *e0, hipEvent_t *e1, hipStream_t st) {
  if (capturing(st)) return false;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_timing) return false;
  if (hipEventCreate(e0) != hipSuccess) return false;
  if (hipEventCreate(e1) != hipSuccess) {
    (void)hipEventDestroy(*e0);
    return false;
  }
  g_events.emplace_back(*e0, *e1);
  return true;
}

// ------------------------------------------------------ host staging
int host_call(size_t ni, const double *const *ip, size_t nj,
              const double *const *jp, int nf, double *const *op, int nout,
              const DevFn &fn) {
  if (ni == 0) return TUPAN_OK;
  if (!ip || !jp || !op) return set_error(TUPAN_ERR_ARG, "NULL field array");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1)
    return set_error(TUPAN_ERR_HIP, "no HIP device available (%s)",
                     hipGetErrorString(e));
  hipStream_t st = nullptr;
  std::vector<double *> di(nf, nullptr), dj(nf, nullptr), dout(nout, nullptr);
  std::vector<void *> owned;
  int rc = TUPAN_OK;
  auto stage = [&](const double *h, size_t n, double **d) -> int {
    int r = device_alloc((void **)d, n * sizeof(double), st);
    if (r) return r;
    owned.push_back(*d);
    hipError_t ce =
        hipMemcpyAsync(*d, h, n * sizeof(double), hipMemcpyHostToDevice, st);
    return ce == hipSuccess ? TUPAN_OK : hip_fail(ce, "H2D copy");
  };
  for (int k = 0; k < nf && !rc; ++k) {
    if (!ip[k]) { rc = set_error(TUPAN_ERR_ARG, "NULL i field %d", k); break; }
    rc = stage(ip[k], ni, &di[k]);
  }
  for (int k = 0; k < nf && !rc && nj; ++k) {
    if (!jp[k]) { rc = set_error(TUPAN_ERR_ARG, "NULL j field %d", k); break; }
    if (jp[k] == ip[k] && nj == ni)
      dj[k] = di[k];
    else
      rc = stage(jp[k], nj, &dj[k]);
  }
  for (int k = 0; k < nout && !rc; ++k) {
    if (!op[k]) { rc = set_error(TUPAN_ERR_ARG, "NULL output %d", k); break; }
    rc = device_alloc((void **)&dout[k], ni * sizeof(double), st);
    if (!rc) owned.push_back(dout[k]);
  }
  if (!rc) rc = fn(di.data(), dj.data(), dout.data(), st);
  int rc_keep = rc;
  if (rc == TUPAN_OK || rc == TUPAN_ERR_NOT_CONVERGED) {
    for (int k = 0; k < nout; ++k) {
      hipError_t ce = hipMemcpyAsync(op[k], dout[k], ni * sizeof(double),
                                     hipMemcpyDeviceToHost, st);
      if (ce != hipSuccess) { rc_keep = hip_fail(ce, "D2H copy"); break; }
    }
  }
  for (void *p : owned) device_free(p, st);
  e = hipStreamSynchronize(st);
  if (e != hipSuccess && rc_keep == TUPAN_OK)
    rc_keep = hip_fail(e, "hipStreamSynchronize");
  return rc_keep;
}

}  // namespace tupan

using namespace tupan;

extern "C" {

const char *tupan_version(void) { return "tupan_hip 0.1.0 (gfx950)"; }

const char *tupan_last_error(void) { return g_err; }

int tupan_device_count(int *count) {
  if (!count) return set_error(TUPAN_ERR_ARG, "NULL count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = (e == hipSuccess) ? n : 0;
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
  return TUPAN_OK;
}

int tupan_set_default_mode(int mode) {
  if (mode < 0 || mode >= TUPAN_MODE_COUNT)
    return set_error(TUPAN_ERR_ARG, "unknown mode %d", mode);
  std::lock_guard<std::mutex> lk(g_mu);
  g_default_mode = mode;
  return TUPAN_OK;
}

int tupan_get_default_mode(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_default_mode;
}

int tupan_set_near_field(double c) {
  if (!(c >= 0.0) || c > 1e6)
    return set_error(TUPAN_ERR_ARG, "near-field factor must be in [0, 1e6]");
  std::lock_guard<std::mutex> lk(g_mu);
  g_near_c = c;
  return TUPAN_OK;
}

int tupan_set_sorted(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_sorted = on != 0;
  return TUPAN_OK;
}

int tupan_set_flag_tau(double tau) {
  if (!(tau >= 0.0) || tau > 1e6)
    return set_error(TUPAN_ERR_ARG, "flag threshold must be in [0, 1e6]");
  std::lock_guard<std::mutex> lk(g_mu);
  g_flag_tau = tau;
  return TUPAN_OK;
}

int tupan_last_flagged(int *count) {
  if (!count) return set_error(TUPAN_ERR_ARG, "NULL count");
  const int *f;
  hipStream_t st;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    f = g_last_flags;
    st = g_last_flags_stream;
  }
  *count = 0;
  if (!f) return TUPAN_OK;
  hipError_t e = hipMemcpyAsync(count, f, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "reading the flag count");
  return TUPAN_OK;
}

int tupan_set_sakura_budget(int solves) {
  if (solves < 1) return set_error(TUPAN_ERR_ARG, "budget must be >= 1");
  std::lock_guard<std::mutex> lk(g_mu);
  g_sakura_budget = solves;
  return TUPAN_OK;
}

int tupan_release_scratch(void) {
  std::vector<std::pair<int, void *>> all;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &kv : g_scratch) {
      if (kv.second.busy)
        return set_error(TUPAN_ERR_ARG, "a scratch buffer is in use");
    }
    for (auto &kv : g_scratch)
      if (kv.second.p) all.emplace_back(std::get<0>(kv.first), kv.second.p);
    all.insert(all.end(), g_retired.begin(), g_retired.end());
    g_scratch.clear();
    g_retired.clear();
  }
  int cur = 0;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  int rc = TUPAN_OK;
  for (auto &dp : all) {
    if (hipSetDevice(dp.first) != hipSuccess) continue;
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipFree(dp.second);
    if (e != hipSuccess && rc == TUPAN_OK) rc = hip_fail(e, "hipFree");
  }
  (void)hipSetDevice(cur);
  return rc;
}

int tupan_set_timing(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_timing = on != 0;
  return TUPAN_OK;
}

int tupan_take_kernel_times(double *total_ms, int *count) {
  if (!total_ms || !count) return set_error(TUPAN_ERR_ARG, "NULL output");
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ev.swap(g_events);
  }
  double sum = 0.0;
  int rc = TUPAN_OK;
  for (auto &p : ev) {
    float ms = 0.f;
    hipError_t e = hipEventSynchronize(p.second);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.first, p.second);
    if (e != hipSuccess && rc == TUPAN_OK) rc = hip_fail(e, "event timing");
    sum += ms;
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  *total_ms = sum;
  *count = (int)ev.size();
  return rc;
}

int tupan_set_launch(int kernel_id, int mode, int ipl, int splits) {
  if (kernel_id < 0 || kernel_id >= TUPAN_K_COUNT)
    return set_error(TUPAN_ERR_ARG, "bad kernel id %d", kernel_id);
  if (mode < 0 || mode >= TUPAN_MODE_COUNT)
    return set_error(TUPAN_ERR_ARG, "unknown mode %d", mode);
  if (ipl != 0 && ipl != 1 && ipl != 2 && ipl != 4)
    return set_error(TUPAN_ERR_ARG, "ipl must be 0, 1, 2 or 4");
  if (splits < 0) return set_error(TUPAN_ERR_ARG, "splits must be >= 0");
  if ((mode == TUPAN_MODE_FP64_IEEE || mode == TUPAN_MODE_FP32_IEEE) &&
      splits > 1)
    return set_error(TUPAN_ERR_ARG, "the IEEE modes run one j split");
  std::lock_guard<std::mutex> lk(g_mu);
  g_override[kernel_id][mode] = LaunchShape{ipl, splits};
  return TUPAN_OK;
}

}  // extern "C"
