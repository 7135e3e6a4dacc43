"""ctypes binding of libtupan_hip.so (include/tupan_hip.h).

The library is built in-tree by ``tupan_amd.build``.  There is no fallback:
if the shared object is missing, or the HIP runtime reports no device,
every call raises ``TupanError``.
"""
import ctypes
import os
import threading

import numpy as np

from . import spec

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TUPAN_HIP_LIB",
                          os.path.join(HERE, "lib", "libtupan_hip.so"))

# NREG scalar block slots and ops (include/tupan_hip.h)
NREG_E0, NREG_W, NREG_U, NREG_S, NREG_MTOT = 0, 1, 2, 3, 4
NREG_DT, NREG_T, NREG_T0, NREG_SUM, NREG_HSUB = 5, 6, 7, 8, 9
NREG_SCALARS = 10
(NREG_OP_BEGIN, NREG_OP_H, NREG_OP_HS, NREG_OP_XDT, NREG_OP_VDT,
 NREG_OP_XEND, NREG_OP_VEND, NREG_OP_S) = range(8)

MODE_FP64 = 0
MODE_FP32C = 1
MODE_FP64_IEEE = 2
MODE_FP32_IEEE = 3
MODE_COUNT = 4
MODES = {"fp64": MODE_FP64, "fp32c": MODE_FP32C, "fp64_ieee": MODE_FP64_IEEE,
         "fp32_ieee": MODE_FP32_IEEE}

OK = 0
ERR_HIP = 1
ERR_ARG = 2
ERR_NOT_CONVERGED = 3


class TupanError(RuntimeError):
    def __init__(self, code, msg):
        RuntimeError.__init__(self, "tupan_hip error %d: %s" % (code, msg))
        self.code = code


class NotConverged(TupanError):
    pass


_lock = threading.Lock()
_lib = None

_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_SZ = ctypes.c_size_t


def lib():
    """Load (once) and return the CDLL; raise TupanError if not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise TupanError(-1, "HIP library not built: %s (run "
                             "python -m tupan_amd.build)" % LIB_PATH)
        # torch-ROCm bundles its own HIP runtime under the same soname as
        # /opt/rocm's.  Whichever loads first serves the whole process, and
        # torch cannot see the GPU through a runtime other than its own, so
        # let torch load its runtime before ours binds to it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        L.tupan_version.restype = ctypes.c_char_p
        L.tupan_last_error.restype = ctypes.c_char_p
        L.tupan_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.tupan_set_default_mode.argtypes = [ctypes.c_int]
        L.tupan_set_launch.argtypes = [ctypes.c_int] * 4
        L.tupan_set_timing.argtypes = [ctypes.c_int]
        L.tupan_set_near_field.argtypes = [ctypes.c_double]
        L.tupan_set_sorted.argtypes = [ctypes.c_int]
        L.tupan_last_flagged.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.tupan_set_flag_tau.argtypes = [ctypes.c_double]
        L.tupan_check_finite_dev.argtypes = [_SZ, ctypes.c_int, _PP, _P, _P]
        L.tupan_set_sakura_budget.argtypes = [ctypes.c_int]
        L.tupan_take_kernel_times.argtypes = [ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_int)]
        for name in ("phi", "acc", "acc_jerk", "snap_crackle", "acc_adjoint",
                     "phi_adjoint"):
            getattr(L, "tupan_%s_dev" % name).argtypes = [
                ctypes.c_int, _SZ, _PP, _SZ, _PP, _PP, _P]
        for name in ("phi", "acc", "acc_jerk", "snap_crackle"):
            getattr(L, "tupan_%s_host" % name).argtypes = [
                ctypes.c_int, _SZ, _PP, _SZ, _PP, _PP]
        L.tupan_tstep_dev.argtypes = [ctypes.c_int, _SZ, _PP, _SZ, _PP,
                                      ctypes.c_double,
                                      _PP, _P]
        L.tupan_tstep_host.argtypes = [ctypes.c_int, _SZ, _PP, _SZ, _PP,
                                       ctypes.c_double,
                                       _PP]
        if hasattr(L, "tupan_pnacc_dev"):
            L.tupan_pnacc_dev.argtypes = [ctypes.c_int, _SZ, _PP, _SZ, _PP,
                                          ctypes.c_uint, _P, _PP, _P]
            L.tupan_pnacc_host.argtypes = [ctypes.c_int, _SZ, _PP, _SZ, _PP,
                                           ctypes.c_uint, _P, _PP]
        if hasattr(L, "tupan_sakura_dev"):
            L.tupan_sakura_dev.argtypes = [_SZ, _PP, _SZ, _PP,
                                           ctypes.c_double, ctypes.c_int,
                                           _PP, _P, _P]
            L.tupan_sakura_host.argtypes = [_SZ, _PP, _SZ, _PP,
                                            ctypes.c_double, ctypes.c_int,
                                            _PP]
        _D, _I = ctypes.c_double, ctypes.c_int
        L.tupan_hermite_predict_dev.argtypes = [_I, _SZ, _PP, _PP, _D, _P]
        L.tupan_hermite_correct_dev.argtypes = [_I, _SZ, _PP, _PP, _D, _P]
        L.tupan_drift_dev.argtypes = [_SZ, _PP, _PP, _D, _I, _PP, _PP, _P]
        L.tupan_kick_dev.argtypes = [_SZ, _PP, _PP, _D, _P]
        L.tupan_kick_pn_half_dev.argtypes = [_SZ, _PP, _PP, _PP, _D, _P]
        L.tupan_pn_flux_dev.argtypes = [_SZ, _PP, _PP, _D, _P]
        L.tupan_apply_delta_dev.argtypes = [_SZ, _PP, _PP, _P]
        L.tupan_pn_flux_kick_dev.argtypes = [_SZ, _PP, _PP, _PP, _D, _P]
        L.tupan_advance_clock_dev.argtypes = [_SZ, _P, _P, _P, _D, _P]
        L.tupan_divide_dev.argtypes = [_SZ, ctypes.c_int, _PP, _PP,
                                       ctypes.c_double, _P]
        L.tupan_divide_by_dev.argtypes = [_SZ, ctypes.c_int, _PP, _PP, _P, _P]
        L.tupan_half_drift_to_dev.argtypes = [_SZ, _PP, _PP, _D, _PP, _P]
        L.tupan_nreg_scalars_dev.argtypes = [_P, ctypes.c_int,
                                             ctypes.c_double, _P]
        L.tupan_nreg_step_end_dev.argtypes = [_SZ, _P, _P, _P, _P,
                                              ctypes.c_double, _P]
        L.tupan_sum_dev.argtypes = [_SZ, _P, _P, _P]
        L.tupan_sum_pairwise_dev.argtypes = [_SZ, _P, _P, _P]
        L.tupan_energy_terms_dev.argtypes = [_SZ, _PP, _P, _P, _P]
        L.tupan_report_terms_dev.argtypes = [_SZ, _PP, ctypes.c_int, _PP, _P]
        L.tupan_diagnostics_dev.argtypes = [_SZ, _PP, ctypes.c_int, _P, _P]
        for name in ("nreg_x", "nreg_v"):
            getattr(L, "tupan_%s_dev" % name).argtypes = [
                _I, _SZ, _PP, _SZ, _PP, _D, _PP, _P]
            getattr(L, "tupan_%s_devdt" % name).argtypes = [
                _I, _SZ, _PP, _SZ, _PP, _P, _PP, _P]
            getattr(L, "tupan_%s_host" % name).argtypes = [
                _I, _SZ, _PP, _SZ, _PP, _D, _PP]
        L.tupan_kepler_dev.argtypes = [_PP, _D, _PP, _P, _P]
        L.tupan_kepler_host.argtypes = [_PP, _D, _PP]
        L.tupan_acc_far_dev.argtypes = [_SZ, _PP, _SZ, _PP, _D, _PP, _P]
        L.tupan_acc_far_kick_pn_half_dev.argtypes = [_SZ, _PP, _D, _PP, _PP,
                                                     _PP, _D, _P]
        L.tupan_sakura_near_dev.argtypes = [_SZ, _PP, _SZ, _PP, _D, _I, _D,
                                            _PP, _P, _P]
        _PI = ctypes.POINTER(ctypes.c_int)
        _PD = ctypes.POINTER(ctypes.c_double)
        L.tupan_ensemble_create.argtypes = [ctypes.POINTER(_P), _I, _PI, _D,
                                            _D]
        L.tupan_ensemble_destroy.argtypes = [_P]
        L.tupan_ensemble_set_launch.argtypes = [_P, _I, _I, _I]
        L.tupan_ensemble_set_mode.argtypes = [_P, _I]
        L.tupan_ensemble_set_order.argtypes = [_P, _I]
        L.tupan_ensemble_set_state.argtypes = [_P, _PP, _PD]
        L.tupan_ensemble_get_state.argtypes = [
            _P, _PP, _PD, ctypes.POINTER(ctypes.c_longlong)]
        L.tupan_ensemble_step.argtypes = [_P, _I, _I]
        L.tupan_ensemble_poll.argtypes = [_P, _PI, _PD, _PI]
        _lib = L
        return _lib


def check(rc):
    if rc == OK:
        return
    msg = lib().tupan_last_error().decode(errors="replace")
    if rc == ERR_NOT_CONVERGED:
        raise NotConverged(rc, msg)
    raise TupanError(rc, msg)


def device_count():
    n = ctypes.c_int(0)
    rc = lib().tupan_device_count(ctypes.byref(n))
    return n.value if rc == OK else 0


def set_launch(kernel, mode, ipl=0, splits=0):
    """Autotuner hook: fix the launch shape of one kernel/mode."""
    check(lib().tupan_set_launch(spec.KERNEL_IDS[kernel], MODES.get(mode, mode),
                                 int(ipl), int(splits)))


def set_near_field(c):
    """FP32C near-field factor of the single-system kernels (default 5; 0
    leaves only the r^2 = 0 pairs to the FP64 branch)."""
    check(lib().tupan_set_near_field(float(c)))


def set_flag_tau(tau):
    """Sorted FP32C: rows whose estimated FP32 error exceeds tau |result|
    are recomputed in FP64 (default 0.7; 0: every row with FP32 tiles)."""
    check(lib().tupan_set_flag_tau(float(tau)))


def last_flagged():
    """Rows the last sorted FP32C call recomputed in FP64 (synchronises)."""
    n = ctypes.c_int(0)
    check(lib().tupan_last_flagged(ctypes.byref(n)))
    return n.value


def set_sorted(on):
    """Large FP32C calls (nj > 4096) over Morton-sorted particles with FP64
    recomputation of flagged rows (default on); off: the unsorted engine."""
    check(lib().tupan_set_sorted(1 if on else 0))


def set_sakura_budget(solves):
    check(lib().tupan_set_sakura_budget(int(solves)))


def release_scratch():
    """Free the per-stream scratch buffers of the device entry points
    (tupan_release_scratch): graphs captured before must not be replayed
    afterwards; later calls allocate afresh."""
    check(lib().tupan_release_scratch())


def set_timing(on):
    check(lib().tupan_set_timing(1 if on else 0))


def take_kernel_times():
    """(total_ms, launches) of the pair-engine kernels timed so far."""
    ms = ctypes.c_double(0)
    n = ctypes.c_int(0)
    check(lib().tupan_take_kernel_times(ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def ptr_array(ptrs):
    """ctypes void* array of the pointers (a cached one when ``ptrs``
    carries it, see integrator.System.ptrs)."""
    cached = getattr(ptrs, "carr", None)
    if cached is not None:
        return cached
    arr = (ctypes.c_void_p * max(1, len(ptrs)))()
    for k, p in enumerate(ptrs):
        arr[k] = p
    return arr


def _cargs(name, mode, scalars):
    """Leading mode and trailing scalar arguments of one kernel."""
    k = spec.KERNELS[name]
    pre = [ctypes.c_int(MODES.get(mode, mode))] if k.has_mode else []
    post = []
    if name == "tstep":
        post = [ctypes.c_double(float(scalars[0]))]
    elif name == "pnacc":
        order, inv = scalars
        inv = (ctypes.c_double * 7)(*[float(x) for x in inv])
        post = [ctypes.c_uint(int(order)), ctypes.cast(inv, _P)]
        post.append(inv)  # keep alive; stripped before the call
    elif name == "sakura":
        post = [ctypes.c_double(float(scalars[0])), ctypes.c_int(int(scalars[1]))]
    elif name in ("acc_far", "nreg_x", "nreg_v"):
        post = [ctypes.c_double(float(scalars[0]))]
    elif name == "sakura_near":
        post = [ctypes.c_double(float(scalars[0])), ctypes.c_int(int(scalars[1])),
                ctypes.c_double(float(scalars[2]))]
    return pre, post


def call_host(name, mode, ni, iptrs, nj, jptrs, optrs, scalars=()):
    """Synchronous host-pointer call (tupan_<name>_host)."""
    L = lib()
    pre, post = _cargs(name, mode, scalars)
    keep = post[2:] if name == "pnacc" else []
    post = post[:2] if name == "pnacc" else post
    ia, ja, oa = ptr_array(iptrs), ptr_array(jptrs), ptr_array(optrs)
    fn = getattr(L, "tupan_%s_host" % name)
    args = pre + [_SZ(ni), ia, _SZ(nj), ja] + post + [oa]
    rc = fn(*args)
    del keep
    check(rc)


STATUS_NOT_CONVERGED = 1
STATUS_NONFINITE = 2


def check_finite_dev(n, ptrs, status_ptr, stream=0):
    """OR STATUS_NONFINITE into the device int at status_ptr when any of
    the n-element device arrays holds NaN or inf (asynchronous)."""
    check(lib().tupan_check_finite_dev(_SZ(n), len(ptrs), ptr_array(ptrs),
                                       _P(status_ptr), _P(stream or 0)))


def call_dev(name, mode, ni, iptrs, nj, jptrs, optrs, scalars=(), stream=0,
             status_ptr=None):
    """Asynchronous device-pointer call (tupan_<name>_dev) on ``stream``."""
    L = lib()
    pre, post = _cargs(name, mode, scalars)
    keep = post[2:] if name == "pnacc" else []
    post = post[:2] if name == "pnacc" else post
    ia, ja, oa = ptr_array(iptrs), ptr_array(jptrs), ptr_array(optrs)
    fn = getattr(L, "tupan_%s_dev" % name)
    args = pre + [_SZ(ni), ia, _SZ(nj), ja] + post + [oa]
    if name in ("sakura", "sakura_near"):
        args.append(_P(status_ptr or 0))
    args.append(_P(stream or 0))
    rc = fn(*args)
    del keep
    check(rc)


def call_kepler_host(iptrs, dt, optrs):
    """kepler_solver_kernel on particles 0 and 1 (host pointers, in place
    allowed)."""
    check(lib().tupan_kepler_host(ptr_array(iptrs), ctypes.c_double(float(dt)),
                                  ptr_array(optrs)))


def call_update(name, *args):
    """Per-particle update kernel ``tupan_<name>_dev``: field-pointer lists
    become ``double**`` (None -> NULL), the rest passes through."""
    conv = []
    for a in args:
        if isinstance(a, (list, tuple)):
            conv.append(ptr_array(a))
        else:
            conv.append(a)
    check(getattr(lib(), "tupan_%s_dev" % name)(*conv))


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
