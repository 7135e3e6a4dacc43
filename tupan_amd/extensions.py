"""Drop-in for tupan/lib/extensions.py on the MI355X backend.

Same classes, constructor arguments, ``calc(ips, jps, *args)`` call surface,
lazily registered output attributes and module-level singletons as the
reference (tupan/lib/extensions.py:30-666), so that integrator code written
against ``tupan.lib.extensions`` runs unchanged against this module.

Backends.  ``get_kernel(name, backend, prec)`` returns an object speaking
the reference's backend protocol (``cty``, ``set_gsize``, ``set_args``,
``map_buffers``, ``run``; tupan/lib/cffi_backend.py:90-126,
tupan/lib/opencl_backend.py:104-171).  The product provides one backend,
"HIP", which calls libtupan_hip.so through its host-pointer ABI.  Further
backends can be registered with ``register_backend`` (the test-suite uses
this to drive the integrators with its CPU checker); nothing here falls back
to a CPU path: a missing library or device raises ``TupanError``.

Precision.  ``prec`` is the reference's I/O precision ("float64" or
"float32", tupan/lib/utils/ctype.py:13-20); float32 arrays are widened to
FP64 for the device and the results narrowed back in place.  ``mode``
("fp64" or "fp32c") picks the arithmetic of Phi/Acc/AccJerk/SnapCrackle on
the device; Tstep, PNAcc and Sakura always run FP64.
"""
from __future__ import division

import logging
import os
import sys
from collections import namedtuple

import numpy as np

from . import _lib, spec

__all__ = ["Phi", "phi", "Acc", "acc", "AccJerk", "acc_jerk",
           "SnapCrackle", "snap_crackle", "Tstep", "tstep",
           "PNAcc", "pnacc", "Sakura", "sakura", "NREG_X", "nreg_x",
           "NREG_V", "nreg_v", "Kepler", "kepler", "Clight", "clight",
           "get_kernel", "register_backend"]

logger = logging.getLogger(__name__)


class Clight(object):
    """PN order and inverse powers of the speed of light
    (tupan/lib/extensions.py:30-60)."""

    def __init__(self):
        self._pn_order = 0
        self._clight = None

    @property
    def pn_order(self):
        return self._pn_order

    @pn_order.setter
    def pn_order(self, value):
        self._pn_order = int(value)

    @property
    def clight(self):
        return self._clight

    @clight.setter
    def clight(self, value):
        self._clight = float(value)
        self.inv1 = 1.0 / self._clight
        self.inv2 = self.inv1**2
        self.inv3 = self.inv1**3
        self.inv4 = self.inv1**4
        self.inv5 = self.inv1**5
        self.inv6 = self.inv1**6
        self.inv7 = self.inv1**7


_Types = namedtuple("Types", ["c_int", "c_int_p", "c_uint", "c_uint_p",
                              "c_real", "c_real_p"])


class HIPKernel(object):
    """Backend-protocol object for one reference kernel symbol.

    ``set_args`` receives the flat argument list the extension classes
    build (libtupan.h order: ni, i-arrays, nj, j-arrays, scalars, outputs);
    ``run`` issues one synchronous host-pointer call on the device.
    """

    def __init__(self, prec, name, mode=None):
        if name not in spec.SYMBOLS:
            raise ValueError("unknown kernel %r" % name)
        self.prec = prec
        self.name = name
        self.key = spec.SYMBOLS[name]
        self.mode = mode
        self.real = np.dtype(np.float64 if prec == "float64" else np.float32)
        ident = lambda x: x
        self.cty = _Types(c_int=int, c_int_p=ident, c_uint=int,
                          c_uint_p=ident, c_real=float, c_real_p=self._arr)
        self.args = ()

    def _arr(self, a):
        if not isinstance(a, np.ndarray) or a.dtype != self.real:
            raise TypeError("%s: expected %s arrays" % (self.name, self.real))
        return a

    def set_gsize(self, ni, nj):
        pass

    def set_args(self, args, start=0):
        self.args = list(args)

    def map_buffers(self, arrays, buffers):
        return arrays

    def run(self):
        if self.key == "kepler":
            return self._run_kepler()
        k = spec.KERNELS[self.key]
        nf = len(k.fields)
        a = self.args
        ni = int(a[0])
        iarr = a[1:1 + nf]
        nj = int(a[1 + nf])
        jarr = a[2 + nf:2 + 2 * nf]
        rest = a[2 + 2 * nf:]
        nout = len(k.outs)
        scalars = rest[:len(rest) - nout]
        outs = rest[len(rest) - nout:]
        if self.key == "pnacc":
            scalars = (scalars[0], tuple(scalars[1:8]))
        mode = self.mode or default_mode(self.prec)
        if mode == "fp32_ieee" and self.key not in spec.FP32_IEEE:
            mode = "fp64"   # no FP32 IEEE variant
        for x in tuple(iarr) + tuple(outs):
            if len(x) < ni:
                raise ValueError("%s: array shorter than ni = %d"
                                 % (self.name, ni))
        for x in jarr:
            if len(x) < nj:
                raise ValueError("%s: array shorter than nj = %d"
                                 % (self.name, nj))
        iw = [_lib.as_f64(x[:ni]) for x in iarr]
        jw = [_lib.as_f64(x[:nj]) for x in jarr]
        # outputs the device can write directly: contiguous float64; any
        # other (float32, strided views) goes through a contiguous buffer
        direct = [o.dtype == np.float64 and o.flags.c_contiguous
                  for o in outs]
        ow = [o if d else np.empty(ni, np.float64)
              for o, d in zip(outs, direct)]
        _lib.call_host(self.key, mode, ni, [x.ctypes.data for x in iw], nj,
                       [x.ctypes.data for x in jw],
                       [o.ctypes.data for o in ow], scalars)
        for o, w, d in zip(outs, ow, direct):
            if not d:
                o[:ni] = w


    def _run_kepler(self):
        # (8 input arrays, dt, 6 outputs) -- libtupan.h:231-246; particles
        # 0 and 1, outputs may be the input arrays themselves
        a = self.args
        ins = [np.ascontiguousarray(x[:2], np.float64) for x in a[:8]]
        res = [np.empty(2, np.float64) for _ in range(6)]
        _lib.call_kepler_host([x.ctypes.data for x in ins], float(a[8]),
                              [o.ctypes.data for o in res])
        for o, r in zip(a[9:15], res):
            o[:2] = r


_BACKENDS = {"HIP": HIPKernel}


def register_backend(name, factory):
    """Make ``factory(prec, symbol)`` available as backend ``name``."""
    _BACKENDS[name] = factory


def get_kernel(name, backend, prec):
    try:
        Kernel = _BACKENDS[backend]
    except KeyError:
        msg = "Inappropriate 'backend': {}. Supported values: {}"
        raise ValueError(msg.format(backend, sorted(_BACKENDS)))
    logger.debug("Using '%s' from %s precision %s extension module.",
                 name, prec, backend)
    return Kernel(prec, name)


def prepare_args(args, argtypes):
    return [argtype(arg) for (arg, argtype) in zip(args, argtypes)]


# The arithmetic of the HIP backend's calls per I/O precision: float64
# calls default to FP64, float32 calls to the reference FP32 build's
# arithmetic ("fp32_ieee": its results bit for bit) for the kernels that
# have it (spec.FP32_IEEE), the others FP64 rounded to float.
_MODE = {"float64": os.environ.get("TUPAN_MODE", "fp64"),
         "float32": os.environ.get("TUPAN_MODE_FP32", "fp32_ieee")}


def default_mode(prec="float64"):
    return _MODE[prec]


def set_default_mode(mode, prec="float64"):
    if mode not in _lib.MODES:
        raise ValueError("mode must be one of %s" % sorted(_lib.MODES))
    if prec not in _MODE:
        raise ValueError("prec must be 'float64' or 'float32'")
    _MODE[prec] = mode


class AbstractExtension(object):
    def set_args(self, ips, jps, *args):
        raise NotImplementedError

    def run(self):
        self.kernel.run()

    def get_result(self):
        return self.kernel.map_buffers(self._outargs, self.outargs)

    def calc(self, ips, jps, *args):
        self.set_args(ips, jps, *args)
        self.run()
        return self.get_result()

    def _register(self, ips, names):
        for n in names:
            if n not in ips.__dict__:
                ips.register_auxiliary_attribute(n, "real")


class _PairExtension(AbstractExtension):
    """Common body of the seven reference extension classes."""
    symbol = None
    nscalars = ()

    def __init__(self, backend, prec, mode=None):
        self.kernel = get_kernel(self.symbol, backend, prec)
        if mode is not None:
            self.kernel.mode = mode
        cty = self.kernel.cty
        k = spec.KERNELS[spec.SYMBOLS[self.symbol]]
        nf = len(k.fields)
        self.fields = k.fields
        self.outs = k.outs
        self.argtypes = ((cty.c_uint,) + (cty.c_real_p,) * nf + (cty.c_uint,)
                         + (cty.c_real_p,) * nf
                         + tuple(getattr(cty, t) for t in self.nscalars))
        self.restypes = (cty.c_real_p,) * len(k.outs)

    def _scalars(self, *args):
        return tuple(args)

    def set_args(self, ips, jps, *args):
        ni = ips.n
        nj = jps.n
        self.kernel.set_gsize(ni, nj)
        self._register(ips, self.outs)
        self._inargs = ((ni,) + tuple(getattr(ips, f) for f in self.fields)
                        + (nj,) + tuple(getattr(jps, f) for f in self.fields)
                        + self._scalars(*args))
        self._outargs = tuple(getattr(ips, o) for o in self.outs)
        self.inargs = prepare_args(self._inargs, self.argtypes)
        self.outargs = prepare_args(self._outargs, self.restypes)
        self.kernel.set_args(self.inargs + self.outargs)


class Phi(_PairExtension):
    """phi_i = -sum_j m_j / s_ij (tupan/lib/extensions.py:100-137)."""
    symbol = "phi_kernel"

    def set_args(self, ips, jps):
        _PairExtension.set_args(self, ips, jps)


class Acc(_PairExtension):
    """a_i = -sum_j m_j r_ij / s_ij^3 (extensions.py:162-203)."""
    symbol = "acc_kernel"

    def set_args(self, ips, jps):
        _PairExtension.set_args(self, ips, jps)


class AccJerk(_PairExtension):
    """Acceleration and its time derivative (extensions.py:236-285)."""
    symbol = "acc_jerk_kernel"

    def set_args(self, ips, jps):
        _PairExtension.set_args(self, ips, jps)


class SnapCrackle(_PairExtension):
    """Second and third time derivatives of the acceleration; needs
    ax..jz on both sides (extensions.py:288-343)."""
    symbol = "snap_crackle_kernel"

    def set_args(self, ips, jps):
        _PairExtension.set_args(self, ips, jps)


class Tstep(_PairExtension):
    """Individual time steps -> ``tstep`` (sum form) and ``tstepij`` (max
    form) (extensions.py:346-387)."""
    symbol = "tstep_kernel"
    nscalars = ("c_real",)

    def set_args(self, ips, jps, eta):
        _PairExtension.set_args(self, ips, jps, eta)


class PNAcc(_PairExtension):
    """Post-Newtonian pair acceleration; order and 1/c^k come from the
    module-level ``clight`` (extensions.py:390-439)."""
    symbol = "pnacc_kernel"
    nscalars = ("c_uint",) + ("c_real",) * 7

    def set_args(self, ips, jps):
        _PairExtension.set_args(self, ips, jps)

    def _scalars(self):
        return (clight.pn_order, clight.inv1, clight.inv2, clight.inv3,
                clight.inv4, clight.inv5, clight.inv6, clight.inv7)


class Sakura(_PairExtension):
    """Pairwise two-body drift of the Sakura integrator
    (extensions.py:442-507)."""
    symbol = "sakura_kernel"
    nscalars = ("c_real", "c_int")

    def set_args(self, ips, jps, dt, flag):
        _PairExtension.set_args(self, ips, jps, dt, flag)


class NREG_X(_PairExtension):
    """NREG position-stage sums at the drifted separation r + v dt:
    ``mrx, mry, mrz`` (sum_j m_j r), ``ax, ay, az`` and ``u``
    (extensions.py:510-561)."""
    symbol = "nreg_Xkernel"
    nscalars = ("c_real",)

    def set_args(self, ips, jps, dt):
        _PairExtension.set_args(self, ips, jps, dt)


class NREG_V(_PairExtension):
    """NREG velocity-stage sums at the kicked velocity v + a dt: ``mvx, mvy,
    mvz`` and ``mk`` (extensions.py:564-617)."""
    symbol = "nreg_Vkernel"
    nscalars = ("c_real",)

    def set_args(self, ips, jps, dt):
        _PairExtension.set_args(self, ips, jps, dt)


class Kepler(AbstractExtension):
    """Two-body Kepler solver: advances particles 0 and 1 of ``ips`` by dt
    in place (extensions.py:620-651)."""

    def __init__(self, backend, prec, mode=None):
        if backend == "CL":       # as the reference: C is fast enough
            backend = "HIP"
        self.kernel = get_kernel("kepler_solver_kernel", backend, prec)
        cty = self.kernel.cty
        self.argtypes = (cty.c_real_p,) * 8 + (cty.c_real,)
        self.restypes = (cty.c_real_p,) * 6

    def set_args(self, ips, jps, dt):
        self._inargs = tuple(getattr(ips, f) for f in spec.KEPLER_FIELDS) + (dt,)
        self._outargs = tuple(getattr(ips, f) for f in spec.KEPLER_OUTS)
        self.inargs = prepare_args(self._inargs, self.argtypes)
        self.outargs = prepare_args(self._outargs, self.restypes)
        self.kernel.set_args(self.inargs + self.outargs)


backend = os.environ.get("TUPAN_BACKEND", "HIP")
_prec = "float32" if "--use_sp" in sys.argv else "float64"

clight = Clight()
phi = Phi(backend, _prec)
acc = Acc(backend, _prec)
acc_jerk = AccJerk(backend, _prec)
snap_crackle = SnapCrackle(backend, _prec)
tstep = Tstep(backend, _prec)
pnacc = PNAcc(backend, _prec)
sakura = Sakura(backend, _prec)
nreg_x = NREG_X(backend, _prec)
nreg_v = NREG_V(backend, _prec)
kepler = Kepler(backend, _prec)
