"""Array-level entry points: one call per kernel evaluation.

``compute(name, ips, jps, *args, mode=...)`` takes host objects (numpy
arrays reachable by field name, as attributes or dict keys) and goes through
the synchronous host-pointer ABI (tupan_<name>_host).  Tensors that live on
the GPU use ``tupan_amd.device.compute`` instead.

Extra scalar arguments per kernel (as in tupan/lib/extensions.py):
  tstep:  eta
  pnacc:  order, (inv1, ..., inv7)
  sakura: dt, flag
  nreg_x, nreg_v: dt
``kepler(ps, dt)`` runs the two-particle Kepler solver.
"""
import numpy as np

from . import _lib, spec


def _field(obj, name):
    return obj[name] if isinstance(obj, dict) else getattr(obj, name)


def compute(name, ips, jps, *args, mode="fp64", out=None):
    k = spec.KERNELS[name]
    iarr = [_lib.as_f64(_field(ips, f)) for f in k.fields]
    jarr = [_lib.as_f64(_field(jps, f)) for f in k.fields]
    ni = len(iarr[0])
    nj = len(jarr[0])
    for a in iarr:
        if len(a) != ni:
            raise ValueError("ragged i-side fields for %s" % name)
    for a in jarr:
        if len(a) != nj:
            raise ValueError("ragged j-side fields for %s" % name)
    if out is None:
        out = tuple(np.empty(ni, dtype=np.float64) for _ in k.outs)
    else:
        out = tuple(out)
        for o in out:
            if o.dtype != np.float64 or not o.flags.c_contiguous or len(o) != ni:
                raise ValueError("outputs must be contiguous float64[ni]")
    # identical host arrays on both sides are passed as the same pointer so
    # that the library stages them once
    iptr = [a.ctypes.data for a in iarr]
    jptr = [a.ctypes.data for a in jarr]
    optr = [o.ctypes.data for o in out]
    _lib.call_host(name, mode, ni, iptr, nj, jptr, optr, args)
    return out


def kepler(ps, dt):
    """kepler_solver_kernel (libtupan.h:231-246) on particles 0 and 1 of
    ``ps``: returns the new (rx, ry, rz, vx, vy, vz), each of length 2."""
    ins = [_lib.as_f64(_field(ps, f))[:2].copy() for f in spec.KEPLER_FIELDS]
    outs = tuple(np.empty(2, dtype=np.float64) for _ in spec.KEPLER_OUTS)
    _lib.call_kepler_host([a.ctypes.data for a in ins], dt,
                          [o.ctypes.data for o in outs])
    return outs
