"""CPU checker for the Tupan direct-summation kernels -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker; the product package
(``tupan_amd``) never imports it and has no CPU fallback.

Two kinds of library sit behind one call interface:

* ``kind="port"``: ``liboracle_f{64,32}.so`` built from ``oracle_kernels.c``,
  our plain-C restatement of the reference kernels (bit-identical to the
  reference C in both precisions; pinned by tests/test_oracle.py);
* ``kind="reference"``: ``_ref/libtupan_ref_f{64,32}.so``, the reference's own
  C kernels (tupan/lib/src/*.c) compiled from source by ``make ref``; present
  only where the reference checkout was available at build time.

Argument lists follow tupan/lib/src/libtupan.h:2-246; the i/j field order of
each kernel follows tupan/lib/extensions.py (Phi :126-131, Acc :192-197,
AccJerk :273-278, SnapCrackle :329-336, Tstep :375-381, PNAcc :424-433,
Sakura :494-500, NREG_X :551-558, NREG_V :602-608, Kepler :642-646).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

F5 = ("mass", "rx", "ry", "rz", "eps2")
F8 = F5 + ("vx", "vy", "vz")
F14 = F8 + ("ax", "ay", "az", "jx", "jy", "jz")

# name -> (i/j fields, scalar arg kinds, output names)
SPECS = {
    "phi_kernel": (F5, (), ("phi",)),
    "acc_kernel": (F5, (), ("ax", "ay", "az")),
    "acc_jerk_kernel": (F8, (), ("ax", "ay", "az", "jx", "jy", "jz")),
    "snap_crackle_kernel": (F14, (), ("sx", "sy", "sz", "cx", "cy", "cz")),
    "tstep_kernel": (F8, ("real",), ("tstep", "tstepij")),
    "pnacc_kernel": (F8, ("uint",) + ("real",) * 7, ("pnax", "pnay", "pnaz")),
    "sakura_kernel": (F8, ("real", "int"),
                      ("drx", "dry", "drz", "dvx", "dvy", "dvz")),
    "nreg_Xkernel": (F8, ("real",),
                     ("mrx", "mry", "mrz", "ax", "ay", "az", "u")),
    "nreg_Vkernel": (("mass", "vx", "vy", "vz", "ax", "ay", "az"), ("real",),
                     ("mvx", "mvy", "mvz", "mk")),
}

# kepler_solver_kernel (libtupan.h:231-246) has no ni/nj: it advances
# particles 0 and 1 of the given arrays.
KEPLER = (F8, ("rx", "ry", "rz", "vx", "vy", "vz"))


def lib_path(prec="float64", kind="port"):
    bits = "64" if prec == "float64" else "32"
    if kind == "port":
        return os.path.join(HERE, "liboracle_f%s.so" % bits)
    if kind == "reference":
        return os.path.join(HERE, "_ref", "libtupan_ref_f%s.so" % bits)
    raise ValueError("kind must be 'port' or 'reference'")


def available(prec="float64", kind="port"):
    return os.path.exists(lib_path(prec, kind))


def _get(obj, name):
    return obj[name] if isinstance(obj, dict) else getattr(obj, name)


class Oracle(object):
    """ctypes front end: ``Oracle(prec, kind).call(name, ips, jps, *args)``.

    ``ips``/``jps`` are dicts or objects exposing the SoA fields by name.
    Returns a tuple of fresh output arrays (ni each), in libtupan.h order.
    """

    def __init__(self, prec="float64", kind="port"):
        path = lib_path(prec, kind)
        if not os.path.exists(path):
            raise RuntimeError("checker library missing: %s (run make -C oracle%s)"
                               % (path, " ref" if kind == "reference" else ""))
        self.prec = prec
        self.kind = kind
        self.lib = ctypes.CDLL(path)
        self.real = np.float64 if prec == "float64" else np.float32
        self.c_real = ctypes.c_double if prec == "float64" else ctypes.c_float
        self.c_uint = ctypes.c_ulong if prec == "float64" else ctypes.c_uint
        self.c_int = ctypes.c_long if prec == "float64" else ctypes.c_int
        self.prefix = "oracle_" if kind == "port" else ""

    def _arr(self, a):
        return np.ascontiguousarray(a, dtype=self.real)

    def call(self, name, ips, jps, *scalars):
        fields, kinds, outs = SPECS[name]
        if len(scalars) != len(kinds):
            raise TypeError("%s expects %d scalar args" % (name, len(kinds)))
        iarrs = [self._arr(_get(ips, f)) for f in fields]
        jarrs = [self._arr(_get(jps, f)) for f in fields]
        ni = len(iarrs[0])
        nj = len(jarrs[0])
        res = [np.zeros(ni, dtype=self.real) for _ in outs]
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        args = [self.c_uint(ni)] + [ptr(a) for a in iarrs]
        args += [self.c_uint(nj)] + [ptr(a) for a in jarrs]
        for k, v in zip(kinds, scalars):
            if k == "real":
                args.append(self.c_real(v))
            elif k == "uint":
                args.append(self.c_uint(int(v)))
            else:
                args.append(self.c_int(int(v)))
        args += [ptr(a) for a in res]
        fn = getattr(self.lib, self.prefix + name)
        fn.restype = None
        fn(*args)
        return tuple(res)

    def kepler(self, ps, dt):
        """kepler_solver_kernel on particles 0 and 1 of ``ps``: returns the
        new (rx, ry, rz, vx, vy, vz) arrays of length 2."""
        fields, outs = KEPLER
        arrs = [self._arr(_get(ps, f))[:2].copy() for f in fields]
        res = [np.zeros(2, dtype=self.real) for _ in outs]
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        fn = getattr(self.lib, self.prefix + "kepler_solver_kernel")
        fn.restype = None
        fn(*([ptr(a) for a in arrs] + [self.c_real(dt)]
             + [ptr(a) for a in res]))
        return tuple(res)

    def twobody_evolve(self, dt, flag, m, e2, state):
        """Single-pair Sakura evolution (port only): returns the new (r, v)."""
        if self.kind != "port":
            raise RuntimeError("twobody_evolve is exported by the port only")
        st = np.ascontiguousarray(state, dtype=self.real).copy()
        fn = self.lib.oracle_twobody_evolve
        fn.restype = None
        fn(self.c_real(dt), self.c_int(int(flag)), self.c_real(m),
           self.c_real(e2), st.ctypes.data_as(ctypes.c_void_p))
        return st
