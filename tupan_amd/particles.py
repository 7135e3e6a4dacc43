"""Minimal SoA particle container with the reference's caller contract.

Mirrors what tupan/integrator and tupan/lib/extensions.py need from
``ParticleSystem`` (tupan/particles/allparticles.py:24-206) and
``NbodyMethods`` (tupan/particles/body.py:20-461):

* per-particle SoA arrays named ``id, mass, eps2, rx, ry, rz, vx, vy, vz,
  time, nstep, tstep`` (body.py:26-39);
* ``register_auxiliary_attribute(name, sctype)`` creating a zeroed array,
  raising if it already exists (allparticles.py:64-76) -- the extension
  classes use it to attach outputs lazily (``phi``, ``ax``.. ``jz``, ...);
* ``set_phi/set_acc/set_acc_jerk/set_snap_crackle/set_tstep/set_pnacc``
  forwarding to the extension singletons (body.py:265-361);
* energies and centre-of-mass helpers used by the IC generator and the
  drivers (body.py:77-321, 400-461), written with the same numpy
  expressions so that initial conditions reproduce the reference's bits.

``eps2`` holds *half* the pair softening: a pair uses eps2_i + eps2_j
(tupan/ics/plummer.py:62-63, tupan/lib/src/phi_kernel_common.h:23).
"""
import copy

import numpy as np

REAL = np.dtype(np.float64)
UINT = np.dtype(np.uint64)
CTYPES = {"real": REAL, "uint": UINT, "int": np.dtype(np.int64)}

ATTRS = (("id", "uint"), ("mass", "real"), ("eps2", "real"),
         ("rx", "real"), ("ry", "real"), ("rz", "real"),
         ("vx", "real"), ("vy", "real"), ("vz", "real"),
         ("time", "real"), ("nstep", "uint"), ("tstep", "real"))


class ParticleSystem(object):
    """SoA container of n bodies (FP64 host arrays)."""

    include_pn_corrections = False
    t_curr = 0.0

    def __init__(self, n=0):
        self.n = int(n)
        for name, sct in ATTRS:
            setattr(self, name, np.zeros(self.n, CTYPES[sct]))

    @classmethod
    def from_arrays(cls, **arrays):
        """Build from named arrays; missing base attributes are zero."""
        n = len(next(iter(arrays.values())))
        ps = cls(n)
        for k, v in arrays.items():
            setattr(ps, k, np.array(v, dtype=getattr(ps, k, np.zeros(0)).dtype
                                    if hasattr(ps, k) else REAL))
        return ps

    def register_auxiliary_attribute(self, attr, sctype):
        if attr in self.__dict__:
            raise ValueError("'{0}' is already a registered "
                             "attribute.".format(attr))
        setattr(self, attr, np.zeros(self.n, CTYPES[sctype]))

    def __len__(self):
        return self.n

    def copy(self):
        return copy.deepcopy(self)

    def __getitem__(self, slc):
        """Slice/mask every per-particle array (copies, as numpy fancy
        indexing does; basic slices are copied too so that results do not
        alias the parent)."""
        out = ParticleSystem.__new__(ParticleSystem)
        n = None
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray) and v.ndim == 1 and len(v) == self.n:
                sub = np.array(v[slc])
                setattr(out, k, sub)
                n = len(sub)
            elif k != "n":
                setattr(out, k, copy.deepcopy(v))
        out.n = 0 if n is None else n
        return out

    # ---- forwarding to the extension singletons (body.py:265-361) ----
    def set_tstep(self, ps, eta):
        from . import extensions
        extensions.tstep.calc(self, ps, eta)

    def set_phi(self, ps):
        from . import extensions
        extensions.phi.calc(self, ps)

    def set_acc(self, ps):
        from . import extensions
        extensions.acc.calc(self, ps)

    def set_pnacc(self, ps):
        from . import extensions
        extensions.pnacc.calc(self, ps)

    def set_acc_jerk(self, ps):
        from . import extensions
        extensions.acc_jerk.calc(self, ps)

    def set_snap_crackle(self, ps):
        from . import extensions
        extensions.snap_crackle.calc(self, ps)

    # ---- diagnostics (body.py:77-321) ----
    @property
    def total_mass(self):
        return float(self.mass.sum())

    @property
    def com_r(self):
        mr = np.array([self.mass * self.rx, self.mass * self.ry,
                       self.mass * self.rz]).T
        return mr.sum(0) / self.total_mass

    @property
    def com_v(self):
        mv = np.array([self.mass * self.vx, self.mass * self.vy,
                       self.mass * self.vz]).T
        return mv.sum(0) / self.total_mass

    def com_move_to(self, com_r, com_v):
        self.rx += com_r[0]
        self.ry += com_r[1]
        self.rz += com_r[2]
        self.vx += com_v[0]
        self.vy += com_v[1]
        self.vz += com_v[2]

    def com_to_origin(self):
        self.com_move_to(-self.com_r, -self.com_v)

    @property
    def linear_momentum(self):
        return np.array([self.mass * self.vx, self.mass * self.vy,
                         self.mass * self.vz]).T.sum(0)

    @property
    def angular_momentum(self):
        px, py, pz = self.mass * self.vx, self.mass * self.vy, self.mass * self.vz
        am = np.array([self.ry * pz - self.rz * py,
                       self.rz * px - self.rx * pz,
                       self.rx * py - self.ry * px]).T
        return am.sum(0)

    @property
    def ke(self):
        return 0.5 * self.mass * (self.vx**2 + self.vy**2 + self.vz**2)

    @property
    def kinetic_energy(self):
        return float(self.ke.sum())

    @property
    def pe(self):
        self.set_phi(self)
        return self.mass * self.phi

    @property
    def potential_energy(self):
        return 0.5 * float(self.pe.sum())

    @property
    def virial_radius(self):
        mtot = self.total_mass
        pe = self.potential_energy
        return (mtot**2) / (-2 * pe)

    def dynrescale_total_mass(self, total_mass):
        m_ratio = total_mass / self.total_mass
        self.mass *= m_ratio
        self.rx *= m_ratio
        self.ry *= m_ratio
        self.rz *= m_ratio

    def dynrescale_virial_radius(self, rvir):
        r_scale = rvir / self.virial_radius
        v_scale = 1 / r_scale**0.5
        self.rx *= r_scale
        self.ry *= r_scale
        self.rz *= r_scale
        self.vx *= v_scale
        self.vy *= v_scale
        self.vz *= v_scale

    def scale_to_virial(self):
        ke = self.kinetic_energy
        pe = self.potential_energy
        v_scale = ((-0.5 * pe) / ke)**0.5
        self.vx *= v_scale
        self.vy *= v_scale
        self.vz *= v_scale

    def to_nbody_units(self):
        self.dynrescale_total_mass(1.0)
        self.dynrescale_virial_radius(1.0)

    def min_tstep(self):
        return abs(self.tstep).min()

    def max_tstep(self):
        return abs(self.tstep).max()
