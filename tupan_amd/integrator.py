"""Device-resident drivers of Tupan's integrators on the MI355X kernels.

The reference's integrators (tupan/integrator/) drive the pair kernels
through ``ps.set_*`` and update numpy arrays between calls.  Here the
particle set lives in HBM (``System``: one contiguous float64 tensor per
field) and both halves run on the device: the pair kernels through the
``tupan_<kernel>_dev`` ABI and the updates through the fused update
kernels of integrate.hip, which evaluate the reference's numpy expressions
in the reference's order (no FMA contraction).  Only the block-step choice
(one min over particles) comes back to the host, as in the reference.

Provided (the reference's method names):
  Hermite  hermite2/4/6/8, ahermite2/4/6/8   (tupan/integrator/hermite.py)
  SIA      sia21s.kdk/.dkd, sia21a.kdk/.dkd  (tupan/integrator/sia.py; shared
           steps, with PN corrections when pn_order > 0)
  Sakura   sakura, asakura                    (tupan/integrator/sakura.py)
  SakuraPN sakura_pn: close pairs by the Sakura Kepler drift, the rest and
           the PN terms by kicks (a hybrid the reference does not have)
  NREG     nreg, anreg                        (tupan/integrator/nreg.py)
  FewBody  drift / two-body Kepler solve      (tupan/integrator/fewbody.py)

Differences from the reference, all deliberate:
  * the state is not copied per step (``ps.copy()``, hermite.py:79): the
    predictor writes the snapshot into a preallocated shadow set;
  * hierarchical SIA (``sia21h``, which hands subsystems of n <= 2 to
    FewBody, sia.py:32) is not provided: it raises;
  * a Sakura pair that exhausts its Kepler-solve budget raises
    ``NotConverged`` at the end of ``evolve`` (the reference loops).
"""
from __future__ import division

import math
import sys

import numpy as np
import torch

from . import _lib, device, spec
from .extensions import Clight

BASE = ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz")
DERIV = ("ax", "ay", "az", "jx", "jy", "jz", "sx", "sy", "sz",
         "cx", "cy", "cz")
PN_FIELDS = ("wx", "wy", "wz", "pn_ke", "pn_mrx", "pn_mry", "pn_mrz",
             "pn_mvx", "pn_mvy", "pn_mvz", "pn_amx", "pn_amy", "pn_amz")


class _Ptrs(tuple):
    """A tuple of device pointers with its ctypes array attached."""


class System(object):
    """SoA particle set in HBM: ``ps.<field>`` is a contiguous float64
    device tensor of length n.  Fields beyond the base set are created on
    first use, zero-filled (the reference's register_auxiliary_attribute,
    tupan/particles/allparticles.py:64-76)."""

    def __init__(self, n, device="cuda"):
        object.__setattr__(self, "_f", {})
        self._ptrs = {}
        self.n = int(n)
        self.device = torch.device(device)
        self.t_curr = 0.0
        self.include_pn_corrections = False
        self.mode = "fp64"
        for f in BASE + ("time", "tstep"):
            self.field(f)
        self._f["nstep"] = torch.zeros(self.n, dtype=torch.int64,
                                       device=self.device)

    @classmethod
    def from_host(cls, ps, device="cuda"):
        """Copy a host particle set (attributes or dict of arrays)."""
        get = (lambda k: ps[k]) if isinstance(ps, dict) else (
            lambda k: getattr(ps, k))
        n = len(get("mass"))
        s = cls(n, device)
        for f in BASE:
            s._f[f].copy_(torch.as_tensor(np.ascontiguousarray(get(f), np.float64)))
        return s

    def field(self, name):
        t = self._f.get(name)
        if t is None:
            t = torch.zeros(self.n, dtype=torch.float64, device=self.device)
            self._f[name] = t
        return t

    def __getattr__(self, name):
        f = object.__getattribute__(self, "_f")
        if name in f:
            return f[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in self._f:
            raise AttributeError("field %r is device memory: update it in "
                                 "place" % name)
        object.__setattr__(self, name, value)

    def to_host(self, fields=None):
        """dict of numpy copies of the given fields (default: base set)."""
        fields = BASE if fields is None else fields
        return {f: self._f[f].cpu().numpy().copy() for f in fields}

    # ---- pair kernels (body.py:324-361) --------------------------------
    def _pair(self, name, *args, mode=None, jps=None, status=None):
        k = spec.KERNELS[name]
        out = tuple(self.field(o) for o in k.outs)
        if jps is not None:
            device.compute(name, self, jps, *args, mode=mode or self.mode,
                           out=out, status=status)
            return out
        # the set with itself: its own fields are contiguous float64 device
        # tensors of length n by construction, so the checks of
        # device.compute are skipped and the cached pointer arrays used
        sptr = None
        if status is not None:
            if status.dtype != torch.int32 or not status.is_cuda:
                raise ValueError("status must be an int32 device tensor")
            sptr = status.data_ptr()
        f = self.ptrs(k.fields)
        _lib.call_dev(name, mode or self.mode, self.n, f, self.n, f,
                      self.ptrs(k.outs), args, stream=_stream(self),
                      status_ptr=sptr)
        return out

    def set_phi(self, mode=None):
        return self._pair("phi", mode=mode)

    def set_acc(self, mode=None):
        return self._pair("acc", mode=mode)

    def set_acc_jerk(self, mode=None):
        return self._pair("acc_jerk", mode=mode)

    def set_snap_crackle(self, mode=None):
        return self._pair("snap_crackle", mode=mode)

    def set_tstep(self, eta):
        return self._pair("tstep", eta)

    def set_pnacc(self, clight):
        inv = (clight.inv1, clight.inv2, clight.inv3, clight.inv4,
               clight.inv5, clight.inv6, clight.inv7)
        return self._pair("pnacc", clight.pn_order, inv)

    def set_sakura(self, dt, flag, status=None):
        return self._pair("sakura", dt, flag, status=status)

    # ---- sums and diagnostics (body.py:79-322) ---------------------------
    def sum(self, name):
        """Deterministic device sum of one field (``ps.<name>.sum()``)."""
        out = torch.empty(1, dtype=torch.float64, device=self.device)
        _lib.call_update("sum", self.n, self.field(name).data_ptr(),
                         out.data_ptr(), _stream(self))
        return float(out.item())

    def total_mass(self):
        return self.sum("mass")

    def sum_np(self, name):
        """Device sum of one field in numpy's order: the bits of the
        reference drivers' ``ps.<name>.sum()``."""
        out = torch.empty(1, dtype=torch.float64, device=self.device)
        _lib.call_update("sum_pairwise", self.n, self.field(name).data_ptr(),
                         out.data_ptr(), _stream(self))
        return float(out.item())

    def energies_np(self):
        """(kinetic, potential) energy as the reference's ``ps.kinetic_energy``
        and ``ps.potential_energy`` compute them (body.py:261-306, no PN
        terms): per-particle terms summed in numpy's order, Phi in the
        set's mode."""
        self.set_phi()
        _lib.call_update("energy_terms", self.n,
                         self.ptrs(("mass", "vx", "vy", "vz", "phi")),
                         self.field("_ke").data_ptr(),
                         self.field("_mphi").data_ptr(), _stream(self))
        return self.sum_np("_ke"), 0.5 * self.sum_np("_mphi")

    def diagnostics(self, with_phi=True, np_order=False):
        """The totals of the reference's report (body.py:79-322: total
        mass, kinetic, potential and virial energy, centre-of-mass position
        and velocity, linear and angular momentum, PN accumulators included
        when enabled); ``with_phi`` first evaluates Phi (``ps.set_phi(ps)``,
        body.py:298; FP64, or the set's mode when that is fp64_ieee).  By
        default one fused device pass; ``np_order`` sums each per-particle
        quantity in numpy's order instead, so that with bit-identical
        inputs the totals are the reference's bits."""
        if with_phi:
            self.set_phi(mode="fp64_ieee" if self.mode == "fp64_ieee"
                         else "fp64")
        names = ("mass", "rx", "ry", "rz", "vx", "vy", "vz", "phi")
        pn = bool(self.include_pn_corrections)
        if pn:
            names += ("pn_ke", "pn_mrx", "pn_mry", "pn_mrz", "pn_mvx",
                      "pn_mvy", "pn_mvz", "pn_amx", "pn_amy", "pn_amz")
        if np_order:
            terms = ["_q%d" % k for k in range(12)]
            _lib.call_update("report_terms", self.n, self.ptrs(names),
                             int(pn), self.ptrs(terms), _stream(self))
            s = np.array([self.sum_np(t) for t in terms])
        else:
            out = torch.empty(12, dtype=torch.float64, device=self.device)
            _lib.call_update("diagnostics", self.n, self.ptrs(names),
                             int(pn), out.data_ptr(), _stream(self))
            s = out.cpu().numpy()
        mtot = float(s[0])
        ke = float(s[1])
        pe = 0.5 * float(s[2])
        return {
            "total_mass": mtot,
            "kinetic_energy": ke,
            "potential_energy": pe,
            "virial_energy": 2 * ke + pe,
            "com_r": s[3:6] / mtot,
            "com_v": s[6:9] / mtot,
            "linear_momentum": s[6:9].copy(),
            "angular_momentum": s[9:12].copy(),
        }

    def kinetic_energy(self):
        return self.diagnostics(with_phi=False)["kinetic_energy"]

    def potential_energy(self):
        return self.diagnostics()["potential_energy"]

    def energy(self):
        d = self.diagnostics()
        return d["kinetic_energy"] + d["potential_energy"]

    def min_tstep(self):
        return float(self.tstep.abs().min())

    def ptrs(self, names):
        """The device pointers of the named fields, as a tuple that also
        carries its ctypes array (``carr``).  Fields are created once and
        never replaced, so the result is cached per name tuple."""
        key = tuple(names)
        p = self._ptrs.get(key)
        if p is None:
            p = _Ptrs(self.field(f).data_ptr() for f in key)
            p.carr = _lib.ptr_array(list(p))
            self._ptrs[key] = p
        return p


def _stream(ps):
    return device.current_stream_ptr(ps.device)


class Base(object):
    """Time-step selection and the step loop (tupan/integrator/__init__.py:
    14-99)."""

    PROVIDED_METHODS = []

    def __init__(self, eta, time, ps, method, pn_order=0, clight=None,
                 mode="fp64", graph=False):
        if method not in self.PROVIDED_METHODS:
            raise ValueError("Unexpected method: {0}".format(method))
        self.graph = bool(graph)
        self._graphs = {}
        self._seen = set()
        self._gstream = None
        self.eta = eta
        self.method = method
        ps.t_curr = time
        ps.mode = mode
        self.ps = ps
        self.clight = Clight()
        ps.include_pn_corrections = False
        if pn_order > 0:
            if clight is None:
                raise TypeError("'clight' is not defined. Please set the "
                                "speed of light argument 'clight' when using "
                                "'pn_order' > 0.")
            self.clight.pn_order = pn_order
            self.clight.clight = clight
            ps.include_pn_corrections = True
        self.status = torch.zeros(1, dtype=torch.int32, device=ps.device)

    @property
    def time(self):
        return self.ps.t_curr

    def get_base_tstep(self, t_end):
        t_curr = self.ps.t_curr
        dt = min(abs(t_end) - abs(t_curr), abs(self.eta))
        dt = max(dt, abs(t_end) * (2 * sys.float_info.epsilon))
        return math.copysign(dt, self.eta)

    def get_min_block_tstep(self, ps, tau):
        min_ts = ps.min_tstep()
        power = int(np.log2(min_ts) - 1)
        min_bts = 2.0**power
        t_next = ps.t_curr + min_bts
        while t_next % min_bts != 0:
            min_bts /= 2
        min_bts = math.copysign(min_bts, tau)
        if abs(min_bts) > abs(tau):
            min_bts = tau
        return min_bts

    def evolve_step(self, t_end):
        tau = self.get_base_tstep(t_end)
        if self.graph:
            # the step length is a host decision (adaptive methods read the
            # smallest time step); the step of that length is a graph
            tau = self.decide_tstep(self.ps, tau)
            self.ps = self._graph_step(self.ps, tau)
        else:
            self.ps = self.do_step(self.ps, tau)

    # A step is decide_tstep (host side: the block step of the adaptive
    # methods, tau itself otherwise) followed by fixed_step (device work
    # only, replayable).
    def decide_tstep(self, ps, tau):
        return tau

    def fixed_step(self, ps, tau):
        raise NotImplementedError

    def do_step(self, ps, tau):
        return self.fixed_step(ps, self.decide_tstep(ps, tau))

    def _graph_step(self, ps, tau):
        """One step of length ``tau`` as a HIP graph replay.  The first step
        of a given length runs eagerly on the capture stream, so that every
        field and library scratch buffer it uses exists there; the second
        captures the step (recorded, not executed) and replays it, and later
        ones replay.  The host clock advances outside the graph.  Scalars
        (tau, eta, clight) are recorded into the graph: a step length is
        captured once and replayed as recorded."""
        g = self._graphs.get(tau)
        if g is None:
            cur = torch.cuda.current_stream(ps.device)
            if self._gstream is None:
                self._gstream = torch.cuda.Stream(device=ps.device)
            self._gstream.wait_stream(cur)
            if tau not in self._seen:
                with torch.cuda.stream(self._gstream):
                    ps = self._eager_graph_step(ps, tau)
                cur.wait_stream(self._gstream)
                self._seen.add(tau)
                return ps
            t0 = ps.t_curr
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=self._gstream):
                self._capture_body(ps, tau)
            ps.t_curr = t0
            self._graphs[tau] = g
        g.replay()
        self._after_replay(ps, tau)
        return ps

    # the pieces of a graph-replayed step; NREG overrides them
    def _eager_graph_step(self, ps, tau):
        return self.fixed_step(ps, tau)

    def _capture_body(self, ps, tau):
        self.fixed_step(ps, tau)

    def _after_replay(self, ps, tau):
        ps.t_curr += tau

    def evolve(self, t_end, callback=None):
        """Step until t_end (tupan/simulation.py:190-193); ``callback(ps)``
        after every step."""
        while abs(self.ps.t_curr) < t_end:
            self.evolve_step(t_end)
            if callback is not None:
                callback(self.ps)
        self.check_status()

    def check_status(self):
        if int(self.status.item()) != 0:
            raise _lib.NotConverged(_lib.ERR_NOT_CONVERGED,
                                    "a Sakura pair exhausted its Kepler-solve "
                                    "budget")

    def _advance_clock(self, ps, tau):
        """ps.tstep = tau, ps.time += tau, ps.nstep += 1: one launch."""
        ps.t_curr += tau
        _lib.call_update("advance_clock", ps.n, ps.tstep.data_ptr(),
                         ps.time.data_ptr(), ps.nstep.data_ptr(), float(tau),
                         _stream(ps))


class Hermite(Base):
    """P(EC)^2 Hermite schemes of order 2, 4, 6, 8 (hermite.py:287-410)."""

    PROVIDED_METHODS = ["hermite2", "ahermite2", "hermite4", "ahermite4",
                        "hermite6", "ahermite6", "hermite8", "ahermite8"]

    def __init__(self, eta, time, ps, method, **kw):
        super(Hermite, self).__init__(eta, time, ps, method, **kw)
        self.order = int(method[-1])
        nd = 6 + 3 * (self.order // 2)
        self.names = ("rx", "ry", "rz", "vx", "vy", "vz") + DERIV[:nd - 6]
        self.shadow = {f: torch.empty(ps.n, dtype=torch.float64,
                                      device=ps.device) for f in self.names}
        self._x0 = _lib.ptr_array([self.shadow[f].data_ptr()
                                   for f in self.names])

    def _derivs(self, ps):
        if self.order == 2:
            ps.set_acc()
        else:
            ps.set_acc_jerk()
            if self.order >= 6:
                ps.set_snap_crackle()

    def get_hermite_tstep(self, ps, eta, tau):
        ps.set_tstep(eta)
        return self.get_min_block_tstep(ps, tau)

    def epec(self, n, ps, tau):
        # ps0 = ps.copy(); ps0.set_derivs(ps0): the derivatives of the
        # current state are evaluated in place and the predictor copies the
        # whole snapshot to the shadow set before advancing ps.
        self._derivs(ps)
        x = ps.ptrs(self.names)
        x0 = self._x0
        st = _stream(ps)
        _lib.call_update("hermite_predict", self.order, ps.n, x, x0,
                         float(tau), st)
        for _ in range(n):
            self._derivs(ps)
            _lib.call_update("hermite_correct", self.order, ps.n, x, x0,
                             float(tau), st)
        return ps

    def decide_tstep(self, ps, tau):
        if "ahermite" in self.method:
            tau = self.get_hermite_tstep(ps, self.eta, tau)
        return tau

    def fixed_step(self, ps, tau):
        ps = self.epec(2, ps, tau)
        self._advance_clock(ps, tau)
        return ps


class SIA(Base):
    """Second-order symplectic SIA21 with shared time steps
    (sia.py:291-353, 964-1119)."""

    PROVIDED_METHODS = ["sia21s.dkd", "sia21a.dkd", "sia21s.kdk",
                        "sia21a.kdk"]

    def __init__(self, eta, time, ps, method, **kw):
        super(SIA, self).__init__(eta, time, ps, method, **kw)
        if ps.include_pn_corrections:
            for f in PN_FIELDS:
                ps.field(f)

    def drift(self, ps, tau):
        pn = ps.include_pn_corrections
        _lib.call_update(
            "drift", ps.n, ps.ptrs(("rx", "ry", "rz")),
            ps.ptrs(("vx", "vy", "vz")), float(tau), 0,
            ps.ptrs(("pn_mrx", "pn_mry", "pn_mrz")) if pn else None,
            ps.ptrs(("pn_mvx", "pn_mvy", "pn_mvz")) if pn else None,
            _stream(ps))

    def newton_acc(self, ps):
        ps.set_acc()

    def kick(self, ps, tau):
        """kick_n, or kick_pn (sia.py:142-156): half kick with w, PN
        accelerations at the half-kicked velocities, the flux bookkeeping
        and the second half kick (the last two fused in one launch)."""
        v = ps.ptrs(("vx", "vy", "vz"))
        a = ps.ptrs(("ax", "ay", "az"))
        st = _stream(ps)
        if not ps.include_pn_corrections:
            self.newton_acc(ps)
            _lib.call_update("kick", ps.n, v, a, float(tau), st)
            return
        w = ps.ptrs(("wx", "wy", "wz"))
        self.newton_acc_kick_half(ps, v, a, w, tau, st)
        ps.set_pnacc(self.clight)
        _lib.call_update(
            "pn_flux_kick", ps.n,
            ps.ptrs(("mass", "rx", "ry", "rz", "vx", "vy", "vz",
                     "pnax", "pnay", "pnaz")),
            ps.ptrs(("wx", "wy", "wz", "pn_ke", "pn_mvx", "pn_mvy", "pn_mvz",
                     "pn_amx", "pn_amy", "pn_amz")), a, float(tau), st)

    def newton_acc_kick_half(self, ps, v, a, w, tau, st):
        """Newtonian accelerations, then kick_pn's first half kick."""
        self.newton_acc(ps)
        _lib.call_update("kick_pn_half", ps.n, v, a, w, float(tau), st)

    def kdk(self, ps, tau):
        self.kick(ps, 0.5 * tau)
        self.drift(ps, 1.0 * tau)
        self.kick(ps, 0.5 * tau)

    def dkd(self, ps, tau):
        self.drift(ps, 0.5 * tau)
        self.kick(ps, 1.0 * tau)
        self.drift(ps, 0.5 * tau)

    def decide_tstep(self, ps, tau):
        if ps.n <= 2:
            raise NotImplementedError("n <= 2 needs the few-body solver "
                                      "(tupan/integrator/fewbody.py)")
        if "a." in self.method:
            ps.set_tstep(self.eta)
            tau = self.get_min_block_tstep(ps, tau)
        return tau

    def fixed_step(self, ps, tau):
        # shared steps: every particle is slow, and SIA21.bridge_sf evolves
        # the slow set twice by d[0]*tau = tau/2 (sia.py:341-353)
        ev = self.kdk if "kdk" in self.method else self.dkd
        ev(ps, 0.5 * tau)
        ev(ps, 0.5 * tau)
        self._advance_clock(ps, tau)
        return ps


class Sakura(Base):
    """Sakura pairwise Kepler splitting (sakura.py:22-162)."""

    PROVIDED_METHODS = ["sakura", "asakura"]

    def sakura_step(self, ps, tau):
        r = ps.ptrs(("rx", "ry", "rz"))
        v = ps.ptrs(("vx", "vy", "vz"))
        x = r + v
        d = ps.ptrs(("drx", "dry", "drz", "dvx", "dvy", "dvz"))
        st = _stream(ps)
        _lib.call_update("drift", ps.n, r, v, float(tau), 1, None, None, st)
        ps.set_sakura(tau / 2, -1, status=self.status)
        _lib.call_update("apply_delta", ps.n, x, d, st)
        ps.set_sakura(tau / 2, 1, status=self.status)
        _lib.call_update("apply_delta", ps.n, x, d, st)
        _lib.call_update("drift", ps.n, r, v, float(tau), 1, None, None, st)
        return ps

    def get_sakura_tstep(self, ps, eta, tau):
        ps.set_tstep(eta)
        # torch's scalar / tensor multiplies by a reciprocal; divide
        # elementwise as numpy does
        e = torch.full_like(ps.tstep, eta)
        qa = torch.div(e, ps.tstep)
        qb = torch.div(e, ps.tstepij)
        diw2 = qa * qa - qb * qb
        w2_sakura = np.float64(float(diw2.max()))
        dt_sakura = eta / (1 + w2_sakura)**0.5
        ps.tstep.fill_(float(dt_sakura))
        return self.get_min_block_tstep(ps, tau)

    def decide_tstep(self, ps, tau):
        if "asakura" in self.method:
            tau = self.get_sakura_tstep(ps, self.eta, tau)
        return tau

    def fixed_step(self, ps, tau):
        ps = self.sakura_step(ps, tau)
        self._advance_clock(ps, tau)
        return ps


class SakuraPN(SIA):
    """Hybrid scheme for compact PN clusters (no reference counterpart: the
    reference's Sakura acts on every pair and has no PN terms).  Pairs closer
    than ``rcut`` are drifted by the Sakura two-body Kepler solve; the other
    pairs are kicked together with the PN terms, with kick_pn's w scheme and
    flux bookkeeping (sia.py:142-156):

        K(tau/2) [far Newton + PN]  D(tau) [free drift + close-pair Kepler]
        K(tau/2)

    The pairs are split once per step, by their separation at the free-drift
    midpoint x + v tau / 2 predicted from the step's start (the
    classification positions cx, cy, cz, set first in the step): both kicks
    and both Sakura passes of the drift then treat every pair the same way,
    so that a pair crossing rcut during a step is neither kicked and
    Kepler-drifted nor left out for part of it, and the midpoint makes the
    split the same whether the step is taken forwards or backwards (for
    free motion).  On the compact PN cluster (N = 2048, dt = 1/8192,
    t = 0.1) the energy error goes from 9.0e-4, splitting by the current
    separation at every kick and drift, to 3.5e-5.  Shared fixed steps
    tau = min(t_end - t, eta)."""

    PROVIDED_METHODS = ["sakura_pn"]

    def __init__(self, eta, time, ps, method, rcut=1e-3, **kw):
        super(SIA, self).__init__(eta, time, ps, method, **kw)
        self.rcut = float(rcut)
        for f in PN_FIELDS + spec.CP:
            ps.field(f)

    def classify(self, ps, tau):
        """The step's classification positions c = r + (v tau) / 2 (the
        drift kernel's half drift)."""
        _lib.call_update("half_drift_to", ps.n, ps.ptrs(("rx", "ry", "rz")),
                         ps.ptrs(("vx", "vy", "vz")), float(tau),
                         ps.ptrs(spec.CP), _stream(ps))

    def newton_acc(self, ps):
        # the far pairs only (r >= rcut), in FP64 whatever the set's mode
        ps._pair("acc_far", self.rcut, mode="fp64")

    def newton_acc_kick_half(self, ps, v, a, w, tau, st):
        # the far-pair sums and the first half kick in the same launches
        # (bit-identical to acc_far followed by kick_pn_half)
        _lib.call_update("acc_far_kick_pn_half", ps.n,
                         ps.ptrs(spec.F5 + spec.CP),
                         float(self.rcut), a, v, w, float(tau), st)

    def drift(self, ps, tau):
        r = ps.ptrs(("rx", "ry", "rz"))
        v = ps.ptrs(("vx", "vy", "vz"))
        x = r + v
        d = ps.ptrs(("drx", "dry", "drz", "dvx", "dvy", "dvz"))
        st = _stream(ps)
        _lib.call_update("drift", ps.n, r, v, float(tau), 1, None, None, st)
        for flag in (-1, 1):
            ps._pair("sakura_near", tau / 2, flag, self.rcut,
                     status=self.status)
            _lib.call_update("apply_delta", ps.n, x, d, st)
        # the second half drift, and with PN terms pn_mr += pn_mv * tau
        # (the whole tau) in the same launch
        mr = mv = None
        if ps.include_pn_corrections:
            mr = ps.ptrs(("pn_mrx", "pn_mry", "pn_mrz"))
            mv = ps.ptrs(("pn_mvx", "pn_mvy", "pn_mvz"))
        _lib.call_update("drift", ps.n, r, v, float(tau), 1, mr, mv, st)

    def decide_tstep(self, ps, tau):
        return tau

    def fixed_step(self, ps, tau):
        self.classify(ps, tau)
        self.kick(ps, 0.5 * tau)
        self.drift(ps, tau)
        self.kick(ps, 0.5 * tau)
        self._advance_clock(ps, tau)
        return ps


class NREG(Base):
    """Algorithmic-regularisation leapfrog (tupan/integrator/nreg.py:22-172):
    drift(h/2W) kick(h/U) drift(h/2W) with the pair sums of the NREG kernels
    at the drifted separations and kicked velocities, positions and
    velocities recentred on the centre of mass after each sub-step.  The
    scalars the reference keeps on the particle type (E0, W, U, S) live on
    the integrator; each sub-step's length needs the last one's sums, which
    come back to the host as in the reference.

    With ``graph=True`` they live in a device scalar block instead
    (TUPAN_NREG_* slots): the sub-step lengths are computed by a one-thread
    kernel in the same double arithmetic and read by the pair kernels from
    device memory, so a step makes no host round trip and is replayed as a
    HIP graph; the host reads the physical time once per step.  The same
    bits as the host-driven step (tested).  The host attributes E0, W, U, S
    then keep their initial values."""

    PROVIDED_METHODS = ["nreg", "anreg"]

    def __init__(self, eta, time, ps, method, **kw):
        super(NREG, self).__init__(eta, time, ps, method, **kw)
        if ps.include_pn_corrections:
            raise NotImplementedError("NREG has no post-Newtonian terms")
        self.is_initialized = False

    def initialize(self):
        """nreg.py:114-126: E0 = KE + PE, W = U = S = -PE (sums in numpy's
        order, as every sum of this driver: with the pair kernels in the
        reference's arithmetic a run reproduces the reference's bits)."""
        ke, pe = self.ps.energies_np()
        self.E0 = ke + pe
        self.W = self.U = self.S = -pe
        self.is_initialized = True
        if self.graph:
            ps = self.ps
            init = [0.0] * _lib.NREG_SCALARS
            init[_lib.NREG_E0] = self.E0
            init[_lib.NREG_W] = init[_lib.NREG_U] = init[_lib.NREG_S] = -pe
            init[_lib.NREG_T] = init[_lib.NREG_T0] = float(ps.t_curr)
            self._sc = torch.tensor(init, dtype=torch.float64).to(ps.device)

    def evolve_step(self, t_end):
        if not self.is_initialized:
            self.initialize()
        super(NREG, self).evolve_step(t_end)

    def nreg_x(self, ps, dt):
        """nreg.py:22-33 (the total mass is taken once per step: masses do
        not change within one)."""
        mtot = self._mtot
        ps._pair("nreg_x", dt)
        _lib.call_update("divide", ps.n, 3, ps.ptrs(("mrx", "mry", "mrz")),
                         ps.ptrs(("rx", "ry", "rz")), float(mtot), _stream(ps))
        self.U = 0.5 * ps.sum_np("u")
        ps.t_curr += dt

    def nreg_v(self, ps, dt):
        """nreg.py:45-59."""
        mtot = self._mtot
        ps._pair("nreg_v", dt)
        _lib.call_update("divide", ps.n, 3, ps.ptrs(("mvx", "mvy", "mvz")),
                         ps.ptrs(("vx", "vy", "vz")), float(mtot), _stream(ps))
        kin = 0.25 * ps.sum_np("mk") / mtot
        self.W = kin - self.E0

    def anreg_step(self, ps, h):
        """nreg.py:77-85."""
        self.nreg_x(ps, 0.5 * (h / self.W))
        self.nreg_v(ps, (h / self.U))
        self.nreg_x(ps, 0.5 * (h / self.W))

    def nreg_step(self, ps, h):
        """nreg.py:88-97."""
        self.anreg_step(ps, 0.5 * (h * self.S))
        self.S = 1 / (2 / self.W - 1 / self.S)
        self.anreg_step(ps, 0.5 * (h * self.S))

    # ---- the step with device-resident scalars (graph=True) ----
    def _sc_ptr(self, slot):
        return self._sc.data_ptr() + 8 * slot

    def _scalar(self, ps, op, h=0.0):
        _lib.check(_lib.lib().tupan_nreg_scalars_dev(
            self._sc.data_ptr(), op, float(h), _stream(ps)))

    def _pair_devdt(self, ps, name):
        k = spec.KERNELS[name]
        f = _lib.ptr_array(ps.ptrs(k.fields))
        _lib.check(getattr(_lib.lib(), "tupan_%s_devdt" % name)(
            _lib.MODES.get(ps.mode, ps.mode), ps.n, f, ps.n, f,
            self._sc_ptr(_lib.NREG_DT), _lib.ptr_array(ps.ptrs(k.outs)),
            _stream(ps)))

    def _x_dev(self, ps):
        st = _stream(ps)
        self._scalar(ps, _lib.NREG_OP_XDT)
        self._pair_devdt(ps, "nreg_x")
        _lib.call_update("divide_by", ps.n, 3, ps.ptrs(("mrx", "mry", "mrz")),
                         ps.ptrs(("rx", "ry", "rz")),
                         self._sc_ptr(_lib.NREG_MTOT), st)
        _lib.call_update("sum_pairwise", ps.n, ps.field("u").data_ptr(),
                         self._sc_ptr(_lib.NREG_SUM), st)
        self._scalar(ps, _lib.NREG_OP_XEND)

    def _v_dev(self, ps):
        st = _stream(ps)
        self._scalar(ps, _lib.NREG_OP_VDT)
        self._pair_devdt(ps, "nreg_v")
        _lib.call_update("divide_by", ps.n, 3, ps.ptrs(("mvx", "mvy", "mvz")),
                         ps.ptrs(("vx", "vy", "vz")),
                         self._sc_ptr(_lib.NREG_MTOT), st)
        _lib.call_update("sum_pairwise", ps.n, ps.field("mk").data_ptr(),
                         self._sc_ptr(_lib.NREG_SUM), st)
        self._scalar(ps, _lib.NREG_OP_VEND)

    def _anreg_dev(self, ps):
        self._x_dev(ps)
        self._v_dev(ps)
        self._x_dev(ps)

    def do_step_dev(self, ps, tau):
        """do_step with every scalar on the device: no host round trip."""
        st = _stream(ps)
        _lib.call_update("sum_pairwise", ps.n, ps.field("mass").data_ptr(),
                         self._sc_ptr(_lib.NREG_MTOT), st)
        self._scalar(ps, _lib.NREG_OP_BEGIN)
        if "anreg" in self.method:
            self._scalar(ps, _lib.NREG_OP_H, tau / 2)
            self._anreg_dev(ps)
        else:
            self._scalar(ps, _lib.NREG_OP_HS, tau)
            self._anreg_dev(ps)
            self._scalar(ps, _lib.NREG_OP_S)
            self._scalar(ps, _lib.NREG_OP_HS, tau)
            self._anreg_dev(ps)
        _lib.call_update("nreg_step_end", ps.n, ps.tstep.data_ptr(),
                         ps.time.data_ptr(), ps.nstep.data_ptr(),
                         self._sc.data_ptr(), float(tau), st)
        return ps

    def _eager_graph_step(self, ps, tau):
        ps = self.do_step_dev(ps, tau)
        self._after_replay(ps, tau)
        return ps

    def _capture_body(self, ps, tau):
        self.do_step_dev(ps, tau)

    def _after_replay(self, ps, tau):
        ps.t_curr = float(self._sc[_lib.NREG_T].item())

    def do_step(self, ps, tau):
        """nreg.py:149-163."""
        self._mtot = ps.sum_np("mass")
        t0 = ps.t_curr
        if "anreg" in self.method:
            self.anreg_step(ps, tau / 2)
        else:
            self.nreg_step(ps, tau)
        ps.tstep.fill_(ps.t_curr - t0)
        ps.time.add_(tau)
        ps.nstep.add_(1)
        return ps


class FewBody(object):
    """tupan/integrator/fewbody.py:14-57 on a device System: nothing for
    n = 0, a drift for n = 1, the two-body Kepler solver (particles 0 and 1,
    in place) otherwise."""

    @staticmethod
    def drift(ps, tau):
        """fewbody.py:17-28 (with pn_drift_com_r when PN is on)."""
        pn = ps.include_pn_corrections
        _lib.call_update(
            "drift", ps.n, ps.ptrs(("rx", "ry", "rz")),
            ps.ptrs(("vx", "vy", "vz")), float(tau), 0,
            ps.ptrs(("pn_mrx", "pn_mry", "pn_mrz")) if pn else None,
            ps.ptrs(("pn_mvx", "pn_mvy", "pn_mvz")) if pn else None,
            _stream(ps))
        return ps

    @staticmethod
    def kepler_solver(ps, tau, status=None):
        """fewbody.py:30-43; ``status`` (device int32) is OR-ed with 1 when
        the solve hit its iteration budget."""
        if ps.include_pn_corrections:
            raise NotImplementedError("The current version of the "
                                      "Kepler-solver does not include "
                                      "post-Newtonian corrections.")
        fields = ps.ptrs(spec.KEPLER_FIELDS)
        outs = ps.ptrs(spec.KEPLER_OUTS)
        _lib.call_update("kepler", fields, float(tau), outs,
                         status.data_ptr() if status is not None else None,
                         _stream(ps))
        return ps

    @staticmethod
    def evolve(ps, tau, status=None):
        """fewbody.py:45-57."""
        if ps.n == 0:
            return ps
        if ps.n == 1:
            return FewBody.drift(ps, tau)
        return FewBody.kepler_solver(ps, tau, status)


class Integrator(object):
    """Method-name dispatch (tupan/integrator/__init__.py:101-160)."""

    PROVIDED_METHODS = (Hermite.PROVIDED_METHODS + SIA.PROVIDED_METHODS
                        + Sakura.PROVIDED_METHODS + SakuraPN.PROVIDED_METHODS
                        + NREG.PROVIDED_METHODS)

    def __init__(self, eta, time, ps, method=None, **kw):
        for cls in (Hermite, SIA, Sakura, SakuraPN, NREG):
            if method in cls.PROVIDED_METHODS:
                self.integrator = cls(eta, time, ps, method, **kw)
                break
        else:
            raise ValueError("Unexpected method: {0}".format(method))

    @property
    def ps(self):
        return self.integrator.ps

    def evolve_step(self, t_end):
        self.integrator.evolve_step(t_end)

    def evolve(self, t_end, callback=None):
        self.integrator.evolve(t_end, callback)
