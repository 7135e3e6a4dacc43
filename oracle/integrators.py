"""CPU restatement of Tupan's integrator steps -- TEST INFRASTRUCTURE.

Only ``tests/`` (and the fixture scripts under ``tests/golden/``) use this
module, as the checker of ``tupan_amd.integrator``; the product never
imports it.

State is a dict of float64 numpy arrays named as the reference's particle
attributes (``mass, eps2, rx, ry, rz, vx, vy, vz``; outputs ``ax`` ..
``cz``, ``tstep``, ``tstepij``, ``pnax`` .., ``drx`` ..; PN bookkeeping
``wx``.., ``pn_ke``, ``pn_mrx``.., ``pn_mvx``.., ``pn_amx``..) plus the
scalar ``t_curr``.  Pair kernels come from an ``oracle.Oracle`` (the C
restatement, or the reference C itself); every update is the reference's
numpy expression, evaluated the same way:

  H2/H4/H6/H8 predict/correct   tupan/integrator/hermite.py:20-303
  Hermite.do_step               hermite.py:390-410
  get_base_tstep                tupan/integrator/__init__.py:48-58
  get_min_block_tstep           __init__.py:60-78
  drift_n, kick_n, drift_pn,    tupan/integrator/sia.py:62-186
  kick_pn                       (+ tupan/particles/body.py:471-525)
  SIA21.kdk/dkd, bridge_sf      sia.py:291-353
  SIA.recurse (shared steps)    sia.py:1087-1119
  sakura_step, get_sakura_tstep tupan/integrator/sakura.py:22-50, 100-117
  nreg_x, nreg_v, anreg_step,   tupan/integrator/nreg.py:22-97, 114-134,
  nreg_step, NREG.do_step       149-172
  FewBody.drift/kepler/evolve   tupan/integrator/fewbody.py:17-57
"""
import math
import sys

import numpy as np

DERIV = ("ax", "ay", "az", "jx", "jy", "jz", "sx", "sy", "sz",
         "cx", "cy", "cz")


def _set(ps, names, values):
    for n, v in zip(names, values):
        ps[n] = np.array(v, dtype=np.float64)


class Kernels(object):
    """The ``set_*`` forwarding of NbodyMethods (body.py:324-361) over an
    oracle kernel library."""

    def __init__(self, oracle):
        self.o = oracle
        self.pn_order = 0
        self.inv = (0.0,) * 7

    def set_clight(self, pn_order, clight):
        """extensions.py:30-60: inverse powers from FP64 Python floats."""
        self.pn_order = int(pn_order)
        inv1 = 1.0 / float(clight)
        self.inv = (inv1, inv1**2, inv1**3, inv1**4, inv1**5, inv1**6,
                    inv1**7)

    def phi(self, ps):
        _set(ps, ("phi",), self.o.call("phi_kernel", ps, ps))

    def acc(self, ps):
        _set(ps, DERIV[:3], self.o.call("acc_kernel", ps, ps))

    def acc_jerk(self, ps):
        _set(ps, DERIV[:6], self.o.call("acc_jerk_kernel", ps, ps))

    def snap_crackle(self, ps):
        _set(ps, DERIV[6:], self.o.call("snap_crackle_kernel", ps, ps))

    def tstep(self, ps, eta):
        _set(ps, ("tstep", "tstepij"), self.o.call("tstep_kernel", ps, ps, eta))

    def pnacc(self, ps):
        _set(ps, ("pnax", "pnay", "pnaz"),
             self.o.call("pnacc_kernel", ps, ps, self.pn_order, *self.inv))

    def sakura(self, ps, dt, flag):
        _set(ps, ("drx", "dry", "drz", "dvx", "dvy", "dvz"),
             self.o.call("sakura_kernel", ps, ps, dt, flag))

    def nreg_x(self, ps, dt):
        _set(ps, ("mrx", "mry", "mrz", "ax", "ay", "az", "u"),
             self.o.call("nreg_Xkernel", ps, ps, dt))

    def nreg_v(self, ps, dt):
        _set(ps, ("mvx", "mvy", "mvz", "mk"),
             self.o.call("nreg_Vkernel", ps, ps, dt))

    def kepler(self, ps, dt):
        """kepler_solver_kernel in place on particles 0 and 1
        (extensions.py:642-646)."""
        new = self.o.kepler(ps, dt)
        for name, v in zip(("rx", "ry", "rz", "vx", "vy", "vz"), new):
            ps[name][:2] = v


def copy(ps):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v)
            for k, v in ps.items()}


# ---- time-step selection (tupan/integrator/__init__.py:48-78) ----------

def get_base_tstep(t_curr, t_end, eta):
    dt = min(abs(t_end) - abs(t_curr), abs(eta))
    dt = max(dt, abs(t_end) * (2 * sys.float_info.epsilon))
    return math.copysign(dt, eta)


def get_min_block_tstep(ps, tau):
    min_ts = abs(ps["tstep"]).min()
    power = int(np.log2(min_ts) - 1)
    min_bts = 2.0**power
    t_next = ps["t_curr"] + min_bts
    while t_next % min_bts != 0:
        min_bts /= 2
    min_bts = math.copysign(min_bts, tau)
    if abs(min_bts) > abs(tau):
        min_bts = tau
    return min_bts


# ---- Hermite (hermite.py) ---------------------------------------------

def _derivs(K, ps, order):
    if order == 2:
        K.acc(ps)
    else:
        K.acc_jerk(ps)
        if order >= 6:
            K.snap_crackle(ps)


def hermite_predict(ps1, ps0, tau, order):
    """ps1 (== ps0 on entry) advanced by the Taylor predictor."""
    for d in "xyz":
        r0, v0, a = ps0["r" + d], ps0["v" + d], ps0["a" + d]
        if order == 2:
            ps1["r" + d] += (a * tau / 2 + v0) * tau
            ps1["v" + d] += a * tau
        elif order == 4:
            j = ps0["j" + d]
            ps1["r" + d] += ((j * tau / 3 + a) * tau / 2 + v0) * tau
            ps1["v" + d] += (j * tau / 2 + a) * tau
        elif order == 6:
            j, s = ps0["j" + d], ps0["s" + d]
            ps1["r" + d] += (((s * tau / 4 + j) * tau / 3 + a) * tau / 2
                             + v0) * tau
            ps1["v" + d] += ((s * tau / 3 + j) * tau / 2 + a) * tau
        else:
            j, s, c = ps0["j" + d], ps0["s" + d], ps0["c" + d]
            ps1["r" + d] += ((((c * tau / 5 + s) * tau / 4 + j) * tau / 3
                              + a) * tau / 2 + v0) * tau
            ps1["v" + d] += (((c * tau / 4 + s) * tau / 3 + j) * tau / 2
                             + a) * tau


def hermite_correct(ps1, ps0, tau, order):
    """ps1's derivatives are those at the predicted state."""
    for d in "xyz":
        a0, a1 = ps0["a" + d], ps1["a" + d]
        v0, r0 = ps0["v" + d], ps0["r" + d]
        if order == 2:
            ps1["v" + d][...] = (a0 + a1) * tau / 2 + v0
            ps1["r" + d][...] = (v0 + ps1["v" + d]) * tau / 2 + r0
        elif order == 4:
            j0, j1 = ps0["j" + d], ps1["j" + d]
            ps1["v" + d][...] = ((j0 - j1) * tau / 6 + (a0 + a1)) * tau / 2 + v0
            ps1["r" + d][...] = (((a0 - a1) * tau / 6 + (v0 + ps1["v" + d]))
                                 * tau / 2 + r0)
        elif order == 6:
            j0, j1 = ps0["j" + d], ps1["j" + d]
            s0, s1 = ps0["s" + d], ps1["s" + d]
            ps1["v" + d][...] = ((((s0 + s1) * tau / 12 + (j0 - j1)) * tau / 5
                                  + (a0 + a1)) * tau / 2 + v0)
            ps1["r" + d][...] = ((((j0 + j1) * tau / 12 + (a0 - a1)) * tau / 5
                                  + (v0 + ps1["v" + d])) * tau / 2 + r0)
        else:
            j0, j1 = ps0["j" + d], ps1["j" + d]
            s0, s1 = ps0["s" + d], ps1["s" + d]
            c0, c1 = ps0["c" + d], ps1["c" + d]
            ps1["v" + d][...] = (((((c0 - c1) * tau / 20 + (s0 + s1)) * tau / 3
                                   + 3 * (j0 - j1)) * tau / 14
                                  + (a0 + a1)) * tau / 2 + v0)
            ps1["r" + d][...] = (((((s0 - s1) * tau / 20 + (j0 + j1)) * tau / 3
                                   + 3 * (a0 - a1)) * tau / 14
                                  + (v0 + ps1["v" + d])) * tau / 2 + r0)


def hermite_epec(K, ps, tau, order, n=2):
    ps0 = copy(ps)
    _derivs(K, ps0, order)
    hermite_predict(ps, ps0, tau, order)
    for _ in range(n):
        _derivs(K, ps, order)
        hermite_correct(ps, ps0, tau, order)
    return ps


def _advance_clock(ps, tau):
    ps["t_curr"] += tau
    ps["tstep"] = np.full(len(ps["mass"]), tau)
    ps["time"] = ps.get("time", np.zeros(len(ps["mass"]))) + tau
    ps["nstep"] = ps.get("nstep", np.zeros(len(ps["mass"]), np.uint64)) + 1


def hermite_step(K, ps, tau, order, eta=None, adaptive=False):
    """Hermite.do_step: one P(EC)^2 step; adaptive methods first take the
    block step from Tstep (hermite.py:361-368)."""
    if adaptive:
        K.tstep(ps, eta)
        tau = get_min_block_tstep(ps, tau)
    hermite_epec(K, ps, tau, order)
    _advance_clock(ps, tau)
    return tau


# ---- SIA21 (sia.py) -----------------------------------------------------

def drift(K, ps, tau, pn=False):
    for d in "xyz":
        ps["r" + d] += ps["v" + d] * tau
    if pn:
        for d in "xyz":
            ps["pn_mr" + d] += ps["pn_mv" + d] * tau


def kick(K, ps, tau, pn=False, acc=None):
    (acc or K.acc)(ps)
    if not pn:
        for d in "xyz":
            ps["v" + d] += ps["a" + d] * tau
        return
    for d in "xyz":
        ps["v" + d] += (ps["a" + d] * tau + ps["w" + d]) / 2
    K.pnacc(ps)
    m = ps["mass"]
    pnfx, pnfy, pnfz = m * ps["pnax"], m * ps["pnay"], m * ps["pnaz"]
    ps["pn_ke"] -= (ps["vx"] * pnfx + ps["vy"] * pnfy + ps["vz"] * pnfz) * tau
    ps["pn_mvx"] -= pnfx * tau
    ps["pn_mvy"] -= pnfy * tau
    ps["pn_mvz"] -= pnfz * tau
    ps["pn_amx"] -= (ps["ry"] * pnfz - ps["rz"] * pnfy) * tau
    ps["pn_amy"] -= (ps["rz"] * pnfx - ps["rx"] * pnfz) * tau
    ps["pn_amz"] -= (ps["rx"] * pnfy - ps["ry"] * pnfx) * tau
    for d in "xyz":
        ps["w" + d][...] = 2 * ps["pna" + d] * tau - ps["w" + d]
    for d in "xyz":
        ps["v" + d] += (ps["a" + d] * tau + ps["w" + d]) / 2


PN_FIELDS = ("wx", "wy", "wz", "pn_ke", "pn_mrx", "pn_mry", "pn_mrz",
             "pn_mvx", "pn_mvy", "pn_mvz", "pn_amx", "pn_amy", "pn_amz")


def sia21_kdk(K, ps, tau, pn=False):
    kick(K, ps, 0.5 * tau, pn)
    drift(K, ps, 1.0 * tau, pn)
    kick(K, ps, 0.5 * tau, pn)


def sia21_dkd(K, ps, tau, pn=False):
    drift(K, ps, 0.5 * tau, pn)
    kick(K, ps, 1.0 * tau, pn)
    drift(K, ps, 0.5 * tau, pn)


def sia21_step(K, ps, tau, kind="kdk", eta=None, adaptive=False, pn=False):
    """SIA.do_step for the shared-step methods sia21s/sia21a on n > 2: all
    particles are slow, so bridge_sf evolves the system twice by tau/2."""
    if len(ps["mass"]) <= 2:
        raise NotImplementedError("n <= 2 goes to the few-body solver")
    if adaptive:
        K.tstep(ps, eta)
        tau = get_min_block_tstep(ps, tau)
    ev = sia21_kdk if kind == "kdk" else sia21_dkd
    ev(K, ps, 0.5 * tau, pn)
    ev(K, ps, 0.5 * tau, pn)
    _advance_clock(ps, tau)
    return tau


# ---- Sakura (sakura.py) -------------------------------------------------

def sakura_step(K, ps, tau):
    for d in "xyz":
        ps["r" + d] += ps["v" + d] * tau / 2
    K.sakura(ps, tau / 2, -1)
    for d in "xyz":
        ps["r" + d] += ps["dr" + d]
    for d in "xyz":
        ps["v" + d] += ps["dv" + d]
    K.sakura(ps, tau / 2, 1)
    for d in "xyz":
        ps["r" + d] += ps["dr" + d]
    for d in "xyz":
        ps["v" + d] += ps["dv" + d]
    for d in "xyz":
        ps["r" + d] += ps["v" + d] * tau / 2


def get_sakura_tstep(K, ps, eta, tau):
    K.tstep(ps, eta)
    iw2_a = (eta / ps["tstep"])**2
    iw2_b = (eta / ps["tstepij"])**2
    diw2 = (iw2_a - iw2_b)
    w2_sakura = diw2.max()
    dt_sakura = eta / (1 + w2_sakura)**0.5
    ps["tstep"][...] = dt_sakura
    return get_min_block_tstep(ps, tau)


def sakura_do_step(K, ps, tau, eta=None, adaptive=False):
    if adaptive:
        tau = get_sakura_tstep(K, ps, eta, tau)
    sakura_step(K, ps, tau)
    _advance_clock(ps, tau)
    return tau


# ---- hybrid PN + Sakura scheme ---------------------------------------------
# Not in the reference (its Sakura acts on every pair, its PN runs in SIA):
# close pairs are drifted by the Sakura two-body Kepler solve, the other
# pairs are kicked together with the PN terms (kick_pn's w scheme and flux
# bookkeeping):
#   K(tau/2) [far Newton + PN]  D(tau) [free drift + close-pair Kepler]
#   K(tau/2)
# The pairs are split by their separation at the classification positions
# cx, cy, cz: the free-drift midpoint r + (v tau) / 2 predicted at the start
# of the step.

def _class_r2(ps):
    c = np.stack([ps["cx"], ps["cy"], ps["cz"]])
    d = c[:, :, None] - c[:, None, :]
    return (d * d).sum(0)


def acc_far(ps, rcut):
    """Acc over the pairs whose classification positions are >= rcut apart
    (r^2 > 0), plain numpy."""
    x = np.stack([ps["rx"], ps["ry"], ps["rz"]])
    r = x[:, :, None] - x[:, None, :]
    r2 = (r * r).sum(0)
    s2 = r2 + (ps["eps2"][:, None] + ps["eps2"][None, :])
    keep = (r2 > 0) & (_class_r2(ps) >= rcut * rcut)
    y = np.where(keep, 1.0 / np.sqrt(np.where(keep, s2, 1.0)), 0.0)
    k = ps["mass"][None, :] * y**3
    a = -(k[None] * r).sum(2)
    for n, d in enumerate("xyz"):
        ps["a" + d] = a[n]


def sakura_near(K, ps, dt, flag, rcut):
    """Sakura deltas from the pairs whose classification positions are
    closer than rcut (self pairs included, as in the full kernel), each pair
    through the C kernel."""
    n = len(ps["mass"])
    r2 = _class_r2(ps)
    d = np.zeros((6, n))
    F8 = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")
    for i in range(n):
        for j in np.nonzero(r2[i] < rcut * rcut)[0]:
            out = K.o.call("sakura_kernel", {f: ps[f][i:i + 1] for f in F8},
                           {f: ps[f][j:j + 1] for f in F8}, dt, flag)
            d[:, i] += np.array([o[0] for o in out])
    for k, f in enumerate(("drx", "dry", "drz", "dvx", "dvy", "dvz")):
        ps[f] = d[k]


def hybrid_drift(K, ps, tau, rcut, pn=False):
    for dd in "xyz":
        ps["r" + dd] += ps["v" + dd] * tau / 2
    for flag in (-1, 1):
        sakura_near(K, ps, tau / 2, flag, rcut)
        for dd in "xyz":
            ps["r" + dd] += ps["dr" + dd]
        for dd in "xyz":
            ps["v" + dd] += ps["dv" + dd]
    for dd in "xyz":
        ps["r" + dd] += ps["v" + dd] * tau / 2
    if pn:
        for dd in "xyz":
            ps["pn_mr" + dd] += ps["pn_mv" + dd] * tau


def hybrid_step(K, ps, tau, rcut, pn=True):
    for d in "xyz":
        ps["c" + d] = ps["r" + d] + ps["v" + d] * tau / 2
    far = lambda p: acc_far(p, rcut)
    kick(K, ps, 0.5 * tau, pn, acc=far)
    hybrid_drift(K, ps, tau, rcut, pn)
    kick(K, ps, 0.5 * tau, pn, acc=far)
    _advance_clock(ps, tau)
    return tau


# ---- diagnostics (body.py:259-321) -------------------------------------

def energy(K, ps, pn=False):
    ke = 0.5 * ps["mass"] * (ps["vx"]**2 + ps["vy"]**2 + ps["vz"]**2)
    if pn:
        ke = ke + ps["pn_ke"]
    K.phi(ps)
    return float(ke.sum()) + 0.5 * float((ps["mass"] * ps["phi"]).sum())


# ---- NREG (nreg.py) -----------------------------------------------------
# The scalars the reference keeps as class attributes of the particle type
# (E0, W, U, S) live in the dict ``st``.

def nreg_initialize(K, ps, st):
    """NREG.initialize (nreg.py:114-126): E0 = KE + PE, W = U = S = -PE."""
    ke = float((0.5 * ps["mass"] * (ps["vx"]**2 + ps["vy"]**2
                                    + ps["vz"]**2)).sum())
    K.phi(ps)
    pe = 0.5 * float((ps["mass"] * ps["phi"]).sum())
    st["E0"] = ke + pe
    st["W"] = -pe
    st["U"] = -pe
    st["S"] = -pe


def nreg_x(K, ps, st, dt):
    """nreg.py:22-33: positions at the drifted separations, recentred."""
    mtot = float(ps["mass"].sum())
    K.nreg_x(ps, dt)
    ps["rx"] = ps["mrx"] / mtot
    ps["ry"] = ps["mry"] / mtot
    ps["rz"] = ps["mrz"] / mtot
    st["U"] = 0.5 * ps["u"].sum()
    ps["t_curr"] += dt


def nreg_v(K, ps, st, dt):
    """nreg.py:45-59: velocities after the kick, recentred; W = K - E0."""
    mtot = float(ps["mass"].sum())
    K.nreg_v(ps, dt)
    ps["vx"] = ps["mvx"] / mtot
    ps["vy"] = ps["mvy"] / mtot
    ps["vz"] = ps["mvz"] / mtot
    kin = 0.25 * ps["mk"].sum() / mtot
    st["W"] = kin - st["E0"]


def anreg_step(K, ps, st, h):
    """nreg.py:77-85: drift(h/2W) kick(h/U) drift(h/2W), each step length
    taken from the scalars as they stand when it starts."""
    nreg_x(K, ps, st, 0.5 * (h / st["W"]))
    nreg_v(K, ps, st, (h / st["U"]))
    nreg_x(K, ps, st, 0.5 * (h / st["W"]))


def nreg_step(K, ps, st, h):
    """nreg.py:88-97."""
    anreg_step(K, ps, st, 0.5 * (h * st["S"]))
    st["S"] = 1 / (2 / st["W"] - 1 / st["S"])
    anreg_step(K, ps, st, 0.5 * (h * st["S"]))


def nreg_do_step(K, ps, st, tau, method):
    """NREG.do_step (nreg.py:149-163): tstep = the physical time advanced,
    time += tau (the regularised step), nstep += 1."""
    t0 = ps["t_curr"]
    if "anreg" in method:
        anreg_step(K, ps, st, tau / 2)
    else:
        nreg_step(K, ps, st, tau)
    dt = ps["t_curr"] - t0
    n = len(ps["mass"])
    ps["tstep"] = np.full(n, dt)
    ps["time"] = ps.get("time", np.zeros(n)) + tau
    ps["nstep"] = ps.get("nstep", np.zeros(n, np.uint64)) + 1
    return dt


# ---- FewBody (fewbody.py) ------------------------------------------------

def fewbody_evolve(K, ps, tau, pn=False):
    """FewBody.evolve: nothing for n = 0, a drift for n = 1, the Kepler
    solver on particles 0 and 1 otherwise (fewbody.py:17-57)."""
    n = len(ps["mass"])
    if n == 0:
        return
    if n == 1:
        drift(K, ps, tau, pn)
        return
    if pn:
        raise NotImplementedError("The current version of the "
                                  "Kepler-solver does not include "
                                  "post-Newtonian corrections.")
    K.kepler(ps, tau)
