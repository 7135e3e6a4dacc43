"""CPU restatement of Tupan's energy/momentum report -- TEST INFRASTRUCTURE.

Only ``tests/`` use this module, as the checker of the device diagnostics
(``tupan_amd.integrator.System.diagnostics``, ``tupan_amd.diagnostics``);
the product never imports it.

``ps`` is a dict of float64 numpy arrays named as the reference's particle
attributes.  Each quantity is the reference's numpy expression:

  total_mass, com_r, com_v        tupan/particles/body.py:79-130
  linear_momentum (lm)            body.py:180-217
  angular_momentum (am)           body.py:219-259
  kinetic_energy (ke)             body.py:261-292
  potential_energy (pe)           body.py:294-306
  virial_energy                   body.py:317-322
  the report (eerr, geerr, ...)   tupan/simulation.py:52-69, 96-129
"""
import math

import numpy as np

PN_FIELDS = ("pn_ke", "pn_mrx", "pn_mry", "pn_mrz", "pn_mvx", "pn_mvy",
             "pn_mvz", "pn_amx", "pn_amy", "pn_amz")


def quantities(K, ps, pn=False):
    """The report's totals; ``K`` supplies Phi (``K.phi(ps)`` sets
    ``ps["phi"]``), as ``ps.set_phi(ps)`` does (body.py:298)."""
    m = ps["mass"]
    mtot = float(m.sum())
    px, py, pz = m * ps["vx"], m * ps["vy"], m * ps["vz"]
    mrx, mry, mrz = m * ps["rx"], m * ps["ry"], m * ps["rz"]
    lmx, lmy, lmz = px.copy(), py.copy(), pz.copy()
    amx = (ps["ry"] * pz) - (ps["rz"] * py)
    amy = (ps["rz"] * px) - (ps["rx"] * pz)
    amz = (ps["rx"] * py) - (ps["ry"] * px)
    ke = 0.5 * m * (ps["vx"]**2 + ps["vy"]**2 + ps["vz"]**2)
    if pn:
        mrx = mrx + ps["pn_mrx"]
        mry = mry + ps["pn_mry"]
        mrz = mrz + ps["pn_mrz"]
        lmx = lmx + ps["pn_mvx"]
        lmy = lmy + ps["pn_mvy"]
        lmz = lmz + ps["pn_mvz"]
        amx = amx + ps["pn_amx"]
        amy = amy + ps["pn_amy"]
        amz = amz + ps["pn_amz"]
        ke = ke + ps["pn_ke"]
    K.phi(ps)
    kinetic = float(ke.sum())
    potential = 0.5 * float((m * ps["phi"]).sum())
    mr = np.array([mrx, mry, mrz]).T
    mv = np.array([lmx, lmy, lmz]).T
    am = np.array([amx, amy, amz]).T
    return {
        "total_mass": mtot,
        "kinetic_energy": kinetic,
        "potential_energy": potential,
        "virial_energy": 2 * kinetic + potential,
        "com_r": mr.sum(0) / mtot,
        "com_v": mv.sum(0) / mtot,
        "linear_momentum": mv.sum(0),
        "angular_momentum": am.sum(0),
    }


class Diagnostic(object):
    """The report's bookkeeping (simulation.py:52-69, 96-129), values only."""

    def __init__(self, K, ps, pn=False):
        self.K = K
        self.pn = pn
        q = quantities(K, ps, pn)
        self.te0 = q["kinetic_energy"] + q["potential_energy"]
        self.com_r0 = q["com_r"]
        self.com_v0 = q["com_v"]
        self.lmom0 = q["linear_momentum"]
        self.amom0 = q["angular_momentum"]
        self.count = 0
        self.ceerr = 0.0

    def report(self, ps):
        q = quantities(self.K, ps, self.pn)
        ke, pe = q["kinetic_energy"], q["potential_energy"]
        te = ke + pe
        eerr = (te - self.te0) / (-pe)
        self.count += 1
        self.ceerr += eerr**2
        geerr = math.sqrt(self.ceerr / self.count)
        return {
            "ke": ke, "pe": pe, "te": te, "virial": q["virial_energy"],
            "eerr": eerr, "geerr": geerr,
            "com_r": float((((q["com_r"] - self.com_r0)**2).sum())**0.5),
            "com_v": float((((q["com_v"] - self.com_v0)**2).sum())**0.5),
            "lmom": float((((q["linear_momentum"]
                             - self.lmom0)**2).sum())**0.5),
            "amom": float((((q["angular_momentum"]
                             - self.amom0)**2).sum())**0.5),
        }
