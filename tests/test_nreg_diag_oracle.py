"""Pins of the checker's NREG, FewBody and diagnostic-report restatements.

* NREG (algorithmic-regularisation leapfrog): on an unsoftened Kepler
  binary the anreg step is exact up to round-off for any step length, so
  energy, angular momentum and the Runge-Lenz vector stay constant to
  ~1e-14 over 400 coarse steps (a wrong sub-step length -- kick over h/W
  instead of h/U, or full instead of half drifts -- changes them by O(1));
  the step is time-symmetric (h then -h returns to the start); the nreg
  composition keeps the invariants too.
* FewBody: n = 1 is an exact drift, n = 2 the Kepler kernel on particles 0
  and 1 only.
* Report: closed forms for a symmetric binary (kinetic and softened
  potential energy, zero centre of mass and momentum, L_z = M d v / 2),
  and their exact change under a translation R and a boost V (com_r = R,
  com_v = V, P = M V, L = L0 + M R x V, K = K0 + M V^2 / 2); PN accumulators
  add to their totals; eerr/geerr bookkeeping.
"""
import math

import numpy as np
import pytest

from oracle import Oracle, available
from oracle import diagnostics as D
from oracle import integrators as I

pytestmark = pytest.mark.skipif(not available("float64", "port"),
                                reason="checker library not built")


@pytest.fixture(scope="module")
def K():
    return I.Kernels(Oracle("float64", "port"))


def binary(e=0.8, a=1.0):
    m = np.array([0.5, 0.5])
    rp = a * (1 - e)
    vp = math.sqrt(1.0 * (1 + e) / rp)
    return {"mass": m, "eps2": np.zeros(2),
            "rx": np.array([-rp / 2, rp / 2]), "ry": np.zeros(2),
            "rz": np.zeros(2), "vx": np.zeros(2),
            "vy": np.array([-vp / 2, vp / 2]), "vz": np.zeros(2),
            "t_curr": 0.0}


def kepler_invariants(ps):
    r = np.array([ps["r" + d][1] - ps["r" + d][0] for d in "xyz"])
    v = np.array([ps["v" + d][1] - ps["v" + d][0] for d in "xyz"])
    mtot = float(ps["mass"].sum())
    E = 0.5 * v @ v - mtot / np.linalg.norm(r)
    L = np.cross(r, v)
    A = np.cross(v, L) / mtot - r / np.linalg.norm(r)
    return E, L, A


@pytest.mark.parametrize("method", ["anreg", "nreg"])
def test_nreg_two_body_orbit_is_exact(K, method):
    ps, st = binary(), {}
    I.nreg_initialize(K, ps, st)
    E0, L0, A0 = kepler_invariants(ps)
    for _ in range(400):
        I.nreg_do_step(K, ps, st, 0.25, method)
    E, L, A = kepler_invariants(ps)
    tol = 1e-14 if method == "anreg" else 1e-12
    assert abs(E - E0) <= tol * abs(E0)
    assert np.abs(L - L0).max() <= tol
    assert np.abs(A - A0).max() <= tol
    assert abs(ps["t_curr"]) > 10.0  # the physical clock advanced


def test_nreg_wrong_substep_lengths_break_the_orbit(K):
    """The invariants above are a real check: a kick over h/W (instead of
    h/U) breaks them."""
    ps, st = binary(), {}
    I.nreg_initialize(K, ps, st)
    E0 = kepler_invariants(ps)[0]
    for _ in range(400):
        I.nreg_x(K, ps, st, 0.0625 / st["W"])
        I.nreg_v(K, ps, st, 0.125 / st["W"])
        I.nreg_x(K, ps, st, 0.0625 / st["W"])
    assert abs(kepler_invariants(ps)[0] - E0) > 1e-3 * abs(E0)


def test_anreg_step_is_time_symmetric(K):
    rng = np.random.RandomState(3)
    n = 16
    ps = {f: rng.standard_normal(n) for f in ("rx", "ry", "rz", "vx", "vy",
                                              "vz")}
    ps["mass"] = rng.uniform(0.5, 1.5, n) / n
    ps["eps2"] = np.full(n, 1e-3)
    ps["t_curr"] = 0.0
    # W = K - E0 takes K in the centre-of-mass frame and E0 in the given
    # one (nreg.py:57-58, 121): the scheme assumes the centre-of-mass frame
    for f in ("rx", "ry", "rz", "vx", "vy", "vz"):
        ps[f] = ps[f] - (ps["mass"] * ps[f]).sum() / ps["mass"].sum()
    st = {}
    I.nreg_initialize(K, ps, st)
    ps0 = I.copy(ps)
    I.anreg_step(K, ps, st, 0.05)
    assert abs(ps["t_curr"]) > 1e-3
    I.anreg_step(K, ps, st, -0.05)
    for f in ("rx", "ry", "rz", "vx", "vy", "vz"):
        assert np.abs(ps[f] - ps0[f]).max() < 1e-13
    assert abs(ps["t_curr"]) < 1e-15


def test_fewbody(K):
    one = {f: np.array([v]) for f, v in zip(
        ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz"),
        (1.0, 0.0, 1.0, 2.0, 3.0, 0.5, -0.25, 0.125))}
    I.fewbody_evolve(K, one, 2.0)
    assert (one["rx"][0], one["ry"][0], one["rz"][0]) == (2.0, 1.5, 3.25)
    three = binary()
    for f in list(three):
        if isinstance(three[f], np.ndarray):
            three[f] = np.append(three[f], 7.0)
    ref = Oracle("float64", "port").kepler(three, 0.3)
    I.fewbody_evolve(K, three, 0.3)
    for f, v in zip(("rx", "ry", "rz", "vx", "vy", "vz"), ref):
        assert np.array_equal(three[f][:2], v)
        assert three[f][2] == 7.0
    with pytest.raises(NotImplementedError):
        I.fewbody_evolve(K, binary(), 0.1, pn=True)


def test_report_closed_forms_and_frame_changes(K):
    d, v, e2 = 0.8, 0.6, 1e-2
    ps = {"mass": np.array([0.5, 0.5]), "eps2": np.array([e2, e2]),
          "rx": np.array([-d / 2, d / 2]), "ry": np.zeros(2),
          "rz": np.zeros(2), "vx": np.zeros(2),
          "vy": np.array([-v, v]), "vz": np.zeros(2)}
    q = D.quantities(K, ps)
    assert q["total_mass"] == 1.0
    assert abs(q["kinetic_energy"] - 0.5 * v * v) < 1e-16
    assert abs(q["potential_energy"] + 0.25 / math.sqrt(d * d + 2 * e2)) \
        < 1e-16
    assert abs(q["virial_energy"]
               - (2 * q["kinetic_energy"] + q["potential_energy"])) < 1e-16
    for k in ("com_r", "com_v", "linear_momentum"):
        assert np.abs(q[k]).max() == 0.0
    L0 = np.array([0.0, 0.0, 0.5 * d * v])
    assert np.abs(q["angular_momentum"] - L0).max() < 1e-16
    R, V = np.array([0.3, -1.1, 2.0]), np.array([-0.7, 0.2, 0.4])
    for i, c in enumerate("xyz"):
        ps["r" + c] = ps["r" + c] + R[i]
        ps["v" + c] = ps["v" + c] + V[i]
    q2 = D.quantities(K, ps)
    assert np.abs(q2["com_r"] - R).max() < 1e-15
    assert np.abs(q2["com_v"] - V).max() < 1e-15
    assert np.abs(q2["linear_momentum"] - V).max() < 1e-15
    assert np.abs(q2["angular_momentum"] - (L0 + np.cross(R, V))).max() \
        < 1e-15
    assert abs(q2["kinetic_energy"]
               - (q["kinetic_energy"] + 0.5 * V @ V)) < 1e-15
    assert abs(q2["potential_energy"] - q["potential_energy"]) < 1e-15


def test_report_pn_accumulators_and_bookkeeping(K):
    rng = np.random.RandomState(5)
    n = 50
    ps = {f: rng.standard_normal(n) for f in ("rx", "ry", "rz", "vx", "vy",
                                              "vz")}
    ps["mass"] = rng.uniform(0.5, 1.5, n) / n
    ps["eps2"] = np.full(n, 1e-2)
    q = D.quantities(K, ps)
    for f in D.PN_FIELDS:
        ps[f] = 1e-3 * rng.standard_normal(n)
    qp = D.quantities(K, ps, pn=True)
    M = q["total_mass"]
    assert abs(qp["kinetic_energy"] - q["kinetic_energy"]
               - ps["pn_ke"].sum()) < 1e-14
    for i, c in enumerate("xyz"):
        assert abs((qp["com_r"][i] - q["com_r"][i]) * M
                   - ps["pn_mr" + c].sum()) < 1e-14
        assert abs(qp["linear_momentum"][i] - q["linear_momentum"][i]
                   - ps["pn_mv" + c].sum()) < 1e-14
        assert abs(qp["angular_momentum"][i] - q["angular_momentum"][i]
                   - ps["pn_am" + c].sum()) < 1e-14
    # eerr = (te - te0) / (-pe), geerr = rms of the eerr so far
    dg = D.Diagnostic(K, ps)
    r1 = dg.report(ps)
    assert r1["eerr"] == 0.0 and r1["geerr"] == 0.0
    for c in "xyz":
        ps["v" + c] = ps["v" + c] * math.sqrt(2.0)
    r2 = dg.report(ps)
    e2 = q["kinetic_energy"] / -q["potential_energy"]
    assert abs(r2["eerr"] - e2) < 1e-12
    assert abs(r2["geerr"] - abs(e2) / math.sqrt(2)) < 1e-12
    assert abs(r2["lmom"] - (math.sqrt(2) - 1)
               * np.linalg.norm(q["linear_momentum"])) < 1e-14
