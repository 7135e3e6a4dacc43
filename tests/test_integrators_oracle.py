"""Pins for the integrator checker (oracle/integrators.py), CPU only.

1. The trajectory fixtures (tests/golden/trajectories.npz, made with the
   reference's own C kernels) are reproduced bit for bit by the checker's
   drivers on the bit-identical C restatement.
2. Mathematics independent of the code:
   * the Hermite predictor is the Taylor polynomial and the corrector is
     exact for polynomial trajectories of the scheme's degree (a dropped or
     mis-scaled term fails), and Hermite of order p (2, 4, 6, 8) converges
     as dt^p on a Kepler binary against the exact elliptic solution;
   * SIA21 is time-reversible and second order;
   * a Sakura step on two bodies is the exact Kepler propagation for any
     step size;
   * with PN terms, the flux bookkeeping closes: linear and angular
     momentum plus their PN accumulators are conserved to round-off, the
     centre of mass plus its accumulator moves uniformly, and Newtonian
     energy plus the PN kinetic-energy accumulator converges as dt^2.
"""
import json
import os

import numpy as np
import pytest

import oracle
from oracle import Oracle
from oracle import integrators as oi

from conftest import ROOT
from test_oracle import kepler_exact

TRAJ = os.path.join(ROOT, "tests", "golden", "trajectories.npz")
BASE = ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz")
STATE = ("rx", "ry", "rz", "vx", "vy", "vz")


@pytest.fixture(scope="module")
def K():
    return oi.Kernels(Oracle("float64", "port"))


def load_traj():
    z = np.load(TRAJ, allow_pickle=False)
    return z, json.loads(str(z["manifest"]))["cases"]


def start(z, meta):
    init = z[meta["name"] + ":init"]
    ps = {f: init[k].copy() for k, f in enumerate(BASE)}
    ps["t_curr"] = 0.0
    ps["tstep"] = np.zeros(meta["n"])
    return ps


def test_checker_reproduces_trajectory_fixtures(K):
    z, cases = load_traj()
    assert {c["name"] for c in cases} >= {"c1_hermite4", "sia21s_kdk",
                                          "sakura", "sia21s_kdk_pn"}
    for meta in cases:
        nm = meta["name"]
        ps = start(z, meta)
        pn = meta["pn"] is not None
        K.pn_order = 0
        if pn:
            K.set_clight(*meta["pn"])
            for f in oi.PN_FIELDS:
                ps[f] = np.zeros(meta["n"])
        m = meta["method"]
        adaptive = m.startswith("a") or "21a" in m
        keep = list(z[nm + ":state_steps"])
        taus, states = [], []
        while abs(ps["t_curr"]) < meta["t_end"]:
            tau = oi.get_base_tstep(ps["t_curr"], meta["t_end"], meta["eta"])
            if "hermite" in m:
                tau = oi.hermite_step(K, ps, tau, int(m[-1]), meta["eta"],
                                      adaptive)
            elif m.startswith("sia21"):
                tau = oi.sia21_step(K, ps, tau, m[-3:], meta["eta"], adaptive,
                                    pn=pn)
            else:
                tau = oi.sakura_do_step(K, ps, tau, meta["eta"], adaptive)
            taus.append(tau)
            if len(taus) - 1 in keep:
                states.append(np.stack([ps[f] for f in STATE]))
        assert np.array_equal(np.array(taus), z[nm + ":tau"]), nm
        assert np.array_equal(np.stack(states), z[nm + ":state"]), nm
        if pn:
            assert np.array_equal(np.stack([ps[f] for f in oi.PN_FIELDS]),
                                  z[nm + ":pn"])


def kepler_binary(m1=0.6, m2=0.4, a=1.0, e=0.5):
    """Two bodies at pericentre, centre of mass at rest at the origin."""
    mu = m1 + m2
    rp = a * (1 - e)
    vp = np.sqrt(mu / a * (1 + e) / (1 - e))
    r = np.array([rp, 0.0, 0.0])
    v = np.array([0.0, vp, 0.0])
    ps = {"mass": np.array([m1, m2]), "eps2": np.zeros(2)}
    for k, d in enumerate("xyz"):
        ps["r" + d] = np.array([m2 / mu * r[k], -m1 / mu * r[k]])
        ps["v" + d] = np.array([m2 / mu * v[k], -m1 / mu * v[k]])
    ps["t_curr"] = 0.0
    ps["tstep"] = np.zeros(2)
    return ps, r, v, mu


def rel_orbit(ps):
    r = np.array([ps["r" + d][0] - ps["r" + d][1] for d in "xyz"])
    v = np.array([ps["v" + d][0] - ps["v" + d][1] for d in "xyz"])
    return r, v


@pytest.mark.parametrize("order,steps", [(2, (256, 512)), (4, (128, 256)),
                                         (6, (64, 128)), (8, (48, 96))])
def test_hermite_order_of_convergence(K, order, steps):
    T = 1.7
    errs = []
    for n in steps:
        ps, r0, v0, mu = kepler_binary()
        for _ in range(n):
            oi.hermite_step(K, ps, T / n, order)
        r, v = rel_orbit(ps)
        rx, vx = kepler_exact(r0, v0, mu, T)
        errs.append(np.linalg.norm(r - rx) + np.linalg.norm(v - vx))
    p = np.log2(errs[0] / errs[1])
    assert order - 0.35 < p < order + 0.6, (order, errs, p)


def plummer_dict(n, seed, eps, K):
    from tupan_amd.ics import make_plummer
    ps = make_plummer(n, eps, seed=seed,
                      phi_fn=lambda p: K.o.call("phi_kernel", p, p)[0])
    d = {f: np.array(getattr(ps, f)) for f in BASE}
    d["t_curr"] = 0.0
    d["tstep"] = np.zeros(n)
    return d


def test_sia21_time_reversible_and_second_order(K):
    ps = plummer_dict(16, 3, 0.1, K)
    s0 = np.stack([ps[f] for f in STATE])
    for kind in ("kdk", "dkd"):
        p = oi.copy(ps)
        for _ in range(8):
            oi.sia21_step(K, p, 1.0 / 32, kind)
        for _ in range(8):
            oi.sia21_step(K, p, -1.0 / 32, kind)
        back = np.stack([p[f] for f in STATE])
        assert np.abs(back - s0).max() < 1e-13, kind
    # reference solution: Hermite-8 with a small step
    T = 0.25
    ref = oi.copy(ps)
    for _ in range(64):
        oi.hermite_step(K, ref, T / 64, 8)
    sref = np.stack([ref[f] for f in STATE])
    errs = []
    for n in (8, 16):
        p = oi.copy(ps)
        for _ in range(n):
            oi.sia21_step(K, p, T / n, "kdk")
        errs.append(np.abs(np.stack([p[f] for f in STATE]) - sref).max())
    order = np.log2(errs[0] / errs[1])
    assert 1.7 < order < 2.4, (errs, order)


def test_sakura_two_bodies_is_exact_kepler(K):
    ps, r0, v0, mu = kepler_binary(0.7, 0.3, 1.3, 0.6)
    com_r = sum(ps["mass"][k] * np.array([ps["r" + d][k] for d in "xyz"])
                for k in range(2))
    tau = 0.9
    oi.sakura_do_step(K, ps, tau)
    r, v = rel_orbit(ps)
    rx, vx = kepler_exact(r0, v0, mu, tau)
    assert np.allclose(r, rx, rtol=0, atol=1e-11)
    assert np.allclose(v, vx, rtol=0, atol=1e-11)
    com_r1 = sum(ps["mass"][k] * np.array([ps["r" + d][k] for d in "xyz"])
                 for k in range(2))
    assert np.allclose(com_r1, com_r, atol=1e-14)


def pn_cluster(K, n=6, seed=11, c=40.0):
    rng = np.random.RandomState(seed)
    ps = {"mass": rng.uniform(0.5, 1.5, n), "eps2": np.full(n, 1e-4)}
    ps["mass"] /= ps["mass"].sum()
    for f in ("rx", "ry", "rz"):
        ps[f] = rng.normal(0, 0.3, n)
    for f in ("vx", "vy", "vz"):
        ps[f] = rng.normal(0, 0.6, n)
    for f in oi.PN_FIELDS:
        ps[f] = np.zeros(n)
    ps["t_curr"] = 0.0
    ps["tstep"] = np.zeros(n)
    K.set_clight(5, c)
    return ps


def _mom(ps):
    m = ps["mass"]
    r = np.stack([ps["r" + d] for d in "xyz"], 1)
    v = np.stack([ps["v" + d] for d in "xyz"], 1)
    P = (m[:, None] * v).sum(0) + np.array([ps["pn_mv" + d].sum() for d in "xyz"])
    L = (m[:, None] * np.cross(r, v)).sum(0) + np.array(
        [ps["pn_am" + d].sum() for d in "xyz"])
    R = (m[:, None] * r).sum(0) + np.array([ps["pn_mr" + d].sum() for d in "xyz"])
    return P, L, R


def test_pn_flux_bookkeeping_closes(K):
    ps = pn_cluster(K)
    P0, L0, R0 = _mom(ps)
    E0 = oi.energy(K, ps, pn=True)
    tau = 1.0 / 256
    for _ in range(16):
        oi.sia21_step(K, ps, tau, "kdk", pn=True)
    P1, L1, R1 = _mom(ps)
    t = ps["t_curr"]
    assert np.abs(ps["pn_ke"]).max() > 1e-6      # the PN terms did work
    assert np.abs(P1 - P0).max() < 1e-14
    assert np.abs(L1 - L0).max() < 1e-14
    assert np.abs(R1 - (R0 + P0 * t)).max() < 1e-14
    # Newtonian energy + accumulated PN work: second-order convergence
    errs = []
    for n in (16, 32):
        p = pn_cluster(K)
        for _ in range(n):
            oi.sia21_step(K, p, 0.0625 / n, "kdk", pn=True)
        errs.append(abs(oi.energy(K, p, pn=True) - E0))
    order = np.log2(errs[0] / errs[1])
    assert 1.6 < order < 2.5, (errs, order)


def _poly_state(coef, t):
    """r(t) = sum_k coef[k] t^k per component and its first seven time
    derivatives, evaluated at t: returns dict rx..cz."""
    out = {}
    names = ("r", "v", "a", "j", "s", "c")
    for d, cf in zip("xyz", coef):
        p = np.polynomial.Polynomial(cf)
        for k, nm in enumerate(names):
            out[nm + d] = np.atleast_1d(p.deriv(k)(t)) if k else np.atleast_1d(p(t))
    return out


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_hermite_predictor_and_corrector_are_exact_on_polynomials(order):
    """The predictor is the Taylor polynomial of r to degree order/2 + 1; the
    corrector (two-point Hermite quadrature of a and v) is exact when r is a
    polynomial of degree `order`."""
    rng = np.random.RandomState(order)
    tau = 0.37
    # predictor
    deg = order // 2 + 1
    coef = rng.uniform(-1, 1, (3, deg + 1))
    ps0 = _poly_state(coef, 0.0)
    ps1 = {k: v.copy() for k, v in ps0.items()}
    oi.hermite_predict(ps1, ps0, tau, order)
    want = _poly_state(coef, tau)
    for f in ("rx", "ry", "rz", "vx", "vy", "vz"):
        assert np.allclose(ps1[f], want[f], rtol=1e-14, atol=1e-15), (order, f)
    # corrector
    coef = rng.uniform(-1, 1, (3, order + 1))
    ps0 = _poly_state(coef, 0.0)
    ps1 = _poly_state(coef, tau)
    want = {k: v.copy() for k, v in ps1.items()}
    for f in ("rx", "ry", "rz", "vx", "vy", "vz"):
        ps1[f] = ps1[f] + 0.1          # the corrector must not depend on these
    oi.hermite_correct(ps1, ps0, tau, order)
    for f in ("rx", "ry", "rz", "vx", "vy", "vz"):
        assert np.allclose(ps1[f], want[f], rtol=1e-13, atol=1e-14), (order, f)


# ---- hybrid PN + Sakura scheme (close pairs by Kepler, the rest kicked) --

def test_hybrid_pieces_reduce_to_the_reference_kernels(K):
    ps = pn_cluster(K, n=12, seed=21)
    full = K.o.call("acc_kernel", ps, ps)
    p = oi.copy(ps)
    for d in "xyz":  # classify by the current positions
        p["c" + d] = p["r" + d].copy()
    oi.acc_far(p, 0.0)
    for k, d in enumerate("xyz"):
        assert np.allclose(p["a" + d], full[k], rtol=1e-13, atol=1e-15)
    sak = K.o.call("sakura_kernel", ps, ps, 0.01, 1)
    oi.sakura_near(K, p, 0.01, 1, np.inf)
    for k, f in enumerate(("drx", "dry", "drz", "dvx", "dvy", "dvz")):
        assert np.array_equal(p[f], sak[k])
    # with rcut = 0 the scheme is the SIA21 kick-drift-kick with PN
    a, b = oi.copy(ps), oi.copy(ps)
    oi.hybrid_step(K, a, 1.0 / 64, 0.0, pn=True)
    oi.sia21_kdk(K, b, 1.0 / 64, pn=True)
    for f in STATE + oi.PN_FIELDS:
        assert np.allclose(a[f], b[f], rtol=1e-13, atol=1e-16), f


def test_hybrid_two_bodies_inside_rcut_is_exact_kepler(K):
    ps, r0, v0, mu = kepler_binary(0.7, 0.3, 1.3, 0.6)
    oi.hybrid_step(K, ps, 0.9, rcut=10.0, pn=False)
    r, v = rel_orbit(ps)
    rx, vx = kepler_exact(r0, v0, mu, 0.9)
    assert np.allclose(r, rx, rtol=0, atol=1e-11)
    assert np.allclose(v, vx, rtol=0, atol=1e-11)


def test_hybrid_pn_bookkeeping_closes(K):
    ps = pn_cluster(K, n=8, seed=31, c=30.0)
    rcut = 0.15
    x = np.stack([ps["rx"], ps["ry"], ps["rz"]])
    r2 = ((x[:, :, None] - x[:, None, :])**2).sum(0)
    assert ((r2 > 0) & (r2 < rcut**2)).any()      # some pairs are close
    P0, L0, R0 = _mom(ps)
    E0 = oi.energy(K, ps, pn=True)
    for _ in range(8):
        oi.hybrid_step(K, ps, 1.0 / 512, rcut)
    P1, L1, R1 = _mom(ps)
    assert np.abs(P1 - P0).max() < 1e-13
    assert np.abs(L1 - L0).max() < 1e-13
    assert np.abs(R1 - (R0 + P0 * ps["t_curr"])).max() < 1e-13
    assert abs(oi.energy(K, ps, pn=True) - E0) < 1e-3 * abs(E0)
