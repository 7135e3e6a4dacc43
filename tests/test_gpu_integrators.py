"""Device-resident integrators (tupan_amd.integrator) against the checker.

1. Every update kernel of integrate.hip reproduces the numpy expression of
   the reference (as restated in oracle/integrators.py) bit for bit on random
   inputs.
2. Every trajectory fixture (reference C kernels + restated drivers) is
   followed by the device driver on the FP64 GPU kernels: identical time-step
   sequence, states within 1e-11 of each component's scale (1e-10 for the
   Sakura cases, whose pair kernel runs transcendental functions), energies
   within 1e-12 relative, and for the Hermite-4 run of N=256 (dt = 1/64, one
   time unit) the Tstep outputs of every step within 1e-12.  Injecting
   1e-15 relative noise into the checker's pair kernels moves these runs by
   at most 3.4e-14 of scale, so the bounds leave three decades for the GPU
   kernels' own rounding.
"""
import json
import os

import numpy as np
import pytest

from oracle import integrators as oi

from conftest import ROOT

pytestmark = pytest.mark.gpu

TRAJ = os.path.join(ROOT, "tests", "golden", "trajectories.npz")
BASE = ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz")
STATE = ("rx", "ry", "rz", "vx", "vy", "vz")


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def _dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float64), device="cuda")


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_hermite_update_kernels_bitwise(torch, order):
    from tupan_amd import _lib
    rng = np.random.RandomState(order)
    n = 1000
    names = ("rx", "ry", "rz", "vx", "vy", "vz") + oi.DERIV[:3 * (order // 2)]
    st = {f: rng.normal(0, 1, n) * 10.0**rng.uniform(-3, 3, n) for f in names}
    tau = 0.0123
    # predictor
    ps0 = {k: v.copy() for k, v in st.items()}
    ps1 = {k: v.copy() for k, v in st.items()}
    oi.hermite_predict(ps1, ps0, tau, order)
    d = {f: _dev(torch, st[f]) for f in names}
    d0 = {f: torch.empty(n, dtype=torch.float64, device="cuda") for f in names}
    x = [d[f].data_ptr() for f in names]
    x0 = [d0[f].data_ptr() for f in names]
    _lib.call_update("hermite_predict", order, n, x, x0, tau, 0)
    torch.cuda.synchronize()
    for f in names:
        assert np.array_equal(d[f].cpu().numpy(), ps1[f]), ("predict", f)
        assert np.array_equal(d0[f].cpu().numpy(), st[f]), ("snapshot", f)
    # corrector: new derivatives at the predicted state
    for f in names[6:]:
        ps1[f] = rng.normal(0, 1, n)
        d[f].copy_(_dev(torch, ps1[f]))
    oi.hermite_correct(ps1, ps0, tau, order)
    _lib.call_update("hermite_correct", order, n, x, x0, tau, 0)
    torch.cuda.synchronize()
    for f in STATE:
        assert np.array_equal(d[f].cpu().numpy(), ps1[f]), ("correct", f)


class _Given(object):
    """Checker 'kernels' that leave the given accelerations in place."""

    def __init__(self, pna=None):
        self.pna = pna

    def acc(self, ps):
        pass

    def pnacc(self, ps):
        for k, d in enumerate("xyz"):
            ps["pna" + d] = self.pna[k].copy()


def test_sia_and_sakura_update_kernels_bitwise(torch):
    from tupan_amd import _lib
    rng = np.random.RandomState(5)
    n = 777
    names = STATE + ("mass", "ax", "ay", "az") + oi.PN_FIELDS + (
        "drx", "dry", "drz", "dvx", "dvy", "dvz")
    st = {f: rng.normal(0, 1, n) for f in names}
    pna = rng.normal(0, 1e-3, (3, n))
    tau = -0.0371
    d = {f: _dev(torch, st[f]) for f in names}
    d.update({"pna" + c: _dev(torch, pna[k]) for k, c in enumerate("xyz")})
    P = lambda fs: [d[f].data_ptr() for f in fs]
    R, V, A = ("rx", "ry", "rz"), ("vx", "vy", "vz"), ("ax", "ay", "az")
    W = ("wx", "wy", "wz")

    def check(ps, what):
        torch.cuda.synchronize()
        for f in names:
            assert np.array_equal(d[f].cpu().numpy(), ps[f]), (what, f)

    ps = {k: v.copy() for k, v in st.items()}
    # drift_n, drift_pn
    oi.drift(None, ps, tau)
    _lib.call_update("drift", n, P(R), P(V), tau, 0, None, None, 0)
    check(ps, "drift")
    oi.drift(None, ps, tau, pn=True)
    _lib.call_update("drift", n, P(R), P(V), tau, 0,
                     P(("pn_mrx", "pn_mry", "pn_mrz")),
                     P(("pn_mvx", "pn_mvy", "pn_mvz")), 0)
    check(ps, "drift_pn")
    # kick_n
    oi.kick(_Given(), ps, tau)
    _lib.call_update("kick", n, P(V), P(A), tau, 0)
    check(ps, "kick")
    # kick_pn = half kick, PN flux, half kick
    oi.kick(_Given(pna), ps, tau, pn=True)
    _lib.call_update("kick_pn_half", n, P(V), P(A), P(W), tau, 0)
    _lib.call_update("pn_flux", n,
                     P(("mass",) + R + V + ("pnax", "pnay", "pnaz")),
                     P(W + ("pn_ke", "pn_mvx", "pn_mvy", "pn_mvz",
                            "pn_amx", "pn_amy", "pn_amz")), tau, 0)
    _lib.call_update("kick_pn_half", n, P(V), P(A), P(W), tau, 0)
    for k, c in enumerate("xyz"):
        ps["pna" + c] = pna[k]
    check(ps, "kick_pn")
    # kick_pn again, with the flux and the second half kick fused
    oi.kick(_Given(pna), ps, tau, pn=True)
    _lib.call_update("kick_pn_half", n, P(V), P(A), P(W), tau, 0)
    _lib.call_update("pn_flux_kick", n,
                     P(("mass",) + R + V + ("pnax", "pnay", "pnaz")),
                     P(W + ("pn_ke", "pn_mvx", "pn_mvy", "pn_mvz",
                            "pn_amx", "pn_amy", "pn_amz")), P(A), tau, 0)
    for k, c in enumerate("xyz"):
        ps["pna" + c] = pna[k]
    check(ps, "kick_pn fused")
    # Sakura: half drift and delta application
    for dd in "xyz":
        ps["r" + dd] += ps["v" + dd] * tau / 2
    _lib.call_update("drift", n, P(R), P(V), tau, 1, None, None, 0)
    check(ps, "half drift")
    for dd in "xyz":
        ps["r" + dd] += ps["dr" + dd]
    for dd in "xyz":
        ps["v" + dd] += ps["dv" + dd]
    _lib.call_update("apply_delta", n, P(R + V),
                     P(("drx", "dry", "drz", "dvx", "dvy", "dvz")), 0)
    check(ps, "apply_delta")
    # end-of-step clock
    tstep = _dev(torch, rng.normal(0, 1, n))
    time = _dev(torch, rng.normal(0, 1, n))
    nstep = torch.as_tensor(rng.randint(0, 1 << 40, n), device="cuda")
    t_ref = time.cpu().numpy() + tau
    n_ref = nstep.cpu().numpy() + 1
    _lib.call_update("advance_clock", n, tstep.data_ptr(), time.data_ptr(),
                     nstep.data_ptr(), tau, 0)
    torch.cuda.synchronize()
    assert np.all(tstep.cpu().numpy() == tau)
    assert np.array_equal(time.cpu().numpy(), t_ref)
    assert np.array_equal(nstep.cpu().numpy(), n_ref)


def _cases():
    z = np.load(TRAJ, allow_pickle=False)
    return [c["name"] for c in json.loads(str(z["manifest"]))["cases"]]


@pytest.mark.parametrize("name", _cases())
def test_trajectory_matches_reference_kernels(torch, name):
    from tupan_amd import integrator as ig
    z = np.load(TRAJ, allow_pickle=False)
    meta = [c for c in json.loads(str(z["manifest"]))["cases"]
            if c["name"] == name][0]
    print("[traj] %s n=%d method=%s" % (name, meta["n"], meta["method"]),
          flush=True)
    init = z[name + ":init"]
    ps = ig.System.from_host({f: init[k] for k, f in enumerate(BASE)})
    kw = {}
    if meta["pn"]:
        kw = dict(pn_order=meta["pn"][0], clight=meta["pn"][1])
    integ = ig.Integrator(meta["eta"], 0.0, ps, meta["method"], **kw)
    keep = set(int(k) for k in z[name + ":state_steps"])
    taus, states, energies, tsteps = [], [], [], []

    def record(p):
        taus.append(float(p.tstep[0]))
        if len(taus) - 1 in keep:
            states.append(np.stack([getattr(p, f).cpu().numpy() for f in STATE]))
        energies.append(p.energy())
        if name == "c1_hermite4":
            p.set_tstep(1.0 / 64)
            tsteps.append(np.stack([p.tstep.cpu().numpy(),
                                    p.tstepij.cpu().numpy()]))
            p.tstep.fill_(taus[-1])
    E0 = ps.energy()
    integ.evolve(meta["t_end"], callback=record)
    assert np.array_equal(np.array(taus), z[name + ":tau"]), name
    ref = z[name + ":state"]
    got = np.stack(states)
    scale = np.abs(ref).max(axis=(0, 2), keepdims=True)
    tol = 1e-10 if "sakura" in name else 1e-11
    err = (np.abs(got - ref) / scale).max()
    print("[traj] %s state err %.3e" % (name, err), flush=True)
    assert err < tol, (name, err)
    eref = z[name + ":energy"]
    e = np.abs(np.array([E0] + energies) - eref) / abs(eref[0])
    assert e.max() < 1e-12, (name, e.max())
    if tsteps:
        tr = z[name + ":tstep"]
        assert np.allclose(np.stack(tsteps), tr, rtol=1e-12, atol=0)
    if meta["pn"]:
        pref = z[name + ":pn"]
        pgot = np.stack([ps.field(f).cpu().numpy() for f in oi.PN_FIELDS])
        pscale = np.abs(pref).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(pgot - pref) / pscale).max() < 1e-10


IEEE_CASES = ("c1_hermite4", "ahermite4", "hermite2", "hermite6", "hermite8",
              "sia21s_kdk", "sia21a_dkd", "sia21s_kdk_pn")


@pytest.mark.parametrize("name", IEEE_CASES)
def test_trajectory_bitwise_in_ieee_mode(torch, name):
    """With the pair kernels in the reference's arithmetic (fp64_ieee) and
    the update kernels evaluating the reference's numpy expressions, a run
    reproduces the reference-kernel trajectory bit for bit: every step
    length, every recorded state, (C1) every step's Tstep outputs and (PN)
    the PN accumulators.  (Sakura, whose Kepler solve has no bit-exact
    mode, is covered to round-off above.)"""
    from tupan_amd import integrator as ig
    z = np.load(TRAJ, allow_pickle=False)
    meta = [c for c in json.loads(str(z["manifest"]))["cases"]
            if c["name"] == name][0]
    print("[traj-ieee] %s n=%d method=%s" % (name, meta["n"], meta["method"]),
          flush=True)
    init = z[name + ":init"]
    ps = ig.System.from_host({f: init[k] for k, f in enumerate(BASE)})
    kw = {}
    if meta["pn"]:
        kw = dict(pn_order=meta["pn"][0], clight=meta["pn"][1])
    integ = ig.Integrator(meta["eta"], 0.0, ps, meta["method"],
                          mode="fp64_ieee", **kw)
    keep = set(int(k) for k in z[name + ":state_steps"])
    taus, states, tsteps = [], [], []

    def record(p):
        taus.append(float(p.tstep[0]))
        if len(taus) - 1 in keep:
            states.append(np.stack([getattr(p, f).cpu().numpy()
                                    for f in STATE]))
        if name == "c1_hermite4":
            p.set_tstep(1.0 / 64)
            tsteps.append(np.stack([p.tstep.cpu().numpy(),
                                    p.tstepij.cpu().numpy()]))
            p.tstep.fill_(taus[-1])
    integ.evolve(meta["t_end"], callback=record)
    assert np.array_equal(np.array(taus), z[name + ":tau"]), name
    got, ref = np.stack(states), z[name + ":state"]
    diff = int(np.sum(got.view(np.uint64) != ref.view(np.uint64)))
    print("[traj-ieee] %s: %d of %d state values differ" % (name, diff,
                                                            got.size),
          flush=True)
    assert diff == 0, name
    if tsteps:
        assert np.array_equal(np.stack(tsteps), z[name + ":tstep"])
    if meta["pn"]:
        pgot = np.stack([ps.field(f).cpu().numpy() for f in oi.PN_FIELDS])
        assert np.array_equal(pgot.view(np.uint64),
                              z[name + ":pn"].view(np.uint64))


def test_hybrid_pn_sakura_matches_checker(torch):
    """Compact PN cluster (N=64, masses U[1,30] normalised, half-mass
    radius 0.1, unsoftened, pn_order 5 at c = 100) with close pairs
    (r < 0.02) drifted by Kepler: device driver vs the checker's scheme on
    the C kernels, 16 steps of 1/4096."""
    from tupan_amd import integrator as ig
    from tupan_amd.ics import UniformMass, make_plummer, scale_radius
    from oracle import Oracle
    K = oi.Kernels(Oracle("float64", "port"))
    p = make_plummer(64, 0.0, imf=UniformMass(1.0, 30.0), seed=4,
                     phi_fn=lambda q: K.o.call("phi_kernel", q, q)[0])
    scale_radius(p, 0.1 / (1.3048 * 3 * np.pi / 16))
    init = {f: np.array(getattr(p, f)) for f in BASE}
    rcut, tau, nstep = 0.02, 1.0 / 4096, 16
    x = np.stack([init["rx"], init["ry"], init["rz"]])
    r2 = ((x[:, :, None] - x[:, None, :])**2).sum(0)
    assert ((r2 > 0) & (r2 < rcut**2)).sum() >= 2
    ref = {f: v.copy() for f, v in init.items()}
    for f in oi.PN_FIELDS:
        ref[f] = np.zeros(64)
    ref["t_curr"] = 0.0
    ref["tstep"] = np.zeros(64)
    K.set_clight(5, 100.0)
    for _ in range(nstep):
        oi.hybrid_step(K, ref, tau, rcut)
    ps = ig.System.from_host(init)
    integ = ig.Integrator(tau, 0.0, ps, "sakura_pn", rcut=rcut, pn_order=5,
                          clight=100.0)
    integ.evolve(nstep * tau)
    assert int(ps.nstep[0]) == nstep
    got = ps.to_host(STATE + oi.PN_FIELDS)
    for f in STATE:
        err = np.abs(got[f] - ref[f]).max() / np.abs(ref[f]).max()
        assert err < 1e-11, (f, err)
    for f in oi.PN_FIELDS:
        scale = np.abs(ref[f]).max() + 1e-300
        assert np.abs(got[f] - ref[f]).max() < 1e-9 * scale, f
