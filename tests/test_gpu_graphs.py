"""Fixed-step integrators replayed as HIP graphs (Integrator(..., graph=True)).

Each method runs twice from the same initial state: eagerly, and with every
step of a repeated length captured once and replayed.  Between the replays
the program does eager work of its own on the same system (a report: a Phi
evaluation, device sums, host reads that synchronise), and after every step
the graph run's state equals the eager run's bit for bit -- the first step
(eager on the capture stream), the second (captured and replayed) and the
later replays alike.  The last step of each run is shorter (t_end is not a
multiple of the step), so that step runs eagerly beside the graphs.
Adaptive methods choose each block step eagerly (a host read) and replay
the graph captured for that step length.

NREG (nreg, anreg) replays with its scalars (E0, W, U, S, the physical
clock) in device memory: after every step its state and physical time equal
the host-driven run's bit for bit, and in the reference-arithmetic mode the
checker's NREG on the reference's C kernels.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz")
STATE = ("rx", "ry", "rz", "vx", "vy", "vz")
PN = ("pn_ke", "pn_mrx", "pn_mry", "pn_mrz", "pn_mvx", "pn_mvy", "pn_mvz",
      "pn_amx", "pn_amy", "pn_amz")


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def _ics(n, seed, soft, compact=False):
    from tupan_amd import kernels
    from tupan_amd.ics import UniformMass, make_plummer, scale_radius
    phi = lambda p: kernels.compute("phi", p, p, mode="fp64")[0]
    if compact:
        p = make_plummer(n, 0.0, imf=UniformMass(1.0, 30.0), seed=seed,
                         phi_fn=phi)
        scale_radius(p, 0.1 / (1.3048 * 3 * np.pi / 16))
    else:
        p = make_plummer(n, soft, seed=seed, phi_fn=phi)
    return {f: np.ascontiguousarray(getattr(p, f), np.float64) for f in BASE}


CASES = [
    ("hermite4", dict(), 1.0 / 64, False),
    ("hermite8", dict(mode="fp64"), 1.0 / 64, False),
    ("hermite4", dict(mode="fp32c"), 1.0 / 64, False),
    ("hermite4", dict(mode="fp64_ieee"), 1.0 / 64, False),
    ("sia21s.kdk", dict(), 1.0 / 64, False),
    ("sia21s.dkd", dict(pn_order=5, clight=100.0), 1.0 / 4096, True),
    ("sakura", dict(), 1.0 / 64, False),
    ("sakura_pn", dict(rcut=0.02, pn_order=5, clight=100.0), 1.0 / 4096,
     True),
]


@pytest.mark.parametrize("method,kw,tau,compact", CASES,
                         ids=["%s-%d" % (c[0], k) for k, c in enumerate(CASES)])
def test_graph_replay_equals_eager(torch, method, kw, tau, compact):
    from tupan_amd.integrator import Integrator, System
    n = 64 if compact else 256
    d = _ics(n, 31, 0.01, compact)
    eager_s, graph_s = System.from_host(d), System.from_host(d)
    eager = Integrator(tau, 0.0, eager_s, method, **kw)
    graph = Integrator(tau, 0.0, graph_s, method, graph=True, **kw)
    nstep = 7
    t_end = (nstep - 0.5) * tau
    fields = STATE + (PN if "pn_order" in kw else ())
    for k in range(nstep):
        print("[graphs] %s %s step %d" % (method, kw, k), flush=True)
        eager.evolve_step(t_end)
        graph.evolve_step(t_end)
        # the program's own eager work between replays, on the same system
        rep = graph_s.diagnostics()
        assert np.isfinite(rep["kinetic_energy"])
        torch.cuda.synchronize()
        a, b = eager_s.to_host(fields), graph_s.to_host(fields)
        for f in fields:
            assert np.array_equal(a[f].view(np.uint64),
                                  b[f].view(np.uint64)), (k, f)
        assert graph_s.t_curr == eager_s.t_curr, k
        assert int(graph_s.nstep[0]) == k + 1
        assert float(graph_s.time[0]) == float(eager_s.time[0])
    assert abs(graph_s.t_curr) >= t_end
    assert len(graph.integrator._graphs) == 1
    graph.integrator.check_status()


ADAPTIVE = [("ahermite4", 1.0 / 32, 0.01), ("sia21a.kdk", 1.0 / 32, 0.01),
            ("asakura", 1.0 / 32, 0.01)]


@pytest.mark.parametrize("method,eta,soft", ADAPTIVE,
                         ids=[c[0] for c in ADAPTIVE])
def test_adaptive_graph_replay_equals_eager(torch, method, eta, soft):
    """Adaptive methods: the block step is chosen eagerly every step (one
    host read of the smallest time step), and the step of that length is
    replayed from the graph captured for it."""
    from tupan_amd.integrator import Integrator, System
    d = _ics(128, 32, soft)
    eager_s, graph_s = System.from_host(d), System.from_host(d)
    eager = Integrator(eta, 0.0, eager_s, method)
    graph = Integrator(eta, 0.0, graph_s, method, graph=True)
    taus = []
    for k in range(12):
        print("[graphs] %s step %d" % (method, k), flush=True)
        t0 = eager_s.t_curr
        eager.evolve_step(10.0)
        graph.evolve_step(10.0)
        taus.append(eager_s.t_curr - t0)
        graph_s.diagnostics()
        torch.cuda.synchronize()
        a, b = eager_s.to_host(STATE), graph_s.to_host(STATE)
        for f in STATE:
            assert np.array_equal(a[f].view(np.uint64),
                                  b[f].view(np.uint64)), (k, f)
        assert graph_s.t_curr == eager_s.t_curr, k
        assert np.array_equal(graph_s.tstep.cpu().numpy(),
                              eager_s.tstep.cpu().numpy()), k
    # a step length met twice has been captured and replayed
    assert len(set(taus)) < len(taus)
    assert graph.integrator._graphs
    graph.integrator.check_status()


@pytest.mark.parametrize("method,mode", [("anreg", "fp64"), ("nreg", "fp64"),
                                         ("nreg", "fp64_ieee")])
def test_nreg_graph_equals_host_driven(torch, method, mode):
    from tupan_amd.integrator import Integrator, System
    d = _ics(200, 14, 1e-3)
    host_s, graph_s = System.from_host(d), System.from_host(d)
    host = Integrator(1.0 / 32, 0.0, host_s, method, mode=mode)
    graph = Integrator(1.0 / 32, 0.0, graph_s, method, mode=mode, graph=True)
    for k in range(6):
        print("[graphs] %s %s step %d" % (method, mode, k), flush=True)
        host.evolve_step(10.0)
        graph.evolve_step(10.0)
        graph_s.diagnostics()  # eager work between replays
        a, b = host_s.to_host(STATE), graph_s.to_host(STATE)
        for f in STATE:
            assert np.array_equal(a[f].view(np.uint64),
                                  b[f].view(np.uint64)), (k, f)
        assert graph_s.t_curr == host_s.t_curr, k
        for f in ("tstep", "time"):
            assert np.array_equal(getattr(graph_s, f).cpu().numpy(),
                                  getattr(host_s, f).cpu().numpy()), (k, f)
        assert int(graph_s.nstep[0]) == k + 1
    assert len(graph.integrator._graphs) == 1


def test_nreg_graph_bitwise_vs_checker_in_ieee_mode(torch):
    from oracle import Oracle
    from oracle import integrators as oi
    from tupan_amd.integrator import Integrator, System
    K = oi.Kernels(Oracle("float64", "port"))
    d = _ics(128, 15, 1e-3)
    s = System.from_host(d)
    integ = Integrator(1.0 / 32, 0.0, s, "anreg", mode="fp64_ieee",
                       graph=True)
    host = dict(d)
    host["t_curr"] = 0.0
    st = {}
    oi.nreg_initialize(K, host, st)
    for k in range(5):
        integ.evolve_step(10.0)
        oi.nreg_do_step(K, host, st, 1.0 / 32, "anreg")
        got = s.to_host(STATE)
        for f in STATE:
            assert np.array_equal(got[f].view(np.uint64),
                                  host[f].view(np.uint64)), (k, f)
        assert s.t_curr == host["t_curr"], k
