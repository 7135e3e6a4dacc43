"""The configured multi-step workloads at their own sizes, on the GPU.

* Ragged ensemble: 512 Plummer clusters (seed 2, N_c log-uniform in
  [64, 4096], equal masses summing to 1 per cluster, eps^2 = 1e-4), every
  cluster on the reference's adaptive Hermite-4 (eta = 0.02, t_end = 1).
  A few ensemble steps, replayed as a HIP graph; every 16th cluster then
  equals the single-system FP64 device driver run alone for the same number
  of its own steps: within 1e-11 of each component's scale, and within 1e-7
  with the ensemble's corrector passes in FP32C (same step counts).
* Compact post-Newtonian cluster: N = 2048 (seed 4, masses uniform in
  [1, 30] normalised to 1, half-mass radius 0.1, unsoftened), pn_order 5
  (1PN + 2PN + 2.5PN) at c = 1e3, pairs closer than 1e-3 drifted by the
  Sakura Kepler solve: a few steps of 1/8192 against the checker's scheme
  on the C kernels (oracle/integrators.hybrid_step), eagerly and replayed
  as HIP graphs (the replay equal to the eager run bit for bit).
"""
import numpy as np
import pytest

from oracle import Oracle
from oracle import integrators as oi

pytestmark = pytest.mark.gpu

STATE = ("rx", "ry", "rz", "vx", "vy", "vz")
BASE = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")


def _scale_err(a, b):
    return max(np.abs(a[f] - b[f]).max() / np.abs(b[f]).max() for f in STATE)


@pytest.fixture(scope="module")
def c3():
    from tupan_amd import ensemble, kernels
    sizes = ensemble.cluster_sizes(512, seed=2)
    phi = lambda p: kernels.compute("phi", p, p, mode="fp64")[0]
    fields = ensemble.make_ensemble(sizes, seed=2, eps2=1e-4, phi_fn=phi)
    print("[c3] %d clusters, %d particles" % (len(sizes), sizes.sum()),
          flush=True)
    return sizes, fields


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-11), ("fp32c", 1e-7)])
def test_c3_512_clusters_match_single_cluster_driver(c3, mode, tol):
    from tupan_amd import ensemble
    from tupan_amd import integrator as ig
    sizes, fields = c3
    steps = 6
    ens = ensemble.Ensemble(sizes, 0.02, 1.0, mode=mode)
    ens.set_state(fields)
    ens.step(steps, graph=True)
    left, pairs, err = ens.poll()
    st, tc, ns = ens.get_state()
    ens.close()
    assert err == 0 and left == len(sizes)
    assert np.all(ns == steps)
    off = np.concatenate([[0], np.cumsum(sizes)])
    worst = 0.0
    for c in range(0, len(sizes), 16):
        sl = slice(off[c], off[c + 1])
        ps = ig.System.from_host({f: fields[f][sl] for f in BASE})
        integ = ig.Integrator(0.02, 0.0, ps, "ahermite4")
        for _ in range(steps):
            integ.evolve_step(1.0)
        assert abs(float(ps.time[0]) - tc[c]) <= 1e-15 * max(1.0, tc[c]), c
        e = _scale_err({f: st[f][sl] for f in STATE}, ps.to_host(STATE))
        worst = max(worst, e)
        assert e < tol, (mode, c, sizes[c], e)
    print("[c3] %s: %d clusters checked, worst %.3g, %.3e pairs"
          % (mode, len(range(0, len(sizes), 16)), worst, pairs), flush=True)


def _c5_model():
    from tupan_amd import kernels
    from tupan_amd.ics import UniformMass, make_plummer, scale_radius
    phi = lambda p: kernels.compute("phi", p, p, mode="fp64")[0]
    p = make_plummer(2048, 0.0, imf=UniformMass(1.0, 30.0), seed=4,
                     phi_fn=phi)
    scale_radius(p, 0.1 / (1.3048 * 3 * np.pi / 16))
    return {f: np.array(getattr(p, f)) for f in BASE}


def test_c5_pn_cluster_matches_checker_eager_and_graph():
    from tupan_amd import integrator as ig
    init = _c5_model()
    rcut, tau, nstep, c = 1e-3, 1.0 / 8192, 4, 1e3
    x = np.stack([init["rx"], init["ry"], init["rz"]])
    r2 = ((x[:, :, None] - x[:, None, :])**2).sum(0)
    close = int(((r2 > 0) & (r2 < rcut**2)).sum()) // 2
    print("[c5] %d pairs closer than rcut at the start" % close, flush=True)
    K = oi.Kernels(Oracle("float64", "port"))
    K.set_clight(5, c)
    ref = {f: v.copy() for f, v in init.items()}
    for f in oi.PN_FIELDS:
        ref[f] = np.zeros(2048)
    ref["t_curr"] = 0.0
    ref["tstep"] = np.zeros(2048)
    for _ in range(nstep):
        oi.hybrid_step(K, ref, tau, rcut)
    runs = {}
    for graph in (False, True):
        ps = ig.System.from_host(init)
        integ = ig.Integrator(tau, 0.0, ps, "sakura_pn", rcut=rcut,
                              pn_order=5, clight=c, graph=graph)
        integ.evolve(nstep * tau)
        assert int(ps.nstep[0]) == nstep
        runs[graph] = ps.to_host(STATE + oi.PN_FIELDS)
    for f in STATE + oi.PN_FIELDS:
        assert np.array_equal(runs[True][f].view(np.uint64),
                              runs[False][f].view(np.uint64)), f
    got = runs[False]
    for f in STATE:
        err = np.abs(got[f] - ref[f]).max() / np.abs(ref[f]).max()
        assert err < 1e-11, (f, err)
    for f in oi.PN_FIELDS:
        scale = np.abs(ref[f]).max() + 1e-300
        assert np.abs(got[f] - ref[f]).max() < 1e-9 * scale, f
