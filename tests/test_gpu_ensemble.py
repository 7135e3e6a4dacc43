"""Ragged ensemble (ensemble.hip) against per-cluster integration.

Every cluster of the ensemble must follow the reference's adaptive
Hermite-4 on its own: the same number of steps to t_end and the same final
state as (a) the single-system device driver tupan_amd.integrator on that
cluster alone and (b) the checker's numpy restatement (oracle/integrators.py)
on the C kernels, for the smallest clusters.  States agree to 1e-11 of each
component's scale (the pair sums are split differently, so only rounding
differs).  The captured-graph replay and the launched-kernel path give
identical bits, and the autotuner's launch shapes agree to 1e-11.

FP32C mode (the Tstep + AccJerk start pass and the corrector AccJerk sums in
FP32C arithmetic): every cluster takes the same steps as in FP64 and ends
within 1e-7 of each component's scale; the
graph replay is bitwise the launched kernels; j-chunk lengths change only
the FP64 grouping of the tile sums (1e-11); bad modes and j chunks that are
not whole tiles are refused.  The device storage order (Morton or input
order) changes only the order of the pair sums (1e-11 in FP64; in FP32C
the near field moves with the tiles, so 1e-7); states come back in input
order.
"""
import numpy as np
import pytest

from oracle import Oracle
from oracle import integrators as oi

pytestmark = pytest.mark.gpu

STATE = ("rx", "ry", "rz", "vx", "vy", "vz")
BASE = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")
ETA, T_END = 0.02, 1.0 / 16


@pytest.fixture(scope="module")
def setup():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tupan_amd import ensemble, kernels
    sizes = ensemble.cluster_sizes(10, seed=3, nmin=16, nmax=300)
    phi = lambda p: kernels.compute("phi", p, p, mode="fp64")[0]
    fields = ensemble.make_ensemble(sizes, seed=5, eps2=1e-4, phi_fn=phi)
    return sizes, fields


def _run(sizes, fields, graph=True, launch=None, mode="fp64",
         order="morton"):
    from tupan_amd import ensemble
    ens = ensemble.Ensemble(sizes, ETA, T_END, mode=mode, order=order)
    if launch:
        ens.set_launch(*launch)
    ens.set_state(fields)
    stats = ens.run(batch=16, graph=graph)
    st, tc, ns = ens.get_state()
    ens.close()
    return st, tc, ns, stats


def _scale_err(a, b):
    return max(np.abs(a[f] - b[f]).max() / np.abs(b[f]).max() for f in STATE)


def test_ensemble_matches_single_cluster_driver(setup):
    from tupan_amd import integrator as ig
    sizes, fields = setup
    st, tc, ns, stats = _run(sizes, fields)
    print("[ens] steps %d pairs %.3e" % (stats["steps"], stats["pairs"]),
          flush=True)
    assert np.all(tc >= T_END)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for c in range(len(sizes)):
        sl = slice(off[c], off[c + 1])
        ps = ig.System.from_host({f: fields[f][sl] for f in BASE})
        integ = ig.Integrator(ETA, 0.0, ps, "ahermite4")
        count = [0]
        integ.evolve(T_END, callback=lambda p: count.__setitem__(0, count[0] + 1))
        assert count[0] == ns[c], (c, count[0], ns[c])
        ref = ps.to_host(STATE)
        err = _scale_err({f: st[f][sl] for f in STATE}, ref)
        assert err < 1e-11, (c, sizes[c], err)


def test_ensemble_matches_checker_on_small_clusters(setup):
    sizes, fields = setup
    st, tc, ns, _ = _run(sizes, fields)
    K = oi.Kernels(Oracle("float64", "port"))
    off = np.concatenate([[0], np.cumsum(sizes)])
    for c in np.argsort(sizes)[:3]:
        sl = slice(off[c], off[c + 1])
        ps = {f: fields[f][sl].copy() for f in BASE}
        ps["t_curr"] = 0.0
        ps["tstep"] = np.zeros(sizes[c])
        steps = 0
        while abs(ps["t_curr"]) < T_END:
            tau = oi.get_base_tstep(ps["t_curr"], T_END, ETA)
            oi.hermite_step(K, ps, tau, 4, ETA, adaptive=True)
            steps += 1
        assert steps == ns[c]
        err = _scale_err({f: st[f][sl] for f in STATE}, ps)
        assert err < 1e-11, (c, err)


def test_graph_replay_is_bitwise_the_launched_kernels(setup):
    sizes, fields = setup
    a = _run(sizes, fields, graph=True)
    b = _run(sizes, fields, graph=False)
    for f in STATE:
        assert np.array_equal(a[0][f], b[0][f])
    assert np.array_equal(a[2], b[2])


def test_launch_shapes_and_autotuner(setup, tmp_path, monkeypatch):
    from tupan_amd import ensemble
    sizes, fields = setup
    ref = _run(sizes, fields)
    for launch in ((2, 64, 4), (1, 1024, 16)):
        got = _run(sizes, fields, launch=launch)
        assert np.array_equal(got[2], ref[2])
        assert _scale_err(got[0], ref[0]) < 1e-11
    monkeypatch.setenv("TUPAN_TUNING_CACHE", str(tmp_path / "tune.json"))
    ens = ensemble.Ensemble(sizes, ETA, T_END)
    best, rates = ens.autotune(fields, trial_steps=4,
                               candidates=[(1, 256, 8), (2, 512, 8)])
    assert best in rates and len(rates) == 2
    best2, _ = ens.autotune(fields)          # from the cache
    assert best2 == best
    ens.run(batch=16)
    st, tc, ns = ens.get_state()
    ens.close()
    assert np.array_equal(ns, ref[2])
    assert _scale_err(st, ref[0]) < 1e-11


def test_fp32c_follows_fp64(setup):
    sizes, fields = setup
    ref = _run(sizes, fields)
    got = _run(sizes, fields, mode="fp32c")
    assert np.array_equal(got[2], ref[2])
    err = _scale_err(got[0], ref[0])
    print("[ens fp32c] scale error %.2e" % err, flush=True)
    assert 0.0 < err < 1e-7


def test_fp32c_graph_and_chunks(setup):
    sizes, fields = setup
    a = _run(sizes, fields, graph=True, mode="fp32c")
    b = _run(sizes, fields, graph=False, mode="fp32c")
    for f in STATE:
        assert np.array_equal(a[0][f], b[0][f])
    assert np.array_equal(a[2], b[2])
    for launch in ((2, 64, 4), (1, 1024, 16)):
        got = _run(sizes, fields, launch=launch, mode="fp32c")
        assert np.array_equal(got[2], a[2])
        assert _scale_err(got[0], a[0]) < 1e-11


def test_fp32c_arguments(setup):
    from tupan_amd import _lib, ensemble
    sizes, _ = setup
    ens = ensemble.Ensemble(sizes, ETA, T_END)
    with pytest.raises(ValueError):
        ens.set_mode("fp64_ieee")
    with pytest.raises(_lib.TupanError):
        _lib.check(_lib.lib().tupan_ensemble_set_mode(ens._h,
                                                      _lib.MODE_FP64_IEEE))
    ens.set_launch(1, 48, 8)
    with pytest.raises(_lib.TupanError, match="multiple"):
        ens.set_mode("fp32c")
    ens.set_launch(1, 64, 8)
    ens.set_mode("fp32c")
    with pytest.raises(_lib.TupanError, match="multiple"):
        ens.set_launch(1, 80, 8)
    ens.close()


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-11), ("fp32c", 1e-7)])
def test_storage_order(setup, mode, tol):
    sizes, fields = setup
    a = _run(sizes, fields, mode=mode)
    b = _run(sizes, fields, mode=mode, order="input")
    assert np.array_equal(a[2], b[2])
    err = _scale_err(a[0], b[0])
    assert 0.0 < err < tol, err
