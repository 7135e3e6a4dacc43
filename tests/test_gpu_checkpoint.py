"""Resume from a checkpoint (tupan_amd.checkpoint) on the device.

For each method a run of 10 steps is compared with a run that stops after
4 steps, writes a checkpoint, is rebuilt from the file (a new System and a
new integrator) and takes the remaining 6 steps: the two end states are the
same bits -- positions, velocities, derivatives, PN accumulators, clocks and
step counters -- since every kernel is deterministic.  Graph runs resume
into graph runs (their graphs are captured again) and into eager runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz")


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
    ("hermite4", {}, 1.0 / 64, False, None),
    ("ahermite4", {}, 1.0 / 32, False, None),
    ("hermite4", {"graph": True}, 1.0 / 64, False, None),
    ("hermite4", {"graph": True}, 1.0 / 64, False, False),
    ("sakura_pn", {"rcut": 0.02, "pn_order": 5, "clight": 100.0},
     1.0 / 4096, True, None),
    ("anreg", {}, 1.0 / 32, False, None),
    ("nreg", {"graph": True}, 1.0 / 32, False, None),
    ("nreg", {"graph": True}, 1.0 / 32, False, False),
]


@pytest.mark.parametrize("method,kw,eta,compact,regraph", CASES,
                         ids=["%s-%d" % (c[0], k) for k, c in enumerate(CASES)])
def test_resume_is_bitwise(torch, tmp_path, method, kw, eta, compact,
                           regraph):
    from tupan_amd import checkpoint
    from tupan_amd.integrator import Integrator, System
    d = _ics(96 if compact else 128, 41, 0.01, compact)
    t_end = 10.0
    full = Integrator(eta, 0.0, System.from_host(d), method, **kw)
    for _ in range(10):
        full.evolve_step(t_end)
    part = Integrator(eta, 0.0, System.from_host(d), method, **kw)
    for _ in range(4):
        part.evolve_step(t_end)
    path = str(tmp_path / "run.npz")
    checkpoint.save(path, part, snap_number=4)
    del part
    resumed = checkpoint.load(path, graph=regraph)
    for _ in range(6):
        resumed.evolve_step(t_end)
    torch.cuda.synchronize()
    a, b = full.ps, resumed.ps
    assert b.t_curr == a.t_curr
    for k, v in a._f.items():
        w = b._f[k]
        assert np.array_equal(v.cpu().numpy().view(np.uint64)
                              if v.dtype == torch.float64
                              else v.cpu().numpy(),
                              w.cpu().numpy().view(np.uint64)
                              if w.dtype == torch.float64
                              else w.cpu().numpy()), k
