"""Checkpoint files (tupan_amd.checkpoint) on the host: every field and the
integrator's state survive a save/load round trip exactly, the file loads
with allow_pickle=False, and the snapshot keys follow the reference's
Snapshot[_NNNNNN]/<system>/<member> naming (tupan/io/hdf5io.py:49-64).
Resuming a device run bit for bit is tested in test_gpu_checkpoint.py."""
import json

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tupan_amd import checkpoint  # noqa: E402
from tupan_amd import integrator as ig  # noqa: E402


def _system(n, seed):
    rng = np.random.RandomState(seed)
    s = ig.System(n, "cpu")
    for f in ig.BASE + ig.PN_FIELDS + ("ax", "ay", "az", "tstep", "time"):
        s.field(f).copy_(torch.as_tensor(rng.standard_normal(n)))
    s.nstep.copy_(torch.as_tensor(rng.randint(0, 1000, n)))
    return s


@pytest.mark.parametrize("method,kw", [
    ("hermite4", {}),
    ("sakura_pn", {"rcut": 0.02, "pn_order": 5, "clight": 100.0}),
    ("asakura", {"mode": "fp64_ieee"}),
    ("anreg", {}),
])
def test_round_trip(tmp_path, method, kw):
    s = _system(37, 1)
    integ = ig.Integrator(1.0 / 64, 0.25, s, method, **kw)
    it = integ.integrator
    if method == "anreg":
        it.E0, it.W, it.U, it.S = -0.25, 0.5, 0.49, 0.51
        it.is_initialized = True
    it.status.fill_(1)
    path = str(tmp_path / "ck.npz")
    checkpoint.save(path, integ, snap_number=12)
    with np.load(path, allow_pickle=False) as z:
        keys = set(z.files)
        state = json.loads(str(z[checkpoint.STATE_KEY]))
    assert "Snapshot_000012/system/rx" in keys
    assert state["method"] == method
    back = checkpoint.load(path, device="cpu")
    b = back.integrator
    assert b.method == method and b.eta == it.eta
    assert b.ps.t_curr == 0.25 and b.ps.mode == s.mode
    assert b.ps.include_pn_corrections == s.include_pn_corrections
    assert int(b.status.item()) == 1
    for k, v in s._f.items():
        assert torch.equal(b.ps._f[k], v), k
    if method == "sakura_pn":
        assert b.rcut == 0.02 and b.clight.pn_order == 5
        assert b.clight.clight == 100.0
    if method == "anreg":
        assert (b.E0, b.W, b.U, b.S) == (-0.25, 0.5, 0.49, 0.51)
        assert b.is_initialized
