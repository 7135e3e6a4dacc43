"""Sakura restricted to close pairs (sakura_near) on the device.

* against the checker's sakura_near (the reference's C pair kernel on
  every pair with r < rcut, self pairs included) on clusters with planted
  close pairs, exact duplicates (the zero state the kernel skips, which the
  reference evaluates to exactly zero) and sizes that are not multiples of
  the 32-record tile or the 64-lane wave: within 1e-11 of each output's
  scale (host and device run different sin/cos/sinh/cosh);
* with a cut-off above every separation, sakura_near evaluates every pair
  through the near mask and equals the full Sakura kernel (which runs the
  dense loop) bit for bit: the mask selects exactly the contributing pairs
  and keeps their order;
* a cut-off of zero selects nothing: zero outputs, clean status;
* the pairs are selected by the classification positions (cx, cy, cz),
  not by the current ones: classification positions spread far apart
  select no close pair (sakura_near zero, acc_far the full Acc), and
  classification positions all equal select every pair (sakura_near the
  full Sakura, acc_far zero);
* acc_far_kick_pn_half (far-pair Acc with kick_pn's first half kick in the
  same launches) gives the bits of acc_far followed by kick_pn_half.
The tests other than the classification test pass the current positions
as classification positions.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F8 = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")
OUT = ("drx", "dry", "drz", "dvx", "dvy", "dvz")


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


@pytest.fixture(scope="module")
def K():
    from oracle import Oracle
    from oracle import integrators as oi
    return oi.Kernels(Oracle("float64", "port"))


def _cluster(n, seed, close=6, dup=2):
    rng = np.random.RandomState(seed)
    d = {"mass": rng.uniform(0.5, 1.5, n) / n,
         "eps2": np.zeros(n)}
    for c in "xyz":
        d["r" + c] = rng.standard_normal(n)
        d["v" + c] = 0.3 * rng.standard_normal(n)
    # close pairs: j placed at distance ~1e-3 of i
    for k in range(close):
        i, j = 2 * k, 2 * k + 1
        off = rng.standard_normal(3)
        off *= rng.uniform(2e-4, 8e-4) / np.linalg.norm(off)
        for q, c in enumerate("xyz"):
            d["r" + c][j] = d["r" + c][i] + off[q]
            d["v" + c][j] = d["v" + c][i] + 0.05 * rng.standard_normal()
    # exact duplicates of particle 20 + k
    for k in range(dup):
        for f in F8:
            d[f][n - 1 - k] = d[f][20 + k]
    return d


def _dev(torch, d, scale=1.0):
    """Device fields, with classification positions = scale * positions."""
    t = {f: torch.as_tensor(np.ascontiguousarray(d[f]), device="cuda")
         for f in F8}
    for c in "xyz":
        t["c" + c] = t["r" + c] if scale == 1.0 else t["r" + c] * scale
    return t


def _near(torch, t, dt, flag, rcut):
    from tupan_amd import device
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    out = device.compute("sakura_near", t, t, dt, flag, rcut, status=st)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out], int(st.item())


@pytest.mark.parametrize("n,flag", [(77, 1), (300, -1), (1000, 1)])
def test_sakura_near_matches_checker(torch, K, n, flag):
    from oracle import integrators as oi
    d = _cluster(n, n, close=6, dup=2)
    rcut, dt = 1e-3, 1.0 / 512
    got, status = _near(torch, _dev(torch, d), dt, flag, rcut)
    assert status == 0
    ref = {f: d[f].copy() for f in F8}
    for c in "xyz":
        ref["c" + c] = ref["r" + c].copy()
    oi.sakura_near(K, ref, dt, flag, rcut)
    touched = 0
    for k, f in enumerate(OUT):
        scale = np.abs(ref[f]).max()
        assert scale > 0, f
        assert np.abs(got[k] - ref[f]).max() <= 1e-11 * scale, f
        touched = max(touched, int((ref[f] != 0).sum()))
    assert touched >= 12  # the planted pairs moved, both members
    for k in range(2):    # duplicates: zero delta from their twin
        i = 20 + k
        assert all(got[q][n - 1 - k] == got[q][i] for q in range(6))


@pytest.mark.parametrize("n", [64, 333, 2048])
def test_sakura_near_all_pairs_equals_full_sakura(torch, n):
    from tupan_amd import device
    d = _cluster(n, 7 + n, close=4, dup=2)
    t = _dev(torch, d)
    dt = 1.0 / 1024
    got, status = _near(torch, t, dt, 1, 1e150)
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    full = device.compute("sakura", t, t, dt, 1, status=st)
    torch.cuda.synchronize()
    assert status == 0 and int(st.item()) == 0
    for k in range(6):
        a = full[k].cpu().numpy()
        assert np.array_equal(a.view(np.uint64), got[k].view(np.uint64)), k


def test_sakura_near_zero_cut_selects_nothing(torch):
    d = _cluster(500, 3)
    got, status = _near(torch, _dev(torch, d), 1.0 / 512, 1, 0.0)
    assert status == 0
    for k in range(6):
        assert not np.any(got[k])


def test_selection_follows_the_classification_positions(torch):
    from tupan_amd import device
    d = _cluster(700, 11, close=6, dup=2)
    rcut, dt = 1e-3, 1.0 / 512
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    full_s = device.compute("sakura", _dev(torch, d), _dev(torch, d), dt, 1,
                            status=st)
    full_a = device.compute("acc", _dev(torch, d), _dev(torch, d))
    # classification positions 1e6 times further apart: no close pair
    far = _dev(torch, d, scale=1e6)
    got, status = _near(torch, far, dt, 1, rcut)
    assert status == 0 and not any(np.any(g) for g in got)
    a = device.compute("acc_far", far, far, rcut)
    for x, y in zip(a, full_a):
        assert np.allclose(x.cpu().numpy(), y.cpu().numpy(), rtol=1e-12,
                           atol=0)
    # classification positions all equal: every pair is close
    same = _dev(torch, d, scale=0.0)
    got, status = _near(torch, same, dt, 1, rcut)
    torch.cuda.synchronize()
    assert status == 0 and int(st.item()) == 0
    for k in range(6):
        a_ = full_s[k].cpu().numpy()
        assert np.array_equal(a_.view(np.uint64), got[k].view(np.uint64)), k
    a = device.compute("acc_far", same, same, rcut)
    assert not any(bool(x.any()) for x in a)


@pytest.mark.parametrize("n", [1000, 2048, 70000])
def test_acc_far_kick_half_is_bitwise_the_two_calls(torch, n):
    from tupan_amd import _lib
    rng = np.random.RandomState(n)
    f5 = [torch.as_tensor(x, device="cuda") for x in (
        rng.uniform(1, 2, n) / n, rng.standard_normal(n),
        rng.standard_normal(n), rng.standard_normal(n), np.full(n, 1e-6))]
    v0 = [torch.as_tensor(rng.standard_normal(n), device="cuda")
          for _ in range(3)]
    w = [torch.as_tensor(rng.standard_normal(n), device="cuda")
         for _ in range(3)]
    f5 = f5 + f5[1:4]  # classification positions: the positions
    ptr = lambda ts: [t.data_ptr() for t in ts]
    rcut, tau = 0.05, 1.0 / 1024
    a1 = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3)]
    v1 = [t.clone() for t in v0]
    _lib.call_update("acc_far", n, ptr(f5), n, ptr(f5), rcut, ptr(a1), None)
    _lib.call_update("kick_pn_half", n, ptr(v1), ptr(a1), ptr(w), tau, None)
    a2 = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3)]
    v2 = [t.clone() for t in v0]
    _lib.call_update("acc_far_kick_pn_half", n, ptr(f5), rcut, ptr(a2),
                     ptr(v2), ptr(w), tau, None)
    torch.cuda.synchronize()
    for x, y in zip(a1 + v1, a2 + v2):
        assert np.array_equal(x.cpu().numpy().view(np.uint64),
                              y.cpu().numpy().view(np.uint64))
    assert not np.array_equal(v2[0].cpu().numpy(), v0[0].cpu().numpy())
