"""FP32C accuracy on the configured single-evaluation workloads, and the
sorted path's own cases, on the GPU.

Workloads: Plummer models with N = 16384, 65536 and 262144 (seed 1,
Salpeter masses on [0.1, 100] normalised to 1, eps = 4/N, Henon units).
The bar is the benchmark's: per particle, |x32 - x64| / |x64| <= 1e-6 for
the acceleration and the jerk, for every particle, against

* the FP64 kernel on every row (itself pinned to the reference's C
  kernels by tests/test_gpu_parity.py), and
* the checker's C kernels (the reference's own C when oracle/_ref was
  built, else the bit-identical restatement) on sampled rows: random rows
  plus the rows where FP32C and FP64 differ most.

FP32C calls with nj > 4096 run over Morton-sorted particles with FP64
recomputation of the rows whose estimated error exceeds 0.7 |result|
(tupan_set_flag_tau); the remaining tests cover that path's edges: the
threshold set to 0 (every row through the FP64 recomputation, also past
the rows its grid holds), i and j different arrays, coincident particles,
sizes around the threshold, the unsorted engine, run-to-run bit identity
and graph replay.
"""
import numpy as np
import pytest

import oracle
from oracle import Oracle
from tupan_amd import _lib, kernels, spec

pytestmark = pytest.mark.gpu

C2_SIZES = (16384, 65536, 262144)
_MODELS = {}


def c2_model(n):
    """The C2 Plummer model of size n (host arrays), built once."""
    if n not in _MODELS:
        from tupan_amd.ics import PowerLaw, make_plummer
        phi = lambda ps: kernels.compute("phi", ps, ps, mode="fp64")[0]
        ps = make_plummer(n, 4.0 / n, imf=PowerLaw(2.35, 0.1, 100.0), seed=1,
                          phi_fn=phi)
        _MODELS[n] = {f: np.ascontiguousarray(getattr(ps, f))
                      for f in spec.F8}
    return _MODELS[n]


def checker():
    kind = "reference" if oracle.available("float64", "reference") else "port"
    return Oracle("float64", kind)


def rel_rows(got, ref, group):
    """|got - ref| / |ref| per row over the components in group."""
    g = np.stack([np.asarray(got[k], np.float64) for k in group])
    r = np.stack([np.asarray(ref[k], np.float64) for k in group])
    return np.linalg.norm(g - r, axis=0) / np.linalg.norm(r, axis=0)


@pytest.fixture(autouse=True)
def defaults():
    yield
    _lib.set_sorted(True)
    _lib.set_flag_tau(0.7)
    _lib.set_near_field(5.0)


@pytest.mark.parametrize("n", C2_SIZES)
def test_c2_acc_jerk_fp32c_within_1e6(n):
    import torch
    from tupan_amd import device
    h = c2_model(n)
    d = device.to_device(h, spec.F8)
    a64 = [t.cpu().numpy() for t in device.compute("acc_jerk", d, d,
                                                    mode="fp64")]
    a32 = [t.cpu().numpy() for t in device.compute("acc_jerk", d, d,
                                                    mode="fp32c")]
    flagged = _lib.last_flagged()
    torch.cuda.synchronize()
    ea, ej = rel_rows(a32, a64, (0, 1, 2)), rel_rows(a32, a64, (3, 4, 5))
    print("N=%d fp32c vs fp64: acc max %.3g rms %.3g, jerk max %.3g rms %.3g,"
          " %d rows recomputed in FP64" % (
              n, ea.max(), np.sqrt(np.mean(ea**2)), ej.max(),
              np.sqrt(np.mean(ej**2)), flagged))
    assert ea.max() <= 1e-6 and ej.max() <= 1e-6
    assert np.sqrt(np.mean(ej**2)) <= 1e-7
    assert 0 < flagged <= 0.01 * n
    # the checker's C kernels on random rows and the worst rows
    rng = np.random.RandomState(11)
    rows = np.unique(np.concatenate([rng.choice(n, 48, replace=False),
                                     np.argsort(ej)[-32:],
                                     np.argsort(ea)[-16:]]))
    hi = {f: np.ascontiguousarray(v[rows]) for f, v in h.items()}
    ref = checker().call("acc_jerk_kernel", hi, h)
    got = [v[rows] for v in a32]
    e64 = [v[rows] for v in a64]
    for grp in ((0, 1, 2), (3, 4, 5)):
        assert rel_rows(got, ref, grp).max() <= 1e-6, grp
        assert rel_rows(e64, ref, grp).max() <= 1e-11, grp


def test_c2_16k_every_kernel_both_modes():
    """N = 16384: Phi, Acc, AccJerk and SnapCrackle (its a and j inputs from
    the FP64 AccJerk) in FP64 against the checker on sampled rows (1e-11
    per particle), and FP32C against FP64 on every row: Phi, acceleration
    and jerk within 1e-6; snap within 1e-6 RMS and 2e-5 worst, crackle
    within 1e-4 RMS and 3e-3 worst (third derivatives whose pair terms
    cancel far more than the jerk's; FP32 arithmetic does not resolve
    them, which is why Hermite 6/8 want the FP64 mode)."""
    n = 16384
    h = dict(c2_model(n))
    aj = kernels.compute("acc_jerk", h, h, mode="fp64")
    h.update(zip(("ax", "ay", "az", "jx", "jy", "jz"), aj))
    rows = np.sort(np.random.RandomState(3).choice(n, 64, replace=False))
    hi = {f: np.ascontiguousarray(v[rows]) for f, v in h.items()}
    ck = checker()
    for name, sym in (("phi", "phi_kernel"), ("acc", "acc_kernel"),
                      ("acc_jerk", "acc_jerk_kernel"),
                      ("snap_crackle", "snap_crackle_kernel")):
        f = spec.KERNELS[name].fields
        i64 = {k: hi[k] for k in f}
        j64 = {k: h[k] for k in f}
        g64 = kernels.compute(name, j64, j64, mode="fp64")
        g32 = kernels.compute(name, j64, j64, mode="fp32c")
        ref = ck.call(sym, i64, j64)
        groups = [(0,)] if name == "phi" else [(0, 1, 2), (3, 4, 5)][
            :len(g64) // 3]
        for grp in groups:
            e = rel_rows([v[rows] for v in g64], ref, grp).max()
            assert e <= 1e-11, (name, grp, e)
        for q, grp in enumerate(groups):
            e = rel_rows(g32, g64, grp)
            mx, rms = e.max(), np.sqrt(np.mean(e**2))
            print("N=16384 %s group %d fp32c vs fp64: max %.3g rms %.3g"
                  % (name, q, mx, rms))
            if name == "snap_crackle":
                assert rms <= (1e-6, 1e-4)[q] and mx <= (2e-5, 3e-3)[q]
            else:
                assert mx <= 1e-6, (name, q, mx)


@pytest.mark.parametrize("ni,nj", [(8192, 8192), (20000, 4160)])
def test_flag_threshold_zero_recomputes_every_row_in_fp64(ni, nj):
    """tau = 0: every row with an FP32 tile is recomputed in FP64, so
    FP32C equals the FP64 kernel to FP64 rounding (sums in another order).
    8192 rows fit the recomputation grid's chunked rows (64 per block of
    16 * recompute_groups(nj)); with nj = 4160 it holds 17408 rows, so
    2592 of the 20000 are evaluated whole by the finishing kernel."""
    h = c2_model(16384)
    hi = {f: v[:ni].copy() for f, v in h.items()} if ni <= 16384 else \
        {f: np.concatenate([v, v[:ni - 16384] * (1.0 + 1e-3)])
         for f, v in h.items()}
    hj = {f: v[:nj].copy() for f, v in h.items()}
    hj = hi if (ni == nj) else hj
    ref = kernels.compute("acc_jerk", hi, hj, mode="fp64")
    _lib.set_flag_tau(0.0)
    try:
        got = kernels.compute("acc_jerk", hi, hj, mode="fp32c")
        flagged = _lib.last_flagged()
    finally:
        _lib.set_flag_tau(0.7)
    assert flagged == ni, flagged
    for grp in ((0, 1, 2), (3, 4, 5)):
        assert rel_rows(got, ref, grp).max() <= 1e-12, grp


def test_sorted_path_other_sets_duplicates_and_reproducibility():
    """i and j different arrays (sorted separately), coincident particles
    (their r^2 = 0 pairs masked as the reference masks them), and two runs
    of the same call bit for bit; the unsorted engine agrees within the
    two paths' errors."""
    n = 9000
    h = {f: v[:n].copy() for f, v in c2_model(16384).items()}
    # 64 particles placed exactly on others, with their own velocities
    src = np.arange(64) * 7
    dst = n - 1 - np.arange(64)
    for f in ("rx", "ry", "rz"):
        h[f][dst] = h[f][src]
    ck = checker()
    rows = np.unique(np.concatenate([src, dst, np.arange(0, n, 97)]))
    for ni in (n, 4321, 1):
        ip = {f: v[:ni].copy() for f, v in h.items()}
        got = kernels.compute("acc_jerk", ip, h, mode="fp32c")
        again = kernels.compute("acc_jerk", ip, h, mode="fp32c")
        for a, b in zip(got, again):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
        r = rows[rows < ni]
        ref = ck.call("acc_jerk_kernel",
                      {f: v[r] for f, v in ip.items()}, h)
        for grp in ((0, 1, 2), (3, 4, 5)):
            e = rel_rows([v[r] for v in got], ref, grp)
            assert np.all(np.isfinite(e)) and e.max() <= 1e-6, (ni, grp)
    _lib.set_sorted(False)
    unsorted = kernels.compute("acc_jerk", h, h, mode="fp32c")
    _lib.set_sorted(True)
    got = kernels.compute("acc_jerk", h, h, mode="fp32c")
    for grp in ((0, 1, 2), (3, 4, 5)):
        assert rel_rows(unsorted, got, grp).max() <= 5e-6, grp


@pytest.mark.parametrize("where", ["everywhere", "half"])
def test_sorted_path_nonuniform_softening(where):
    """Per-particle softenings (the general far loop, which adds each j's
    e2; a tile whose j's share one e2 takes a loop with the sum formed once
    per tile) against the checker, for every FP32C kernel: random e2
    everywhere, or only for x > 0, so that a call mixes both kinds of far
    tiles."""
    n = 8192
    h = {f: v[:n].copy() for f, v in c2_model(16384).items()}
    scale = np.random.RandomState(5).uniform(0.25, 4.0, n)
    if where == "half":
        scale = np.where(h["rx"] > 0.0, scale, 1.0)
    h["eps2"] = h["eps2"] * scale
    rows = np.arange(0, n, 37)
    ck = checker()
    for name, sym in (("phi", "phi_kernel"), ("acc", "acc_kernel"),
                      ("acc_jerk", "acc_jerk_kernel")):
        f = spec.KERNELS[name].fields
        got = kernels.compute(name, {k: h[k] for k in f},
                              {k: h[k] for k in f}, mode="fp32c")
        ref = ck.call(sym, {k: h[k][rows] for k in f}, {k: h[k] for k in f})
        groups = [(0,)] if name == "phi" else [(0, 1, 2), (3, 4, 5)][
            :len(got) // 3]
        for grp in groups:
            e = rel_rows([v[rows] for v in got], ref, grp).max()
            assert e <= 1e-6, (name, grp, e)


@pytest.mark.parametrize("nj", [4097, 4127, 4128, 4129, 12345])
@pytest.mark.parametrize("ni", [1, 63, 65, 1000])
def test_sorted_path_sizes(ni, nj):
    """Sizes around the sorted path's threshold and tile boundaries, for
    every FP32C kernel, against the checker."""
    h = c2_model(16384)
    jp = {f: v[:nj].copy() for f, v in h.items()}
    ip = {f: v[-ni:].copy() for f, v in h.items()}
    ck = checker()
    for name, sym in (("phi", "phi_kernel"), ("acc", "acc_kernel"),
                      ("acc_jerk", "acc_jerk_kernel")):
        f = spec.KERNELS[name].fields
        got = kernels.compute(name, {k: ip[k] for k in f},
                              {k: jp[k] for k in f}, mode="fp32c")
        ref = ck.call(sym, {k: ip[k] for k in f}, {k: jp[k] for k in f})
        groups = [(0,)] if name == "phi" else [(0, 1, 2), (3, 4, 5)][
            :len(got) // 3]
        for grp in groups:
            e = rel_rows(got, ref, grp).max()
            assert e <= 1e-6, (name, ni, nj, grp, e)


def test_sorted_path_graph_replay_matches_eager():
    """A sorted FP32C call captured into a HIP graph (it records no
    memset) replays to the eager call's bits, replay after replay, with
    eager work and synchronisation between the replays."""
    import torch
    from tupan_amd import device
    h = c2_model(16384)
    d = device.to_device(h, spec.F8)
    eager = [t.clone() for t in device.compute("acc_jerk", d, d,
                                                mode="fp32c")]
    out = tuple(torch.empty_like(eager[0]) for _ in range(6))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        device.compute("acc_jerk", d, d, mode="fp32c", out=out)  # scratch
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        device.compute("acc_jerk", d, d, mode="fp32c", out=out)
    for _ in range(3):
        for o in out:
            o.zero_()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(out, eager):
            assert torch.equal(a, b)
        # eager work between replays
        device.compute("acc", d, d, mode="fp32c")
        torch.cuda.synchronize()
