"""The per-stream scratch cache of the device entry points.

* growing and shrinking calls on one stream (the cache grows, later smaller
  calls reuse the larger buffers) give the same bits as calls on a fresh
  stream;
* two streams working at once each use their own buffers;
* host threads calling on the same stream at once: one takes the cached
  buffers, the others allocate their own (stream-ordered), and every
  result is right;
* a call that would have to grow a buffer while its stream is being
  captured fails with a message, before launching anything, and the
  capture ends cleanly; after the same call has run once eagerly on that
  stream it captures, and the replays match the eager result;
* tupan_release_scratch frees the cache, and later calls rebuild it and
  give the same bits.
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F8 = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def _set(torch, n, seed):
    rng = np.random.RandomState(seed)
    d = {f: rng.standard_normal(n) for f in F8}
    d["mass"] = np.abs(d["mass"]) / n
    d["eps2"] = np.full(n, 1e-4)
    return {f: torch.as_tensor(v, device="cuda") for f, v in d.items()}


def _acc_jerk(torch, d, mode, stream=None):
    from tupan_amd import device
    st = stream.cuda_stream if stream is not None else None
    out = device.compute("acc_jerk", d, d, mode=mode, stream=st)
    return out


def _bits(ts):
    return [t.cpu().numpy().view(np.uint64) for t in ts]


@pytest.mark.parametrize("mode", ["fp64", "fp32c", "fp64_ieee"])
def test_grow_and_shrink_on_one_stream(torch, mode):
    sizes = (1000, 5000, 1000, 300, 5000, 7001)
    sets = {n: _set(torch, n, n) for n in set(sizes)}
    s = torch.cuda.Stream()
    got = []
    with torch.cuda.stream(s):
        for n in sizes:
            got.append(_bits(_acc_jerk(torch, sets[n], mode, s)))
    s.synchronize()
    for n, g in zip(sizes, got):
        fresh = torch.cuda.Stream()   # a stream with no cached buffers
        with torch.cuda.stream(fresh):
            want = _bits(_acc_jerk(torch, sets[n], mode, fresh))
        fresh.synchronize()
        for a, b in zip(g, want):
            assert np.array_equal(a, b), (mode, n)


def test_two_streams_at_once(torch):
    a, b = _set(torch, 6000, 1), _set(torch, 4000, 2)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    ref_a = _bits(_acc_jerk(torch, a, "fp64"))
    ref_b = _bits(_acc_jerk(torch, b, "fp64"))
    outs = []
    for _ in range(4):
        with torch.cuda.stream(s1):
            oa = _acc_jerk(torch, a, "fp64", s1)
        with torch.cuda.stream(s2):
            ob = _acc_jerk(torch, b, "fp64", s2)
        outs.append((oa, ob))
    torch.cuda.synchronize()
    for oa, ob in outs:
        assert all(np.array_equal(x, y) for x, y in zip(_bits(oa), ref_a))
        assert all(np.array_equal(x, y) for x, y in zip(_bits(ob), ref_b))


def test_host_threads_on_one_stream(torch):
    sets = [_set(torch, 3000 + 500 * k, 10 + k) for k in range(4)]
    refs = [_bits(_acc_jerk(torch, d, "fp64")) for d in sets]
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    results, errors = [None] * 4, []

    def work(k):
        try:
            outs = []
            for _ in range(5):
                outs.append(_acc_jerk(torch, sets[k], "fp64", s))
            results[k] = outs
        except Exception as e:  # reported below
            errors.append(repr(e))

    th = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    s.synchronize()
    assert not errors, errors
    for k in range(4):
        for o in results[k]:
            assert all(np.array_equal(x, y) for x, y in zip(_bits(o), refs[k]))


def test_capture_needs_an_eager_call_first(torch):
    from tupan_amd import _lib, device
    d = _set(torch, 2500, 7)
    s = torch.cuda.Stream()
    out = tuple(torch.empty(2500, dtype=torch.float64, device="cuda")
                for _ in range(6))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with pytest.raises(_lib.TupanError, match="capture"):
        with torch.cuda.graph(g, stream=s):
            device.compute("acc_jerk", d, d, mode="fp64", out=out,
                           stream=s.cuda_stream)
    # the same call eagerly on s, then captured and replayed
    with torch.cuda.stream(s):
        device.compute("acc_jerk", d, d, mode="fp64", out=out,
                       stream=s.cuda_stream)
    s.synchronize()
    want = _bits(out)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        device.compute("acc_jerk", d, d, mode="fp64", out=out,
                       stream=s.cuda_stream)
    for _ in range(3):
        for o in out:
            o.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert all(np.array_equal(x, y) for x, y in zip(_bits(out), want))


def test_release_then_reuse(torch):
    from tupan_amd import _lib
    d = _set(torch, 4000, 9)
    before = [_bits(_acc_jerk(torch, d, m)) for m in ("fp64", "fp32c",
                                                       "fp64_ieee")]
    torch.cuda.synchronize()
    _lib.release_scratch()
    for _ in range(2):
        after = [_bits(_acc_jerk(torch, d, m)) for m in ("fp64", "fp32c",
                                                          "fp64_ieee")]
        for x, y in zip(before, after):
            assert all(np.array_equal(a, b) for a, b in zip(x, y))
