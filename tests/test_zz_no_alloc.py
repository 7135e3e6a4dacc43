"""tupan_acc_jerk_static_dev, the AccJerk entry point that allocates
nothing, against tupan_acc_jerk_dev in FP64 mode, through the C ABI.

Three calls in a row on the same device, each smaller than the last:
N = 65536 (records and the default 16 block rows of partials, 52 MB,
exceed the 32 MiB static array, so the j split shrinks: the sums regroup,
within 1e-13 of the largest value), N = 8192 (both fit: the same launch
shape, the same bits) and N = 3000 (records built in the kernel: the same
bits).
A call whose j records alone exceed the array is refused before any
device work.
"""
import ctypes

import numpy as np
import pytest

from tupan_amd import _lib, spec

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p


def _model(n, seed):
    rng = np.random.RandomState(seed)
    h = {f: rng.normal(0, 1, n) for f in spec.F8}
    h["mass"] = rng.uniform(0.1, 1.0, n) / n
    h["eps2"] = np.full(n, 1e-4)
    return h


def _ptrs(ts):
    return (P * len(ts))(*[t.data_ptr() for t in ts])


def test_static_scratch_matches_the_allocating_entry_point():
    import torch
    L = _lib.lib()
    L.tupan_acc_jerk_static_dev.argtypes = [ctypes.c_size_t, P,
                                            ctypes.c_size_t, P, P, P]
    L.tupan_acc_jerk_dev.argtypes = [ctypes.c_int, ctypes.c_size_t, P,
                                     ctypes.c_size_t, P, P, P]
    st = torch.cuda.current_stream().cuda_stream
    for n, seed, bitwise in ((65536, 1, False), (8192, 2, True),
                             (3000, 3, True)):
        print("[no-alloc] n = %d" % n, flush=True)
        h = _model(n, seed)
        d = [torch.tensor(h[f], dtype=torch.float64, device="cuda")
             for f in spec.F8]
        ref = [torch.empty(n, dtype=torch.float64, device="cuda")
               for _ in range(6)]
        got = [torch.full((n,), float("nan"), dtype=torch.float64,
                          device="cuda") for _ in range(6)]
        ip = _ptrs(d)
        assert L.tupan_acc_jerk_dev(_lib.MODE_FP64, n, ip, n, ip, _ptrs(ref),
                                    P(st)) == _lib.OK
        rc = L.tupan_acc_jerk_static_dev(n, ip, n, ip, _ptrs(got), P(st))
        assert rc == _lib.OK, L.tupan_last_error()
        torch.cuda.synchronize()
        for g, r in zip(got, ref):
            if bitwise:
                assert torch.equal(g, r), n
            else:
                rel = (g - r).abs().max() / r.abs().max()
                assert float(rel) <= 1e-13, (n, float(rel))
    # j records alone beyond the 32 MiB array: refused, nothing launched
    big = 600000
    rc = L.tupan_acc_jerk_static_dev(1, _ptrs(d), big, _ptrs(d),
                                     _ptrs(got), P(st))
    assert rc == _lib.ERR_ARG
    assert b"static scratch" in L.tupan_last_error()
