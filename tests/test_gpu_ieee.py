"""The reference-arithmetic mode (fp64_ieee) on the GPU: bit for bit.

Phi, Acc, AccJerk, SnapCrackle, Tstep, PNAcc and NREG X/V in fp64_ieee
perform the reference C's
operations in its order with its rounding (no FMA contraction, correctly
rounded division and sqrt, j summed in ascending order), so their outputs
must equal the reference C's bits exactly:

* against the golden fixtures (reference C outputs, edge cases included);
* against the CPU checker (bit-identical to the reference C) on fresh
  inputs with ni != nj and sizes straddling the wave, tile and block
  boundaries;
* through the libtupan.h drop-in with the library default set to it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import Oracle  # noqa: E402
from tupan_amd import _lib, kernels, spec  # noqa: E402

IEEE_KERNELS = ("phi", "acc", "acc_jerk", "snap_crackle", "tstep", "pnacc",
                "nreg_x", "nreg_v")
NREG_V_FIELDS = ("mass", "vx", "vy", "vz", "ax", "ay", "az")
FIELDS = {"phi": spec.F5, "acc": spec.F5, "acc_jerk": spec.F8,
          "snap_crackle": spec.F14, "tstep": spec.F8, "pnacc": spec.F8,
          "nreg_x": spec.F8, "nreg_v": NREG_V_FIELDS}
ORACLE_NAME = {"nreg_x": "nreg_Xkernel", "nreg_v": "nreg_Vkernel"}
# every PN order term active (7: 1PN .. 3.5PN) at c = 10
_INV = tuple(10.0**-k for k in range(1, 8))
ARGS = {"tstep": ((0.03125,), (0.03125,)),
        "pnacc": ((7, _INV), (7,) + _INV),
        "nreg_x": ((0.01,), (0.01,)), "nreg_v": ((0.01,), (0.01,))}


@pytest.fixture(scope="module", autouse=True)
def device_present():
    n = _lib.device_count()
    assert n >= 1, "no HIP device: %s" % _lib.lib().tupan_last_error()


def same_bits(a, b):
    """Equal bit patterns, or NaN in both."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    nan = np.isnan(a)
    if not np.array_equal(nan, np.isnan(b)):
        return False
    return np.array_equal(a[~nan].view(np.uint64), b[~nan].view(np.uint64))


def test_golden_bitwise(golden):
    n = 0
    for case in golden.cases:
        if case["prec"] != "float64" or case["kernel"] not in IEEE_KERNELS:
            continue
        ip, jp = golden.inputs(case)
        got = kernels.compute(case["kernel"], ip, jp, *golden.args(case),
                              mode="fp64_ieee")
        for g, r in zip(got, golden.outputs(case)):
            assert same_bits(g, r), case
        n += 1
    assert n > 0


def _random_set(n, seed):
    rng = np.random.RandomState(seed)
    d = {f: rng.standard_normal(n) for f in spec.F14}
    d["mass"] = rng.uniform(0.1, 2.0, n) / n
    d["eps2"] = rng.uniform(0.0, 1e-3, n)
    d["eps2"][::7] = 0.0
    # a coincident pair: r^2 = 0 off the diagonal
    for f in ("rx", "ry", "rz"):
        d[f][5] = d[f][3]
    return d


@pytest.mark.parametrize("name", IEEE_KERNELS)
def test_fresh_inputs_bitwise_vs_checker(name):
    """Sizes straddle the 64-record j tile, the 4-pair groups, the 64-lane
    wave and the 4-wave block (idle waves in the last block)."""
    o = Oracle("float64", "port")
    d = _random_set(2113, 11)
    for ni, nj in ((1, 2113), (63, 65), (65, 63), (257, 2111), (300, 4),
                   (2113, 1), (2113, 2113), (128, 0), (129, 127)):
        ip = {f: d[f][:ni] for f in FIELDS[name]}
        jp = {f: d[f][2113 - nj:] for f in FIELDS[name]}
        args, oargs = ARGS.get(name, ((), ()))
        got = kernels.compute(name, ip, jp, *args, mode="fp64_ieee")
        ref = o.call(ORACLE_NAME.get(name, name + "_kernel"), ip, jp, *oargs)
        for g, r in zip(got, ref):
            assert same_bits(g, r), (name, ni, nj)


def test_large_n_rows_bitwise():
    """N = 32768 self-gravity: a sample of rows against the checker."""
    from tupan_amd.ics import make_plummer
    phi = lambda ps: kernels.compute("phi", ps, ps, mode="fp64")[0]
    ps = make_plummer(32768, 1.0 / 32768, seed=9, phi_fn=phi)
    d = {f: np.ascontiguousarray(getattr(ps, f)) for f in spec.F8}
    got = kernels.compute("acc_jerk", d, d, mode="fp64_ieee")
    rows = np.sort(np.random.RandomState(0).choice(32768, 256, replace=False))
    ref = Oracle("float64", "port").call(
        "acc_jerk_kernel", {f: v[rows] for f, v in d.items()}, d)
    for g, r in zip(got, ref):
        assert same_bits(g[rows], r)


def test_dropin_default_mode_reproduces_reference_bits(golden):
    from test_gpu_parity import _call_dropin, _load_dropin
    lib = _load_dropin("64")
    _lib.check(_lib.lib().tupan_set_default_mode(_lib.MODE_FP64_IEEE))
    try:
        n = 0
        for case in golden.cases:
            if (case["prec"] != "float64"
                    or case["kernel"] not in IEEE_KERNELS):
                continue
            ip, jp = golden.inputs(case)
            got = _call_dropin(lib, case["kernel"], ip, jp, golden.args(case),
                               np.float64)
            for g, r in zip(got, golden.outputs(case)):
                assert same_bits(g, r), case
            n += 1
        assert n > 0
    finally:
        _lib.check(_lib.lib().tupan_set_default_mode(_lib.MODE_FP64))
