"""The reference-arithmetic modes (fp64_ieee, fp32_ieee) on the GPU: bit
for bit.

Phi, Acc, AccJerk, SnapCrackle, Tstep, PNAcc and NREG X/V in fp64_ieee,
and Phi, Acc, AccJerk, SnapCrackle and Tstep in fp32_ieee, perform the
reference C's operations in its order with its rounding (no FMA
contraction, correctly rounded division and sqrt, j summed in ascending
order) in double or in float, so their outputs must equal the bits of the
reference's FP64 or FP32 build exactly:

* against the golden fixtures (reference C outputs, edge cases included);
* against the CPU checker (bit-identical to the reference C in both
  precisions) on fresh inputs with ni != nj and sizes straddling the wave,
  tile and block boundaries;
* through the libtupan.h drop-ins: libtupan_f64.so with the library
  default set to fp64_ieee, libtupan_f32.so in its own default mode;
* through the extension classes with float32 I/O.
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


# ---- the FP32 build ---------------------------------------------------

FP32_KERNELS = spec.FP32_IEEE


def _f32(d):
    return {f: np.ascontiguousarray(v, np.float32) for f, v in d.items()}


def test_golden_fp32_bitwise(golden):
    """The reference FP32 build's outputs in the fixtures (inputs rounded to
    float32 as it received them), equal in every bit."""
    n = 0
    for case in golden.cases:
        if case["prec"] != "float32" or case["kernel"] not in FP32_KERNELS:
            continue
        ip, jp = golden.inputs(case, np.float32)
        got = kernels.compute(case["kernel"], ip, jp, *golden.args(case),
                              mode="fp32_ieee")
        ref = golden.outputs(case)
        assert len(got) == len(ref)
        for g, r in zip(got, ref):
            assert r.dtype == np.float32
            # float results, widened exactly
            assert np.array_equal(g.astype(np.float32).astype(np.float64),
                                  g, equal_nan=True), case
            assert same_bits(g, r), case
        n += 1
    assert n >= 15


@pytest.mark.parametrize("name", FP32_KERNELS)
def test_fresh_inputs_fp32_bitwise_vs_checker(name):
    """The float32 checker (bit-identical to the reference FP32 build) on
    the sizes of the FP64 test: wave, group, tile and block boundaries."""
    o = Oracle("float32", "port")
    d = _f32(_random_set(2113, 12))
    for ni, nj in ((1, 2113), (63, 65), (65, 63), (257, 2111), (300, 4),
                   (2113, 1), (2113, 2113), (128, 0), (129, 127)):
        ip = {f: d[f][:ni] for f in FIELDS[name]}
        jp = {f: d[f][2113 - nj:] for f in FIELDS[name]}
        args, oargs = ARGS.get(name, ((), ()))
        got = kernels.compute(name, ip, jp, *args, mode="fp32_ieee")
        ref = o.call(name + "_kernel", ip, jp, *oargs)
        for g, r in zip(got, ref):
            assert same_bits(g, r), (name, ni, nj)


def test_fp32_large_n_rows_bitwise():
    """N = 32768 self-gravity in float: a sample of rows against the
    float32 checker."""
    from tupan_amd.ics import make_plummer
    phi = lambda ps: kernels.compute("phi", ps, ps, mode="fp64")[0]
    ps = make_plummer(32768, 1.0 / 32768, seed=9, phi_fn=phi)
    d = _f32({f: getattr(ps, f) for f in spec.F8})
    got = kernels.compute("acc_jerk", d, d, mode="fp32_ieee")
    rows = np.sort(np.random.RandomState(1).choice(32768, 256, replace=False))
    ref = Oracle("float32", "port").call(
        "acc_jerk_kernel", {f: v[rows] for f, v in d.items()}, d)
    for g, r in zip(got, ref):
        assert same_bits(g[rows], r)


def test_dropin_f32_reproduces_reference_fp32_bits(golden):
    """libtupan_f32.so in its default mode: the reference FP32 build's bits
    for the kernels that have the mode; FP64 rounded to float for the
    others (their bound is in test_gpu_parity)."""
    from test_gpu_parity import _call_dropin, _load_dropin
    lib = _load_dropin("32")
    assert lib.tupan_abi_get_mode() == _lib.MODE_FP32_IEEE
    n = 0
    for case in golden.cases:
        if case["prec"] != "float32" or case["kernel"] not in FP32_KERNELS:
            continue
        ip, jp = golden.inputs(case, np.float32)
        got = _call_dropin(lib, case["kernel"], ip, jp, golden.args(case),
                           np.float32)
        for g, r in zip(got, golden.outputs(case)):
            assert g.dtype == np.float32
            assert np.array_equal(g.view(np.uint32), r.view(np.uint32)), case
        n += 1
    assert n >= 15


def test_extension_classes_float32_reproduce_reference_fp32_bits(golden):
    """extensions.<Kernel>("HIP", "float32"): the float32 I/O path of the
    reference's integrators, default mode."""
    from tupan_amd import extensions
    from tupan_amd.particles import ParticleSystem
    classes = {"phi": extensions.Phi, "acc": extensions.Acc,
               "acc_jerk": extensions.AccJerk,
               "snap_crackle": extensions.SnapCrackle,
               "tstep": extensions.Tstep}
    assert extensions.default_mode("float32") == "fp32_ieee"
    n = 0
    for case in golden.cases:
        if case["prec"] != "float32" or case["kernel"] not in FP32_KERNELS:
            continue
        k = spec.KERNELS[case["kernel"]]
        ip, jp = golden.inputs(case, np.float32)

        def pset(d):
            ps = ParticleSystem(len(d["mass"]))
            for f in k.fields:
                setattr(ps, f, d[f].copy())
            return ps
        ips = pset(ip)
        for o in k.outs:
            setattr(ips, o, np.zeros(len(ip["mass"]), np.float32))
        jps = ips if (case["ipre"] is None and case["jpre"] is None) \
            else pset(jp)
        res = classes[case["kernel"]]("HIP", "float32").calc(
            ips, jps, *case["args"])
        for g, r in zip(res, golden.outputs(case)):
            assert g.dtype == np.float32
            assert np.array_equal(g.view(np.uint32), r.view(np.uint32)), case
        n += 1
    assert n >= 15
