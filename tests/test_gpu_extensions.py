"""The Python drop-in on the HIP backend, on the GPU.

``extensions.<Kernel>("HIP", prec).calc(ips, jps, *args)`` -- the call
surface the reference's integrators use (tupan/lib/extensions.py:94-97,
118-137, 424-433, 462-507, 642-646) -- for all ten libtupan.h kernels, in
the reference's two I/O precisions, on every golden case: the outputs
land in the particle sets' attributes and equal, bit for bit, the array
API's results for the same inputs (FP64 I/O) or those results rounded to
float32 (FP32 I/O), and agree with the reference's own outputs in the
fixtures.  PNAcc takes its order and 1/c^k from the module-level
``clight``, Tstep its eta, Sakura dt and flag, NREG dt, and Kepler updates
particles 0 and 1 in place.  Outputs that are strided views are written
through a contiguous buffer, leaving their neighbours alone.

Also the error paths: a Sakura pair whose Kepler sub-stepping exhausts the
solve budget (reported, results finite), and NaN outputs reported by the
device status word and by the drop-in libraries' last status.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT
from tupan_amd import _lib, extensions, kernels, spec
from tupan_amd.particles import ParticleSystem

pytestmark = pytest.mark.gpu

CLASSES = {"phi": extensions.Phi, "acc": extensions.Acc,
           "acc_jerk": extensions.AccJerk,
           "snap_crackle": extensions.SnapCrackle,
           "tstep": extensions.Tstep, "pnacc": extensions.PNAcc,
           "sakura": extensions.Sakura, "nreg_x": extensions.NREG_X,
           "nreg_v": extensions.NREG_V, "kepler": extensions.Kepler}


def particles(d, dtype, fields, outs=()):
    """A ParticleSystem holding the given fields as `dtype` arrays, with
    float32 output attributes pre-registered for FP32 I/O."""
    n = len(d["mass"])
    ps = ParticleSystem(n)
    for f in fields:
        # copies: Kepler updates its particles in place
        setattr(ps, f, np.array(d[f], dtype=dtype, copy=True))
    if dtype == np.float32:
        for o in outs:
            setattr(ps, o, np.zeros(n, np.float32))
    return ps


def ext_args(case):
    a = case["args"]
    k = case["kernel"]
    if k == "pnacc":
        extensions.clight.pn_order = int(a[0])
        extensions.clight.clight = 1.0 / a[1]
        c = extensions.clight
        return (), (int(a[0]), (c.inv1, c.inv2, c.inv3, c.inv4, c.inv5,
                                c.inv6, c.inv7))
    if k == "sakura":
        return (a[0], int(a[1])), (a[0], int(a[1]))
    return tuple(a), tuple(a)


@pytest.mark.parametrize("prec", ["float64", "float32"])
def test_every_kernel_through_the_extension_classes(golden, prec):
    dtype = np.float64 if prec == "float64" else np.float32
    seen = set()
    for case in golden.cases:
        if case["prec"] != "float64":
            continue
        k = case["kernel"]
        seen.add(k)
        ip, jp = golden.inputs(case, dtype)
        ref = golden.outputs(case)
        if k == "kepler":
            ps = particles(ip, dtype, spec.KEPLER_FIELDS)
            before = {f: getattr(ps, f).copy() for f in spec.KEPLER_FIELDS}
            rx = ps.rx
            extensions.Kepler("HIP", prec).calc(ps, ps, case["args"][0])
            assert ps.rx is rx  # in place
            want = kernels.kepler({f: np.asarray(before[f], np.float64)
                                   for f in spec.KEPLER_FIELDS},
                                  case["args"][0])
            for f, w, r in zip(spec.KEPLER_OUTS, want, ref):
                got = getattr(ps, f)
                assert np.array_equal(got[:2], w.astype(dtype)), (case, f)
                assert np.array_equal(got[2:], before[f][2:])
                if prec == "float64":
                    # the fixtures' bound: 1e-11 of the component's size
                    # (the Laguerre solve stops at a 2^-42 step)
                    assert (np.abs(got[:2] - r).max()
                            <= 1e-11 * (np.abs(r).max() + 1e-300)), (case, f)
            continue
        kk = spec.KERNELS[k]
        ips = particles(ip, dtype, kk.fields, kk.outs)
        jps = ips if (case["ipre"] is None and case["jpre"] is None) else \
            particles(jp, dtype, kk.fields)
        xargs, cargs = ext_args(case)
        res = CLASSES[k]("HIP", prec).calc(ips, jps, *xargs)
        assert len(res) == len(kk.outs)
        for o, r in zip(kk.outs, res):
            assert r is getattr(ips, o)  # results are the attributes
        # the array API on the same (widened) inputs, in the mode of the
        # I/O precision: float32 calls default to the reference FP32
        # build's arithmetic where it exists
        mode = "fp32_ieee" if (prec == "float32"
                               and k in spec.FP32_IEEE) else "fp64"
        want = kernels.compute(k, {f: np.asarray(getattr(ips, f), np.float64)
                                   for f in kk.fields},
                               {f: np.asarray(getattr(jps, f), np.float64)
                                for f in kk.fields}, *cargs, mode=mode)
        for g, w in zip(res, want):
            assert np.array_equal(g, w.astype(dtype), equal_nan=True), case
        if prec == "float64":
            for g, r in zip(res, ref):
                s = np.max(np.abs(r)) + 1e-300
                assert np.nanmax(np.abs(g - r)) <= 1e-9 * s, case
    assert seen == set(CLASSES)


def test_strided_outputs_leave_neighbours_alone(golden):
    case = golden.find(kernel="acc_jerk", set="P", prec="float64")[0]
    d = golden.set("P")
    n = len(d["mass"])
    ps = particles(d, np.float64, spec.F8)
    backing = {o: np.full(2 * n, 7.0) for o in spec.KERNELS["acc_jerk"].outs}
    for o, b in backing.items():
        setattr(ps, o, b[::2])  # strided float64 views
    extensions.AccJerk("HIP", "float64").calc(ps, ps)
    want = kernels.compute("acc_jerk", ps, ps)
    for o, w in zip(spec.KERNELS["acc_jerk"].outs, want):
        assert np.array_equal(backing[o][::2], w)
        assert np.all(backing[o][1::2] == 7.0)
    for o, r in zip(spec.KERNELS["acc_jerk"].outs, golden.outputs(case)):
        assert np.allclose(backing[o][::2], r, rtol=1e-10, atol=0)


def test_sakura_budget_exhaustion_is_reported_with_finite_results(golden):
    """The softened tight binary of the fixtures (set K1, particles 0 and 1:
    its energy-check sub-stepping needs ~2^19 Kepler solves at dt = 1/64 in
    the reference itself) under a budget of one solve per pair: the host
    call raises NotConverged after writing finite results, and the device
    call ORs bit 1 into the status word.  The budget is restored after."""
    import torch
    from tupan_amd import device
    d = golden.set("K1")
    h = {f: np.ascontiguousarray(d[f][:2]) for f in spec.F8}
    try:
        _lib.set_sakura_budget(1)
        out = tuple(np.empty(2) for _ in range(6))
        with pytest.raises(_lib.NotConverged):
            kernels.compute("sakura", h, h, 1.0 / 64, -1, out=out)
        assert all(np.all(np.isfinite(o)) for o in out)
        dd = device.to_device(h, spec.F8)
        st = torch.zeros(1, dtype=torch.int32, device="cuda")
        got = device.compute("sakura", dd, dd, 1.0 / 64, -1, status=st)
        torch.cuda.synchronize()
        assert int(st.item()) == _lib.STATUS_NOT_CONVERGED
        assert all(bool(torch.isfinite(g).all()) for g in got)
    finally:
        _lib.set_sakura_budget(1 << 20)
    # a pair that needs one solve is unaffected by the smallest budget
    _lib.set_sakura_budget(1)
    try:
        p = golden.set("P")
        h = {f: np.ascontiguousarray(p[f][:2]) for f in spec.F8}
        kernels.compute("sakura", h, h, 1.0 / 64, -1)
    finally:
        _lib.set_sakura_budget(1 << 20)


def test_nonfinite_outputs_are_reported():
    """A NaN mass makes the accelerations NaN: the device status word gets
    bit 2 (tupan_check_finite_dev), and the drop-in libraries record
    TUPAN_ERR_NONFINITE as their last status; clean inputs leave both
    clear."""
    import torch
    from tupan_amd import device
    rng = np.random.RandomState(2)
    h = {f: rng.normal(0, 1, 500) for f in spec.F8}
    h["mass"] = rng.uniform(0.5, 1.5, 500) / 500
    h["eps2"] = np.full(500, 1e-4)
    d = device.to_device(h, spec.F8)
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    for mode in ("fp64", "fp32c"):
        device.compute("acc", d, d, mode=mode, status=st)
    torch.cuda.synchronize()
    assert int(st.item()) == 0
    d["mass"][17] = float("nan")
    device.compute("acc", d, d, mode="fp64", status=st)
    torch.cuda.synchronize()
    assert int(st.item()) == _lib.STATUS_NONFINITE
    # the drop-in FP64 library
    lib = ctypes.CDLL(os.path.join(ROOT, "tupan_amd", "lib",
                                   "libtupan_f64.so"))
    P = ctypes.c_void_p
    ptr = lambda a: a.ctypes.data_as(P)
    n = ctypes.c_ulong(500)
    outs = [np.empty(500) for _ in range(3)]
    f5 = [h[f] for f in spec.F5]
    lib.acc_kernel(n, *[ptr(a) for a in f5], n, *[ptr(a) for a in f5],
                   *[ptr(o) for o in outs])
    assert lib.tupan_abi_last_status() == 0
    bad = dict(h)
    bad["mass"] = h["mass"].copy()
    bad["mass"][3] = float("nan")
    f5 = [bad[f] for f in spec.F5]
    lib.acc_kernel(n, *[ptr(a) for a in f5], n, *[ptr(a) for a in f5],
                   *[ptr(o) for o in outs])
    assert lib.tupan_abi_last_status() == 4  # TUPAN_ERR_NONFINITE
