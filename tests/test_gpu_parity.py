"""Parity of the HIP kernels with the reference, on the GPU.

Bars:
* FP64 mode vs the golden fixtures (reference C outputs): per particle,
  1e-12 relative, or a few ulps of the sum of the pair-term magnitudes
  where those cancel (FP64_TOL);
* FP32C mode vs the same fixtures: 1e-6 relative or ~3 FP32 ulps of that
  sum (FP32C_TOL);
* the FP32 drop-in library: at least as close to the FP64 truth as the
  reference's own FP32 build (within one float ulp of slack);
* larger sizes against the CPU checker, plus size-independent properties
  (Newton's third law, run-to-run bit reproducibility, launch-shape
  independence).  The benchmark sizes' FP32C accuracy (<= 1e-6 per
  particle) is tests/test_gpu_fp32c.py's.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import OUT_GROUPS, ROOT, bound_ratio, rel_norm_err

pytestmark = pytest.mark.gpu

from oracle import Oracle  # noqa: E402
from tupan_amd import _lib, kernels, spec  # noqa: E402

# (rel, abs): per particle and output vector, ||got - ref|| must stay below
# rel * ||ref|| + abs * S, S = sum_j ||f_ij|| from the fixtures (the
# condition of the sum).  FP64: 1e-12 relative where the sum is well
# conditioned and a few FP64 ulps of the absolute sum where the pair terms
# cancel.  Sakura: its Laguerre solve stops at a 2^-42 step, so pair terms
# agree to ~1e-13 of the orbit size and to ~1e-10 of their own (small,
# dt-proportional) size.  FP32C: 1e-6 relative / ~3 FP32 ulps of S.
FP64_TOL = {"phi": (1e-12, 1e-14), "acc": (1e-12, 1e-14),
            "acc_jerk": (1e-12, 1e-14), "snap_crackle": (1e-12, 1e-14),
            "pnacc": (1e-12, 1e-14), "sakura": (1e-12, 1e-10),
            "nreg_x": (1e-12, 1e-14), "nreg_v": (1e-12, 1e-14)}
# Kepler: the Laguerre solve stops at a 2^-42 step: new states agree to
# ~1e-13 of the orbit's size
KEPLER_TOL = 1e-11
TSTEP_TOL = 1e-12
FP32C_TOL = (1e-6, 2e-7)
# SnapCrackle's FP32 pair recurrences (alpha, beta, gamma) round ~10 times
# in series and cancel inside one pair (beta - 4 alpha^2): a single pair
# term carries up to ~2e-6 relative error
FP32C_TOL_SNAP = (4e-6, 1e-6)
FP32C_KERNELS = ("phi", "acc", "acc_jerk", "snap_crackle")


@pytest.fixture(scope="module", autouse=True)
def device_present():
    n = _lib.device_count()
    assert n >= 1, "no HIP device: %s" % _lib.lib().tupan_last_error()


def _err(kernel, got, ref):
    g = {k: v for k, v in enumerate(got)}
    r = {k: v for k, v in enumerate(ref)}
    return rel_norm_err(g, r, OUT_GROUPS[kernel])


def _sakura_ratio(got, ref, S, ip, jp, dt):
    """Sakura pair terms are differences of two propagated relative states
    (Kepler and free drift) of size ~|r0| + |v0| dt, so besides the
    FP64_TOL bound they may carry ~2^-42 of that size: T_i = sum_j
    m_j/(m_i+m_j) (|r0_ij| + |v0_ij| |dt|) for dr, sum_j w_ij |v0_ij| for
    dv.  (Fast unbound pairs make the true terms ~1e-17 from operands ~1,
    in the reference as here.)"""
    x = np.stack([ip[f] for f in ("rx", "ry", "rz")], 1)[:, None, :] - \
        np.stack([jp[f] for f in ("rx", "ry", "rz")], 1)[None, :, :]
    v = np.stack([ip[f] for f in ("vx", "vy", "vz")], 1)[:, None, :] - \
        np.stack([jp[f] for f in ("vx", "vy", "vz")], 1)[None, :, :]
    w = jp["mass"][None, :] / (ip["mass"][:, None] + jp["mass"][None, :])
    rn, vn = np.linalg.norm(x, axis=2), np.linalg.norm(v, axis=2)
    T = [np.sum(w * (rn + vn * abs(dt)), 1), np.sum(w * vn, 1)]
    rel, absol = FP64_TOL["sakura"]
    worst = 0.0
    for g, grp in enumerate(OUT_GROUPS["sakura"]):
        d = np.sqrt(sum((got[k] - ref[k])**2 for k in grp))
        r = np.sqrt(sum(ref[k]**2 for k in grp))
        worst = max(worst, float(np.max(
            d / (rel * r + absol * S[g] + 1e-12 * T[g] + 1e-300))))
    return worst


def test_golden_fp64(golden):
    worst = {}
    for case in golden.cases:
        if case["prec"] != "float64":
            continue
        k = case["kernel"]
        ip, jp = golden.inputs(case)
        ref = golden.outputs(case)
        if k == "kepler":
            got = kernels.kepler(ip, *golden.args(case))
            e = max(np.abs(g - r).max() / (np.abs(r).max() + 1e-300)
                    for g, r in zip(got, ref)) / KEPLER_TOL
            worst[k] = max(worst.get(k, 0.0), e)
            assert e <= 1.0, (case, e)
            continue
        got = kernels.compute(k, ip, jp, *golden.args(case), mode="fp64")
        if k == "tstep":
            e = _err(k, got, ref) / TSTEP_TOL
        elif k == "sakura":
            e = _sakura_ratio(got, ref, golden.abs_sums(case), ip, jp,
                              case["args"][0])
        else:
            rel, absol = FP64_TOL[k]
            e = bound_ratio(got, ref, golden.abs_sums(case), OUT_GROUPS[k],
                            rel, absol)
        worst[k] = max(worst.get(k, 0.0), e)
        assert e <= 1.0, (case, e)
    print("fp64 worst error/bound per kernel:", worst)


def test_golden_fp32c(golden):
    """FP32C against the reference outputs, bound FP32C_TOL (see above);
    the N = 65536 Plummer model is checked at 1e-6 strictly below."""
    worst = {}
    for case in golden.cases:
        if case["prec"] != "float64" or case["kernel"] not in FP32C_KERNELS:
            continue
        ip, jp = golden.inputs(case)
        got = kernels.compute(case["kernel"], ip, jp, mode="fp32c")
        tol = FP32C_TOL_SNAP if case["kernel"] == "snap_crackle" else FP32C_TOL
        e = bound_ratio(got, golden.outputs(case), golden.abs_sums(case),
                        OUT_GROUPS[case["kernel"]], *tol)
        key = (case["kernel"], case["set"])
        worst[key] = max(worst.get(key, 0.0), e)
        assert e <= 1.0, (case, e)
    print("fp32c worst error/bound:", worst)


def _load_dropin(bits):
    return ctypes.CDLL(os.path.join(ROOT, "tupan_amd", "lib",
                                    "libtupan_f%s.so" % bits))


DROPIN_SYMBOL = {"nreg_x": "nreg_Xkernel", "nreg_v": "nreg_Vkernel"}


def _call_dropin(lib, name, ip, jp, args, real):
    if name == "kepler":
        return _call_dropin_kepler(lib, ip, args[0], real)
    k = spec.KERNELS[name]
    ia = [np.ascontiguousarray(ip[f], real) for f in k.fields]
    ja = [np.ascontiguousarray(jp[f], real) for f in k.fields]
    ni, nj = len(ia[0]), len(ja[0])
    outs = [np.zeros(ni, real) for _ in k.outs]
    c_real = ctypes.c_double if real == np.float64 else ctypes.c_float
    c_uint = ctypes.c_ulong if real == np.float64 else ctypes.c_uint
    c_int = ctypes.c_long if real == np.float64 else ctypes.c_int
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    cargs = [c_uint(ni)] + [p(a) for a in ia] + [c_uint(nj)] + [p(a) for a in ja]
    if name == "tstep":
        cargs.append(c_real(args[0]))
    elif name == "pnacc":
        cargs += [c_uint(args[0])] + [c_real(x) for x in args[1]]
    elif name == "sakura":
        cargs += [c_real(args[0]), c_int(args[1])]
    elif name in ("nreg_x", "nreg_v"):
        cargs.append(c_real(args[0]))
    cargs += [p(o) for o in outs]
    fn = getattr(lib, DROPIN_SYMBOL.get(name, name + "_kernel"))
    fn.restype = None
    fn(*cargs)
    assert lib.tupan_abi_last_status() == 0
    return outs


def _call_dropin_kepler(lib, ps, dt, real):
    """In place, as the reference's Kepler extension calls it."""
    arrs = {f: np.ascontiguousarray(ps[f], real).copy()
            for f in spec.KEPLER_FIELDS}
    c_real = ctypes.c_double if real == np.float64 else ctypes.c_float
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    fn = lib.kepler_solver_kernel
    fn.restype = None
    fn(*([p(arrs[f]) for f in spec.KEPLER_FIELDS] + [c_real(dt)]
         + [p(arrs[f]) for f in spec.KEPLER_OUTS]))
    assert lib.tupan_abi_last_status() == 0
    return [arrs[f][:2] for f in spec.KEPLER_OUTS]


def test_dropin_f64_matches_c_abi(golden):
    lib = _load_dropin("64")
    for case in golden.cases:
        if case["prec"] != "float64" or case["ipre"] or case["jpre"]:
            continue
        ip, jp = golden.inputs(case)
        a = _call_dropin(lib, case["kernel"], ip, jp, golden.args(case),
                         np.float64)
        if case["kernel"] == "kepler":
            b = kernels.kepler(ip, *golden.args(case))
        else:
            b = kernels.compute(case["kernel"], ip, jp, *golden.args(case),
                                mode="fp64")
        for x, y in zip(a, b):
            assert np.array_equal(x, y, equal_nan=True), case


@pytest.mark.parametrize("abi_mode", [_lib.MODE_FP32_IEEE, _lib.MODE_FP64])
def test_dropin_f32_at_least_as_accurate_as_reference_f32(golden, abi_mode):
    """libtupan_f32.so against the FP64 truth: in its default mode (the
    reference FP32 build's arithmetic where it exists, FP64 rounded to float
    elsewhere) and with FP64 arithmetic throughout (tupan_abi_set_mode)."""
    lib = _load_dropin("32")
    assert lib.tupan_abi_set_mode(abi_mode) == 0
    try:
        _f32_vs_truth(lib, golden)
    finally:
        assert lib.tupan_abi_set_mode(_lib.MODE_FP32_IEEE) == 0


def _f32_vs_truth(lib, golden):
    for case in golden.cases:
        if case["prec"] != "float32":
            continue
        ip, jp = golden.inputs(case, np.float32)
        ours = _call_dropin(lib, case["kernel"], ip, jp, golden.args(case),
                            np.float32)
        ref32 = golden.outputs(case)
        # FP64 truth on the same float32-rounded inputs
        ip64 = {f: v.astype(np.float64) for f, v in ip.items()}
        jp64 = {f: v.astype(np.float64) for f, v in jp.items()}
        if case["kernel"] == "kepler":
            truth = kernels.kepler(ip64, *golden.args(case))
        else:
            truth = kernels.compute(case["kernel"], ip64, jp64,
                                    *golden.args(case), mode="fp64")
        for o, r, t in zip(ours, ref32, truth):
            scale = np.max(np.abs(t)) + 1e-300
            e_ours = np.abs(o.astype(np.float64) - t) / scale
            e_ref = np.abs(r.astype(np.float64) - t) / scale
            assert np.all(e_ours <= e_ref + 2.0**-22), case


def _plummer(n, seed, soft):
    from tupan_amd.ics import make_plummer
    phi = lambda ps: kernels.compute("phi", ps, ps, mode="fp64")[0]
    return make_plummer(n, soft, seed=seed, phi_fn=phi)


@pytest.mark.parametrize("name,args", [("phi", ()), ("acc", ()),
                                       ("acc_jerk", ()), ("tstep", (0.02,))])
def test_ragged_sizes_vs_oracle(name, args):
    """ni != nj, sizes straddling wave (64), tile (32) and split
    boundaries, against the CPU checker (sizes it finishes in seconds)."""
    ps = _plummer(3001, 5, 0.01)
    o = Oracle("float64", "port")
    d = {f: getattr(ps, f) for f in spec.F8}
    for ni, nj in ((1, 3001), (63, 65), (65, 63), (257, 2999), (3001, 1),
                   (3001, 3001), (128, 4)):
        ip = {f: v[:ni] for f, v in d.items()}
        jp = {f: v[3001 - nj:] for f, v in d.items()}
        ref = o.call(name + "_kernel", ip, jp, *args)
        got = kernels.compute(name, ip, jp, *args, mode="fp64")
        e = _err(name, got, ref)
        assert e <= 1e-12, (name, ni, nj, e)


def test_empty_j_gives_empty_sums():
    d = {f: np.ones(5) for f in spec.F8}
    e = {f: np.zeros(0) for f in spec.F8}
    got = kernels.compute("acc_jerk", d, e, mode="fp64")
    assert all(np.all(g == 0) for g in got)
    dt = kernels.compute("tstep", d, e, 0.125, mode="fp64")
    assert np.all(dt[0] == 0.125) and np.all(dt[1] == 0.125)
    got = kernels.compute("acc_jerk", e, d, mode="fp32c")
    assert all(len(g) == 0 for g in got)


def test_launch_shapes_agree_and_runs_are_bitwise_reproducible():
    ps = _plummer(20000, 6, 0.002)
    d = {f: getattr(ps, f) for f in spec.F8}
    base = kernels.compute("acc_jerk", d, d, mode="fp64")
    again = kernels.compute("acc_jerk", d, d, mode="fp64")
    for a, b in zip(base, again):
        assert np.array_equal(a, b)
    try:
        for mode in ("fp64", "fp32c"):
            ref = kernels.compute("acc_jerk", d, d, mode=mode) if mode == "fp32c" else base
            for ipl in (1, 2, 4):
                for splits in (1, 7, 64):
                    _lib.set_launch("acc_jerk", mode, ipl, splits)
                    got = kernels.compute("acc_jerk", d, d, mode=mode)
                    e = _err("acc_jerk", got, ref)
                    # FP32C: another shape regroups the FP32 tile sums, so
                    # results move within the mode's own error (below)
                    assert e <= (1e-13 if mode == "fp64" else 4e-6), (mode, ipl, splits, e)
    finally:
        _lib.set_launch("acc_jerk", "fp64", 0, 0)
        _lib.set_launch("acc_jerk", "fp32c", 0, 0)


_CACHE = {}


@pytest.mark.parametrize("ni,nj", [(300, 4096), (300, 4097), (65537, 200)])
def test_fp32c_context_paths(ni, nj):
    """FP32C builds its near-field context, tile origins and records in one
    block for nj <= 4096 and ni <= 65536, else in four launches: both sides
    of each bound stay within the mode's error of FP64, per particle (Phi
    and acc 1e-6; jerk RMS 3e-7, worst 3e-5; snap RMS 1e-6, worst 1e-4;
    crackle RMS 1e-5, worst 1e-3), for every FP32C kernel."""
    if "ctx" not in _CACHE:
        ps = _plummer(65537, 8, 0.002)
        d = {f: getattr(ps, f) for f in spec.F8}
        aj = kernels.compute("acc_jerk", d, d, mode="fp64")
        _CACHE["ctx"] = d, dict(zip(("ax", "ay", "az", "jx", "jy", "jz"), aj))
    d, extra = _CACHE["ctx"]
    ip = {f: v[:ni] for f, v in d.items()}
    jp = {f: v[:nj] for f, v in d.items()}
    for k in ("phi", "acc", "acc_jerk", "snap_crackle"):
        if k == "snap_crackle":
            ip2 = dict(ip, **{f: v[:ni] for f, v in extra.items()})
            jp2 = dict(jp, **{f: v[:nj] for f, v in extra.items()})
        else:
            ip2, jp2 = ip, jp
        a = kernels.compute(k, ip2, jp2, mode="fp32c")
        b = kernels.compute(k, ip2, jp2, mode="fp64")
        if k == "phi":
            rel = np.abs(a[0] - b[0]) / np.abs(b[0])
            assert rel.max() < 1e-6, (k, ni, nj, rel.max())
            continue
        # per particle: |x32 - x64| / |x64| of each 3-vector output
        for g in range(len(a) // 3):
            x = np.stack(a[3 * g:3 * g + 3])
            y = np.stack(b[3 * g:3 * g + 3])
            rel = (np.linalg.norm(x - y, axis=0) /
                   np.linalg.norm(y, axis=0))
            rms = np.sqrt(np.mean(rel ** 2))
            if k != "snap_crackle":
                # acceleration, jerk
                ok = rel.max() < 1e-6 if g == 0 else (rms < 3e-7 and
                                                     rel.max() < 3e-5)
            else:
                # snap, crackle
                ok = rms < (1e-6, 1e-5)[g] and rel.max() < (1e-4, 1e-3)[g]
            assert ok, (k, ni, nj, g, rms, rel.max())


def test_newton_third_law_large_n():
    ps = _plummer(32768, 7, 4.0 / 32768)
    d = {f: getattr(ps, f) for f in spec.F8}
    for mode, tol in (("fp64", 1e-13), ("fp32c", 1e-7)):
        got = kernels.compute("acc", d, d, mode=mode)
        for a in got:
            s = np.sum(ps.mass * a)
            assert abs(s) < tol * np.sum(np.abs(ps.mass * a)), (mode, s)


def test_device_api_matches_host_api():
    import torch
    from tupan_amd import device
    ps = _plummer(1000, 8, 0.01)
    hd = {f: getattr(ps, f) for f in spec.F8}
    host = kernels.compute("acc_jerk", hd, hd, mode="fp32c")
    d = device.to_device(ps, spec.F8)
    dev = device.compute("acc_jerk", d, d, mode="fp32c")
    torch.cuda.synchronize()
    for a, b in zip(host, dev):
        assert np.array_equal(a, b.cpu().numpy())
