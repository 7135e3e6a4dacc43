"""CPU-only checks of the host side: the C ABI library loads and exports
every symbol its headers declare, the drop-in libraries export the
reference's libtupan.h symbols, the product has no silent fallback, and the
host logic (particle container, initial conditions, extension classes)
behaves like the reference's.  No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import oracle
from oracle import Oracle
from tupan_amd import _lib, extensions, spec
from tupan_amd.ics import PowerLaw, make_plummer
from tupan_amd.particles import ParticleSystem

LIBDIR = os.path.join(ROOT, "tupan_amd", "lib")


def header_symbols(path):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tupan_\w+)\s*\(", src)))


def exported(path):
    lib = ctypes.CDLL(path)
    return lib


def test_core_library_exports_header():
    syms = header_symbols(os.path.join(ROOT, "include", "tupan_hip.h"))
    assert "tupan_acc_jerk_dev" in syms and len(syms) >= 20
    lib = exported(os.path.join(LIBDIR, "libtupan_hip.so"))
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


REF_SYMBOLS = ["phi_kernel", "acc_kernel", "acc_jerk_kernel",
               "snap_crackle_kernel", "tstep_kernel", "pnacc_kernel",
               "sakura_kernel"]


@pytest.mark.parametrize("bits", ["64", "32"])
def test_dropin_libraries_export_libtupan_symbols(bits):
    path = os.path.join(LIBDIR, "libtupan_f%s.so" % bits)
    lib = exported(path)
    hdr = header_symbols(os.path.join(ROOT, "include", "libtupan_abi.h"))
    assert hdr == ["tupan_abi_get_mode", "tupan_abi_last_status",
                   "tupan_abi_set_mode"]
    for s in REF_SYMBOLS + hdr:
        assert hasattr(lib, s), s


def test_dropin_library_modes_need_no_device():
    """libtupan_f64.so follows the process default mode (-1);
    libtupan_f32.so starts in the reference FP32 build's arithmetic."""
    for bits, start in (("64", -1), ("32", _lib.MODE_FP32_IEEE)):
        lib = ctypes.CDLL(os.path.join(LIBDIR, "libtupan_f%s.so" % bits))
        assert lib.tupan_abi_get_mode() == start
        assert lib.tupan_abi_set_mode(_lib.MODE_COUNT) == _lib.ERR_ARG
        assert lib.tupan_abi_set_mode(-2) == _lib.ERR_ARG
        assert lib.tupan_abi_get_mode() == start
        for m in (-1, _lib.MODE_FP64, _lib.MODE_FP64_IEEE, start):
            assert lib.tupan_abi_set_mode(m) == _lib.OK
            assert lib.tupan_abi_get_mode() == m


def test_dropin_header_declares_reference_symbols():
    src = open(os.path.join(ROOT, "include", "libtupan_abi.h")).read()
    for s in REF_SYMBOLS:
        assert re.search(r"\bvoid\s+%s\s*\(" % s, src), s


def test_version_and_mode_calls_need_no_device():
    L = _lib.lib()
    assert b"gfx950" in L.tupan_version()
    assert L.tupan_get_default_mode() == 0
    assert L.tupan_set_default_mode(7) == _lib.ERR_ARG
    assert L.tupan_set_launch(2, 1, 3, 0) == _lib.ERR_ARG
    assert L.tupan_set_launch(2, 1, 0, 0) == _lib.OK
    # the reference-arithmetic mode keeps the reference's single j sum
    assert L.tupan_set_launch(2, _lib.MODE_FP64_IEEE, 2, 1) == _lib.OK
    assert L.tupan_set_launch(2, _lib.MODE_FP64_IEEE, 0, 4) == _lib.ERR_ARG
    assert L.tupan_set_launch(2, _lib.MODE_FP64_IEEE, 0, 0) == _lib.OK
    assert L.tupan_set_launch(2, _lib.MODE_FP32_IEEE, 0, 2) == _lib.ERR_ARG
    assert L.tupan_set_launch(2, _lib.MODE_FP32_IEEE, 1, 1) == _lib.OK
    assert L.tupan_set_launch(2, _lib.MODE_FP32_IEEE, 0, 0) == _lib.OK
    assert L.tupan_set_launch(2, _lib.MODE_COUNT, 0, 0) == _lib.ERR_ARG
    assert L.tupan_set_default_mode(_lib.MODE_COUNT) == _lib.ERR_ARG
    assert L.tupan_set_default_mode(_lib.MODE_FP64_IEEE) == _lib.OK
    assert L.tupan_set_default_mode(_lib.MODE_FP64) == _lib.OK


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libtupan_hip.so")
    with pytest.raises(_lib.TupanError):
        _lib.lib()
    ps = ParticleSystem(4)
    with pytest.raises(_lib.TupanError):
        extensions.acc_jerk.calc(ps, ps)


def test_no_oracle_import_in_product():
    pkg = os.path.join(ROOT, "tupan_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".c", ".cpp", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle", txt, re.M), f
                assert "liboracle" not in txt, f


def test_spec_matches_checker_specs():
    assert len(spec.SYMBOLS) == 10     # every symbol of libtupan.h
    for sym, key in spec.SYMBOLS.items():
        if key == "kepler":
            assert (spec.KEPLER_FIELDS, spec.KEPLER_OUTS) == oracle.KEPLER
            continue
        fields, _, outs = oracle.SPECS[sym]
        assert spec.KERNELS[key].fields == fields
        assert spec.KERNELS[key].outs == outs


# ---------------------------------------------------------------- host logic
def test_register_auxiliary_attribute():
    ps = ParticleSystem(3)
    ps.register_auxiliary_attribute("phi", "real")
    assert ps.phi.dtype == np.float64 and ps.phi.shape == (3,)
    with pytest.raises(ValueError):
        ps.register_auxiliary_attribute("phi", "real")


def _oracle_phi(ps):
    return Oracle("float64", "port").call("phi_kernel", ps, ps)[0]


def test_plummer_henon_units():
    ps = make_plummer(512, 0.01, seed=0, phi_fn=_oracle_phi)
    assert ps.total_mass == pytest.approx(1.0, abs=1e-15)
    assert np.all(ps.mass == ps.mass[0])
    assert np.all(ps.eps2 == 0.01**2 / 2)
    pe = 0.5 * float((ps.mass * _oracle_phi(ps)).sum())
    ke = ps.kinetic_energy
    # positions are rescaled to r_vir = 1 with the softening left as is
    # (body.py:430-443), so U is -1/2 only up to the softening's effect
    assert pe == pytest.approx(-0.5, rel=2e-3)
    assert 2 * ke + pe == pytest.approx(0.0, abs=1e-12)  # virial, exact
    assert ke + pe == pytest.approx(-0.25, rel=2e-3)
    assert np.allclose(ps.com_r, 0, atol=1e-15)
    assert np.allclose(ps.com_v, 0, atol=1e-15)


def test_plummer_is_seeded():
    a = make_plummer(128, 0.01, seed=7, phi_fn=_oracle_phi)
    b = make_plummer(128, 0.01, seed=7, phi_fn=_oracle_phi)
    c = make_plummer(128, 0.01, seed=8, phi_fn=_oracle_phi)
    assert np.array_equal(a.rx, b.rx) and np.array_equal(a.vz, b.vz)
    assert not np.array_equal(a.rx, c.rx)


def test_powerlaw_imf():
    np.random.seed(1)
    m = PowerLaw(2.35, 0.1, 100.0).sample(200000)
    assert m.min() >= 0.1 and m.max() <= 100.0
    # fraction above 1 Msun for dN/dm ~ m^-2.35 on [0.1, 100]
    k = -1.35
    frac = (100.0**k - 1.0) / (100.0**k - 0.1**k)
    assert np.mean(m > 1.0) == pytest.approx(frac, abs=3e-3)


class OracleKernel(extensions.HIPKernel):
    """Test backend: the reference protocol served by the CPU checker."""

    def run(self):
        if self.key == "kepler":
            a = self.args
            ps = dict(zip(spec.KEPLER_FIELDS, a[:8]))
            res = Oracle(self.prec, "port").kepler(ps, a[8])
            for o, r in zip(a[9:15], res):
                o[:2] = r
            return
        k = spec.KERNELS[self.key]
        nf = len(k.fields)
        a = self.args
        ni, nj = int(a[0]), int(a[1 + nf])
        ip = dict(zip(k.fields, a[1:1 + nf]))
        jp = dict(zip(k.fields, a[2 + nf:2 + 2 * nf]))
        rest = a[2 + 2 * nf:]
        nout = len(k.outs)
        res = Oracle(self.prec, "port").call(self.name, ip, jp,
                                             *rest[:len(rest) - nout])
        for o, r in zip(rest[len(rest) - nout:], res):
            o[:ni] = r


extensions.register_backend("ORACLE", OracleKernel)


def test_extension_protocol_registers_outputs(golden):
    case = golden.find(kernel="acc_jerk", set="P", prec="float64")[0]
    d = golden.set("P")
    ps = ParticleSystem(len(d["mass"]))
    for f in spec.F8:
        getattr(ps, f)[...] = d[f]
    ext = extensions.AccJerk("ORACLE", "float64")
    res = ext.calc(ps, ps)
    assert res[0] is ps.ax and res[5] is ps.jz
    for got, ref in zip(res, golden.outputs(case)):
        assert np.array_equal(got, ref)


def test_extension_pnacc_uses_module_clight(golden):
    case = golden.find(kernel="pnacc", set="R", prec="float64", ipre=None,
                       jpre=None)[-1]
    d = golden.set("R")
    ps = ParticleSystem(len(d["mass"]))
    for f in spec.F8:
        getattr(ps, f)[...] = d[f]
    extensions.clight.pn_order = case["args"][0]
    extensions.clight.clight = 1.0 / case["args"][1]
    res = extensions.PNAcc("ORACLE", "float64").calc(ps, ps)
    for got, ref in zip(res, golden.outputs(case)):
        assert np.allclose(got, ref, rtol=1e-15, atol=0)


def test_unknown_backend_rejected():
    with pytest.raises(ValueError):
        extensions.Acc("CL", "float64")


def test_extension_nreg_and_kepler_surface(golden):
    """NREG_X / NREG_V register the reference's output names and take dt;
    Kepler updates particles 0 and 1 in place (extensions.py:510-651)."""
    d = golden.set("P")
    ps = ParticleSystem(len(d["mass"]))
    for f in spec.F8:
        getattr(ps, f)[...] = d[f]
    case = golden.find(kernel="nreg_x", set="P", prec="float64")[0]
    res = extensions.NREG_X("ORACLE", "float64").calc(ps, ps, case["args"][0])
    assert res[0] is ps.mrx and res[6] is ps.u
    for got, ref in zip(res, golden.outputs(case)):
        assert np.array_equal(got, ref)
    for f in ("ax", "ay", "az"):
        getattr(ps, f)[...] = d[f]
    case = golden.find(kernel="nreg_v", set="P", prec="float64")[0]
    res = extensions.NREG_V("ORACLE", "float64").calc(ps, ps, case["args"][0])
    assert res[3] is ps.mk
    for got, ref in zip(res, golden.outputs(case)):
        assert np.array_equal(got, ref)
    case = golden.find(kernel="kepler", set="P", prec="float64")[0]
    rx = ps.rx
    extensions.Kepler("ORACLE", "float64").calc(ps, ps, case["args"][0])
    assert ps.rx is rx
    for f, ref in zip(spec.KEPLER_OUTS, golden.outputs(case)):
        assert np.array_equal(getattr(ps, f)[:2], ref)
        assert np.array_equal(getattr(ps, f)[2:], d[f][2:])


class _Clock(object):
    """Stand-in particle set for the block-step logic (host only)."""

    def __init__(self, t, tsteps):
        self.t_curr = t
        self.tstep = np.asarray(tsteps)

    def min_tstep(self):
        return float(np.abs(self.tstep).min())


def test_block_step_logic_matches_checker():
    """integrator.Base.get_base_tstep / get_min_block_tstep restate
    tupan/integrator/__init__.py:48-78 exactly as the checker does."""
    from oracle import integrators as oi
    torch = pytest.importorskip("torch")  # noqa: F841
    from tupan_amd import integrator as ig
    rng = np.random.RandomState(3)
    base = ig.Base.__new__(ig.Base)
    for _ in range(2000):
        t = float(rng.choice([0.0, rng.uniform(0, 3), 0.125 * rng.randint(0, 40)]))
        ts = 10.0**rng.uniform(-7, 1, rng.randint(1, 9))
        tau = float(rng.choice([1.0 / 64, 0.02, rng.uniform(1e-4, 1.0)]))
        eta = float(rng.choice([0.02, 1.0 / 32, -0.01]))
        t_end = t + float(rng.uniform(0, 2))
        base.eta = eta
        p = _Clock(t, ts)
        base.ps = p
        assert base.get_base_tstep(t_end) == oi.get_base_tstep(t, t_end, eta)
        d = {"tstep": np.asarray(ts), "t_curr": t}
        assert base.get_min_block_tstep(p, tau) == oi.get_min_block_tstep(d, tau)


def test_integrator_rejects_unknown_methods():
    torch = pytest.importorskip("torch")  # noqa: F841
    from tupan_amd import integrator as ig
    with pytest.raises(ValueError):
        ig.Integrator(0.01, 0.0, None, "sia43s.kdk")
    assert "ahermite4" in ig.Integrator.PROVIDED_METHODS
    assert "sakura_pn" in ig.Integrator.PROVIDED_METHODS
