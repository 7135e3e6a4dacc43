"""Argument validation of the C ABI (no device needed).

Every entry point checks its arguments before it touches the device and
reports a bad one as TUPAN_ERR_ARG with a message in tupan_last_error();
empty work (ni = 0, n = 0) succeeds without a device.
"""
import ctypes

import pytest

from tupan_amd import _lib

NULL = None
BIG = (1 << 62)


@pytest.fixture(scope="module")
def L():
    return _lib.lib()


def _pp(n):
    return (ctypes.c_void_p * n)()  # an array of NULL field pointers


def _err(L, rc):
    assert rc == _lib.ERR_ARG, rc
    msg = L.tupan_last_error()
    assert msg and len(msg) > 0
    return msg


def test_unknown_modes_rejected(L):
    for name in ("phi", "acc", "acc_jerk", "snap_crackle"):
        fn = getattr(L, "tupan_%s_dev" % name)
        _err(L, fn(7, 4, _pp(14), 4, _pp(14), _pp(8), NULL))
    _err(L, L.tupan_tstep_dev(7, 4, _pp(8), 4, _pp(8), 0.1, _pp(2), NULL))
    for name in ("nreg_x", "nreg_v"):
        fn = getattr(L, "tupan_%s_dev" % name)
        _err(L, fn(7, 4, _pp(8), 4, _pp(8), 0.1, _pp(8), NULL))


def test_fp32_ieee_mode_where_it_exists(L):
    """Phi, Acc, AccJerk, SnapCrackle and Tstep take TUPAN_MODE_FP32_IEEE
    (empty work: no device); PNAcc and NREG X/V refuse it."""
    m = _lib.MODE_FP32_IEEE
    for name, nf, no in (("phi", 5, 1), ("acc", 5, 3), ("acc_jerk", 8, 6),
                         ("snap_crackle", 14, 6)):
        fn = getattr(L, "tupan_%s_dev" % name)
        assert fn(m, 0, _pp(nf), 0, _pp(nf), _pp(no), NULL) == _lib.OK
    assert L.tupan_tstep_dev(m, 0, _pp(8), 0, _pp(8), 0.1, _pp(2),
                             NULL) == _lib.OK
    inv = (ctypes.c_double * 7)(*([0.1] * 7))
    _err(L, L.tupan_pnacc_dev(m, 0, _pp(8), 0, _pp(8), 7, inv, _pp(3),
                              NULL))
    for name in ("nreg_x", "nreg_v"):
        fn = getattr(L, "tupan_%s_dev" % name)
        _err(L, fn(m, 0, _pp(8), 0, _pp(8), 0.1, _pp(8), NULL))


def test_null_fields_and_sizes_rejected(L):
    _err(L, L.tupan_acc_dev(0, 4, NULL, 4, NULL, NULL, NULL))
    _err(L, L.tupan_acc_dev(0, 4, _pp(5), 4, _pp(5), _pp(3), NULL))
    _err(L, L.tupan_acc_jerk_dev(1, 4, _pp(8), 4, _pp(8), _pp(6), NULL))
    _err(L, L.tupan_acc_dev(0, BIG, _pp(5), 4, _pp(5), _pp(3), NULL))
    _err(L, L.tupan_phi_dev(2, 4, _pp(5), 4, _pp(5), _pp(1), NULL))
    _err(L, L.tupan_pnacc_dev(0, 4, _pp(8), 4, _pp(8), 5, NULL, _pp(3),
                              NULL))


def test_empty_work_needs_no_device(L):
    for mode in (0, 1, 2):
        assert L.tupan_acc_jerk_dev(mode, 0, _pp(8), 0, _pp(8), _pp(6),
                                    NULL) == _lib.OK
    assert L.tupan_tstep_dev(0, 0, _pp(8), 0, _pp(8), 0.1, _pp(2),
                             NULL) == _lib.OK
    assert L.tupan_divide_dev(0, 3, _pp(3), _pp(3), 2.0, NULL) == _lib.ERR_ARG
    assert L.tupan_energy_terms_dev(0, _pp(5), NULL, NULL, NULL) \
        == _lib.ERR_ARG  # NULL outputs are rejected even for n = 0


def test_update_and_reduction_arguments(L):
    _err(L, L.tupan_divide_dev(10, 0, _pp(3), _pp(3), 2.0, NULL))
    _err(L, L.tupan_divide_dev(10, 10, _pp(10), _pp(10), 2.0, NULL))
    _err(L, L.tupan_sum_dev(10, NULL, NULL, NULL))
    _err(L, L.tupan_sum_pairwise_dev(10, NULL, NULL, NULL))
    _err(L, L.tupan_diagnostics_dev(10, NULL, 0, NULL, NULL))
    _err(L, L.tupan_report_terms_dev(10, _pp(18), 1, _pp(12), NULL))
    _err(L, L.tupan_report_terms_dev(BIG, _pp(18), 1, _pp(12), NULL))


def test_tuning_knobs_validated(L):
    _err(L, L.tupan_set_near_field(-1.0))
    _err(L, L.tupan_set_near_field(1e7))
    _err(L, L.tupan_set_sakura_budget(0))
    _err(L, L.tupan_set_launch(99, 0, 0, 0))
    _err(L, L.tupan_set_launch(2, 0, 3, 0))
    _err(L, L.tupan_set_launch(2, 0, 0, -1))
    _err(L, L.tupan_set_flag_tau(-0.1))
    _err(L, L.tupan_set_flag_tau(float("nan")))
    _err(L, L.tupan_last_flagged(NULL))
    n = ctypes.c_int(-1)
    # no sorted call yet in this process: nothing to read, no device needed
    assert L.tupan_last_flagged(ctypes.byref(n)) == _lib.OK and n.value == 0
    # restore the defaults
    assert L.tupan_set_near_field(5.0) == _lib.OK
    assert L.tupan_set_flag_tau(0.7) == _lib.OK
    assert L.tupan_set_sorted(1) == _lib.OK
    assert L.tupan_set_sakura_budget(1 << 20) == _lib.OK
