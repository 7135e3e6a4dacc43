"""The PNAcc flop count per PN order (tupan_amd.spec.PNACC_FLOPS) is the
operation count of the PN terms as compiled: tests/pnacc_flops.cpp
evaluates tupan_amd/csrc/pn_terms.inc with a counting number type."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT
from tupan_amd import spec


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pnacc_flops_match_the_terms(tmp_path):
    exe = str(tmp_path / "pnacc_flops")
    subprocess.check_call(["g++", "-std=c++17", "-O0",
                           "-I", os.path.join(ROOT, "tupan_amd", "csrc"),
                           os.path.join(ROOT, "tests", "pnacc_flops.cpp"),
                           "-o", exe])
    for order, want in spec.PNACC_FLOPS.items():
        out = subprocess.check_output([exe, str(order)]).decode().split()
        assert int(out[1]) == want, (order, out)
    assert spec.KERNELS["pnacc"].flops == spec.PNACC_FLOPS[5]
