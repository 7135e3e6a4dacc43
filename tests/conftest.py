import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "kernels.npz")


def pytest_configure(config):
    config.addinivalue_line("markers",
                            "gpu: needs an MI355X (runs through libtupan_hip)")


class Golden(object):
    """Access to tests/golden/kernels.npz (inputs + reference-C outputs)."""

    def __init__(self, path=GOLDEN):
        self.z = np.load(path, allow_pickle=False)
        man = json.loads(str(self.z["manifest"]))
        self.meta = man["meta"]
        self.cases = man["cases"]
        self._sets = {}

    def set(self, name):
        if name not in self._sets:
            pre = "in:%s:" % name
            self._sets[name] = {k[len(pre):]: self.z[k] for k in self.z.files
                                if k.startswith(pre)}
        return self._sets[name]

    def inputs(self, case, dtype=np.float64):
        d = self.set(case["set"])
        ip = {f: v[:case["ipre"]] if case["ipre"] else v for f, v in d.items()}
        jp = {f: v[:case["jpre"]] if case["jpre"] else v for f, v in d.items()}
        ip = {f: np.ascontiguousarray(v, dtype) for f, v in ip.items()}
        jp = {f: np.ascontiguousarray(v, dtype) for f, v in jp.items()}
        return ip, jp

    def outputs(self, case):
        out = []
        k = 0
        while "out:%d:%d" % (case["id"], k) in self.z.files:
            out.append(self.z["out:%d:%d" % (case["id"], k)])
            k += 1
        return out

    def abs_sums(self, case):
        """S_i = sum_j ||f_ij|| per output group, or None (Tstep, FP32)."""
        out = []
        g = 0
        while "abs:%d:%d" % (case["id"], g) in self.z.files:
            out.append(self.z["abs:%d:%d" % (case["id"], g)])
            g += 1
        return out or None

    @staticmethod
    def args(case):
        """Scalar args in the form tupan_amd.kernels.compute takes."""
        a = case["args"]
        if case["kernel"] == "pnacc":
            return (int(a[0]), tuple(a[1:8]))
        return tuple(a)

    @staticmethod
    def oracle_args(case):
        return tuple(case["args"])

    def find(self, **kw):
        return [c for c in self.cases
                if all(c.get(k) == v for k, v in kw.items())]


@pytest.fixture(scope="session")
def golden():
    return Golden()


def rel_norm_err(got, ref, groups):
    """Per-particle relative error ||got - ref|| / ||ref|| over component
    groups (e.g. (ax, ay, az) and (jx, jy, jz)); returns max over particles
    and groups.  Particles whose reference norm is below 1e-300 of the
    largest are compared in absolute terms against that largest norm."""
    worst = 0.0
    for g in groups:
        r = np.sqrt(sum(np.asarray(ref[k], np.float64)**2 for k in g))
        d = np.sqrt(sum((np.asarray(got[k], np.float64) -
                         np.asarray(ref[k], np.float64))**2 for k in g))
        scale = np.maximum(r, 1e-300 * max(r.max(), 1e-300))
        if len(r):
            worst = max(worst, float(np.max(d / scale)))
    return worst


def bound_ratio(got, ref, S, groups, rel, absol):
    """max_i ||got_i - ref_i|| / (rel ||ref_i|| + absol S_i) over groups."""
    worst = 0.0
    for g, grp in enumerate(groups):
        d = np.sqrt(sum((np.asarray(got[k], np.float64) - ref[k])**2 for k in grp))
        r = np.sqrt(sum(np.asarray(ref[k], np.float64)**2 for k in grp))
        if len(d):
            worst = max(worst, float(np.max(d / (rel * r + absol * S[g] + 1e-300))))
    return worst


OUT_GROUPS = {
    "phi": [(0,)],
    "acc": [(0, 1, 2)],
    "acc_jerk": [(0, 1, 2), (3, 4, 5)],
    "snap_crackle": [(0, 1, 2), (3, 4, 5)],
    "tstep": [(0,), (1,)],
    "pnacc": [(0, 1, 2)],
    "sakura": [(0, 1, 2), (3, 4, 5)],
    "nreg_x": [(0, 1, 2), (3, 4, 5), (6,)],
    "nreg_v": [(0, 1, 2), (3,)],
    "kepler": [(0, 1, 2), (3, 4, 5)],
}
