"""Generate the golden fixtures from the reference's own C kernels.

    make -C oracle ref            # compiles tupan/lib/src/*.c -> oracle/_ref/
    python tests/golden/make_golden.py

Inputs are our own seeded arrays; outputs come from the compiled reference
C (oracle/_ref/libtupan_ref_f{64,32}.so, built from the reference sources
with the reference's flags).  Only inputs and outputs are stored: no
reference source travels with them.  Cases follow the reference test's
shapes (tupan/tests/test_extensions.py:39-127: random sets, prefix slices
ps[:k] on either side, eta = 1/64, PN order 7 at clight = 128) and add the
edge cases that test leaves open (coincident particles, softened sets, PN
orders 0..7, every Sakura flag with close, wide, bound and unbound pairs,
the Plummer N=256 seed-0 model, the NREG pair sums at several dt and the
two-body Kepler solver on bound, unbound and softened pairs).  For every FP64 sum kernel the file also
holds, per particle and output vector, S_i = sum_j ||f_ij||, the magnitude
scale of its pair terms (the condition of the sum), from the same library.
"""
import json
import os
import platform
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import Oracle, SPECS  # noqa: E402
from tupan_amd.ics import make_plummer  # noqa: E402

F14 = SPECS["snap_crackle_kernel"][0]
SHORT = {"phi_kernel": "phi", "acc_kernel": "acc",
         "acc_jerk_kernel": "acc_jerk", "snap_crackle_kernel": "snap_crackle",
         "tstep_kernel": "tstep", "pnacc_kernel": "pnacc",
         "sakura_kernel": "sakura", "nreg_Xkernel": "nreg_x",
         "nreg_Vkernel": "nreg_v", "kepler_solver_kernel": "kepler"}


def random_set(seed, n, soft):
    rng = np.random.RandomState(seed)
    d = {}
    d["mass"] = rng.random_sample(n)
    d["eps2"] = (rng.uniform(1e-4, 1e-2, n) / 2) if soft else np.zeros(n)
    for f in ("rx", "ry", "rz", "vx", "vy", "vz"):
        d[f] = rng.random_sample(n) * 10
    for f in ("ax", "ay", "az", "jx", "jy", "jz"):
        d[f] = rng.uniform(-1, 1, n)
    # a coincident pair with different velocities exercises the r^2 > 0
    # mask (tupan/lib/src/smoothing.h:16-18)
    for f in ("rx", "ry", "rz"):
        d[f][5] = d[f][3]
    return d


def cluster_set(seed, n, eps2):
    rng = np.random.RandomState(seed)
    d = {"mass": rng.uniform(0.5, 1.5, n) / n,
         "eps2": np.full(n, eps2 / 2)}
    for f in ("rx", "ry", "rz"):
        d[f] = rng.normal(0, 0.1, n)
    for f in ("vx", "vy", "vz"):
        d[f] = rng.normal(0, 0.5, n)
    # a tight bound binary and a fast hyperbolic encounter
    d["rx"][1] = d["rx"][0] + 2e-3
    d["ry"][1] = d["ry"][0]
    d["rz"][1] = d["rz"][0]
    d["vx"][1] = d["vx"][0]
    d["vy"][1] = d["vy"][0] + 0.9 * np.sqrt((d["mass"][0] + d["mass"][1]) / 2e-3)
    d["vz"][1] = d["vz"][0]
    d["rx"][3] = d["rx"][2] + 5e-3
    d["vx"][3] = d["vx"][2] - 30.0
    for f in F14[8:]:
        d[f] = np.zeros(n)
    return d


def plummer_set(ref):
    phi = lambda ps: ref.call("phi_kernel", ps, ps)[0]
    ps = make_plummer(256, 0.01, seed=0, phi_fn=phi)
    d = {f: getattr(ps, f).copy() for f in F14[:8]}
    aj = ref.call("acc_jerk_kernel", d, d)
    for f, v in zip(("ax", "ay", "az", "jx", "jy", "jz"), aj):
        d[f] = v
    return d


# output vector groups per kernel (Tstep's outputs are not plain sums)
GROUPS = {"phi": [(0,)], "acc": [(0, 1, 2)],
          "acc_jerk": [(0, 1, 2), (3, 4, 5)],
          "snap_crackle": [(0, 1, 2), (3, 4, 5)],
          "pnacc": [(0, 1, 2)], "sakura": [(0, 1, 2), (3, 4, 5)],
          "nreg_x": [(0, 1, 2), (3, 4, 5), (6,)],
          "nreg_v": [(0, 1, 2), (3,)]}


def abs_sums(lib, name, ips, jps, args):
    """Per i and output group, sum_j ||f_ij||: the magnitude scale of the
    pair terms, i.e. the condition of the sum (one j at a time)."""
    groups = GROUPS[SHORT[name]]
    ni = len(ips["mass"])
    S = [np.zeros(ni) for _ in groups]
    for j in range(len(jps["mass"])):
        one = {f: v[j:j + 1] for f, v in jps.items()}
        c = lib.call(name, ips, one, *args)
        for g, grp in enumerate(groups):
            S[g] += np.sqrt(sum(np.asarray(c[k], np.float64)**2 for k in grp))
    return S


def sl(d, k, n):
    if k is None:
        return d
    return {f: v[:k] for f, v in d.items()}


def main():
    ref64 = Oracle("float64", "reference")
    ref32 = Oracle("float32", "reference")
    sets = {
        "R": random_set(11, 256, soft=False),
        "S": random_set(12, 256, soft=True),
        "P": plummer_set(ref64),
        "K0": cluster_set(13, 48, 0.0),
        "K1": cluster_set(13, 48, 1e-6),
    }
    c = 128.0
    inv_c = tuple(c**-k for k in range(1, 8))
    c3 = 1e3
    inv_c3 = tuple(c3**-k for k in range(1, 8))

    cases = []

    def add(kernel, setname, prec, args=(), ipre=None, jpre=None):
        cases.append(dict(kernel=kernel, set=setname, prec=prec,
                          args=list(args), ipre=ipre, jpre=jpre))

    grav = [("phi_kernel", ()), ("acc_kernel", ()), ("acc_jerk_kernel", ()),
            ("snap_crackle_kernel", ()), ("tstep_kernel", (1.0 / 64,))]
    for name, args in grav:
        add(name, "R", "float64", args)
        for k in (1, 2, 3, 63, 64, 65, 127, 255):
            add(name, "R", "float64", args, jpre=k)
        for k in (1, 3, 65):
            add(name, "R", "float64", args, ipre=k)
        add(name, "S", "float64", args)
        add(name, "P", "float64", args)
        add(name, "R", "float32", args)
        add(name, "P", "float32", args)
        add(name, "R", "float32", args, jpre=65)
    add("tstep_kernel", "P", "float64", (0.02,))
    for order in (0, 2, 4, 5, 6, 7):
        add("pnacc_kernel", "R", "float64", (order,) + inv_c)
        add("pnacc_kernel", "K1", "float64", (order,) + inv_c3)
    for k in (1, 64, 65):
        add("pnacc_kernel", "R", "float64", (7,) + inv_c, jpre=k)
    add("pnacc_kernel", "R", "float32", (7,) + inv_c)
    add("pnacc_kernel", "R", "float32", (4,) + inv_c)
    # With softening the reference's energy-check sub-stepping
    # (universal_kepler_solver.h:580-606) does not terminate in practical
    # time for the tight binary of the K sets beyond dt ~ 1/64, so the
    # softened set is only sampled at small dt.
    for s, dts in (("K0", (1.0 / 64, 0.25)), ("K1", (1.0 / 64, 1.0 / 128))):
        for dt in dts:
            for flag in (-2, -1, 0, 1, 2):
                add("sakura_kernel", s, "float64", (dt, flag))
    add("sakura_kernel", "R", "float64", (1.0 / 64, -1))
    add("sakura_kernel", "P", "float64", (1.0 / 64, 2))
    add("sakura_kernel", "K1", "float32", (1.0 / 64, -1))
    add("sakura_kernel", "K1", "float32", (1.0 / 64, 2))

    # NREG pair sums at drifted / kicked separations (libtupan.h:152-201)
    for s_, dt in (("R", 1.0 / 64), ("S", 1.0 / 64), ("P", 1.0 / 64),
                   ("S", 0.0), ("P", -0.125)):
        add("nreg_Xkernel", s_, "float64", (dt,))
        add("nreg_Vkernel", s_, "float64", (dt,))
    for k in (1, 63, 65):
        add("nreg_Xkernel", "R", "float64", (1.0 / 64,), jpre=k)
        add("nreg_Vkernel", "R", "float64", (1.0 / 64,), jpre=k)
    for s_ in ("R", "P"):
        add("nreg_Xkernel", s_, "float32", (1.0 / 64,))
        add("nreg_Vkernel", s_, "float32", (1.0 / 64,))
    # Kepler two-body solver on particles 0 and 1 (libtupan.h:231-246):
    # K0/K1 hold a tight bound binary there, R an unbound pair, P a wide
    # Plummer pair; softened K1 only at small dt (see Sakura above)
    for s_, dts in (("K0", (1.0 / 64, 0.25, -0.1)), ("K1", (1.0 / 64, 1.0 / 128)),
                    ("R", (1.0 / 64, 0.25)), ("P", (1.0 / 64, 0.25))):
        for dt in dts:
            add("kepler_solver_kernel", s_, "float64", (dt,))
    add("kepler_solver_kernel", "K0", "float32", (1.0 / 64,))
    add("kepler_solver_kernel", "R", "float32", (0.25,))

    out = {}
    for sname, d in sets.items():
        for f, v in d.items():
            out["in:%s:%s" % (sname, f)] = np.asarray(v, np.float64)
    for cid, cs in enumerate(cases):
        lib = ref64 if cs["prec"] == "float64" else ref32
        d = sets[cs["set"]]
        ips = sl(d, cs["ipre"], None)
        jps = sl(d, cs["jpre"], None)
        if cs["prec"] == "float32":
            ips = {f: v.astype(np.float32) for f, v in ips.items()}
            jps = {f: v.astype(np.float32) for f, v in jps.items()}
        if cs["kernel"] == "kepler_solver_kernel":
            res = lib.kepler(ips, *cs["args"])
        else:
            res = lib.call(cs["kernel"], ips, jps, *cs["args"])
        if cs["prec"] == "float64" and SHORT[cs["kernel"]] in GROUPS:
            for g, v in enumerate(abs_sums(lib, cs["kernel"], ips, jps,
                                           cs["args"])):
                out["abs:%d:%d" % (cid, g)] = v
        print(cid, cs["kernel"], cs["set"], cs["prec"], flush=True)
        cs["id"] = cid
        cs["kernel"] = SHORT[cs["kernel"]]
        for k, v in enumerate(res):
            out["out:%d:%d" % (cid, k)] = v

    meta = dict(
        generator="tests/golden/make_golden.py",
        oracle="compiled reference C (tupan/lib/src/*.c), gcc -O3 -std=c99",
        gcc=subprocess.run(["gcc", "--version"], stdout=subprocess.PIPE,
                           universal_newlines=True).stdout.splitlines()[0],
        glibc=" ".join(platform.libc_ver()),
        numpy=np.__version__,
    )
    out["manifest"] = np.array(json.dumps(dict(meta=meta, cases=cases)))
    path = os.path.join(HERE, "kernels.npz")
    np.savez_compressed(path, **out)
    print("wrote %s: %d cases, %.0f KB" % (path, len(cases),
                                         os.path.getsize(path) / 1024.0))


if __name__ == "__main__":
    main()
