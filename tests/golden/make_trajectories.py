"""Generate the integrator trajectory fixtures from the reference's C kernels.

    make -C oracle ref
    python tests/golden/make_trajectories.py

Each case runs one of the reference's integrators, restated step for step in
numpy (oracle/integrators.py, citing tupan/integrator/*.py), on the pair
kernels of the reference's own C (oracle/_ref, compiled from
tupan/lib/src), and stores the initial state, the per-step time steps and
energies and the state after every step.  The device drivers of
tupan_amd.integrator are checked against these on the GPU.

Cases
  c1_hermite4   Plummer N=256 seed 0, equal masses, eps^2 = 1e-4, Hermite-4
                (P(EC)^2) with shared dt = 1/64 for 1 time unit (64 steps);
                Tstep (eta = 1/64) recorded every step
  ahermite4     Plummer N=128 seed 5, eps^2 = 1e-4, adaptive block steps
                eta = 1/32 up to t = 1/8
  hermite2/6/8  Plummer N=64 seed 6, dt = 1/128, 8 steps
  sia21s_kdk    Plummer N=128 seed 7, dt = 1/64, 16 steps
  sia21a_dkd    Plummer N=128 seed 7, eta = 1/32, adaptive, t = 1/8
  sakura        Plummer N=64 seed 8, eps^2 = 1e-4, dt = 1/128, 8 steps
  asakura       same, eta = 1/64, adaptive, t = 1/32
  sia21s_kdk_pn compact cluster N=64 seed 4, masses U[1,30] normalised,
                half-mass radius 0.1, eps^2 = 1e-4, pn_order 5 at c = 1e3
                (1PN+2PN+2.5PN), dt = 1/1024, 16 steps, with the PN energy /
                momentum flux bookkeeping

The state is stored after every step for c1_hermite4 and after at most 16
evenly spaced steps (always including the last) for the others.
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

from oracle import Oracle  # noqa: E402
from oracle import integrators as oi  # noqa: E402
from tupan_amd.ics import UniformMass, make_plummer, scale_radius  # noqa: E402

BASE = ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz")
STATE = ("rx", "ry", "rz", "vx", "vy", "vz")


def plummer(K, n, seed, eps, imf=None):
    ps = make_plummer(n, eps, imf=imf, seed=seed,
                      phi_fn=lambda p: K.o.call("phi_kernel", p, p)[0])
    return {f: np.array(getattr(ps, f), dtype=np.float64) for f in BASE}


CASES = [
    # name, n, seed, eps, method, eta/dt, t_end, pn
    ("c1_hermite4", 256, 0, 0.01, "hermite4", 1.0 / 64, 1.0, None),
    ("ahermite4", 128, 5, 0.01, "ahermite4", 1.0 / 32, 0.125, None),
    ("hermite2", 64, 6, 0.01, "hermite2", 1.0 / 128, 1.0 / 16, None),
    ("hermite6", 64, 6, 0.01, "hermite6", 1.0 / 128, 1.0 / 16, None),
    ("hermite8", 64, 6, 0.01, "hermite8", 1.0 / 128, 1.0 / 16, None),
    ("sia21s_kdk", 128, 7, 0.01, "sia21s.kdk", 1.0 / 64, 0.25, None),
    ("sia21a_dkd", 128, 7, 0.01, "sia21a.dkd", 1.0 / 32, 0.125, None),
    ("sakura", 64, 8, 0.01, "sakura", 1.0 / 128, 1.0 / 16, None),
    ("asakura", 64, 8, 0.01, "asakura", 1.0 / 64, 1.0 / 32, None),
    ("sia21s_kdk_pn", 64, 4, 1e-2, "sia21s.kdk", 1.0 / 1024, 1.0 / 64,
     (5, 1e3)),
]


def step(K, ps, method, eta, tau):
    adaptive = method.startswith("a") or "21a" in method
    if "hermite" in method:
        return oi.hermite_step(K, ps, tau, int(method[-1]), eta, adaptive)
    if method.startswith("sia21"):
        return oi.sia21_step(K, ps, tau, method[-3:], eta, adaptive,
                             pn="pn_ke" in ps)
    return oi.sakura_do_step(K, ps, tau, eta, adaptive)


def run_case(K, name, n, seed, eps, method, eta, t_end, pn):
    if pn:
        ps = plummer(K, n, seed, eps, imf=UniformMass(1.0, 30.0))
        # half-mass radius of a Plummer sphere of virial radius 1 is
        # 0.7686 (r_h = 1.3048 a, r_vir = 16/(3 pi) a): scale it to 0.1
        p = type("P", (), {})()
        for f in BASE:
            setattr(p, f, ps[f])
        scale_radius(p, 0.1 / (1.3048 * 3 * np.pi / 16))
        K.set_clight(*pn)
        for f in oi.PN_FIELDS:
            ps[f] = np.zeros(n)
    else:
        ps = plummer(K, n, seed, eps)
    ps["t_curr"] = 0.0
    ps["tstep"] = np.zeros(n)
    init = {f: ps[f].copy() for f in BASE}
    E0 = oi.energy(K, ps, pn=bool(pn))
    taus, energies, states, tsteps = [], [E0], [], []
    while abs(ps["t_curr"]) < t_end:
        tau = oi.get_base_tstep(ps["t_curr"], t_end, eta)
        tau = step(K, ps, method, eta, tau)
        taus.append(tau)
        energies.append(oi.energy(K, ps, pn=bool(pn)))
        states.append(np.stack([ps[f] for f in STATE]))
        if name == "c1_hermite4":
            K.tstep(ps, 1.0 / 64)
            tsteps.append(np.stack([ps["tstep"], ps["tstepij"]]))
    nsteps = len(states)
    if name == "c1_hermite4" or nsteps <= 16:
        keep = np.arange(nsteps)
    else:
        keep = np.unique(np.r_[np.arange(nsteps // 16 - 1, nsteps,
                                         nsteps // 16), nsteps - 1])
    out = {"init": np.stack([init[f] for f in BASE]),
           "tau": np.array(taus), "energy": np.array(energies),
           "state_steps": keep, "state": np.stack([states[k] for k in keep])}
    if tsteps:
        out["tstep"] = np.stack(tsteps)
    if pn:
        out["pn"] = np.stack([ps[f] for f in oi.PN_FIELDS])
    meta = {"name": name, "n": n, "seed": seed, "eps": eps, "method": method,
            "eta": eta, "t_end": t_end, "pn": list(pn) if pn else None,
            "steps": len(taus)}
    return meta, out


def main():
    o = Oracle("float64", "reference")
    K = oi.Kernels(o)
    arrays = {}
    manifest = []
    for case in CASES:
        K.pn_order = 0
        meta, out = run_case(K, *case)
        print("%-14s steps %3d  dE/E %.3e" % (
            meta["name"], meta["steps"],
            (out["energy"][-1] - out["energy"][0]) / abs(out["energy"][0])),
            flush=True)
        manifest.append(meta)
        for k, v in out.items():
            arrays["%s:%s" % (meta["name"], k)] = v
    gcc = subprocess.run(["gcc", "--version"], stdout=subprocess.PIPE,
                         universal_newlines=True).stdout.splitlines()[0]
    info = {"cases": manifest, "meta": {
        "kernels": "reference C tupan/lib/src via oracle/_ref (-O3 -std=c99)",
        "drivers": "oracle/integrators.py", "gcc": gcc,
        "glibc": platform.libc_ver()[1], "numpy": np.__version__}}
    arrays["manifest"] = np.array(json.dumps(info))
    path = os.path.join(HERE, "trajectories.npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
