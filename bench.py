#!/usr/bin/env python3
"""AccJerk single-evaluation benchmark on one MI355X.

Workload (the single-GPU kernel configuration the metric is quoted on):
Plummer sphere, N = 65536, seed 1, Salpeter masses (dN/dm ~ m^-2.35 on
[0.1, 100], normalised to total mass 1), softening eps = 4/N (per-particle
eps2 field = eps^2/2), Henon units.  One "step" is one AccJerk evaluation of
all N^2 (i, j) pairs (self pairs included and masked, as the reference
counts them) in the FP32-compensated mode, inputs resident in HBM.

Reported:
  value       N^2 * steps / elapsed, in Ginteractions/s
  roofline    42 flop per interaction (the reference's count,
              tupan/lib/src/acc_jerk_kernel_common.h:56) / the pair-engine
              kernel's mean duration, measured live with HIP events on the
              launch stream; bound = vector ALU (FP32), peak 157.3 TFLOP/s
  accuracy    per-particle |a - a64|/|a64| and |j - j64|/|j64| (max, rms)
              of the timed steps' outputs against the reference's C kernels
              on the rows the cpu_baseline leg evaluates (all N when the
              host threads finish them within its sample; the metric's
              <= 1e-6 bar); accuracy_vs_fp64_gpu: the same against the FP64
              kernel on every row
  fp32c_rows_recomputed_fp64
              rows the last FP32C step flagged and recomputed in FP64
  fp64        throughput of the FP64 kernel on the same workload
  host_abi    end-to-end rate through the host-pointer entry point (H2D +
              kernel + D2H per call; never the value)
  fp64_ieee   throughput of the reference-arithmetic mode (the reference C's
              operations, order and rounding on the GPU), whose outputs on
              the cpu_baseline's sampled rows must equal the reference C's
              bit for bit
  cpu_baseline  the reference's C AccJerk (oracle/_ref, or our bit-identical
              C restatement when _ref is absent) on a sample of i-rows x all
              N j-particles, on the host threads this process may use (at
              most 16; the reference itself is single-threaded, and that
              figure is reported beside it)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Acc+Jrk Ginteractions/s at N=64K, FP32-comp vs FP64 at <=1e-6 "
          "rel err; %peak")
FLOPS_PER_PAIR = 42
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=30)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=65536)
    p.add_argument("--mode", default="fp32c",
                   choices=["fp32c", "fp64", "fp64_ieee"])
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="target CPU work for the cpu_baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--graph", action="store_true",
                   help="pn and nreg workloads: replay each step as a HIP "
                        "graph")
    p.add_argument("--no-fp64", action="store_true")
    p.add_argument("--no-full", action="store_true",
                   help="ensemble workload: skip the run of every cluster "
                        "to t = 1")
    p.add_argument("--sizes", default="16384,65536,262144",
                   help="sizes workload: the N values (comma-separated)")
    p.add_argument("--workload", default="accjerk",
                   choices=["accjerk", "c1", "sizes", "ensemble", "pn", "diag",
                            "nreg", "adjoint"],
                   help="accjerk: the headline single-evaluation metric; "
                        "c1: Hermite-4 steps at N = 256 per mode, eager and "
                        "graph-replayed; sizes: Acc, AccJerk, SnapCrackle at N = 16K, 64K, "
                        "256K in both modes with errors vs FP64; "
                        "ensemble: the 512-cluster ragged ensemble (adaptive "
                        "Hermite-4) against looping the single-system path; "
                        "pn: the compact post-Newtonian cluster with Sakura "
                        "close pairs; diag: the energy/momentum report at "
                        "N = 65536; nreg: NREG steps at N = 1024; adjoint: "
                        "the differentiation workload at N = 16384")
    p.add_argument("--pmc-target", action="store_true",
                   help="one evaluation per mode plus a 256 MiB device copy "
                        "(the byte-counter calibration kernel), for "
                        "rocprofv3 --pmc runs; prints no JSON")
    return p.parse_args()


PMC_PROFILE = os.path.join(ROOT, "profiles", "acc_jerk_n65536_pmc.md")


def measured_traffic(mode):
    """HBM bytes per pair-engine launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel factor);
    None when that summary is absent."""
    key = "traffic_bytes_%s:" % mode
    try:
        for line in open(PMC_PROFILE):
            if line.strip().startswith(key):
                return float(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def make_workload(n, seed=1):
    from tupan_amd import kernels
    from tupan_amd.ics import PowerLaw, make_plummer
    phi = lambda ps: kernels.compute("phi", ps, ps, mode="fp64")[0]
    return make_plummer(n, 4.0 / n, imf=PowerLaw(2.35, 0.1, 100.0), seed=seed,
                        phi_fn=phi)


def rel_err(got, ref, rows=None):
    """Per-particle |a - a_ref| / |a_ref| and the same for the jerk (max
    and RMS over the rows); got/ref: 6 host arrays or device tensors."""
    host = lambda x: x.cpu().numpy() if hasattr(x, "cpu") else np.asarray(x)
    out = {}
    for name, sl in (("acc", slice(0, 3)), ("jerk", slice(3, 6))):
        g = np.stack([host(x) for x in got[sl]])
        r = np.stack([host(x) for x in ref[sl]])
        if rows is not None:
            g = g[:, rows]
        e = np.linalg.norm(g - r, axis=0) / np.linalg.norm(r, axis=0)
        out[name + "_max"] = float(e.max())
        out[name + "_rms"] = float(np.sqrt(np.mean(e**2)))
    return out


def time_steps(fn, steps, warmup, torch):
    from tupan_amd import _lib
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    _lib.take_kernel_times()  # drop warm-up launches
    _lib.set_timing(True)
    ev = [(torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.set_timing(False)
    k_ms, k_n = _lib.take_kernel_times()
    # per-step device times (events on the launch stream around each step)
    time_steps.step_ms = sorted(a.elapsed_time(b) for a, b in ev)
    return elapsed, k_ms / max(k_n, 1), k_n


def host_abi_rates(ps, n, torch, calls=5):
    """End-to-end rate through the host-pointer entry point (the path of the
    drop-in libtupan.h symbols): H2D of the 8 input fields from pageable
    numpy arrays, the kernel, D2H of the 6 outputs, synchronous."""
    from tupan_amd import kernels, spec
    host = {f: np.ascontiguousarray(getattr(ps, f)) for f in spec.F8}
    res = {"note": "tupan_acc_jerk_host: H2D 8 x %d FP64 (pageable), kernel, "
                   "D2H 6 x %d, synchronous; %d calls" % (n, n, calls)}
    for mode in ("fp32c", "fp64"):
        kernels.compute("acc_jerk", host, host, mode=mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            kernels.compute("acc_jerk", host, host, mode=mode)
        dt = (time.perf_counter() - t0) / calls
        res[mode] = {"value": float(n) * n / dt / 1e9,
                     "unit": "Ginteractions/s", "ms_per_call": 1e3 * dt}
    return res


def host_threads():
    """Host threads for the CPU baseline: the cores this process may run on,
    at most 16 (one GPU's share of its host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(ps, n, target_s, gpu_ieee=None, threads=None):
    """Reference C AccJerk on a bounded sample of i-rows, on one thread and
    on ``threads`` host threads (the rows dealt in contiguous slices, one
    C call per slice: ctypes drops the GIL for the call, and the C kernel is
    reentrant).  gpu_ieee: host copies of the GPU's fp64_ieee outputs for
    all n rows; the multi-thread sample's outputs are compared with them bit
    for bit."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle
    kind = "reference" if oracle.available("float64", "reference") else "port"
    o = oracle.Oracle("float64", kind)
    fields = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")
    jd = {f: np.ascontiguousarray(getattr(ps, f)) for f in fields}
    rng = np.random.RandomState(0)
    threads = threads or host_threads()

    def run(rows, nthr):
        idx = np.sort(rng.choice(n, rows, replace=False))
        idd = {f: np.ascontiguousarray(v[idx]) for f, v in jd.items()}
        cuts = np.linspace(0, rows, nthr + 1).astype(int)
        parts = [{f: v[cuts[k]:cuts[k + 1]] for f, v in idd.items()}
                 for k in range(nthr)]
        t0 = time.perf_counter()
        if nthr == 1:
            outs = [o.call("acc_jerk_kernel", parts[0], jd)]
        else:
            with ThreadPoolExecutor(nthr) as ex:
                outs = list(ex.map(
                    lambda q: o.call("acc_jerk_kernel", q, jd), parts))
        dt = time.perf_counter() - t0
        res = [np.concatenate([r[k] for r in outs]) for k in range(6)]
        return dt, idx, res

    probe_rows = 64
    dt, _, _ = run(probe_rows, 1)
    rows = int(min(n, max(probe_rows, probe_rows * target_s / max(dt, 1e-6))))
    dt1, _, _ = run(rows, 1)
    rate1 = rows * n / dt1
    rows_mt = int(min(n, rows * threads))
    dtm, idx, res = run(rows_mt, threads)
    rate = rows_mt * n / dtm
    extra = {}
    if gpu_ieee is not None:
        same = all(np.array_equal(r.view(np.uint64), g[idx].view(np.uint64))
                   for r, g in zip(res, gpu_ieee))
        extra["fp64_ieee_bitwise_equal"] = bool(same)
    # the reference outputs of the sampled rows, for the accuracy figure
    cpu_baseline.rows, cpu_baseline.outputs = idx, res
    what = ("reference C tupan/lib/src compiled -O3 -std=c99"
            if kind == "reference" else "bit-identical C restatement")
    return dict(extra, **{
        "value": rate / 1e9, "unit": "Ginteractions/s", "cores": threads,
        "kind": kind,
        "sample": "AccJerk FP64, %d sampled i-rows x all %d j (%.1f s, %d "
                  "host threads, %s)" % (rows_mt, n, dtm, threads, what),
        "single_thread": {"value": rate1 / 1e9, "unit": "Ginteractions/s",
                          "sample": "%d sampled i-rows x all %d j (%.1f s)"
                                    % (rows, n, dt1)}})


ENS_CANDIDATES = [(1, 256, 8), (1, 512, 8), (1, 1024, 8), (2, 512, 8),
                  (1, 512, 4), (1, 512, 16), (1, 1024, 16), (2, 512, 16),
                  (1, 512, 32)]


def bench_ensemble(args, torch):
    """512 Plummer clusters, seed 2, N_c log-uniform in [64, 4096], equal
    masses summing to 1 per cluster, eps^2 = 1e-4; every cluster integrated
    with adaptive Hermite-4 (eta = 0.02, t_end = 1) on its own block steps.
    A step advances every cluster by one of its own steps: Tstep + 3 AccJerk
    = 4 N_c^2 pair interactions.  Timed: `steps` ensemble steps after the
    autotuner and `warmup` steps.  Baseline: the single-system device path
    (tupan_amd.integrator, one cluster at a time, kernel launches from the
    host) on every 16th cluster for the same number of steps.  full_run_t1
    (unless --no-full): every cluster to t = 1 in both modes
    (ensemble_full_run)."""
    from tupan_amd import ensemble, kernels
    from tupan_amd import integrator as ig
    sizes = ensemble.cluster_sizes(512, seed=2)
    print("[bench] ensemble: %d clusters, %d particles, sum N^2 = %.3e"
          % (len(sizes), sizes.sum(), float((sizes.astype(float)**2).sum())),
          file=sys.stderr, flush=True)
    phi = lambda p: kernels.compute("phi", p, p, mode="fp64")[0]
    fields = ensemble.make_ensemble(sizes, seed=2, eps2=1e-4, phi_fn=phi)
    ens = ensemble.Ensemble(sizes, 0.02, 1.0)
    print("[bench] autotuning", file=sys.stderr, flush=True)
    best, rates = ens.autotune(fields, trial_steps=4,
                               candidates=ENS_CANDIDATES, use_cache=False)
    ens.set_state(fields)
    ens.step(args.warmup)
    ens.poll()
    _, p0, _ = ens.poll()
    print("[bench] timing %d ensemble steps" % args.steps, file=sys.stderr,
          flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(args.steps)
    left, p1, err = ens.poll()
    elapsed = time.perf_counter() - t0
    ens_rate = (p1 - p0) / elapsed
    ens.close()
    # the same steps with the AccJerk sums in FP32C arithmetic
    print("[bench] fp32c ensemble", file=sys.stderr, flush=True)
    e32 = ensemble.Ensemble(sizes, 0.02, 1.0, mode="fp32c")
    best32, _ = e32.autotune(fields, trial_steps=4,
                             candidates=ENS_CANDIDATES, use_cache=False)
    e32.set_state(fields)
    e32.step(args.warmup)
    _, q0, _ = e32.poll()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e32.step(args.steps)
    _, q1, _ = e32.poll()
    el32 = time.perf_counter() - t0
    e32.close()
    # baseline: looping the single-system path per cluster
    sample = list(range(0, len(sizes), 16))
    off = np.concatenate([[0], np.cumsum(sizes)])
    pairs_b = 0.0
    t_b = 0.0
    print("[bench] per-cluster loop on %d clusters" % len(sample),
          file=sys.stderr, flush=True)
    for c in sample:
        sl = slice(off[c], off[c + 1])
        ps = ig.System.from_host({f: fields[f][sl] for f in ensemble.BASE})
        integ = ig.Integrator(0.02, 0.0, ps, "ahermite4")
        integ.evolve_step(1.0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            integ.evolve_step(1.0)
        torch.cuda.synchronize()
        t_b += time.perf_counter() - t
        pairs_b += 4.0 * float(sizes[c])**2 * args.steps
    loop_rate = pairs_b / t_b
    full = ensemble_full_run(args, sizes, fields, best, best32, torch) \
        if not args.no_full else None
    result = {
        "metric": "ensemble AccJrk+Tstep Ginteractions/s, 512 clusters, "
                  "adaptive Hermite-4",
        "value": ens_rate / 1e9, "unit": "Ginteractions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
        "data": "synthetic",
        "config": {"workload": "ragged ensemble: 512 Plummer clusters seed 2,"
                               " N_c log-uniform [64, 4096], eps^2=1e-4, "
                               "ahermite4 eta=0.02",
                   "particles": int(sizes.sum()),
                   "launch": list(best)},
        "autotune": {"%d,%d,%d" % k: v / 1e9 for k, v in rates.items()},
        "fp32c": {"value": (q1 - q0) / el32 / 1e9,
                  "unit": "Ginteractions/s",
                  "ms_per_step": 1e3 * el32 / args.steps,
                  "launch": list(best32)},
        "per_cluster_loop": {"value": loop_rate / 1e9,
                             "unit": "Ginteractions/s",
                             "clusters": len(sample),
                             "speedup": ens_rate / loop_rate},
        "full_run_t1": full,
    }
    print(json.dumps(result))


def ensemble_full_run(args, sizes, fields, best, best32, torch):
    """Every cluster integrated to t_end = 1 (the configured length), in
    FP64 and with FP32C pair passes, from the same start and with each
    mode's autotuned launch shape: wall time, ensemble steps, pair
    interactions, per-cluster energy error |E(1) - E(0)| / |E(0)| (Phi on
    the GPU in FP64), and the FP32C end state against FP64's."""
    from tupan_amd import ensemble, kernels
    off = np.concatenate([[0], np.cumsum(sizes)])

    def energies(st):
        e = np.empty(len(sizes))
        for c in range(len(sizes)):
            sl = slice(off[c], off[c + 1])
            d = {f: np.ascontiguousarray(st[f][sl]) if f in st
                 else np.ascontiguousarray(fields[f][sl])
                 for f in ensemble.BASE}
            phi = kernels.compute("phi", d, d, mode="fp64")[0]
            m = d["mass"]
            kin = 0.5 * (m * (d["vx"]**2 + d["vy"]**2 + d["vz"]**2)).sum()
            e[c] = kin + 0.5 * (m * phi).sum()
        return e

    e0 = energies({f: fields[f] for f in ensemble.STATE})
    out = {}
    for mode, shape in (("fp64", best), ("fp32c", best32)):
        ens = ensemble.Ensemble(sizes, 0.02, 1.0, mode=mode)
        ens.set_launch(*shape)
        ens.set_state(fields)
        torch.cuda.synchronize()

        def progress(k, left, pairs):
            if k % 512 == 0:
                print("[bench] full run %s: %d steps, %d clusters left"
                      % (mode, k, left), file=sys.stderr, flush=True)

        stats = ens.run(batch=64, graph=True, progress=progress)
        st, tc, ns = ens.get_state()
        ens.close()
        de = np.abs(energies(st) - e0) / np.abs(e0)
        out[mode] = {"seconds": stats["seconds"], "ensemble_steps":
                     stats["steps"], "pairs": stats["pairs"],
                     "value": stats["pairs"] / stats["seconds"] / 1e9,
                     "unit": "Ginteractions/s",
                     "cluster_steps_total": int(ns.sum()),
                     "all_at_t1": bool(np.all(np.abs(tc - 1.0) < 1e-12)),
                     "energy_rel_err_max": float(de.max()),
                     "energy_rel_err_median": float(np.median(de)),
                     "_state": st, "_ns": ns}
    a, b = out["fp32c"].pop("_state"), out["fp64"].pop("_state")
    na, nb = out["fp32c"].pop("_ns"), out["fp64"].pop("_ns")
    same = na == nb
    err = [max(np.abs(a[f][off[c]:off[c + 1]] - b[f][off[c]:off[c + 1]]).max()
               / np.abs(b[f][off[c]:off[c + 1]]).max()
               for f in ensemble.STATE) for c in range(len(sizes))]
    err = np.array(err)
    out["fp32c_vs_fp64"] = {
        "speedup": out["fp64"]["seconds"] / out["fp32c"]["seconds"],
        "clusters_same_step_count": int(same.sum()),
        "state_scale_err_max_same_steps": float(err[same].max())
        if same.any() else None,
        "state_scale_err_median": float(np.median(err))}
    return out


def bench_sizes(args, torch):
    """The single-evaluation kernels at N in {16384, 65536, 262144}: Plummer
    seed 1, Salpeter masses [0.1, 100], eps = 4/N.  Phi, Acc, AccJerk and
    SnapCrackle (its a, j inputs from the FP64 AccJerk) in FP64 and FP32C,
    Tstep (eta = 0.02) in FP64 (its only mode): interactions/s from the live
    engine timing and from the whole call (events around it), the FP32C
    rows recomputed in FP64, and per-particle relative errors
    |x - x64| / |x64| (max, RMS) of every FP32C output against the FP64
    kernel."""
    from tupan_amd import _lib, device, spec
    rows = []
    for n in [int(x) for x in args.sizes.split(",")]:
        print("[bench] sizes: N=%d" % n, file=sys.stderr, flush=True)
        ps = make_workload(n)
        d = device.to_device(ps, spec.F8)
        aj = device.compute("acc_jerk", d, d, mode="fp64")
        for k, f in enumerate(("ax", "ay", "az", "jx", "jy", "jz")):
            d[f] = aj[k].clone()
        for name in ("phi", "acc", "acc_jerk", "snap_crackle", "tstep"):
            res = {}
            modes = ("fp64",) if name == "tstep" else ("fp64", "fp32c")
            extra = (0.02,) if name == "tstep" else ()
            for mode in modes:
                out = tuple(torch.empty(n, dtype=torch.float64, device="cuda")
                            for _ in spec.KERNELS[name].outs)
                step = lambda: device.compute(name, d, d, *extra, mode=mode,
                                              out=out)
                reps = max(2, min(20, int(2e11 / (n * n))))
                _, k_ms, _ = time_steps(step, reps, 1, torch)
                call_ms = float(np.median(time_steps.step_ms))
                res[mode] = (out, k_ms, call_ms)
            pairs = float(n) * n
            row = {"n": n, "kernel": name,
                   "flops_per_interaction": spec.KERNELS[name].flops}
            for mode in modes:
                row[mode + "_Gint_s"] = pairs / (res[mode][1] * 1e-3) / 1e9
                row[mode + "_kernel_ms"] = res[mode][1]
                # the whole call: for FP32C the sort, packing, engine,
                # finish and the FP64 recomputation of the flagged rows
                row[mode + "_call_ms"] = res[mode][2]
                row[mode + "_call_Gint_s"] = pairs / (res[mode][2] * 1e-3) / 1e9
            if "fp32c" in res:
                row["fp32c_rows_recomputed_fp64"] = _lib.last_flagged()
            if "fp32c" not in res:
                rows.append(row)
                print(json.dumps(row), file=sys.stderr, flush=True)
                continue
            got, ref = res["fp32c"][0], res["fp64"][0]
            if name == "phi":
                e = (got[0] - ref[0]).abs() / ref[0].abs()
                row["phi_max"] = float(e.max())
                row["phi_rms"] = float(e.pow(2).mean().sqrt())
            for g in range(len(got) // 3):
                G = torch.stack(got[3 * g:3 * g + 3])
                R = torch.stack(ref[3 * g:3 * g + 3])
                e = (G - R).norm(dim=0) / R.norm(dim=0)
                tag = ("acc", "jerk", "snap", "crackle")[
                    g + (2 if name == "snap_crackle" else 0)]
                row[tag + "_max"] = float(e.max())
                row[tag + "_rms"] = float(e.pow(2).mean().sqrt())
            rows.append(row)
            print(json.dumps(row), file=sys.stderr, flush=True)
    big = [r for r in rows if r["kernel"] == "acc_jerk"][-1]
    for r in rows:
        if r["n"] == 65536 and r["kernel"] == "acc_jerk":
            big = r
    result = {
        "metric": "single-evaluation kernels at N=16K/64K/256K, FP32C and "
                  "FP64, Ginteractions/s and errors vs FP64",
        "value": big["fp32c_Gint_s"], "unit": "Ginteractions/s",
        "n_gpus": 1, "steps": None, "warmup": 1, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "Plummer seed 1, Salpeter [0.1,100], eps=4/N;"
                               " Phi, Acc, AccJerk, SnapCrackle, Tstep"},
        "rows": rows,
    }
    print(json.dumps(result))


def bench_pn(args, torch):
    """Compact PN cluster: N=2048, seed 4, masses uniform in [1, 30]
    normalised to 1, Plummer scaled to half-mass radius 0.1, unsoftened,
    pn_order 5 (1PN + 2PN + 2.5PN) at c = 1e3; pairs closer than 1e-3 are
    drifted by the FP64 Sakura Kepler solve, the rest and the PN terms are
    kicked (integrator 'sakura_pn', dt = 1/8192) for 0.1 time units.  A
    step evaluates far Acc, PNAcc and near Sakura twice each: 6 N^2 pair
    visits.  Reported: energy error including the PN work accumulator,
    closure of momentum and angular momentum with their PN accumulators, and
    pair throughput against the reference's C kernels for the same pair
    visits (Acc and PNAcc, and Sakura over every pair as the reference
    applies it) on a bounded sample of host steps.  pnacc_roofline: the
    PNAcc kernel alone (order 5, PNACC_FLOPS[5] flop per pair) against the
    FP64 vector peak, on the final state and on the N = 65536 benchmark
    model."""
    from tupan_amd import device, kernels, spec
    from tupan_amd import integrator as ig
    from tupan_amd.ics import UniformMass, make_plummer, scale_radius
    n, dt, t_end, rcut, c = 2048, 1.0 / 8192, 0.1, 1e-3, 1e3
    phi = lambda p: kernels.compute("phi", p, p, mode="fp64")[0]
    p = make_plummer(n, 0.0, imf=UniformMass(1.0, 30.0), seed=4, phi_fn=phi)
    scale_radius(p, 0.1 / (1.3048 * 3 * np.pi / 16))
    ps = ig.System.from_host(p)
    integ = ig.Integrator(dt, 0.0, ps, "sakura_pn", rcut=rcut, pn_order=5,
                          clight=c, graph=args.graph)

    def invariants(s):
        d = s.diagnostics()   # PN accumulators included
        return (d["kinetic_energy"] + d["potential_energy"],
                d["linear_momentum"], d["angular_momentum"])

    E0, P0, L0 = invariants(ps)
    x = np.stack([p.rx, p.ry, p.rz])
    close0 = 0
    for i in range(0, n, 256):
        d2 = ((x[:, i:i + 256, None] - x[:, None, :])**2).sum(0)
        close0 += int(((d2 > 0) & (d2 < rcut * rcut)).sum())
    print("[bench] pn: N=%d, %d close pairs at t=0, %d steps"
          % (n, close0 // 2, int(round(t_end / dt))), file=sys.stderr,
          flush=True)
    # warm-up: with graphs, the first step runs eagerly on the capture
    # stream and the second is captured and replayed
    warm = 2 if args.graph else 1
    for _ in range(warm):
        integ.evolve_step(t_end)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nstep = 0
    while abs(ps.t_curr) < t_end:
        integ.evolve_step(t_end)
        nstep += 1
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    integ.integrator.check_status()
    E1, P1, L1 = invariants(ps)
    visits = 6.0 * n * n * nstep
    gpu_rate = visits / elapsed
    # the reference's C kernels for the same pair visits, on the host
    import oracle
    kind = "reference" if oracle.available("float64", "reference") else "port"
    o = oracle.Oracle("float64", kind)
    h = ps.to_host()
    inv = tuple(c**-k for k in range(1, 8))
    t = time.perf_counter()
    reps = 0
    while time.perf_counter() - t < args.cpu_seconds / 2 or reps == 0:
        o.call("acc_kernel", h, h)
        o.call("pnacc_kernel", h, h, 5, *inv)
        o.call("sakura_kernel", h, h, dt / 2, 1)
        reps += 1
    cpu_rate = 3.0 * n * n * reps / (time.perf_counter() - t)
    # the PNAcc kernel alone against the FP64 vector peak: on the final
    # state, and on the N = 65536 Plummer model with the same c and order
    roof = {}
    for tag, hs in (("n2048_final_state", h), ("n65536_plummer", None)):
        if hs is None:
            hs = make_workload(65536)
        dd = device.to_device(hs, spec.F8)
        nn = dd["mass"].numel()
        out = tuple(torch.empty(nn, dtype=torch.float64, device="cuda")
                    for _ in range(3))
        step = lambda: device.compute("pnacc", dd, dd, 5, inv, out=out)
        _, k_ms, _ = time_steps(step, 20 if nn <= 4096 else 5, 2, torch)
        ach = spec.PNACC_FLOPS[5] * float(nn) * nn / (k_ms * 1e-3) / 1e12
        roof[tag] = {"bound": "valu", "achieved": ach,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP64_TFLOPS, "kernel_ms": k_ms,
                     "flops_per_interaction": spec.PNACC_FLOPS[5]}
    result = {
        "metric": "PN cluster pair visits/s (far Acc + PNAcc + near Sakura),"
                  " N=2048",
        "value": gpu_rate / 1e9, "unit": "Gpairs/s", "n_gpus": 1,
        "steps": nstep, "warmup": warm, "ms_per_step": 1e3 * elapsed / nstep,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64", "data": "synthetic",
        "config": {"workload": "compact PN cluster N=2048 seed 4, m~U[1,30],"
                               " r_h=0.1, eps=0, pn_order 5, c=1e3, Sakura "
                               "pairs r<1e-3, dt=1/8192, t=0.1",
                   "close_pairs_t0": close0 // 2, "graph": bool(args.graph)},
        "pnacc_roofline": roof,
        "energy_rel_err": abs(E1 - E0) / abs(E0),
        "momentum_closure": float(np.abs(P1 - P0).max()),
        "angular_momentum_closure": float(np.abs(L1 - L0).max()),
        "cpu_baseline": {"value": cpu_rate / 1e9, "unit": "Gpairs/s",
                         "cores": 1, "kind": kind,
                         "sample": "%d x (Acc + PNAcc order 5 + Sakura over "
                                   "all pairs), N=2048, 1 thread" % reps},
    }
    print(json.dumps(result))


def bench_diag(args, torch):
    """The energy/momentum report (tupan/simulation.py:96-129) on the N =
    65536 benchmark model: one FP64 Phi evaluation plus the fused device
    reduction (System.diagnostics), against the reference's path on one
    host core -- its C Phi kernel (timed on a sample of i-rows, scaled to N)
    plus the report's numpy reductions (timed in full)."""
    from tupan_amd import integrator as ig
    n = 65536
    p = make_workload(n)
    ps = ig.System.from_host(p)
    for _ in range(2):
        ps.diagnostics()
    torch.cuda.synchronize()
    reps = max(3, args.steps // 3)
    t0 = time.perf_counter()
    for _ in range(reps):
        d = ps.diagnostics()
    elapsed = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        ps.diagnostics(with_phi=False)
    red = (time.perf_counter() - t0) / reps
    import oracle
    from oracle import diagnostics as OD
    kind = "reference" if oracle.available("float64", "reference") else "port"
    o = oracle.Oracle("float64", kind)
    h = {f: np.ascontiguousarray(getattr(p, f)) for f in
         ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz")}
    rows = 2048
    idx = np.sort(np.random.RandomState(0).choice(n, rows, replace=False))
    t = time.perf_counter()
    o.call("phi_kernel", {f: v[idx] for f, v in h.items()}, h)
    phi_cpu = (time.perf_counter() - t) * n / rows

    phi_host = ps.phi.cpu().numpy()

    class _K(object):  # Phi is timed above: here the reductions only
        def phi(self, q):
            q["phi"] = phi_host
    t = time.perf_counter()
    q = OD.quantities(_K(), dict(h))
    numpy_cpu = time.perf_counter() - t
    cpu = phi_cpu + numpy_cpu
    result = {
        "metric": "energy/momentum report time, N=65536 (Phi + reductions)",
        "value": 1e3 * elapsed, "unit": "ms", "n_gpus": 1, "steps": reps,
        "warmup": 2, "ms_per_step": 1e3 * elapsed, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
        "data": "synthetic",
        "config": {"workload": "diagnostic report on the Plummer N=65536 "
                               "seed 1 Salpeter model (FP64 Phi + fused "
                               "reduction of 12 totals)"},
        "reduction_ms": 1e3 * red,
        "energy": d["kinetic_energy"] + d["potential_energy"],
        "energy_host": q["kinetic_energy"] + q["potential_energy"],
        "cpu_baseline": {"value": 1e3 * cpu, "unit": "ms", "cores": 1,
                         "kind": kind,
                         "sample": "C phi_kernel on %d sampled i-rows x all "
                                   "%d j (scaled: %.1f s) + numpy report "
                                   "reductions in full (%.2f ms), 1 thread"
                                   % (rows, n, phi_cpu, 1e3 * numpy_cpu)},
    }
    print(json.dumps(result))


def bench_nreg(args, torch):
    """NREG (anreg, the algorithmic-regularisation leapfrog of
    tupan/integrator/nreg.py) on a Plummer N=1024 model (seed 1, equal
    masses, eps = 4/N): regularised steps per second on the device against
    the reference's C NREG kernels driven by the reference's numpy step
    (the checker's restatement) on one host core."""
    from tupan_amd import integrator as ig
    from tupan_amd import kernels
    from tupan_amd.ics import make_plummer
    n, h, steps = 1024, 1.0 / 64, max(16, args.steps)
    phi = lambda q: kernels.compute("phi", q, q, mode="fp64")[0]
    p = make_plummer(n, 4.0 / n, seed=1, phi_fn=phi)
    ps = ig.System.from_host(p)
    integ = ig.Integrator(h, 0.0, ps, "anreg", graph=args.graph)
    warm = 2 if args.graph else 1   # graph: eager step, then capture
    for _ in range(warm):
        integ.evolve_step(1e9)
    torch.cuda.synchronize()
    e0 = integ.integrator.E0
    t0 = time.perf_counter()
    for _ in range(steps):
        integ.evolve_step(1e9)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    d = ps.diagnostics()
    e1 = d["kinetic_energy"] + d["potential_energy"]
    import oracle
    from oracle import integrators as OI
    kind = "reference" if oracle.available("float64", "reference") else "port"
    K = OI.Kernels(oracle.Oracle("float64", kind))
    host = {f: np.ascontiguousarray(getattr(p, f)) for f in
            ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz")}
    host["t_curr"] = 0.0
    st = {}
    OI.nreg_initialize(K, host, st)
    t = time.perf_counter()
    cpu_steps = 0
    while time.perf_counter() - t < args.cpu_seconds / 2 or cpu_steps == 0:
        OI.nreg_do_step(K, host, st, h, "anreg")
        cpu_steps += 1
    cpu_rate = cpu_steps / (time.perf_counter() - t)
    result = {
        "metric": "NREG (anreg) steps/s, Plummer N=1024",
        "value": steps / elapsed, "unit": "steps/s", "n_gpus": 1,
        "steps": steps, "warmup": warm, "ms_per_step": 1e3 * elapsed / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64", "data": "synthetic",
        "config": {"workload": "anreg, Plummer N=1024 seed 1, equal masses, "
                               "eps=4/N, h=1/64 (2 NREG X + 1 NREG V pair "
                               "sums per step)", "graph": bool(args.graph)},
        "energy_rel_err": abs(e1 - e0) / abs(e0),
        "cpu_baseline": {"value": cpu_rate, "unit": "steps/s", "cores": 1,
                         "kind": kind,
                         "sample": "%d anreg steps, C NREG kernels + numpy "
                                   "step, 1 thread" % cpu_steps},
    }
    print(json.dumps(result))


def bench_c1(args, torch):
    """The Hermite-4 trajectory case: Plummer N=256, seed 0, equal masses,
    eps^2 = 1e-4 (eps2 field 5e-5), P(EC)^2 with dt = 1/64: 64 steps per time
    unit, each 3 AccJerk evaluations.  Per mode (fp64, fp32c, fp64_ieee),
    eager and replayed as a HIP graph, one time unit is run to warm up
    (scratch, capture) and the next one is timed.  Reference: the checker's
    Hermite driver (hermite.py's update formulas in numpy) on the reference
    C kernels, one host thread, for the first time unit; the fp64_ieee run's
    state after it must equal the reference run's bit for bit."""
    from tupan_amd import integrator as ig
    from tupan_amd import kernels
    from tupan_amd.ics import make_plummer
    n, dt = 256, 1.0 / 64
    base = ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz")
    phi = lambda q: kernels.compute("phi", q, q, mode="fp64")[0]
    p = make_plummer(n, 0.01, seed=0, phi_fn=phi)
    init = {f: np.ascontiguousarray(getattr(p, f), dtype=np.float64)
            for f in base}
    runs, ieee_state = {}, None
    for mode in ("fp64", "fp32c", "fp64_ieee"):
        for graph in (False, True):
            ps = ig.System.from_host(init)
            integ = ig.Integrator(dt, 0.0, ps, "hermite4", mode=mode,
                                  graph=graph)
            integ.evolve(1.0)
            if mode == "fp64_ieee" and not graph:
                ieee_state = ps.to_host(base)
            el = None
            for unit in (2.0, 3.0):  # best of two timed units
                torch.cuda.synchronize()
                nst = int(ps.nstep[0].item())
                t0 = time.perf_counter()
                integ.evolve(unit)
                torch.cuda.synchronize()
                e = time.perf_counter() - t0
                steps = int(ps.nstep[0].item()) - nst
                el = e if el is None else min(el, e)
            runs["%s%s" % (mode, "_graph" if graph else "")] = {
                "steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps}
            print("[bench] c1 %s graph=%s: %.3f ms/step" % (
                mode, graph, 1e3 * el / steps), file=sys.stderr, flush=True)
    import oracle
    from oracle import integrators as OI
    kind = "reference" if oracle.available("float64", "reference") else "port"
    K = OI.Kernels(oracle.Oracle("float64", kind))
    host = {f: v.copy() for f, v in init.items()}
    host["t_curr"] = 0.0
    host["tstep"] = np.zeros(n)
    t = time.perf_counter()
    cpu_steps = 0
    while abs(host["t_curr"]) < 1.0:
        tau = OI.get_base_tstep(host["t_curr"], 1.0, dt)
        OI.hermite_step(K, host, tau, 4)
        cpu_steps += 1
    cpu_el = time.perf_counter() - t
    same = all(np.array_equal(ieee_state[f].view(np.uint64),
                              host[f].view(np.uint64)) for f in base)
    best = runs["fp64_graph"]
    result = {
        "metric": "Hermite-4 steps/s, Plummer N=256 (FP64)",
        "value": best["steps_per_s"], "unit": "steps/s", "n_gpus": 1,
        "steps": 64, "warmup": 128, "ms_per_step": best["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64", "data": "synthetic",
        "config": {"workload": "hermite4 P(EC)^2, Plummer N=256 seed 0, "
                               "equal masses, eps^2=1e-4, dt=1/64, one time "
                               "unit (3 AccJerk evaluations per step)",
                   "graph": True},
        "modes": runs,
        "fp64_ieee_bitwise_equal": bool(same),
        "cpu_baseline": {"value": cpu_steps / cpu_el, "unit": "steps/s",
                         "cores": 1, "kind": kind,
                         "sample": "%d Hermite-4 steps (numpy driver + C "
                                   "AccJerk), 1 thread" % cpu_steps},
    }
    print(json.dumps(result))


def bench_adjoint(args, torch):
    """The differentiation workload: Plummer N=16384, seed 3, eps^2 = 1e-4,
    target field from the masses perturbed by a seeded 5% Gaussian,
    L = sum_i |a_i - a_target,i|^2; one gradient = forward Acc + the
    adjoint pair sum (tupan_amd.autograd), FP32C and FP64; and the 64-step
    leapfrog rollout (dt = 1/128, checkpoint every 8) with the loss on the
    final positions.  The FP32C gradient's error is taken against the FP64
    one (the FP64 one is pinned to central differences by the tests).
    cpu_baseline: the adjoint pair sum restated in numpy (the reference has
    no differentiation), one host thread, the full sum."""
    from tupan_amd import autograd as ag
    from tupan_amd import kernels
    from tupan_amd.ics import make_plummer
    n, seed, eps2 = 16384, 3, 1e-4
    phi = lambda p: kernels.compute("phi", p, p, mode="fp64")[0]
    ps = make_plummer(n, np.sqrt(eps2), seed=seed, phi_fn=phi)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    pos = torch.stack([dev(ps.rx), dev(ps.ry), dev(ps.rz)]).contiguous()
    vel = torch.stack([dev(ps.vx), dev(ps.vy), dev(ps.vz)]).contiguous()
    rng = np.random.RandomState(seed + 100)
    mass, e2 = dev(ps.mass), dev(ps.eps2)
    mt = dev(ps.mass * (1 + 0.05 * rng.standard_normal(n)))
    at = ag.acc(pos, mt, e2, "fp64").detach()

    def grad(mode):
        p = pos.clone().requires_grad_(True)
        m = mass.clone().requires_grad_(True)
        ((ag.acc(p, m, e2, mode) - at)**2).sum().backward()
        return p.grad, m.grad

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, out

    reps = max(5, args.steps // 3)
    t32, (gp32, gm32) = timed(lambda: grad("fp32c"), reps)
    t64, (gp64, gm64) = timed(lambda: grad("fp64"), reps)
    err_p = float((gp32 - gp64).abs().max() / gp64.abs().max())
    err_m = float((gm32 - gm64).abs().max() / gm64.abs().max())

    def rollout():
        p = pos.clone().requires_grad_(True)
        m = mass.clone().requires_grad_(True)
        pf, _ = ag.leapfrog_rollout(p, vel, m, e2, 1.0 / 128, 64,
                                    checkpoint_every=8, mode="fp32c")
        (pf**2).sum().backward()
        return p.grad
    troll, _ = timed(rollout, 1)
    from oracle import adjoint as oad
    d = {f: np.ascontiguousarray(getattr(ps, f)) for f in
         ("mass", "eps2", "rx", "ry", "rz")}
    g = np.random.RandomState(0).standard_normal((3, n))
    t = time.perf_counter()
    oad.acc_adjoint(d, g)
    cpu = time.perf_counter() - t
    pairs = float(n) * n
    result = {
        "metric": "Acc-loss gradient (forward + adjoint), N=16384, FP32C",
        "value": 1e3 * t32, "unit": "ms", "n_gpus": 1, "steps": reps,
        "warmup": 1, "ms_per_step": 1e3 * t32, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "Plummer N=16384 seed 3, eps^2=1e-4, target "
                               "masses 5% Gaussian (seed 103), "
                               "L=sum|a-a_t|^2; 64-step leapfrog rollout "
                               "dt=1/128, checkpoint 8"},
        "pair_sums_Gpairs_s": 2 * pairs / t32 / 1e9,
        "fp64_gradient_ms": 1e3 * t64,
        "fp32c_vs_fp64_gradient_err": {"positions": err_p, "masses": err_m},
        "rollout_fwd_bwd_ms": 1e3 * troll,
        "cpu_baseline": {"value": 1e3 * cpu, "unit": "ms", "cores": 1,
                         "kind": "port",
                         "sample": "the adjoint pair sum alone (numpy "
                                   "restatement, all %d x %d pairs), 1 "
                                   "thread" % (n, n)},
    }
    print(json.dumps(result))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from tupan_amd import _lib, device, spec

    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "ensemble":
        bench_ensemble(args, torch)
        return
    if args.workload == "pn":
        bench_pn(args, torch)
        return
    if args.workload == "sizes":
        bench_sizes(args, torch)
        return
    if args.workload == "diag":
        bench_diag(args, torch)
        return
    if args.workload == "nreg":
        bench_nreg(args, torch)
        return
    if args.workload == "c1":
        bench_c1(args, torch)
        return
    if args.workload == "adjoint":
        bench_adjoint(args, torch)
        return
    n = args.n
    print("[bench] building Plummer N=%d workload" % n, file=sys.stderr, flush=True)
    ps = make_workload(n)
    d = device.to_device(ps, spec.F8)
    outs = tuple(torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(6))
    if args.pmc_target:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # 256 MiB read + 256 MiB written
        for mode in ("fp32c", "fp64"):
            device.compute("acc_jerk", d, d, mode=mode, out=outs)
        torch.cuda.synchronize()
        print("[bench] pmc target done", file=sys.stderr, flush=True)
        return

    def step():
        device.compute("acc_jerk", d, d, mode=args.mode, out=outs)

    print("[bench] timing %s x%d" % (args.mode, args.steps), file=sys.stderr, flush=True)
    elapsed, k_ms, k_n = time_steps(step, args.steps, args.warmup, torch)
    flagged = _lib.last_flagged() if args.mode == "fp32c" else None
    step_ms = time_steps.step_ms
    pairs = float(n) * n
    value = pairs * args.steps / elapsed / 1e9
    ms_per_step = 1e3 * elapsed / args.steps
    achieved = FLOPS_PER_PAIR * pairs / (k_ms * 1e-3) / 1e12

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "Ginteractions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.mode == "fp32c" else "fp64",
        "data": "synthetic",
        "config": {
            "workload": "AccJerk single evaluation, Plummer N=%d seed 1, "
                        "Salpeter masses [0.1,100], eps=4/N, %s mode"
                        % (n, args.mode),
            "n": n, "mode": args.mode,
            "interactions_per_step": pairs,
        },
        "fp32c_rows_recomputed_fp64": flagged,
        "step_ms": {"median": float(np.median(step_ms)),
                    "min": float(step_ms[0]), "max": float(step_ms[-1]),
                    "note": "device time of each timed step (HIP events on "
                            "the launch stream)"},
        "roofline": {
            "bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": measured_traffic(args.mode) if n == 65536 else None,
            "kernel_ms": k_ms, "kernel_launches": k_n,
            "flops_per_interaction": FLOPS_PER_PAIR,
        },
    }
    if args.mode == "fp32c" and not args.no_fp64:
        ref = tuple(torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(6))

        def step64():
            device.compute("acc_jerk", d, d, mode="fp64", out=ref)

        print("[bench] fp64 reference + timing", file=sys.stderr, flush=True)
        e64, k64, _ = time_steps(step64, max(3, args.steps // 3), 1, torch)
        result["fp64"] = {"value": pairs * max(3, args.steps // 3) / e64 / 1e9,
                          "unit": "Ginteractions/s", "kernel_ms": k64}
        result["accuracy_vs_fp64_gpu"] = rel_err(outs, ref)
        result["accuracy_vs_fp64_gpu"]["rows"] = n
    gpu_ieee = None
    if args.mode == "fp32c" and not args.no_fp64:
        ieee = tuple(torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(6))

        def step_ieee():
            device.compute("acc_jerk", d, d, mode="fp64_ieee", out=ieee)

        print("[bench] fp64_ieee timing", file=sys.stderr, flush=True)
        ei, ki, _ = time_steps(step_ieee, 3, 1, torch)
        result["fp64_ieee"] = {"value": pairs * 3 / ei / 1e9,
                               "unit": "Ginteractions/s", "kernel_ms": ki}
        gpu_ieee = [t.cpu().numpy() for t in ieee]
    if args.mode == "fp32c" and not args.no_fp64:
        print("[bench] host-pointer ABI", file=sys.stderr, flush=True)
        result["host_abi"] = host_abi_rates(ps, n, torch)
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(ps, n, args.cpu_seconds,
                                              gpu_ieee)
        if args.mode == "fp32c":
            # the timed steps' outputs against the reference's C kernels
            # on the rows cpu_baseline evaluated (every row when the host
            # threads finish them all in the sample)
            acc = rel_err(outs, cpu_baseline.outputs, cpu_baseline.rows)
            acc["vs"] = ("%s C AccJerk (cpu_baseline outputs)"
                         % result["cpu_baseline"]["kind"])
            acc["rows"] = int(len(cpu_baseline.rows))
            result["accuracy"] = acc
    print(json.dumps(result))


if __name__ == "__main__":
    main()
