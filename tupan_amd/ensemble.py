"""Ragged ensembles of independent clusters (ensemble.hip) and their
launch-shape autotuner.

Each cluster follows the reference's adaptive Hermite-4 ("ahermite4",
tupan/integrator/hermite.py:340-410) on its own clock; the library batches
the pair sums of all clusters into persistent kernels and replays one
captured HIP graph per ensemble step.  ``Ensemble.run`` queues steps in
batches and polls the count of unfinished clusters between batches.

Autotuning: ``autotune`` times a few ensemble steps for each candidate
launch shape (i-particles per lane, j-chunk length of a work item, resident
waves per SIMD) from the same starting state and keeps the fastest; results
are cached in a JSON file keyed by the device and the cluster-size
signature (``TUPAN_TUNING_CACHE``, default ~/.cache/tupan_amd/ensemble.json).
"""
import ctypes
import hashlib
import json
import os
import time

import numpy as np

from . import _lib

BASE = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")
STATE = ("rx", "ry", "rz", "vx", "vy", "vz")

CANDIDATES = [(ipl, jc, w) for ipl in (1, 2) for jc in (256, 512, 1024)
              for w in (4, 8, 16)]


def _cache_path():
    return os.environ.get("TUPAN_TUNING_CACHE", os.path.join(
        os.path.expanduser("~"), ".cache", "tupan_amd", "ensemble.json"))


class Ensemble(object):
    """``sizes``: particles per cluster (clusters are concatenated in this
    order in every field array)."""

    ORDERS = {"input": 0, "morton": 1}

    def __init__(self, sizes, eta, t_end, mode="fp64", order="morton"):
        self.sizes = np.asarray(sizes, dtype=np.int32)
        if self.sizes.ndim != 1 or len(self.sizes) == 0:
            raise ValueError("sizes must be a non-empty 1-D list")
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)[:-1]])
        self.ntot = int(self.sizes.sum())
        self.eta = float(eta)
        self.t_end = float(t_end)
        self.launch = None
        h = ctypes.c_void_p()
        arr = (ctypes.c_int * len(self.sizes))(*[int(x) for x in self.sizes])
        _lib.check(_lib.lib().tupan_ensemble_create(
            ctypes.byref(h), len(self.sizes), arr, self.eta, self.t_end))
        self._h = h
        self.mode = "fp64"
        if mode != "fp64":
            self.set_mode(mode)
        self.set_order(order)

    def set_order(self, order):
        """Device storage order of each cluster's particles from the next
        set_state: 'morton' (neighbours share waves and j tiles, which
        the FP32C near field needs) or 'input'.  Results come back in
        input order either way."""
        if order not in self.ORDERS:
            raise ValueError("ensemble order: 'morton' or 'input'")
        _lib.check(_lib.lib().tupan_ensemble_set_order(self._h,
                                                       self.ORDERS[order]))
        self.order = order

    def set_mode(self, mode):
        """'fp64' or 'fp32c' (the three pair passes of a step -- Tstep +
        AccJerk at the start state, two corrector AccJerk sums -- in FP32C
        arithmetic, one i-particle per lane; the updates stay FP64)."""
        if mode not in ("fp64", "fp32c"):
            raise ValueError("ensemble mode: 'fp64' or 'fp32c'")
        _lib.check(_lib.lib().tupan_ensemble_set_mode(self._h,
                                                      _lib.MODES[mode]))
        self.mode = mode

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().tupan_ensemble_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, ipl, jchunk, waves_per_simd):
        _lib.check(_lib.lib().tupan_ensemble_set_launch(
            self._h, int(ipl), int(jchunk), int(waves_per_simd)))
        self.launch = (int(ipl), int(jchunk), int(waves_per_simd))

    def set_state(self, fields, t0=None):
        """fields: name -> host array of length ntot (the base set)."""
        arrs = [np.ascontiguousarray(fields[f], dtype=np.float64) for f in BASE]
        for a in arrs:
            if a.shape != (self.ntot,):
                raise ValueError("every field must have the total length")
        self._keep = arrs
        ptrs = _lib.ptr_array([a.ctypes.data for a in arrs])
        t = None
        if t0 is not None:
            t0 = np.ascontiguousarray(t0, dtype=np.float64)
            t = t0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        _lib.check(_lib.lib().tupan_ensemble_set_state(self._h, ptrs, t))

    def get_state(self):
        out = {f: np.empty(self.ntot) for f in STATE}
        tc = np.empty(len(self.sizes))
        ns = np.empty(len(self.sizes), dtype=np.int64)
        ptrs = _lib.ptr_array([out[f].ctypes.data for f in STATE])
        _lib.check(_lib.lib().tupan_ensemble_get_state(
            self._h, ptrs, tc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            ns.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))))
        return out, tc, ns

    def step(self, steps=1, graph=True):
        _lib.check(_lib.lib().tupan_ensemble_step(self._h, int(steps),
                                                  1 if graph else 0))

    def poll(self):
        """(unfinished clusters, pair interactions so far, error flag)."""
        u = ctypes.c_int(0)
        p = ctypes.c_double(0)
        e = ctypes.c_int(0)
        _lib.check(_lib.lib().tupan_ensemble_poll(
            self._h, ctypes.byref(u), ctypes.byref(p), ctypes.byref(e)))
        return u.value, p.value, e.value

    def run(self, batch=32, graph=True, max_steps=None, progress=None):
        """Step until every cluster reaches t_end (or max_steps ensemble
        steps).  Returns dict(steps, pairs, seconds)."""
        t0 = time.perf_counter()
        steps = 0
        while True:
            k = batch if max_steps is None else min(batch, max_steps - steps)
            if k <= 0:
                break
            self.step(k, graph)
            steps += k
            left, pairs, err = self.poll()
            if err:
                raise _lib.TupanError(_lib.ERR_ARG, "a cluster met a zero or "
                                      "non-finite time step")
            if progress is not None:
                progress(steps, left, pairs)
            if left == 0:
                break
        left, pairs, _ = self.poll()
        return {"steps": steps, "pairs": pairs, "unfinished": left,
                "seconds": time.perf_counter() - t0}

    # ---- autotuning -----------------------------------------------------
    def signature(self):
        import torch
        name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "?"
        hist = np.bincount(np.minimum(np.log2(self.sizes).astype(int), 30))
        key = "%s|%s|%d|%s" % (name, self.mode, len(self.sizes),
                               ",".join(map(str, hist)))
        return hashlib.sha1(key.encode()).hexdigest()[:16], key

    def autotune(self, fields, trial_steps=8, candidates=None, use_cache=True):
        """Pick the fastest launch shape on ``trial_steps`` ensemble steps
        from ``fields``; leaves the state reset to ``fields``.  Returns
        (best shape, {shape: pairs/s})."""
        sig, key = self.signature()
        path = _cache_path()
        cache = {}
        if use_cache and os.path.exists(path):
            try:
                cache = json.load(open(path))
            except (OSError, ValueError):
                cache = {}
        if use_cache and sig in cache:
            best = tuple(cache[sig]["launch"])
            self.set_launch(*best)
            self.set_state(fields)
            return best, cache[sig].get("rates", {})
        rates = {}
        cands = candidates or CANDIDATES
        if self.mode == "fp32c":   # one i-particle per lane
            cands = sorted({(1, jc, w) for _, jc, w in cands})
        for cand in cands:
            self.set_launch(*cand)
            self.set_state(fields)
            self.step(1)            # plan, capture, warm up
            self.poll()
            self.set_state(fields)
            t = time.perf_counter()
            self.step(trial_steps)
            _, pairs, _ = self.poll()
            rates[cand] = pairs / (time.perf_counter() - t)
        best = max(rates, key=rates.get)
        self.set_launch(*best)
        self.set_state(fields)
        if use_cache:
            cache[sig] = {"key": key, "launch": list(best),
                          "rates": {"%d,%d,%d" % c: r for c, r in rates.items()}}
            try:
                os.makedirs(os.path.dirname(path), exist_ok=True)
                json.dump(cache, open(path, "w"), indent=1)
            except OSError:
                pass
        return best, rates


def cluster_sizes(ncl=512, seed=2, nmin=64, nmax=4096):
    """N_c = round(exp(U(ln nmin, ln nmax))) from RandomState(seed)."""
    rng = np.random.RandomState(seed)
    return np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), ncl))).astype(int)


def make_ensemble(sizes, seed=2, eps2=1e-4, phi_fn=None):
    """Concatenated Plummer clusters (equal masses summing to 1 per cluster,
    Henon units, eps^2 per pair = eps2); cluster c is make_plummer with
    seed 1000 * seed + c."""
    from .ics import make_plummer
    parts = []
    for c, n in enumerate(sizes):
        ps = make_plummer(int(n), np.sqrt(eps2), seed=1000 * seed + c,
                          phi_fn=phi_fn)
        parts.append({f: np.array(getattr(ps, f)) for f in BASE})
    return {f: np.concatenate([p[f] for p in parts]) for f in BASE}
