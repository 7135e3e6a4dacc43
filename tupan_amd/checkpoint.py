"""Checkpoint and resume of a device-resident run.

The reference restarts from a pickle of its whole Simulation object
(tupan/simulation.py:178-185, 222-228) and dumps snapshots as HDF5 groups
``Snapshot[_NNNNNN]/<system>/<member>``, one structured dataset per member
type (tupan/io/hdf5io.py:49-64).  Here a checkpoint is one numpy ``.npz``
file that loads with ``allow_pickle=False`` (nothing in it executes on
load), holding the same snapshot naming in a per-field variant (one array
per field of the one particle set, not one structured dataset per member):

  ``Snapshot[_NNNNNN]/system/<field>``  every field of the device System
                                        (state, derivatives, PN accumulators,
                                        clocks), copied to the host
  ``Integrator/state``                  a JSON document: method, eta, mode,
                                        physical time, PN order and speed of
                                        light, graph replay, the Sakura
                                        status word and the method's own
                                        scalars (the Sakura-PN cut-off,
                                        NREG's E0, W, U, S)

Every kernel is deterministic, so a run resumed from a checkpoint continues
bit for bit as the uninterrupted run would have.  HIP graphs are not
stored: a resumed graph run captures its steps again.
"""
import json

import numpy as np
import torch

from . import _lib
from . import integrator as ig

STATE_KEY = "Integrator/state"


def _snapshot_name(snap_number):
    if isinstance(snap_number, int):
        return "Snapshot_" + str(snap_number).zfill(6)
    return "Snapshot"


def save(path, integ, snap_number=None):
    """Write ``integ`` (an Integrator or one of its method classes) and its
    particle set to ``path`` (numpy adds ``.npz`` when missing)."""
    it = getattr(integ, "integrator", integ)
    ps = it.ps
    if ps.device.type == "cuda":
        torch.cuda.synchronize(ps.device)
    snap = _snapshot_name(snap_number)
    arrays = {"%s/system/%s" % (snap, k): v.cpu().numpy()
              for k, v in ps._f.items()}
    pn = bool(ps.include_pn_corrections)
    state = {"snapshot": snap, "n": ps.n, "method": it.method,
             "eta": it.eta, "mode": ps.mode, "t_curr": float(ps.t_curr),
             "pn_order": it.clight.pn_order if pn else 0,
             "clight": it.clight.clight if pn else None,
             "graph": bool(it.graph), "status": int(it.status.item())}
    if isinstance(it, ig.SakuraPN):
        state["rcut"] = it.rcut
    if isinstance(it, ig.NREG) and it.is_initialized:
        if it.graph:
            sc = it._sc.cpu().numpy()
            state["nreg"] = {k: float(sc[getattr(_lib, "NREG_" + k)])
                             for k in ("E0", "W", "U", "S")}
        else:
            state["nreg"] = {"E0": it.E0, "W": it.W, "U": it.U, "S": it.S}
    arrays[STATE_KEY] = np.array(json.dumps(state))
    np.savez(path, **arrays)


def load(path, device="cuda", graph=None):
    """Rebuild the Integrator saved in ``path`` on ``device``; ``graph``
    overrides the saved graph-replay choice."""
    with np.load(path, allow_pickle=False) as z:
        state = json.loads(str(z[STATE_KEY]))
        base = state["snapshot"] + "/system/"
        fields = {k[len(base):]: z[k] for k in z.files if k.startswith(base)}
    ps = ig.System(state["n"], device)
    for k, v in fields.items():
        t = ps._f["nstep"] if k == "nstep" else ps.field(k)
        t.copy_(torch.as_tensor(v))
    kw = {"mode": state["mode"],
          "graph": state["graph"] if graph is None else bool(graph)}
    if state["pn_order"]:
        kw.update(pn_order=state["pn_order"], clight=state["clight"])
    if "rcut" in state:
        kw["rcut"] = state["rcut"]
    integ = ig.Integrator(state["eta"], state["t_curr"], ps, state["method"],
                          **kw)
    it = integ.integrator
    it.status.fill_(state["status"])
    if "nreg" in state:
        s = state["nreg"]
        it.E0, it.W, it.U, it.S = s["E0"], s["W"], s["U"], s["S"]
        it.is_initialized = True
        if it.graph:
            init = [0.0] * _lib.NREG_SCALARS
            for k in ("E0", "W", "U", "S"):
                init[getattr(_lib, "NREG_" + k)] = s[k]
            init[_lib.NREG_T] = init[_lib.NREG_T0] = float(ps.t_curr)
            it._sc = torch.tensor(init, dtype=torch.float64).to(ps.device)
    return integ
