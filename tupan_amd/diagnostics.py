"""The energy/momentum report of tupan/simulation.py:39-129 on the device.

``Diagnostic(fname, time, ps, report_freq, pn_order)`` takes a device
``System`` and prints the reference's columns (time, dtime, ke, pe, te,
virial, eerr, geerr, com_r, com_v, lmom, amom, wct) with the reference's
format.  Each report costs one Phi evaluation and the report's twelve
totals summed in numpy's order on the device (``System.diagnostics``), so
that a run in the reference-arithmetic mode prints the reference's numbers
bit for bit; the reference evaluates Phi and then about a dozen numpy
reductions on the host.
"""
from __future__ import print_function

import math
import sys
import time as _time

HEADER = ('{0:13s} {1:10s} '
          '{2:10s} {3:10s} {4:15s} '
          '{5:10s} {6:10s} {7:10s} '
          '{8:10s} {9:10s} {10:10s} '
          '{11:10s} {12:13s}')
ROW = ('{time:< 13.6e} {dtime:< 10.3e} '
       '{ke:< 10.3e} {pe:< 10.3e} {te:< 15.8e} '
       '{virial:< 10.3e} {eerr:< 10.3e} {geerr:< 10.3e} '
       '{com_r:< 10.3e} {com_v:< 10.3e} {lmom:< 10.3e} '
       '{amom:< 10.3e} {wct:< 13.6e}')


def myprint(data, fname, fmode):
    """simulation.py:29-36."""
    if fname is None:
        return
    if fname == '<stdout>':
        print(data, file=sys.stdout)
    elif fname == '<stderr>':
        print(data, file=sys.stderr)
    else:
        with open(fname, fmode) as fobj:
            print(data, file=fobj)


def _norm(a, b):
    return float((((a - b)**2).sum())**0.5)


class Diagnostic(object):
    """simulation.py:39-129.  ``fname`` None computes without printing;
    ``print_diagnostic`` also returns the row's values as a dict."""

    def __init__(self, fname, time, ps, report_freq=4, pn_order=0):
        self.fname = fname
        self.time = time
        self.report_freq = report_freq
        self.include_pn_corrections = bool(pn_order)
        self.nreport = 0
        self.is_initialized = False

    def initialize(self, ps):
        d = ps.diagnostics(np_order=True)
        self.ke0 = d["kinetic_energy"]
        self.pe0 = d["potential_energy"]
        self.te0 = self.ke0 + self.pe0
        self.com_r0 = d["com_r"]
        self.com_v0 = d["com_v"]
        self.lmom0 = d["linear_momentum"]
        self.amom0 = d["angular_momentum"]
        self.count = 0
        self.ceerr = 0.0
        self.t_start = _time.time()
        self.print_header()
        self.is_initialized = True

    def print_header(self):
        myprint(HEADER.format('#00:time', '#01:dtime',
                              '#02:ke', '#03:pe', '#04:te',
                              '#05:virial', '#06:eerr', '#07:geerr',
                              '#08:com_r', '#09:com_v', '#10:lmom',
                              '#11:amom', '#12:wct'),
                self.fname, 'w')

    def diagnostic_report(self, ps):
        if not self.is_initialized:
            self.initialize(ps)
        t_curr = ps.t_curr
        row = None
        if self.nreport % self.report_freq == 0:
            row = self.print_diagnostic(t_curr, t_curr - self.time, ps)
        self.nreport += 1
        return row

    def print_diagnostic(self, time, dtime, ps):
        self.time = time
        d = ps.diagnostics(np_order=True)
        ke = d["kinetic_energy"]
        pe = d["potential_energy"]
        te = ke + pe
        eerr = (te - self.te0) / (-pe)
        self.count += 1
        self.ceerr += eerr**2
        geerr = math.sqrt(self.ceerr / self.count)
        row = dict(time=time, dtime=dtime, ke=ke, pe=pe, te=te,
                   virial=d["virial_energy"], eerr=eerr, geerr=geerr,
                   com_r=_norm(d["com_r"], self.com_r0),
                   com_v=_norm(d["com_v"], self.com_v0),
                   lmom=_norm(d["linear_momentum"], self.lmom0),
                   amom=_norm(d["angular_momentum"], self.amom0),
                   wct=_time.time() - self.t_start)
        myprint(ROW.format(**row), self.fname, 'a')
        return row
