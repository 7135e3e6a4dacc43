"""Kernel table: field lists, outputs and flop conventions.

Field order is the reference's argument order (tupan/lib/src/libtupan.h);
output attribute names are the ones tupan/lib/extensions.py attaches to the
i-particles (Phi :132, Acc :198, AccJerk :279-280, SnapCrackle :337-338,
Tstep :382, PNAcc :434, Sakura :501-502, NREG_X :555-557, NREG_V :609-610).

``flops`` is the reference's own per-interaction count from the comments
of each *_kernel_common.h (phi :32, acc :38, acc_jerk :56, snap_crackle :95,
tstep :60).  PNAcc has none ("40+???", pnacc_kernel_common.h:70): its count
here is PNACC_FLOPS[pn_order], the operations one pair executes (each +, -,
*, / once; negations, rational constants and additions of unused PN slots
free), counted by tests/pnacc_flops.cpp on pn_terms.inc plus the 33 of the
set-up and accumulation lines; the "pnacc" entry carries the order-5 count
(1PN + 2PN + 2.5PN).  Sakura has none ("36 + ?") and none is defined: its
cost per pair depends on the Kepler solve's iterations.
"""

# floating-point operations of one PNAcc pair per PN order
PNACC_FLOPS = {0: 50, 1: 50, 2: 82, 3: 82, 4: 179, 5: 197, 6: 452, 7: 626}
from collections import namedtuple

F5 = ("mass", "rx", "ry", "rz", "eps2")
F8 = F5 + ("vx", "vy", "vz")
F14 = F8 + ("ax", "ay", "az", "jx", "jy", "jz")
# classification positions of the pair-selected kernels (positions fixed
# for a hybrid step)
CP = ("cx", "cy", "cz")

Kernel = namedtuple("Kernel", "name kid fields outs has_mode flops")

KERNELS = {
    "phi": Kernel("phi", 0, F5, ("phi",), True, 14),
    "acc": Kernel("acc", 1, F5, ("ax", "ay", "az"), True, 20),
    "acc_jerk": Kernel("acc_jerk", 2, F8,
                       ("ax", "ay", "az", "jx", "jy", "jz"), True, 42),
    "snap_crackle": Kernel("snap_crackle", 3, F14,
                           ("sx", "sy", "sz", "cx", "cy", "cz"), True, 114),
    "tstep": Kernel("tstep", 4, F8, ("tstep", "tstepij"), True, 42),
    "pnacc": Kernel("pnacc", 5, F8, ("pnax", "pnay", "pnaz"), True, 197),
    "sakura": Kernel("sakura", 6, F8,
                     ("drx", "dry", "drz", "dvx", "dvy", "dvz"), False, None),
    "nreg_x": Kernel("nreg_x", 9, F8,
                     ("mrx", "mry", "mrz", "ax", "ay", "az", "u"), True, 37),
    "nreg_v": Kernel("nreg_v", 10, ("mass", "vx", "vy", "vz", "ax", "ay", "az"),
                     ("mvx", "mvy", "mvz", "mk"), True, 25),
    # pair-selected variants for hybrid schemes (no reference counterpart):
    # Acc over the pairs whose classification positions CP are >= rcut
    # apart, Sakura over the pairs < rcut apart
    "acc_far": Kernel("acc_far", 1, F5 + CP, ("ax", "ay", "az"), False, 20),
    "sakura_near": Kernel("sakura_near", 6, F8 + CP,
                          ("drx", "dry", "drz", "dvx", "dvy", "dvz"), False,
                          None),
    # reverse-mode derivatives (no reference counterpart): inputs F5 plus the
    # upstream gradient dL/da (gx, gy, gz) or dL/dphi (h); outputs dL/dr and
    # dL/dm.  Flops per pair count every IEEE operation of the FP64 pair
    # (rsq as one).
    "acc_adjoint": Kernel("acc_adjoint", 7, F5 + ("gx", "gy", "gz"),
                          ("grx", "gry", "grz", "gm"), True, 48),
    "phi_adjoint": Kernel("phi_adjoint", 8, F5 + ("h",),
                          ("grx", "gry", "grz", "gm"), True, 25),
}

KERNEL_IDS = {k: v.kid for k, v in KERNELS.items()}

# kernels with the reference FP32 build's arithmetic (mode "fp32_ieee")
FP32_IEEE = ("phi", "acc", "acc_jerk", "snap_crackle", "tstep")

# reference C symbol -> kernel key
SYMBOLS = {
    "phi_kernel": "phi",
    "acc_kernel": "acc",
    "acc_jerk_kernel": "acc_jerk",
    "snap_crackle_kernel": "snap_crackle",
    "tstep_kernel": "tstep",
    "pnacc_kernel": "pnacc",
    "sakura_kernel": "sakura",
    "nreg_Xkernel": "nreg_x",
    "nreg_Vkernel": "nreg_v",
    "kepler_solver_kernel": "kepler",
}

# kepler_solver_kernel (libtupan.h:231-246): two particles, updated in place
KEPLER_FIELDS = F8
KEPLER_OUTS = ("rx", "ry", "rz", "vx", "vy", "vz")
