"""CPU statement of the Acc / Phi adjoints -- TEST INFRASTRUCTURE.

Only ``tests/`` use this module, as the checker of the adjoint kernels
(tupan_amd/csrc/adjoint.hip); the product never imports it.

The reference has no differentiation, so there is nothing to restate: these
are the gradients of its forward sums (tupan/lib/src/acc_kernel_common.h:
7-37, phi_kernel_common.h:7-31, softening of smoothing.h:10-35) written
directly as pair sums in numpy, in row blocks so that memory stays bounded.
They are pinned (tests/test_adjoint_oracle.py) by central finite
differences of the forward kernels of the C checker and by the closed-form
tidal tensor.
"""
import numpy as np


def _pairs(d, rows):
    x = np.stack([d["rx"], d["ry"], d["rz"]])
    r = x[:, rows, None] - x[:, None, :]           # r_kj = x_k - x_j
    r2 = (r * r).sum(0)
    s2 = r2 + (d["eps2"][rows, None] + d["eps2"][None, :])
    y = np.where(r2 > 0, 1.0 / np.sqrt(np.where(r2 > 0, s2, 1.0)), 0.0)
    return r, y


def acc_adjoint(d, g, block=256):
    """d: dict of the F5 fields; g: (3, n) upstream dL/da.  Returns
    (dL/dx (3, n), dL/dm (n,))."""
    n = len(d["mass"])
    m = d["mass"]
    gx = np.zeros((3, n))
    gm = np.zeros(n)
    for b in range(0, n, block):
        rows = np.arange(b, min(n, b + block))
        r, y = _pairs(d, rows)
        y3 = y**3
        u = m[rows][None, :, None] * g[:, None, :] - m[None, None, :] * g[:, rows, None]
        ru = (r * u).sum(0)
        gx[:, rows] = (y3[None] * (u - 3.0 * (ru * y * y)[None] * r)).sum(2)
        gm[rows] = (y3 * (g[:, None, :] * r).sum(0)).sum(1)
    return gx, gm


def phi_adjoint(d, h, block=256):
    """h: (n,) upstream dL/dphi.  Returns (dL/dx (3, n), dL/dm (n,))."""
    n = len(d["mass"])
    m = d["mass"]
    gx = np.zeros((3, n))
    gm = np.zeros(n)
    for b in range(0, n, block):
        rows = np.arange(b, min(n, b + block))
        r, y = _pairs(d, rows)
        w = (h[rows, None] * m[None, :] + h[None, :] * m[rows, None]) * y**3
        gx[:, rows] = (w[None] * r).sum(2)
        gm[rows] = -(h[None, :] * y).sum(1)
    return gx, gm


def tidal_tensor(d, k):
    """T_k = d a_k / d x_k = -sum_j m_j (I/s^3 - 3 r r^T / s^5), r = x_k - x_j
    (closed form of the derivative of the softened Acc sum)."""
    x = np.stack([d["rx"], d["ry"], d["rz"]])
    r = x[:, k, None] - x
    r2 = (r * r).sum(0)
    mask = r2 > 0
    s2 = r2 + d["eps2"][k] + d["eps2"]
    T = np.zeros((3, 3))
    for j in np.nonzero(mask)[0]:
        s = np.sqrt(s2[j])
        T -= d["mass"][j] * (np.eye(3) / s**3
                             - 3.0 * np.outer(r[:, j], r[:, j]) / s**5)
    return T
