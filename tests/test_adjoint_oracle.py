"""Pins for the adjoint checker (oracle/adjoint.py), CPU only.

The adjoints are gradients of the reference's forward sums, so they are
fixed by (1) central finite differences of the C checker's Acc and Phi
kernels (bit-identical to the reference C) with respect to every position
and mass of a small softened set, and (2) the closed-form tidal tensor
d a_k / d x_k.  A dropped term (the m_k g_j or the m_j g_k half of u), a
sign or a swapped index fails (1).
"""
import numpy as np
import pytest

from oracle import Oracle
from oracle import adjoint as oa


@pytest.fixture(scope="module")
def o64():
    return Oracle("float64", "port")


def small_set(seed, n=24):
    rng = np.random.RandomState(seed)
    d = {"mass": rng.uniform(0.2, 1.0, n), "eps2": rng.uniform(1e-3, 1e-2, n)}
    for f in ("rx", "ry", "rz"):
        d[f] = rng.normal(0, 1, n)
    return d


def _fd(fun, d, keys, h=1e-6):
    out = {}
    for f in keys:
        g = np.zeros(len(d[f]))
        for k in range(len(d[f])):
            dp = {kk: v.copy() for kk, v in d.items()}
            dm = {kk: v.copy() for kk, v in d.items()}
            dp[f][k] += h
            dm[f][k] -= h
            g[k] = (fun(dp) - fun(dm)) / (2 * h)
        out[f] = g
    return out


def test_acc_adjoint_matches_finite_differences(o64):
    d = small_set(1)
    rng = np.random.RandomState(2)
    g = rng.normal(0, 1, (3, len(d["mass"])))
    loss = lambda p: float((np.stack(o64.call("acc_kernel", p, p)) * g).sum())
    fd = _fd(loss, d, ("rx", "ry", "rz", "mass"))
    gx, gm = oa.acc_adjoint(d, g)
    for k, f in enumerate(("rx", "ry", "rz")):
        assert np.allclose(gx[k], fd[f], rtol=1e-6, atol=1e-7 * np.abs(fd[f]).max())
    assert np.allclose(gm, fd["mass"], rtol=1e-6, atol=1e-7 * np.abs(fd["mass"]).max())


def test_phi_adjoint_matches_finite_differences(o64):
    d = small_set(3)
    rng = np.random.RandomState(4)
    h = rng.normal(0, 1, len(d["mass"]))
    loss = lambda p: float((o64.call("phi_kernel", p, p)[0] * h).sum())
    fd = _fd(loss, d, ("rx", "ry", "rz", "mass"))
    gx, gm = oa.phi_adjoint(d, h)
    for k, f in enumerate(("rx", "ry", "rz")):
        assert np.allclose(gx[k], fd[f], rtol=1e-6, atol=1e-7 * np.abs(fd[f]).max())
    assert np.allclose(gm, fd["mass"], rtol=1e-6, atol=1e-7 * np.abs(fd["mass"]).max())


def test_acc_adjoint_single_source_is_tidal_tensor():
    d = small_set(5)
    n = len(d["mass"])
    for k in (0, 7, n - 1):
        g = np.zeros((3, n))
        g[:, k] = [0.3, -1.2, 0.7]
        gx, _ = oa.acc_adjoint(d, g)
        T = oa.tidal_tensor(d, k)
        assert np.allclose(T, T.T, rtol=1e-13, atol=0)
        assert np.allclose(gx[:, k], T @ g[:, k], rtol=1e-12, atol=1e-14)


def test_adjoint_drops_r2_zero_pairs_like_the_forward():
    """Pairs with r^2 = 0 (the self pair, and distinct coincident particles)
    are dropped as in the forward kernels: without softening the sums stay
    finite.  (At exactly coincident distinct particles the forward Acc is
    differentiable with the pair's term m_j/s^3 I; the adjoint, like the
    forward mask, leaves it out.)"""
    d = small_set(6, 8)
    d["eps2"][:] = 0.0
    for f in ("rx", "ry", "rz"):
        d[f][5] = d[f][3]
    g = np.ones((3, 8))
    for gx, gm in (oa.acc_adjoint(d, g), oa.phi_adjoint(d, g[0])):
        assert np.isfinite(gx).all() and np.isfinite(gm).all()
