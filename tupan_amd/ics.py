"""Initial conditions: Plummer sphere and mass functions.

``make_plummer`` replays tupan/ics/plummer.py:21-133 call for call, so that a
given seed produces the reference's particles bit for bit (legacy numpy
global RNG, masses drawn under a saved/restored RNG state, then
permutation, radii, angles, potential, rejection-sampled speeds, angles;
then centre of mass to origin, Henon units and virial scaling).

The three potential evaluations inside (``set_phi`` for the speeds, and the
potential energy in ``to_nbody_units`` and ``scale_to_virial``) go through
``phi_fn(ps) -> phi``; by default that is the product Phi kernel on the GPU
(``ps.set_phi(ps)``).

Mass functions:
* ``EqualMass``: tupan/ics/imf.py:74-82.  With a constant density the
  reference's rejection sampler accepts every draw and returns 1.0 for all
  masses; it runs under a saved RNG state, so only the values matter.
* ``PowerLaw(alpha, mlow, mhigh)``: dN/dm ~ m^-alpha on [mlow, mhigh] by
  inverse CDF (Salpeter: alpha = 2.35).  The reference's
  ``IMF.salpeter1955`` rejects mlow < 0.4 (imf.py:84-90), so the [0.1, 100]
  range the kernel benchmarks use needs this sampler.
* ``UniformMass(lo, hi)``: masses uniform in [lo, hi].
"""
import numpy as np

from .particles import ParticleSystem


class EqualMass(object):
    def sample(self, n):
        return np.ones(n)


class PowerLaw(object):
    def __init__(self, alpha, mlow, mhigh):
        self.alpha, self.mlow, self.mhigh = float(alpha), float(mlow), float(mhigh)

    def sample(self, n):
        u = np.random.random(n)
        k = 1.0 - self.alpha
        lo, hi = self.mlow**k, self.mhigh**k
        return (lo + u * (hi - lo))**(1.0 / k)


class UniformMass(object):
    def __init__(self, lo, hi):
        self.lo, self.hi = float(lo), float(hi)

    def sample(self, n):
        return np.random.uniform(self.lo, self.hi, size=n)


def _default_phi(ps):
    ps.set_phi(ps)
    return ps.phi


def make_plummer(n, eps, imf=None, seed=None, mfrac=0.999, phi_fn=None):
    """Plummer sphere in Henon units (M=1, E=-1/4), softening type 0
    (uniform): every particle gets eps2 = eps^2/2 (plummer.py:34-63)."""
    imf = EqualMass() if imf is None else imf
    phi_fn = _default_phi if phi_fn is None else phi_fn
    ps = ParticleSystem(n)
    np.random.seed(seed)
    ilist = np.arange(n)
    ps.id[...] = ilist

    srand = np.random.get_state()
    ps.mass[...] = imf.sample(n)
    ps.mass /= ps.total_mass
    eps2 = np.ones(n)
    eps2 *= (eps * eps) / np.mean(eps2)
    ps.eps2[...] = eps2 / 2
    np.random.set_state(srand)

    # positions, plummer.py:65-75
    irand = np.random.permutation(ilist)
    mrand = (irand + np.random.random(n)) * mfrac / n
    radius = 1.0 / np.sqrt(np.power(mrand, -2.0 / 3.0) - 1.0)
    theta = np.arccos(np.random.uniform(-1.0, 1.0, size=n))
    phi = np.random.uniform(0.0, 2.0 * np.pi, size=n)
    ps.rx[...] = radius * np.sin(theta) * np.cos(phi)
    ps.ry[...] = radius * np.sin(theta) * np.sin(phi)
    ps.rz[...] = radius * np.cos(theta)

    pot = np.array(phi_fn(ps), dtype=np.float64)

    # speeds, plummer.py:77-93 (the reference's r2 < r1 rejection)
    count = 0
    rnd = np.empty(n)
    while count < n:
        r1 = np.random.random()
        r2 = np.random.random()
        if r2 < r1:
            rnd[count] = r2
            count += 1
    velocity = np.sqrt(-2 * rnd * pot)
    theta = np.arccos(np.random.uniform(-1.0, 1.0, size=n))
    phi = np.random.uniform(0.0, 2.0 * np.pi, size=n)
    ps.vx[...] = velocity * np.sin(theta) * np.cos(phi)
    ps.vy[...] = velocity * np.sin(theta) * np.sin(phi)
    ps.vz[...] = velocity * np.cos(theta)

    def potential_energy():
        return 0.5 * float((ps.mass * np.asarray(phi_fn(ps))).sum())

    # com_to_origin, to_nbody_units, scale_to_virial (body.py:172-176, 400-461)
    ps.com_to_origin()
    ps.dynrescale_total_mass(1.0)
    mtot = ps.total_mass
    rvir = (mtot**2) / (-2 * potential_energy())
    r_scale = 1.0 / rvir
    v_scale = 1 / r_scale**0.5
    ps.rx *= r_scale
    ps.ry *= r_scale
    ps.rz *= r_scale
    ps.vx *= v_scale
    ps.vy *= v_scale
    ps.vz *= v_scale
    ke = ps.kinetic_energy
    pe = potential_energy()
    v_scale = ((-0.5 * pe) / ke)**0.5
    ps.vx *= v_scale
    ps.vy *= v_scale
    ps.vz *= v_scale
    return ps


def scale_radius(ps, factor):
    """Scale positions by ``factor`` and velocities by 1/sqrt(factor): keeps
    the system in virial equilibrium (used for the compact PN cluster,
    whose half-mass radius is set to 0.1)."""
    ps.rx *= factor
    ps.ry *= factor
    ps.rz *= factor
    vs = 1.0 / np.sqrt(factor)
    ps.vx *= vs
    ps.vy *= vs
    ps.vz *= vs
    return ps
