"""Scenes shared by the CPU and GPU tests (built on the CPU, seeded)."""

import os

import numpy

from sphere_amd import scenes
from sphere_amd.sim import sim

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    """A committed .bin fixture, read (and upgraded to v2.15 when older)."""
    s = sim()
    s.readbin(os.path.join(GOLDEN, name), verbose=False)
    return s


def settling_case(n, seed):
    """A closed box with downward and sideways initial velocities, so that
    particles reach the floor and the side walls within a short test."""
    box = (0.12, 0.12, 0.3) if n <= 400 else (0.16, 0.16, 0.45)
    s = scenes.settling_box(n=n, seed=seed, box=box)
    rng = numpy.random.default_rng(seed)
    s.vel[:, 2] = -rng.uniform(0.5, 2.0, n)
    s.vel[:, 0] = rng.uniform(-0.5, 0.5, n)
    return s


def crowded_case(n_small=44, seed=9):
    """One large sphere overlapped by more small spheres than a contact row
    holds (NC = 32), so that the slot policy's overflow rules (slot nc is
    overwritten, the rest of the cell is skipped) decide the contact rows."""
    from sphere_amd.sim import sim
    rng = numpy.random.default_rng(seed)
    s = sim(np=1 + n_small, sid="crowded")
    s.radius[0] = 1.0
    s.radius[1:] = rng.uniform(0.25, 0.3, n_small)
    s.x[0, :] = [3.0, 3.0, 3.0]
    # Fibonacci sphere directions, each small sphere 0.05 inside contact.
    k = numpy.arange(n_small) + 0.5
    phi = numpy.arccos(1.0 - 2.0 * k / n_small)
    th = numpy.pi * (1.0 + 5.0 ** 0.5) * k
    d = numpy.stack([numpy.cos(th) * numpy.sin(phi), numpy.sin(th) * numpy.sin(phi),
                     numpy.cos(phi)], axis=1)
    s.x[1:, :] = s.x[0, :] + d * (1.0 + s.radius[1:, None] - 0.05)
    s.vel[1:, :] = rng.normal(0.0, 0.01, (n_small, 3))
    s.defineWorldBoundaries(L=[6.0, 6.0, 6.0])
    s.initTemporal(total=1.0, file_dt=0.1)
    return s
