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
