"""Every scene the GPU tests use is built (and stepped on the oracle) here."""

import pytest

import oracle
from sphere_amd import scenes

from . import cases


@pytest.mark.parametrize("n,seed", [(300, 11), (1200, 12)])
def test_settling_case_builds(n, seed):
    s = cases.settling_case(n, seed)
    o = oracle.Oracle(s.tobytes())
    o.step(2)


@pytest.mark.parametrize("n,seed,shrink", [(500, 21, 0.97), (2000, 22, 0.95), (800, 31, 0.96),
                                           (800, 32, 0.96), (400, 41, 0.97), (600, 7, 0.97)])
def test_shear_case_builds(n, seed, shrink):
    s = scenes.shear_box(n=n, seed=seed, phi=0.3, shrink=shrink)
    o = oracle.Oracle(s.tobytes())
    o.step(2)
    assert (o.get("contacts") != n).sum() > 0


@pytest.mark.parametrize("name", ["shear1k.bin", "cube-init.output00029.bin", "init-1e4-slab.bin"])
def test_fixture_loads_and_is_contact_rich(name):
    s = cases.fixture(name)
    n = int(s.np[0])
    o = oracle.Oracle(s.tobytes())
    o.step(1)
    assert (o.get("contacts") != n).sum() > n
