"""sph_porosity_noalloc (no counterpart in the reference): the variant of
sph_porosity that stages its per-slice results in the library's fixed
device array instead of a per-call allocation.  Called through the C ABI
and compared bit for bit with sph_porosity, on consecutive calls whose
inputs shrink (fewer slices, then a smaller simulation), and in a batch."""

import ctypes

import numpy
import pytest

from sphere_amd import engine

from . import cases

pytestmark = pytest.mark.gpu

SPH_E_INPUT, SPH_E_ARG = -3, -5  # include/sphere_hip.h
MAX_SLICES = 65536


def _noalloc(ctx, slices, sim=0):
    L = engine.load()
    z = numpy.full(slices, numpy.nan)
    por = numpy.full(slices, numpy.nan)
    rc = L.sph_porosity_noalloc(ctx._h, sim, slices, z.ctypes.data, por.ctypes.data)
    return rc, z, por


def _check(ctx, slices, sim=0):
    rc, za, pa = _noalloc(ctx, slices, sim)
    assert rc == 0
    zb, pb = ctx.porosity(slices, sim)
    assert numpy.array_equal(za, zb) and numpy.array_equal(pa, pb)
    assert numpy.isfinite(pa).all() and (pa < 1.0).any()


def test_noalloc_porosity_matches_sph_porosity():
    big = cases.fixture("shear1k.bin")
    small = cases.settling_case(300, 11)
    print("noalloc porosity: shear1k (1000 particles) 8 then 4 slices", flush=True)
    with engine.Context(big.tobytes()) as ctx:
        ctx.step(20)
        _check(ctx, 8)
        _check(ctx, 4)  # the second call: fewer slices in the same array
    print("noalloc porosity: settling box (300 particles) 3 slices", flush=True)
    with engine.Context(small.tobytes()) as ctx:
        ctx.step(20)
        _check(ctx, 3)  # a smaller simulation


def test_noalloc_porosity_in_a_batch_and_its_limits():
    a, b = cases.fixture("shear1k.bin"), cases.settling_case(300, 11)
    print("noalloc porosity: batch of shear1k and a 300-particle box", flush=True)
    with engine.Context(images=[a.tobytes(), b.tobytes()]) as ctx:
        ctx.step(10)
        _check(ctx, 6, sim=0)
        _check(ctx, 3, sim=1)
        # Beyond the array: refused before any device work.
        rc, _, _ = _noalloc(ctx, MAX_SLICES + 1)
        assert rc == SPH_E_ARG
        assert b"maximum" in engine.load().sph_last_error(ctx._h)
        # Slices thinner than twice the last radius: the reference's error.
        rc, _, _ = _noalloc(ctx, 10000)
        assert rc == SPH_E_INPUT
        # The array still serves the next call.
        _check(ctx, 5, sim=0)
