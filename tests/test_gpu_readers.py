"""Output-side readers on the GPU against the oracle's restatement of the
reference: DEM::findOverlaps (`sphere --contacts`) and DEM::porosity
(`sphere --porosity`), bit for bit, on the current state of a context."""

import numpy
import pytest

import oracle
from sphere_amd import engine

from . import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["shear1k", "settling", "crowded"])
def test_find_overlaps_matches_oracle(case):
    if case == "shear1k":
        s = cases.fixture("shear1k.bin")
    elif case == "settling":
        s = cases.settling_case(1200, 12)
    else:
        s = cases.crowded_case()
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    ref.step(30)
    ctx.step(30)
    pa, da = ctx.find_overlaps()
    pb, db = ref.find_overlaps()
    assert len(pb) > 0
    assert numpy.array_equal(pa, pb) and numpy.array_equal(da, db)
    ctx.close()


def test_find_overlaps_in_a_batch_and_with_a_short_buffer():
    a, b = cases.fixture("shear1k.bin"), cases.settling_case(300, 11)
    ctx = engine.Context(images=[a.tobytes(), b.tobytes()])
    for k, s in enumerate((a, b)):
        pa, da = ctx.find_overlaps(sim=k)
        pb, db = oracle.Oracle(s.tobytes()).find_overlaps()
        assert numpy.array_equal(pa, pb) and numpy.array_equal(da, db)
    # a buffer shorter than the list: the first pairs, and the full count
    import ctypes
    L = engine.load()
    n = ctypes.c_uint64()
    pairs = numpy.zeros((5, 2), dtype=numpy.uint32)
    delta = numpy.zeros(5)
    assert L.sph_find_overlaps(ctx._h, 0, pairs.ctypes.data, delta.ctypes.data, 5,
                               ctypes.byref(n)) == 0
    pb, db = oracle.Oracle(a.tobytes()).find_overlaps()
    assert n.value == len(pb)
    assert numpy.array_equal(pairs, pb[:5]) and numpy.array_equal(delta, db[:5])
    ctx.close()


@pytest.mark.parametrize("case,slices", [("shear1k", 8), ("settling", 5)])
def test_porosity_matches_oracle(case, slices):
    s = cases.fixture("shear1k.bin") if case == "shear1k" else cases.settling_case(1200, 12)
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    ref.step(20)
    ctx.step(20)
    za, pa = ctx.porosity(slices)
    zb, pb = ref.porosity(slices)
    assert numpy.array_equal(za, zb) and numpy.array_equal(pa, pb)
    with pytest.raises(engine.EngineError):
        ctx.porosity(100000)
    ctx.close()
