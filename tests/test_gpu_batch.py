"""Batched contexts: several independent simulations (own np, grid, walls,
bonds, parameters, dt) stepped by the same launches.  Every simulation of the
batch must match the oracle run on its own .bin, bit for bit."""

import numpy
import pytest

import oracle
from sphere_amd import engine

from . import cases
from .test_gpu_bonds import _bonded
from .test_gpu_parity import EXACT_FIELDS

pytestmark = pytest.mark.gpu

WALL_FIELDS = ("walls_nx", "walls_mvfd", "walls_acc", "walls_force_pp")


def _compare_sim(ctx, sim, ref, fields):
    bad = []
    for name in fields:
        a, b = ctx.get(name, sim=sim), ref.get(name)
        if a.shape != b.shape or not numpy.array_equal(a, b):
            bad.append("sim %d: %s" % (sim, name))
    return bad


def _run_batch(scenes, steps, chunks=3, **scalars):
    images = [s.tobytes() for s in scenes]
    refs = [oracle.Oracle(im) for im in images]
    ctx = engine.Context(images=images)
    assert ctx.nsims == len(images)
    for k, v in scalars.items():
        ctx.set_scalar(k, v)
    per = steps // chunks
    for _ in range(chunks):
        ctx.step(per)
        for r in refs:
            r.step(per)
    return ctx, refs


def _mixed_scenes():
    a = cases.settling_case(300, 11)
    b = cases.fixture("shear1k.bin")
    c = cases.crowded_case()
    d = _bonded(frac=0.3, seed=4, sigma_b=300.0)
    e = cases.fixture("shear1k.bin")
    e.shear(shear_stress=2.0e3)
    return [a, b, c, d, e]


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_matches_oracle_per_simulation(mode):
    scenes = _mixed_scenes()
    ctx, refs = _run_batch(scenes, 120, sort_mode=mode)
    bad = []
    for k, (s, r) in enumerate(zip(scenes, refs)):
        fields = EXACT_FIELDS + WALL_FIELDS
        if int(s.nb0[0]) > 0:
            fields = fields + ("bonds", "bonds_delta", "bonds_omega")
        bad += _compare_sim(ctx, k, r, fields)
        assert ctx.tobytes(sim=k) == r.tobytes(), k
    assert bad == []
    ctx.close()


def test_batch_with_adaptive_grid_and_lane_counts():
    scenes = _mixed_scenes()[:3]
    s = cases.fixture("shear1k.bin")
    s.adaptiveGrid()
    scenes.append(s)
    for lanes in (1, 8):
        ctx, refs = _run_batch(scenes, 90, contact_lanes=lanes)
        bad = []
        for k, r in enumerate(refs):
            bad += _compare_sim(ctx, k, r, EXACT_FIELDS + WALL_FIELDS)
        assert bad == [], lanes
        ctx.close()


def test_batch_fields_and_scalars_are_per_simulation():
    scenes = _mixed_scenes()[:2]
    ctx = engine.Context(images=[s.tobytes() for s in scenes])
    for k, s in enumerate(scenes):
        assert int(ctx.scalar("np", sim=k)) == int(s.np[0])
        assert ctx.scalar("dt", sim=k) == float(s.time_dt[0])
        assert numpy.array_equal(ctx.get("x", sim=k)[:, :3], s.x)
    ctx.step(10)
    t = [ctx.scalar("time", sim=k) for k in range(2)]
    assert t[0] != t[1]  # each simulation advances by its own dt
    with pytest.raises(engine.EngineError):
        ctx.get("x", sim=2)
    ctx.close()


def test_graph_replay_bit_exact():
    """Steady-state chunks replayed from a captured HIP graph, with reads of
    the state between replays, match the oracle on every replay."""
    scenes = _mixed_scenes()[:3]
    images = [s.tobytes() for s in scenes]
    refs = [oracle.Oracle(im) for im in images]
    ctx = engine.Context(images=images)
    ctx.set_scalar("graph_steps", 8)
    for chunk in (5, 16, 24, 8, 3):
        ctx.step(chunk)
        bad = []
        for k, r in enumerate(refs):
            r.step(chunk)
            bad += _compare_sim(ctx, k, r, EXACT_FIELDS)
        assert bad == [], chunk
    ctx.close()


def test_batch_of_sheared_beds_with_own_stiffness_and_dt():
    """The benchmark's batch workload at small sizes: tiled sheared beds,
    each with its own friction, normal stress, stiffness, damping and dt,
    stepped together (with graph replay), each bit-exact with the oracle."""
    from sphere_amd import scenes as sc
    base = cases.fixture("shear1k.bin")
    cells = sc.batch_shear_cells([1000, 2000, 3000, 1000, 4000, 2000], [base], seed=11,
                                 kn_range=(1.0e6, 1.0e8))
    assert len({float(c.time_dt[0]) for c in cells}) == len(cells)
    ctx, refs = _run_batch(cells, 150, graph_steps=25)
    bad = []
    for k, r in enumerate(refs):
        bad += _compare_sim(ctx, k, r, EXACT_FIELDS + WALL_FIELDS)
    assert bad == []
    ctx.close()
