"""The full-size workloads: the 4,096-sphere gravity settling box, a
1e5-particle direct-shear bed (the benchmark's size and stiffness) and a
1e6-particle bonded shear bed (2% of the touching pairs bonded, 20 kPa,
5e-4 m/s), stepped by the engine and by the oracle from the same state and
compared bit for bit; then the properties that do not depend on size, at
full size: the same result for every contact lane count, with graph
replay, and for a batch holding the bed twice."""

import numpy
import pytest

import oracle
from sphere_amd import engine

from . import cases
from .test_gpu_parity import EXACT_FIELDS, _compare

pytestmark = pytest.mark.gpu

WALL_FIELDS = ("walls_nx", "walls_mvfd", "walls_acc", "walls_force_pp")
STATE = ("x", "vel", "angvel", "force", "torque", "contacts", "delta_t", "es", "ev", "p")


def _same_state(a, b, fields=STATE):
    return [f for f in fields if not numpy.array_equal(a.get(f), b.get(f))]


def test_shear_bed_1e5_bit_exact():
    s = cases.lattice_bed(100000, seed=2001, median=0.01, sigma=0.2, k_n=1.16e7,
                          normal_stress=10.0e3, shear_vel=1.0e-3)
    image = s.tobytes()
    steps = 20
    print("1e5 bed: %d particles, %d steps" % (int(s.np[0]), steps), flush=True)
    ref = oracle.Oracle(image)
    ref.step(steps)
    ctx = engine.Context(image)
    ctx.step(steps)
    bad = _compare(ctx, ref) + _compare(ctx, ref, WALL_FIELDS)
    assert bad == []
    n = int(s.np[0])
    assert (ctx.get("contacts") != n).sum() > 2 * n
    # size-independent: every lane count, and graph replay, give the same
    for lanes in (1, 8, 16):
        with engine.Context(image) as other:
            other.set_scalar("contact_lanes", lanes)
            other.step(steps)
            assert _same_state(other, ctx) == [], lanes
    with engine.Context(image) as other:
        other.set_scalar("graph_steps", 5)
        other.step(steps)
        assert _same_state(other, ctx) == []
    ctx.close()


def test_gravity_settling_4096_bit_exact():
    """The 4,096-sphere gravity settling workload (radii uniform in
    [0.005, 0.01] m, seeded random sequential insertion in a 0.2 x 0.2 x
    0.6 m box with frictional walls), 3,000 steps, every field."""
    from sphere_amd import scenes
    s = scenes.settling_box(4096, seed=1001)
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    for _ in range(3):
        ref.step(1000)
        ctx.step(1000)
        assert _compare(ctx, ref) == []
    ctx.close()


def test_bonded_shear_bed_1e6_bit_exact():
    s = cases.lattice_bed(1000000, seed=5001, median=0.005, sigma=0.25, k_n=1.0e8,
                          normal_stress=20.0e3, shear_vel=5.0e-4, bond_frac=0.02)
    image = s.tobytes()
    steps = 4
    print("1e6 bonded bed: %d particles, %d bonds, %d steps" % (int(s.np[0]), int(s.nb0[0]),
                                                                 steps), flush=True)
    ref = oracle.Oracle(image)
    ref.step(steps)
    ctx = engine.Context(image)
    ctx.step(steps)
    bad = _compare(ctx, ref) + _compare(ctx, ref, WALL_FIELDS)
    bad += _compare(ctx, ref, ("bonds", "bonds_delta", "bonds_omega"))
    assert bad == []
    n = int(s.np[0])
    assert (ctx.get("contacts") != n).sum() > 2 * n
    assert int(s.nb0[0]) > 20000
    ctx.close()


def test_bonded_shear_bed_1e6_breaking_bit_exact():
    """The large bonded shear bed with its configured bonds: 1e6 spheres
    (lognormal radii, median 0.005 m, sigma 0.25), 20 kPa, 5e-4 m/s, 2% of
    the touching pairs bonded with a stiffness of 1e8 N/m that break at a
    normal extension of 0.05 of the smaller radius (SPH_S_BOND_K,
    SPH_S_BOND_BREAK_EXT).  Bonds break throughout the 300 steps; every
    field and every bond bit-exact with the oracle at steps 100, 200, 300."""
    s = cases.lattice_bed(1000000, seed=5001, median=0.005, sigma=0.25, k_n=1.0e8,
                          normal_stress=20.0e3, shear_vel=5.0e-4, bond_frac=0.02)
    image = s.tobytes()
    nb = int(s.nb0[0])
    ref = oracle.Oracle(image)
    ref.set_bond_model(1.0e8, 0.05)
    ctx = engine.Context(image)
    ctx.set_scalar("bond_k", 1.0e8)
    ctx.set_scalar("bond_break_ext", 0.05)
    broken = []
    for k in range(3):
        print("1e6 bonded bed with breaking bonds: steps %d..%d" % (100 * k, 100 * (k + 1)),
              flush=True)
        ref.step(100)
        ctx.step(100)
        bad = _compare(ctx, ref) + _compare(ctx, ref, WALL_FIELDS)
        bad += _compare(ctx, ref, ("bonds", "bonds_delta", "bonds_omega"))
        assert bad == [], (k, bad)
        broken.append(int((ref.get("bonds")[:, 0] == nb).sum()))
    print("bonds broken after 100, 200, 300 steps: %s of %d" % (broken, nb), flush=True)
    assert 0 < broken[0] < broken[1] < broken[2] < nb
    ctx.close()


def test_batch_of_two_full_beds_matches_single():
    """A batch holding the 1e5 bed twice steps each copy exactly as a
    context of its own."""
    s = cases.lattice_bed(100000, seed=2001, median=0.01, sigma=0.2, k_n=1.16e7,
                          normal_stress=10.0e3, shear_vel=1.0e-3)
    image = s.tobytes()
    with engine.Context(image) as one, engine.Context(images=[image, image]) as two:
        one.step(10)
        two.step(10)
        for k in (0, 1):
            for f in ("x", "vel", "angvel", "force", "contacts"):
                assert numpy.array_equal(two.get(f, sim=k), one.get(f)), (k, f)


def test_batch_of_256_beds_bit_exact():
    """The batched sweep at full size: 256 sheared beds of log-uniform size
    in [500, 20000] (1.36e6 particles), each with its own stiffness
    (log-uniform 1e6..1e8 N/m, its own dt), friction (0.3..0.7) and normal
    stress (5..160 kPa), in one context; every simulation bit-exact with
    its own oracle after 20 steps."""
    from sphere_amd import scenes
    sizes = scenes.batch_sizes(256, seed=7)
    rng = numpy.random.default_rng(7)
    images = []
    for k, n in enumerate(sizes):
        s = cases.lattice_bed(int(n), seed=100 + k, median=0.0075, sigma=0.2,
                              k_n=float(10.0 ** rng.uniform(6.0, 8.0)),
                              normal_stress=float(rng.uniform(5.0e3, 160.0e3)),
                              mu=float(rng.uniform(0.3, 0.7)))
        images.append(s.tobytes())
    print("batch: %d simulations, %d particles" % (len(images), int(sum(sizes))), flush=True)
    with engine.Context(images=images) as ctx:
        ctx.step(20)
        for k, image in enumerate(images):
            ref = oracle.Oracle(image)
            ref.step(20)
            for f in ("x", "vel", "angvel", "force", "contacts", "delta_t"):
                assert numpy.array_equal(ctx.get(f, sim=k), ref.get(f)), (k, f)
            assert numpy.array_equal(ctx.get("walls_nx", sim=k), ref.get("walls_nx")), k


def test_benchmark_bed_1000_steps_bit_exact():
    """The benchmark's own workload over a long horizon: the 1e5-particle
    direct-shear bed settled, consolidated under 10 kPa and sheared at
    1e-3 m/s exactly as the benchmark builds it (settled_shear_cell with
    seed 2001 and its default settling, stepped by the engine), then 1,000
    steps on the GPU and on the oracle, every field and the walls compared
    after each 250.  The interval holds re-sorts of the particle order and
    rebuilds of the candidate lists."""
    from sphere_amd import scenes
    s = scenes.settled_shear_cell(100000, 2001, open_stepper=engine.Context)
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    sorts0, builds0 = ctx.scalar("sorts"), ctx.scalar("vl_builds")
    for k in range(4):
        ref.step(250)
        ctx.step(250)
        bad = _compare(ctx, ref) + _compare(ctx, ref, WALL_FIELDS)
        assert bad == [], (250 * (k + 1), bad)
    sorts = int(ctx.scalar("sorts") - sorts0)
    builds = int(ctx.scalar("vl_builds") - builds0)
    print("1e5 benchmark bed, 1000 steps: %d re-sorts, %d list rebuilds" % (sorts, builds))
    assert sorts >= 1 and builds >= 1
    ctx.close()


def test_consolidation_batch_of_256_500_steps_bit_exact():
    """The benchmark's batched sweep over a long horizon: the 256
    consolidation beds it times (consolidation_batch with seed 2001: settled
    500-particle cells with radii uniform in [0.005, 0.01] m tiled to
    log-uniform sizes in [500, 20000], each with its own stiffness, dt,
    friction and normal stress), 500 steps in one context, every simulation
    bit-exact with its own oracle.  The oracles run on a pool of threads,
    one OpenMP thread each."""
    from concurrent.futures import ThreadPoolExecutor
    import os
    from sphere_amd import scenes
    cells = scenes.consolidation_batch(256, 2001, engine.Context)
    images = [c.tobytes() for c in cells]
    fields = ("x", "vel", "angvel", "force", "torque", "contacts", "delta_t", "es", "ev")
    with engine.Context(images=images) as ctx:
        ctx.step(500)
        got = [{f: ctx.get(f, sim=k) for f in fields + ("walls_nx",)}
               for k in range(len(images))]

    def check(k):
        oracle.set_threads(1)
        ref = oracle.Oracle(images[k])
        ref.step(500)
        bad = [f for f in fields + ("walls_nx",) if not numpy.array_equal(got[k][f], ref.get(f))]
        ref.close()
        return bad

    threads = oracle.get_threads()
    workers = max(1, min(16, os.cpu_count() or 1))
    try:
        with ThreadPoolExecutor(workers) as pool:
            results = list(pool.map(check, range(len(images))))
    finally:
        oracle.set_threads(threads)
    bad = [(k, b) for k, b in enumerate(results) if b]
    assert bad == []
