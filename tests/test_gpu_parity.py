"""HIP engine vs CPU oracle, through the C ABI, on identical .bin images.

Integer and index outputs (cell keys, permutation, cell ranges, contact
lists) must be bit-exact; the FP64 state is also bit-exact because both
sides evaluate the reference's expressions in the same order without FMA
contraction and with correctly rounded square roots.  Where a test allows
a tolerance it says so.
"""

import numpy
import pytest

import oracle
from sphere_amd import engine, scenes

from . import cases

pytestmark = pytest.mark.gpu

EXACT_FIELDS = ("perm", "cellid", "contacts", "x", "vel", "angvel", "angpos", "force",
                "torque", "xyzsum", "acc", "angacc", "es_dot", "es", "ev_dot", "ev", "p",
                "delta_t")


def _compare(ctx, ref, fields=EXACT_FIELDS):
    bad = []
    for name in fields:
        a, b = ctx.get(name), ref.get(name)
        if a.shape != b.shape or not numpy.array_equal(a, b):
            if a.dtype.kind == "f" and a.shape == b.shape:
                d = numpy.abs(a - b)
                bad.append("%s max|d|=%g at %s" % (name, d.max(), numpy.unravel_index(
                    numpy.argmax(d), d.shape)))
            else:
                bad.append(name)
    return bad


def _occupied_cells_equal(ctx, ref):
    cs_a, cs_b = ctx.get("cellstart"), ref.get("cellstart")
    assert numpy.array_equal(cs_a, cs_b)
    occ = cs_b != 0xFFFFFFFF
    assert numpy.array_equal(ctx.get("cellend")[occ], ref.get("cellend")[occ])


def _run_pair(s, steps, chunks=1):
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    per = steps // chunks
    for _ in range(chunks):
        ref.step(per)
        ctx.step(per)
    return ctx, ref


@pytest.mark.parametrize("n,seed", [(300, 11), (1200, 12)])
def test_settling_box_bit_exact(n, seed):
    s = cases.settling_case(n, seed)
    ctx, ref = _run_pair(s, 200, chunks=4)
    assert _compare(ctx, ref) == []
    _occupied_cells_equal(ctx, ref)
    ctx.close()


@pytest.mark.parametrize("name,steps", [("shear1k.bin", 300), ("cube-init.output00029.bin", 200),
                                        ("init-1e4-slab.bin", 200)])
def test_fixture_bit_exact(name, steps):
    """Contact-rich beds: the reference's own settled outputs (upgraded to
    v2.15) and a grown shear cell with a stress-controlled top wall."""
    s = cases.fixture(name)
    n = int(s.np[0])
    ctx, ref = _run_pair(s, steps, chunks=4)
    bad = _compare(ctx, ref)
    bad += _compare(ctx, ref, ("walls_nx", "walls_mvfd", "walls_acc", "walls_force_pp"))
    assert bad == []
    _occupied_cells_equal(ctx, ref)
    c = ctx.get("contacts")
    assert (c != n).sum() > n  # contact-rich: more than one partner per particle
    ctx.close()


def test_settling_full_trajectory_bit_exact():
    """The 4,096-sphere settling workload, all 2,000 steps."""
    s = scenes.settling_box()
    ctx, ref = _run_pair(s, 2000, chunks=8)
    assert _compare(ctx, ref) == []
    _occupied_cells_equal(ctx, ref)
    ctx.close()


def test_contact_model_1_bit_exact():
    s = cases.fixture("shear1k.bin")
    s.contactmodel[0] = 1
    ctx, ref = _run_pair(s, 40, chunks=2)
    assert _compare(ctx, ref, [f for f in EXACT_FIELDS if f not in ("contacts", "delta_t")]) == []
    ctx.close()


def test_capillary_cohesion_within_tolerance():
    """Model 1 with liquid bridges (db > 0, capillaryCohesion_exp).  The
    bridge force goes through FP32 expf, whose last bit is library-specific,
    so this one case is compared with a tolerance: relative 1e-6 (absolute
    1e-6 of the field's largest magnitude) after 40 steps."""
    s = cases.fixture("shear1k.bin")
    s.contactmodel[0] = 1
    s.kappa[0] = 0.5
    s.db[0] = 2.0e-4
    s.V_b[0] = 1.0e-9
    ctx, ref = _run_pair(s, 40, chunks=2)
    for name in ("x", "vel", "angvel", "force", "torque"):
        a, b = ctx.get(name), ref.get(name)
        scale = numpy.abs(b).max()
        assert numpy.allclose(a, b, rtol=1e-6, atol=1e-6 * scale), name
    ctx.close()


def test_contact_model_3_bit_exact():
    s = cases.fixture("init-1e4-slab.bin")
    s.contactmodel[0] = 3
    s.gamma_r[0] = 1.0
    s.mu_r[0] = 0.1
    ctx, ref = _run_pair(s, 40, chunks=2)
    assert _compare(ctx, ref) == []
    ctx.close()


def test_write_bin_matches_oracle():
    s = scenes.shear_box(n=400, seed=41, phi=0.3, shrink=0.97)
    ctx, ref = _run_pair(s, 25)
    assert ctx.tobytes() == ref.tobytes()
    ctx.close()


def test_host_writes_between_steps_keep_parity():
    """Positions written back between steps (a re-hash before the next step)
    and an iteration counter of the other parity leave results unchanged."""
    s = cases.fixture("shear1k.bin")
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    ref.step(60)
    ctx.step(60)
    ctx.set("x", ctx.get("x"))
    ref.step(60)
    ctx.step(60)
    it = ctx.scalar("iter")
    ctx.set_scalar("iter", it + 1)
    ctx.step(61)
    ref.step(61)
    assert _compare(ctx, ref) == []
    # New velocities on both sides (the contact search reads sorted copies).
    for name in ("vel", "angvel"):
        v = ref.get(name).copy()
        v[:, :3] *= 1.01
        ref.set(name, v)
        ctx.set(name, v)
    ctx.step(40)
    ref.step(40)
    assert _compare(ctx, ref) == []
    ctx.close()


@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
@pytest.mark.parametrize("case", ["settling", "shear1k", "crowded", "hertz"])
def test_contact_lanes_bit_exact(case, lanes):
    """Every lane count of the contact kernel gives the oracle's contact rows
    and state: settling (re-sorts and list rebuilds), a dense sheared bed, a
    contact row that overflows NC, and the Hertz model."""
    if case == "settling":
        s, steps = cases.settling_case(1200, 12), 200
    elif case == "crowded":
        s, steps = cases.crowded_case(), 40
    else:
        s, steps = cases.fixture("shear1k.bin"), 150
        if case == "hertz":
            s.contactmodel[0] = 3
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    ctx.set_scalar("contact_lanes", lanes)
    assert ctx.scalar("contact_lanes") == lanes
    for _ in range(2):
        ref.step(steps // 2)
        ctx.step(steps // 2)
    assert _compare(ctx, ref) == []
    ctx.close()


@pytest.mark.parametrize("case", ["settling", "shear1k", "hertz"])
def test_fused_contact_integration_bit_exact(case):
    """The contact kernel that also integrates (one vector component per
    lane) gives the oracle's state, order and rows."""
    if case == "settling":
        s, steps = cases.settling_case(1200, 12), 200
    else:
        s, steps = cases.fixture("shear1k.bin"), 150
        if case == "hertz":
            s.contactmodel[0] = 3
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    ctx.set_scalar("fused", 1)
    assert ctx.scalar("fused") == 1
    for _ in range(2):
        ref.step(steps // 2)
        ctx.step(steps // 2)
    assert _compare(ctx, ref) == []
    ctx.close()


@pytest.mark.parametrize("lanes", [1, 4])
def test_xcd_workgroup_order_bit_exact(lanes):
    """The contact kernel with each XCD given one contiguous range of
    workgroups gives the oracle's rows and state as well."""
    image = cases.fixture("shear1k.bin").tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    assert ctx.scalar("xcd_order") == 0
    ctx.set_scalar("contact_lanes", lanes)
    ctx.set_scalar("xcd_order", 1)
    assert ctx.scalar("xcd_order") == 1
    ref.step(100)
    ctx.step(100)
    assert _compare(ctx, ref) == []
    ctx.close()


@pytest.mark.parametrize("mode,merge_max", [(0, 2048), (1, 2048), (1, 3), (1, 0), (2, 2048)])
@pytest.mark.parametrize("case", ["settling", "shear1k"])
def test_resort_placement_bit_exact(case, mode, merge_max):
    """The re-sort at each step start (mode 0), in one workgroup at the end of
    the previous step (mode 1: merging the particles that changed cell into
    the old order, or sorting anew when more than merge_max changed), and the
    automatic choice (mode 2) all give the oracle's order and state."""
    if case == "settling":
        s, steps = cases.settling_case(1200, 12), 240
    else:
        s = cases.fixture("shear1k.bin")
        s.vel[:, :2] += 5.0  # a uniform drift: particles keep crossing cell faces
        steps = 240
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    ctx.set_scalar("sort_mode", mode)
    ctx.set_scalar("resort_merge_max", merge_max)
    for _ in range(4):
        ref.step(steps // 4)
        ctx.step(steps // 4)
    assert ctx.scalar("sorts") > 2
    assert _compare(ctx, ref) == []
    _occupied_cells_equal(ctx, ref)
    ctx.close()


@pytest.mark.parametrize("per_wall", [False, True])
def test_triaxial_walls_bit_exact(per_wall):
    """Five dynamic walls (triaxial: wall 0 at constant velocity, the side
    walls under a normal stress), with the reference's wall-0 force sum for
    every wall, and with each wall's own sum (SPH_S_WALL_SUMS = 1)."""
    s = cases.settling_case(1200, 12)
    s.triaxial(wvel=-0.01, normal_stress=2.0e3)
    image = s.tobytes()
    try:
        oracle.set_wall_sums_per_wall(per_wall)
        ref = oracle.Oracle(image)
        ctx = engine.Context(image)
        ctx.set_scalar("wall_sums", 1 if per_wall else 0)
        for _ in range(2):
            ref.step(80)
            ctx.step(80)
    finally:
        oracle.set_wall_sums_per_wall(False)
    bad = _compare(ctx, ref)
    for f in ("walls_nx", "walls_mvfd", "walls_acc"):
        if not numpy.array_equal(ctx.get(f), ref.get(f)):
            bad.append(f)
    assert bad == []
    mv = ctx.get("walls_mvfd").reshape(-1, 4)
    if per_wall:
        assert len({float(v) for v in mv[1:, 2]}) > 1  # the side walls' own forces
    ctx.close()


@pytest.mark.parametrize("case", ["shear1k", "batch"])
def test_autotune_keeps_results(case):
    """sph_autotune's trials are ordinary steps under other settings: the
    simulation it leaves, and every step after it, are bit-exact with the
    oracle's; and it leaves valid settings."""
    a = cases.fixture("shear1k.bin")
    if case == "shear1k":
        images = [a.tobytes()]
        ctx = engine.Context(images[0])
    else:
        images = [a.tobytes(), cases.settling_case(1200, 12).tobytes()]
        ctx = engine.Context(images=images)
    us = ctx.autotune(20)
    assert us > 0.0
    assert ctx.scalar("contact_lanes") in (4, 8, 16)
    assert ctx.scalar("contact_wg") in (128, 256, 512)
    assert ctx.scalar("vl_cap") in (16, 24, 32)
    assert ctx.scalar("radix_bits") in (4, 6, 8)
    # 3 (lanes) + 3 (workgroup size, with 4 lanes) + 3 (list capacity) + 2
    # (workgroup order) trials of 30 steps; the digit width times sorts.
    done = int(ctx.scalar("iter"))
    assert done % 30 == 0 and 8 * 30 <= done <= 11 * 30
    ctx.step(30)
    for k, image in enumerate(images):
        ref = oracle.Oracle(image)
        ref.step(done + 30)
        for name in ("x", "vel", "angvel", "force", "torque", "contacts", "delta_t"):
            assert numpy.array_equal(ctx.get(name, sim=k), ref.get(name)), (k, name)
    ctx.close()


@pytest.mark.parametrize("knob,value,case", [
    ("contact_wg", 128, "shear1k"), ("contact_wg", 512, "shear1k"), ("contact_wg", 512, "batch"),
    ("vl_cap", 8, "shear1k"), ("vl_cap", 16, "batch"), ("vl_cap", 24, "settling"),
    ("radix_bits", 4, "settling"), ("radix_bits", 6, "settling"), ("radix_bits", 4, "batch"),
    ("radix_bits", 6, "shear1k")])
@pytest.mark.parametrize("sort_mode", [0, 1])
def test_tuning_knobs_bit_exact(knob, value, case, sort_mode):
    """Every setting the autotuner tries gives the oracle's results bit for
    bit: the contact workgroup size, the candidate-list capacity (particles
    beyond it scan their 27 cells) and the radix digit width, with the sort
    as per-step launches (sort mode 0) and as the end-of-step one-workgroup
    re-sort forced to sort anew (sort mode 1, merge limit 0)."""
    a = cases.fixture("shear1k.bin")
    if case == "shear1k":
        images = [a.tobytes()]
    elif case == "settling":
        images = [cases.settling_case(1200, 12).tobytes()]
    else:
        images = [a.tobytes(), cases.settling_case(1200, 12).tobytes()]
    ctx = engine.Context(images=images) if len(images) > 1 else engine.Context(images[0])
    ctx.set_scalar("sort_mode", sort_mode)
    for k in range(len(images)):
        ctx.set_scalar("resort_merge_max", 0, sim=k)
        if knob == "vl_cap":
            ctx.set_scalar(knob, value, sim=k)
    if knob != "vl_cap":
        ctx.set_scalar(knob, value)
    assert ctx.scalar(knob) == value
    refs = [oracle.Oracle(image) for image in images]
    for _ in range(3):
        ctx.step(40)
        for ref in refs:
            ref.step(40)
    if case == "settling":
        assert ctx.scalar("sorts") > 1  # re-sorts after the first step
    for k, ref in enumerate(refs):
        bad = _compare_sim(ctx, k, ref)
        assert bad == [], (k, bad)
    ctx.close()


def _compare_sim(ctx, k, ref, fields=EXACT_FIELDS):
    bad = []
    for name in fields:
        if not numpy.array_equal(ctx.get(name, sim=k), ref.get(name)):
            bad.append(name)
    return bad


@pytest.mark.parametrize("case", ["shear1k", "settling", "bonded", "batch"])
def test_integrate_lanes_bit_exact(case):
    """The integration with four work-items per particle (one vector
    component each, SPH_S_INTEGRATE_LANES = 4) and its wall-partials pass:
    bit-exact with the oracle, walls included, single and batched."""
    from .test_gpu_bonds import _bonded
    if case == "shear1k":
        images = [cases.fixture("shear1k.bin").tobytes()]
    elif case == "settling":
        images = [cases.settling_case(1200, 12).tobytes()]
    elif case == "bonded":
        images = [_bonded(frac=0.3, seed=4, sigma_b=300.0).tobytes()]
    else:
        images = [cases.fixture("shear1k.bin").tobytes(), cases.settling_case(300, 11).tobytes()]
    ctx = engine.Context(images=images) if len(images) > 1 else engine.Context(images[0])
    ctx.set_scalar("integrate_lanes", 4)
    ctx.step(120)
    for k, image in enumerate(images):
        ref = oracle.Oracle(image)
        ref.step(120)
        bad = []
        for name in EXACT_FIELDS + ("walls_nx", "walls_mvfd", "walls_acc"):
            a, b = ctx.get(name, sim=k), ref.get(name)
            if a.shape != b.shape or not numpy.array_equal(a, b):
                bad.append(name)
        assert bad == [], (k, bad)
    ctx.close()


def _tiny(n, seed=3):
    """n spheres (n may be 0) falling in a closed 0.1 m box: for the edge
    sizes, the empty simulation (which the reference runs: checkValues only
    notes np = 0, src/sphere.cpp:168-172) and one particle."""
    from sphere_amd.sim import sim
    s = sim(np=n, sid="tiny%d" % n)
    rng = numpy.random.default_rng(seed)
    if n:
        s.radius[:] = rng.uniform(0.004, 0.006, n)
        s.x[:] = rng.uniform(0.01, 0.09, (n, 3))
        s.vel[:, :3] = rng.normal(0.0, 0.1, (n, 3))
    s.g[:] = [0.0, 0.0, -9.81]
    s.origo[:] = 0.0
    s.L[:] = [0.1, 0.1, 0.1]
    s.num[:] = [8, 8, 8]
    s.time_dt[0] = 1.0e-5
    s.time_total[0] = 1.0
    s.time_file_dt[0] = 0.1
    return s


@pytest.mark.parametrize("n", [0, 1])
def test_edge_sizes_bit_exact(n):
    s = _tiny(n)
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    ctx.step(3000)
    ref.step(3000)
    assert ctx.tobytes() == ref.tobytes()
    if n:
        assert _compare(ctx, ref) == []
        assert ctx.get("x")[0, 2] < s.x[0, 2]  # it fell
    ctx.close()


def test_ragged_batch_bit_exact():
    """Simulation sizes that are not multiples of a workgroup, one of a single
    particle, batched together."""
    images = [_tiny(1).tobytes(), cases.settling_case(257, 5).tobytes(),
              _tiny(7, seed=9).tobytes(), cases.settling_case(300, 11).tobytes()]
    ctx = engine.Context(images=images)
    ctx.step(200)
    for k, image in enumerate(images):
        ref = oracle.Oracle(image)
        ref.step(200)
        for name in ("x", "vel", "angvel", "force", "torque", "contacts", "delta_t"):
            assert numpy.array_equal(ctx.get(name, sim=k), ref.get(name)), (k, name)
    ctx.close()


@pytest.mark.parametrize("lanes,wg", [(4, 128), (4, 256), (4, 512), (8, 256), (16, 256), (1, 256)])
def test_dense_packing_many_contacts_per_wavefront(lanes, wg):
    """The contact interaction takes a wavefront's occupied slots one per
    lane, 64 (or as many lanes as are running) at a time, and adds each
    particle's contributions in slot order.  A face-centred cubic packing
    (interior spheres with 12 contacts, every wavefront of 16 particles
    holding 68-176 contacts: two or three rounds) whose last wavefront is
    partly empty (379 particles): bit-exact with the oracle over 300 steps,
    for every contact lane count and workgroup size."""
    s = cases.fcc_case()
    n = int(s.np[0])
    assert n % 16 != 0
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    ctx.set_scalar("contact_lanes", lanes)
    if lanes == 4:
        ctx.set_scalar("contact_wg", wg)
    ctx.step(1)
    ref.step(1)
    z = (ref.get("contacts") != n).sum(1)[ref.get("perm").astype(numpy.int64)]
    per_wave = numpy.add.reduceat(z, numpy.arange(0, n, 16))
    assert per_wave.min() > 64
    assert _compare(ctx, ref) == []
    for _ in range(3):
        ctx.step(100)
        ref.step(100)
        assert _compare(ctx, ref) == []
    ctx.close()


@pytest.mark.parametrize("lanes", [4, 8])
def test_dense_packing_hertz_bit_exact(lanes):
    """The dense face-centred cubic packing under the Hertzian law (contact
    model 3): two or three interaction rounds per wavefront, bit-exact with
    the oracle over 300 steps."""
    s = cases.fcc_case()
    s.contactmodel[0] = 3
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    ctx.set_scalar("contact_lanes", lanes)
    for _ in range(3):
        ctx.step(100)
        ref.step(100)
        assert _compare(ctx, ref) == []
    ctx.close()
