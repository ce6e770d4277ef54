"""The adjoint (sph_adjoint) against central finite differences of the
oracle.  The adjoint differentiates lengths as exact roots, so the
differences are taken on the oracle with FP64 lengths (oracle.set_exact_length).
Walls in these cases are fixed (wmode 0 or none); the stress-controlled
walls, bonds and batches are checked in test_gpu_adjoint_full.py."""

import numpy
import pytest

import oracle
from sphere_amd import engine, scenes
from sphere_amd.sim import sim

from . import cases

pytestmark = pytest.mark.gpu


def _loss_parts(o):
    return o.get("x")[:, :3], o.get("vel")[:, :3], o.get("angvel")


def _fd_dir(image, steps, perturb, h, seeds):
    """Central difference of L = sum c . state along `perturb` (which edits
    a sim copy by +-h), on the oracle with exact lengths."""
    outs = []
    for sign in (1.0, -1.0):
        s = sim().frombytes(image)
        perturb(s, sign * h)
        o = oracle.Oracle(s.tobytes())
        o.step(steps)
        outs.append(_loss_parts(o))
    return sum(float(numpy.sum(c * (a - b))) for c, a, b in zip(seeds, outs[0], outs[1])) / (2 * h)


def _adjoint(image, steps, seeds):
    ctx = engine.Context(image)
    g = ctx.adjoint(steps, *seeds)
    ctx.close()
    return g


@pytest.fixture(autouse=True)
def exact_lengths():
    oracle.set_exact_length(True)
    yield
    oracle.set_exact_length(False)


def test_equal_mass_head_on_collision_swaps_velocities():
    """Two equal elastic spheres: after the collision v1' = v2 and v2' = v1,
    so d v2'_z / d v1_z = 1 and d v2'_z / d v2_z = 0 in continuous time (the
    discrete collision, about 30 steps long, departs by a few per cent);
    and the adjoint matches central differences of the oracle."""
    s = sim(np=2, sid="adj2")
    s.radius[:] = 1.0
    s.x[0, :] = [5.0, 5.0, 2.0]
    s.x[1, :] = [5.0, 5.0, 4.05]
    s.vel[0, 2] = 1.0
    s.defineWorldBoundaries(L=[10, 10, 10])
    s.initTemporal(total=1.0, file_dt=0.1)
    s.g[:] = 0.0
    s.gamma_n[0] = 0.0
    s.gamma_t[0] = 0.0
    dt = float(s.time_dt[0])
    steps = int(0.08 / dt)
    seeds = [numpy.zeros((2, 3)), numpy.zeros((2, 3)), numpy.zeros((2, 3))]
    seeds[1][1, 2] = 1.0
    image = s.tobytes()
    g = _adjoint(image, steps, seeds)
    assert abs(g["v"][0, 2] - 1.0) < 0.05
    assert abs(g["v"][1, 2]) < 0.05
    for i in (0, 1):
        def pert(s2, eps, i=i):
            s2.vel[i, 2] += eps
        fd = _fd_dir(image, steps, pert, 1e-7, seeds)
        assert abs(g["v"][i, 2] - fd) < 1e-4, (i, g["v"][i, 2], fd)


def _check_params(image, steps, names, seeds, rtol, h=1e-6):
    g = _adjoint(image, steps, seeds)
    scale = max(abs(g[n]) for n in names)
    bad = []
    for name in names:
        v0 = float(getattr(sim().frombytes(image), name)[0])

        def pert(s, d, name=name, v0=v0):
            getattr(s, name)[0] = v0 * (1.0 + d)

        fd = _fd_dir(image, steps, pert, h, seeds) / v0
        if abs(g[name] - fd) > rtol * max(abs(fd), 1e-3 * scale):
            bad.append("%s adjoint %.9e fd %.9e" % (name, g[name], fd))
    return bad, g


def _seeds(n, seed):
    rng = numpy.random.default_rng(seed)
    return [rng.standard_normal((n, 3)) for _ in range(3)]


def test_settling_box_parameter_gradients():
    s = cases.settling_case(300, 11)
    s.gamma_wt[0] = 0.5 * float(s.gamma_wn[0])
    image = s.tobytes()
    bad, g = _check_params(image, 40, ("k_n", "k_t", "gamma_n", "gamma_t", "mu_s", "mu_d",
                                       "gamma_wn", "gamma_wt", "mu_wd"), _seeds(300, 1), 1e-4)
    assert bad == []


def test_settling_box_state_gradients():
    s = cases.settling_case(300, 11)
    image = s.tobytes()
    steps = 40
    seeds = _seeds(300, 2)
    g = _adjoint(image, steps, seeds)
    rng = numpy.random.default_rng(3)
    for key, field, h in (("x", "x", 1e-8), ("v", "vel", 1e-6)):
        d = rng.standard_normal((300, 3))

        def pert(s2, eps, field=field, d=d):
            getattr(s2, field)[:, :3] += eps * d

        fd = _fd_dir(image, steps, pert, h, seeds)
        ad = float(numpy.sum(g[key] * d))
        assert abs(ad - fd) <= 1e-4 * abs(fd), (key, ad, fd)


def test_sheared_bed_parameter_gradients():
    """Periodic images, fixed-velocity layers, a fixed top wall, a dense bed.
    In a settled bed the overlaps are a few hundred FP32 roundings of the
    contact length, so the reference's FP32 lengths and exact lengths give
    materially different trajectories (and many contacts sit exactly at the
    Coulomb limit, where they choose different branches); the adjoint follows
    the recorded FP32 trajectory.  It is compared here with central
    differences of the oracle with its FP32 lengths over 5 steps, within 2 %
    (the differences themselves carry the FP32 quantisation)."""
    oracle.set_exact_length(False)
    s = cases.fixture("shear1k.bin")
    s.wmode[0] = 0
    bad, g = _check_params(s.tobytes(), 5, ("k_n", "k_t", "gamma_n", "gamma_t", "mu_s"),
                           _seeds(1000, 4), 2e-2, h=1e-5)
    assert bad == []


def test_loose_periodic_box_parameter_gradients():
    """Fresh contacts in a periodic box (images), every parameter."""
    s = scenes.shear_box(n=600, seed=7, phi=0.3, shrink=0.97)
    s.wmode[0] = 0
    names = [n for n in ("k_n", "k_t", "gamma_n", "gamma_t", "mu_s", "mu_d", "gamma_wn",
                         "gamma_wt", "mu_wd") if float(getattr(s, n)[0]) != 0.0]
    bad, g = _check_params(s.tobytes(), 30, names, _seeds(600, 5), 1e-4)
    assert bad == []


def test_adjoint_leaves_the_forward_state():
    s = cases.settling_case(300, 11)
    image = s.tobytes()
    ctx = engine.Context(image)
    ctx.adjoint(25, numpy.ones((300, 3)))
    oracle.set_exact_length(False)
    ref = oracle.Oracle(image)
    ref.step(25)
    assert numpy.array_equal(ctx.get("x"), ref.get("x"))
    ctx.close()


def test_unsupported_cases_are_refused():
    s = cases.fixture("shear1k.bin")
    s.contactmodel[0] = 3
    ctx = engine.Context(s.tobytes())
    with pytest.raises(engine.EngineError):
        ctx.adjoint(5, numpy.ones((1000, 3)))
    ctx.close()


@pytest.mark.gpu
def test_overflowing_contact_row_is_refused():
    """A particle overlapped by more spheres than a row holds: some of them
    list it while it cannot list them, and the adjoint (which gathers each
    contact's partner part from the mirror slot) refuses the result."""
    s = cases.crowded_case()
    s.contactmodel[0] = 2
    n = int(s.np[0])
    ctx = engine.Context(s.tobytes())
    with pytest.raises(engine.EngineError) as e:
        ctx.adjoint(2, None, numpy.ones((n, 3)), None)
    assert e.value.code == -6  # SPH_E_UNSUPPORTED
    assert "partner's row" in str(e.value)
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case,steps,K", [("settling", 40, 7), ("shear1k", 30, 10)])
def test_checkpointed_adjoint_matches_recorded(case, steps, K):
    """Segments recomputed from whole-context checkpoints give the gradients
    of recording every step bit for bit (state cotangents and the parameter
    sums, which are formed in a fixed order), and leave the context at the
    same final state."""
    if case == "settling":
        s = cases.settling_case(1200, 12)
    else:
        s = cases.fixture("shear1k.bin")
    image = s.tobytes()
    n = int(s.np[0])
    rng = numpy.random.default_rng(3)
    seed_x, seed_v = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    out = []
    for k in (0, K):
        ctx = engine.Context(image)
        ctx.set_scalar("adj_checkpoint", k)
        g = ctx.adjoint(steps, seed_x, seed_v, None)
        out.append((g, ctx.get("x"), ctx.get("vel"), ctx.get("delta_t")))
        ctx.close()
    (ga, xa, va, da), (gb, xb, vb, db) = out
    for f in ("x", "v", "w"):
        assert numpy.array_equal(ga[f], gb[f]), f
    for name in engine.ADJOINT_PARAMS:
        assert gb[name] == ga[name], name
    assert numpy.array_equal(xa, xb) and numpy.array_equal(va, vb)
    assert numpy.array_equal(da, db)


@pytest.mark.parametrize("case,steps", [("settling", 40), ("shear1k", 10)])
def test_adjoint_matches_tangent_oracle(case, steps):
    """Reverse mode against forward mode: the engine's adjoint of
    L = c . x(steps) against the directional derivatives of the oracle's
    dual-number build along each parameter and an initial-velocity direction,
    on the same FP32-length trajectory.  Both differentiate lengths as exact
    roots; they differ only in how the FP32-rounded length enters the
    linearisation (the adjoint normalises by it), about 1e-7 per contact."""
    oracle.set_exact_length(False)
    if case == "settling":
        s = cases.settling_case(300, 11)
    else:
        s = cases.fixture("shear1k.bin")
        s.wmode[0] = 0
    n = int(s.np[0])
    image = s.tobytes()
    rng = numpy.random.default_rng(12)
    c = rng.standard_normal((n, 3))
    d = rng.standard_normal((n, 3))
    g = _adjoint(image, steps, [c, None, None])
    worst = 0.0
    for kind in ("k_n", "k_t", "gamma_n", "gamma_t", "mu_s", "mu_d", "vel"):
        t = oracle.TangentOracle(image)
        if kind == "vel":
            t.seed_field("vel", d)
            ad = float(numpy.sum(g["v"] * d))
        else:
            t.seed_param(kind, 1.0)
            ad = g[kind]
        t.step(steps)
        tan = float(numpy.sum(c * t.tangent("x")))
        if tan == 0.0 and ad == 0.0:
            continue
        rel = abs(ad - tan) / abs(tan)
        worst = max(worst, rel)
        assert rel <= 1e-6, (kind, ad, tan)
    print("%s: adjoint vs tangent oracle, worst relative difference %.2e" % (case, worst))
