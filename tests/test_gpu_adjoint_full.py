"""The adjoint's coverage beyond contacts: the walls' equation of motion
(stress-controlled walls), bonds (including bonds that break), batches of
simulations, and the column-collapse calibration at full size; checked
against the oracle's dual-number build (forward-mode derivatives of the
same FP32-length trajectory, oracle.TangentOracle) to 1e-6 relative, and
batched gradients against single-context gradients bit for bit."""

import numpy
import pytest

import oracle
from sphere_amd import engine, scenes
from sphere_amd.sim import sim

from . import cases

pytestmark = pytest.mark.gpu

PARAMS = ("k_n", "k_t", "gamma_n", "gamma_t", "mu_s", "mu_d", "gamma_wn", "gamma_wt", "mu_wd")


def _tangent_dot(image, steps, seed_kind, seed_value, c_x, c_v=None):
    """sum c_x . dx(steps) (+ c_v . dv) along one seeded direction."""
    t = oracle.TangentOracle(image)
    if seed_kind in ("vel", "x"):
        t.seed_field(seed_kind, seed_value)
    else:
        for name in seed_kind:
            t.seed_param(name, 1.0)
    t.step(steps)
    val = float(numpy.sum(c_x * t.tangent("x")))
    if c_v is not None:
        val += float(numpy.sum(c_v * t.tangent("vel")))
    return val


def _check_against_tangent(image, steps, names, rtol=1e-6, with_x=True, seed=12):
    s = sim().frombytes(image)
    n = int(s.np[0])
    rng = numpy.random.default_rng(seed)
    c_x, c_v = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    d_v, d_x = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    with engine.Context(image) as ctx:
        g = ctx.adjoint(steps, c_x, c_v, None)
    checks = [(name, g[name], (name,)) for name in names]
    checks.append(("v0", float(numpy.sum(g["v"] * d_v)), "vel"))
    if with_x:
        checks.append(("x0", float(numpy.sum(g["x"] * d_x)), "x"))
    worst, bad = 0.0, []
    for label, ad, kind in checks:
        seed_value = d_v if kind == "vel" else d_x if kind == "x" else None
        tan = _tangent_dot(image, steps, kind, seed_value, c_x, c_v)
        if tan == 0.0 and ad == 0.0:
            continue
        rel = abs(ad - tan) / abs(tan)
        worst = max(worst, rel)
        if rel > rtol:
            bad.append("%s adjoint %.12e tangent %.12e rel %.2e" % (label, ad, tan, rel))
    return bad, worst, g


def test_stress_controlled_wall_gradients():
    """The consolidated, sheared 1,000-particle cell with its
    stress-controlled top wall (wmode 1) loaded by the bed: the wall's
    acceleration is (sum of the particles' wall forces - sigma0 Lx Ly) / m,
    so every particle pressing on it feels every other through the wall."""
    s = cases.fixture("shear1k.bin")
    assert int(s.wmode[0]) == 1
    image = s.tobytes()
    with engine.Context(image) as ctx:
        ctx.step(1)
        assert (ctx.get("walls_force_pp") != 0.0).any()  # the wall is loaded
    bad, worst, g = _check_against_tangent(image, 12, ("k_n", "k_t", "gamma_n", "gamma_t",
                                                       "mu_s"))
    print("wmode-1 wall: worst relative difference %.2e" % worst)
    assert bad == []


def test_stress_controlled_wall_changes_the_gradient():
    """The same cell with the wall held (wmode 0) gives different gradients:
    the wall's response to the particles is part of the derivative."""
    s = cases.fixture("shear1k.bin")
    rng = numpy.random.default_rng(2)
    c = rng.standard_normal((int(s.np[0]), 3))
    out = []
    for mode in (1, 0):
        s.wmode[0] = mode
        with engine.Context(s.tobytes()) as ctx:
            out.append(ctx.adjoint(12, c)["k_n"])
    assert out[0] != out[1]


@pytest.mark.parametrize("breaking", [False, True])
def test_bonded_bed_gradients(breaking):
    """A sheared lattice bed with 30% of its touching pairs bonded
    (bondsLinear, src/cohesion.cuh:8-233), under a stress-controlled wall;
    bonds strong enough that none breaks, or with `breaking` strengths of
    1e7 Pa, at which about a fifth of them break during the run (the broken
    bonds' branches are frozen as recorded)."""
    s = cases.lattice_bed(600, 17, 0.01, 0.2, bond_frac=0.3)
    rng = numpy.random.default_rng(4)
    s.vel[:, :3] += rng.normal(0.0, 0.02, (600, 3))
    s.sigma_b[0] = s.tau_b[0] = 1.0e7 if breaking else 1.0e12
    first = s.bonds[:, 0].copy()
    image = s.tobytes()
    steps = 15
    o = oracle.Oracle(image)
    o.step(steps)
    # A broken bond's first index becomes nb0 (src/cohesion.cuh:205); a bond
    # whose first particle is nb0 from the start is not counted.
    nb0 = int(s.nb0[0])
    broken = int(((o.get("bonds")[:, 0] == nb0) & (first != nb0)).sum())
    if breaking:
        assert 0 < broken < int(s.nb0[0])
    else:
        assert broken == 0
    bad, worst, g = _check_against_tangent(image, steps, ("k_n", "k_t", "gamma_n", "gamma_t"))
    print("bonds (breaking=%s, %d broken): worst relative difference %.2e"
          % (breaking, broken, worst))
    assert bad == []


def _batch_images():
    a = cases.settling_case(300, 11)
    b = cases.fixture("shear1k.bin")
    c = cases.lattice_bed(400, 19, 0.01, 0.2, bond_frac=0.3)
    return [a.tobytes(), b.tobytes(), c.tobytes()]


def test_batched_adjoint_equals_single_contexts():
    """Each simulation's gradients in a batch (sph_adjoint_batch) are its
    single-context gradients, bit for bit: parameters, and the state
    cotangents of its own rows."""
    images = _batch_images()
    sizes = [int(sim().frombytes(im).np[0]) for im in images]
    rows = sum(sizes)
    rng = numpy.random.default_rng(9)
    c_x, c_v = rng.standard_normal((rows, 3)), rng.standard_normal((rows, 3))
    steps = 12
    with engine.Context(images=images) as ctx:
        gb = ctx.adjoint(steps, c_x, c_v, None)
    off = 0
    for k, (im, n) in enumerate(zip(images, sizes)):
        with engine.Context(im) as ctx:
            g = ctx.adjoint(steps, c_x[off:off + n], c_v[off:off + n], None)
        for f in ("x", "v", "w"):
            assert numpy.array_equal(gb[f][off:off + n], g[f]), (k, f)
        for name in PARAMS:
            assert gb[name][k] == g[name], (k, name)
        off += n


def test_adjoint_is_reproducible_bit_for_bit():
    s = cases.fixture("shear1k.bin")
    rng = numpy.random.default_rng(5)
    c = rng.standard_normal((1000, 3))
    out = []
    for _ in range(2):
        with engine.Context(s.tobytes()) as ctx:
            out.append(ctx.adjoint(10, c, c, None))
    for key in ("x", "v", "w") + PARAMS:
        assert numpy.array_equal(out[0][key], out[1][key]), key


@pytest.mark.parametrize("edit", ["E", "wmode3", "model3", "rate_state"])
def test_unsupported_paths_are_refused(edit):
    s = cases.fixture("shear1k.bin")
    if edit == "E":
        s.E[0] = 1.0e7
    elif edit == "wmode3":
        s.wmode[0] = 3
    elif edit == "model3":
        s.contactmodel[0] = 3
    else:
        s.fixvel[0] = 7.0
    with engine.Context(s.tobytes()) as ctx:
        with pytest.raises(engine.EngineError) as e:
            ctx.adjoint(3, numpy.ones((1000, 3)))
        assert e.value.code == -6


def test_column_collapse_gradient_at_full_size():
    """The column-collapse calibration at full size: 2,000 spheres (radii
    uniform in [0.008, 0.012] m), 500 steps sampled every 50, the gradient
    of the position MSE against a target made with k_n* = 2e7 and
    gamma_n* = 0.3 critical, from k_n = 1e7, gamma_n = 0.2 critical,
    recorded in checkpointed segments of 50 steps; against the dual-number
    oracle for k_n (tied to k_t), gamma_n, mu (mu_s and mu_d) and an
    initial-velocity direction, to 1e-6 relative."""
    col = scenes.collapse_column(2000, seed=3001, k_n=1.0e7, damping_frac=0.2,
                                 open_stepper=lambda img: engine.Context(img))
    every, nsamples = 50, 10
    rmin, rho = float(col.radius.min()), float(col.rho[0])
    star = sim().frombytes(col.tobytes())
    star.k_n[0] = star.k_t[0] = 2.0e7
    star.gamma_n[0] = star.gamma_t[0] = 0.3 * scenes.critical_damping(rmin, rho, 2.0e7)

    def samples(img):
        out = []
        with engine.Context(img) as ctx:
            for _ in range(nsamples):
                ctx.step(every)
                out.append(ctx.get("x")[:, :3].copy())
        return numpy.stack(out)

    target = samples(star.tobytes())
    image = col.tobytes()
    X = samples(image)
    seeds = 2.0 * (X - target) / X.size
    with engine.Context(image) as ctx:
        ctx.set_scalar("adj_checkpoint", 50)
        g = ctx.adjoint_samples(every, nsamples, seeds)
    rng = numpy.random.default_rng(31)
    d_v = rng.standard_normal((2000, 3))
    checks = [("k_n", g["k_n"] + g["k_t"], ("k_n", "k_t")),
              ("gamma_n", g["gamma_n"], ("gamma_n",)),
              ("mu", g["mu_s"] + g["mu_d"], ("mu_s", "mu_d")),
              ("v0", float(numpy.sum(g["v"] * d_v)), "vel")]
    worst, bad = 0.0, []
    for label, ad, kind in checks:
        t = oracle.TangentOracle(image)
        if kind == "vel":
            t.seed_field("vel", d_v)
        else:
            for name in kind:
                t.seed_param(name, 1.0)
        tan = 0.0
        for k in range(nsamples):
            t.step(every)
            tan += float(numpy.sum(seeds[k] * t.tangent("x")))
        rel = abs(ad - tan) / abs(tan)
        worst = max(worst, rel)
        if rel > 1e-6:
            bad.append("%s adjoint %.12e tangent %.12e rel %.2e" % (label, ad, tan, rel))
    print("column collapse, full size: worst relative difference %.2e" % worst)
    assert bad == []


def test_batch_trajectory_matches_single_trajectories():
    """The batched torch binding (BatchTrajectory: one batched context, one
    sph_adjoint_samples_batch) against one Trajectory per simulation: the
    same samples and the same gradients (each simulation's own theta row
    and initial-velocity rows), bit for bit."""
    import torch

    from sphere_amd import autograd
    sims = [cases.settling_case(300, 21), cases.settling_case(250, 22),
            cases.fixture("shear1k.bin")]
    every, nsamples = 10, 3
    batch = autograd.BatchTrajectory(sims, every, nsamples, params=("k_n", "gamma_n", "mu"))
    theta0 = numpy.array([[float(s.k_n[0]) * f, float(s.gamma_n[0]), float(s.mu_s[0])]
                          for s, f in zip(sims, (1.0, 1.3, 0.9))])
    rng = numpy.random.default_rng(4)
    weights = [torch.as_tensor(rng.standard_normal((nsamples, int(s.np[0]), 3))) for s in sims]
    theta = torch.tensor(theta0, requires_grad=True)
    v0 = torch.as_tensor(numpy.concatenate([s.vel[:, :3] for s in sims])).requires_grad_(True)
    X = batch(theta, v0=v0)
    (X * torch.cat(weights, dim=1)).sum().backward()
    off = 0
    for k, s in enumerate(sims):
        n = int(s.np[0])
        single = autograd.Trajectory(s, every, nsamples, params=("k_n", "gamma_n", "mu"))
        th = torch.tensor(theta0[k], requires_grad=True)
        v = torch.as_tensor(s.vel[:, :3].copy()).requires_grad_(True)
        Y = single(th, v0=v)
        assert numpy.array_equal(X.detach().numpy()[:, off:off + n], Y.detach().numpy()), k
        (Y * weights[k]).sum().backward()
        assert numpy.array_equal(theta.grad.numpy()[k], th.grad.numpy()), k
        assert numpy.array_equal(v0.grad.numpy()[off:off + n], v.grad.numpy()), k
        off += n


@pytest.mark.parametrize("case", ["single", "checkpointed", "batch"])
def test_record_then_reverse_equals_adjoint_samples(case):
    """sph_adjoint_record + sph_adjoint_reverse (the torch binding's forward
    and backward) give sph_adjoint_samples's gradients bit for bit, the
    recorded samples are a plain run's positions, and the context ends in
    the same state; stepping between the two calls is refused."""
    if case == "batch":
        images = [cases.fixture("shear1k.bin").tobytes(), cases.settling_case(300, 11).tobytes()]
    else:
        images = [cases.fixture("shear1k.bin").tobytes()]

    def open_ctx():
        ctx = engine.Context(images=images) if len(images) > 1 else engine.Context(images[0])
        if case == "checkpointed":
            ctx.set_scalar("adj_checkpoint", 15)  # segments of 15 of the 40 steps
        return ctx

    every, nsamples = 10, 4
    a = open_ctx()
    rows = a.rows()
    rng = numpy.random.default_rng(5)
    sx = rng.standard_normal((nsamples, rows, 3))
    ga = a.adjoint_samples(every, nsamples, sx)
    b = open_ctx()
    samples = b.adjoint_record(every, nsamples)
    gb = b.adjoint_reverse(every, nsamples, sx)
    for key in ga:
        assert numpy.array_equal(numpy.asarray(ga[key]), numpy.asarray(gb[key])), key
    for k in range(len(images)):
        assert numpy.array_equal(a.get("x", sim=k), b.get("x", sim=k))
    c = open_ctx()
    offs = numpy.cumsum([0] + [int(c.scalar("np", sim=k)) for k in range(len(images))])
    for m in range(nsamples):
        c.step(every)
        for k in range(len(images)):
            assert numpy.array_equal(samples[m, offs[k]:offs[k + 1]], c.get("x", sim=k)[:, :3])
    d = open_ctx()
    d.adjoint_record(every, nsamples)
    d.step(1)
    with pytest.raises(engine.EngineError):
        d.adjoint_reverse(every, nsamples, sx)
    for ctx in (a, b, c, d):
        ctx.close()
