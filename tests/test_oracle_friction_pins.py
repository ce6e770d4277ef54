"""Pins of the oracle's friction, Young's-modulus, bond and boundary-condition
paths against closed forms, invariants and the reference's own tests.

Each check is independent of the oracle's code:

* the reference's oblique-collision and Young's-modulus tests
  (tests/contactmodel.py:137-222, tests/contactmodel_young.py:144-232,
  tests/contactmodel_wall_young.py), restated with their tolerances;
* closed forms of a contact held at a fixed overlap with a constant
  tangential speed (two fully fixed spheres, one spinning; fixvel < 0 keeps
  both velocities, src/integration.cuh:151-158): the accumulated
  displacement, the stick force, the first slip step, the Luding reset and
  the stick/slip cycle after it, the model-1 and wall friction laws, and the
  E-scaled stiffnesses k = pi/2 E (r_a + r_b)/2 and pi/2 E r_a;
* per-contact invariants on a sheared bed: the Coulomb identities of the
  sliding and sticking contacts, Newton's third law, the torque of each
  contact about the contact point, and the particle sums;
* bonds: energy conservation and the oscillation frequencies of the shear,
  twist and bend modes (tests/bond_tests.py:109-188), and the step at which
  a bond breaks under the stress criterion (src/cohesion.cuh:193-206);
* the wmode-3 shear-stress boundary, the adaptive grid and the rate-state
  velocity switch against their closed forms.
"""

import math

import numpy
import pytest

import oracle
from oracle import CLOG
from sphere_amd.sim import sim
from tests.cases import fixture

NC = 32
RHO = 2600.0


def _f32len(v):
    """The reference's length(): FP32 root of the FP64 dot product (the rule
    pinned on its own in test_oracle_pins.py)."""
    v = numpy.asarray(v, dtype=numpy.float64)
    return float(numpy.sqrt(numpy.float32(numpy.dot(v, v))))


def _mass(r):
    return RHO * 4.0 / 3.0 * math.pi * r ** 3


def _run_like_sim_run(s):
    """The oracle stepped as many steps as sim.run() takes for s (t += dt
    while t <= t_total, src/device.cu:1031,2536)."""
    o = oracle.Oracle(s.tobytes())
    t, dt, n = float(s.time_current[0]), float(s.time_dt[0]), 0
    while t <= float(s.time_total[0]):
        t += dt
        n += 1
    o.step(n)
    return sim().frombytes(o.tobytes()), n


# ---------------------------------------------------------------------------
# The reference's oblique collisions (tests/contactmodel.py:137-222 and the
# same cases with Young's modulus, tests/contactmodel_young.py:144-232)

def oblique_pair(case, E=0.0):
    s = sim(np=2, sid="contactmodeltest")
    s.radius[:] = [1.0, 1.0]
    s.x[0, :] = [5.0, 5.0, 2.0]
    s.x[1, :] = [5.0, 5.0, 4.05]
    if E > 0.0:
        s.setYoungsModulus(E)
    s.vel[0, 2] = 1.0
    s.vel[0, 0] = 1.0
    if case == "no_slip":
        s.mu_s[0] = 1e9
        s.mu_d[0] = 1e9
    else:
        s.mu_s[0] = 0.3
        s.mu_d[0] = 0.3
    if case == "slip_viscous":
        s.gamma_t[0] = 1.0e3
    s.defineWorldBoundaries(L=[10, 10, 10])
    s.initTemporal(total=0.1, file_dt=0.01)
    return s


def check_oblique(orig, after, case):
    """The reference's assertions for its oblique cases (1/8 .. 8/8)."""
    e0 = orig.totalKineticEnergy()
    e1 = after.totalKineticEnergy() + after.totalRotationalEnergy()
    if case != "no_slip":
        e1 += after.totalFrictionalEnergy()
    if case == "slip_viscous":
        e1 += after.totalViscousEnergy()
    assert abs((e0 - e1) / e0) < 0.05
    assert (after.angvel[:, 1] < 0.0).all()
    if case != "no_slip":
        assert after.totalFrictionalEnergy() > 0.0


@pytest.mark.parametrize("E", [0.0, 7.0e9])
@pytest.mark.parametrize("case", ["no_slip", "slip", "slip_viscous"])
def test_reference_oblique_collisions(case, E):
    orig = oblique_pair(case, E)
    after, _ = _run_like_sim_run(orig)
    check_oblique(orig, after, case)
    # The spheres separated: the collision is over, not merely in progress.
    d = numpy.linalg.norm(after.x[1, :3] - after.x[0, :3])
    assert d > 2.0


def young_pair(r, z, gamma_n=0.0, E=7.0e9):
    s = sim(np=2, sid="contactmodeltest")
    s.radius[:] = r
    s.x[0, :] = [5.0, 5.0, z[0]]
    s.x[1, :] = [5.0, 5.0, z[1]]
    s.setYoungsModulus(E)
    s.vel[0, 2] = 1.0
    s.defineWorldBoundaries(L=[10, 10, 10])
    s.initTemporal(total=0.1, file_dt=0.01)
    s.gamma_n[0] = gamma_n
    return s


@pytest.mark.parametrize("r,z", [((1.0, 1.0), (2.0, 4.05)), ((2.0, 1.0), (2.0, 5.05)),
                                 ((1.0, 2.0), (2.0, 5.05))])
@pytest.mark.parametrize("gamma_n", [0.0, 1.0e6])
def test_reference_young_collisions(r, z, gamma_n):
    """tests/contactmodel_young.py:16-142 (elastic 1/4..4/4, viscoelastic)."""
    orig = young_pair(r, z, gamma_n)
    after, _ = _run_like_sim_run(orig)
    e0 = orig.totalKineticEnergy()
    if gamma_n == 0.0:
        assert abs((e0 - after.totalKineticEnergy()) / e0) < 1e-3
        if r[0] == r[1]:
            assert abs((orig.vel[0, 2] - after.vel[1, 2]) / orig.vel[0, 2]) < 1e-3
    else:
        if r[0] == r[1]:
            assert orig.vel[0, 2] > after.vel[1, 2]
        e1 = after.totalKineticEnergy() + after.totalViscousEnergy()
        assert abs((e0 - e1) / e0) < 0.05


def young_wall(case):
    """tests/contactmodel_wall_young.py's four set-ups."""
    s = sim(np=1, sid="contactmodeltest")
    s.radius[:] = 1.0
    s.x[0, :] = [5.0, 5.0, 1.05]
    s.setYoungsModulus(7.0e9)
    s.vel[0, 2] = -0.1
    if case in ("oblique", "oblique_viscous"):
        s.vel[0, 0] = 0.1
    s.defineWorldBoundaries(L=[10, 10, 10])
    s.gamma_wn[0] = 1.0e6 if case in ("viscous", "oblique_viscous") else 0.0
    s.gamma_wt[0] = 1.0e6 if case == "oblique_viscous" else 0.0
    s.initTemporal(total=0.3 if case == "oblique" else 1.0, file_dt=0.01)
    return s


def check_young_wall(orig, after, case):
    e0 = orig.energy("kin")
    if case == "elastic":
        assert abs((after.vel[0, 2] - 0.1) / after.vel[0, 2]) < 1e-3
        assert abs((e0 - after.totalKineticEnergy()) / e0) < 1e-3
    elif case == "oblique":
        assert abs((e0 - after.totalKineticEnergy()) / e0) < 1e-3
    elif case == "viscous":
        assert e0 > after.energy("kin")
        assert abs((e0 - (after.energy("kin") + after.energy("visc_n"))) / e0) < 0.05
    else:
        assert e0 > after.energy("kin")
        assert after.angvel[0, 0] == 0.0 and after.angvel[0, 1] > 0.0
        assert after.angvel[0, 2] == 0.0


@pytest.mark.parametrize("case", ["elastic", "oblique", "viscous", "oblique_viscous"])
def test_reference_young_wall_collisions(case):
    orig = young_wall(case)
    after, _ = _run_like_sim_run(orig)
    check_young_wall(orig, after, case)


# ---------------------------------------------------------------------------
# Closed forms at a fixed overlap: two fully fixed spheres (fixvel = -1 keeps
# velocity and angular velocity, src/integration.cuh:151-158), the upper one
# spinning about y, so the contact's normal, overlap and tangential velocity
# are constant.

def fixed_spinning_pair(model=2, r=0.01, overlap=1.0e-4, omega=2.0, k_n=1.0e5, k_t=1.0e5,
                        gamma_n=0.0, gamma_t=0.05, mu_s=0.5, mu_d=0.4, E=0.0, dt=1.0e-5):
    s = sim(np=2, sid="stickslip")
    s.radius[:] = r
    s.x[0, :] = [0.5, 0.5, 0.5 + 2 * r - overlap]   # upper sphere a
    s.x[1, :] = [0.5, 0.5, 0.5]                      # lower sphere b
    s.fixvel[:] = -1.0
    s.angvel[0, 1] = omega
    s.contactmodel[0] = model
    s.k_n[0], s.k_t[0] = k_n, k_t
    s.E[0] = E
    s.gamma_n[0], s.gamma_t[0] = gamma_n, gamma_t
    s.mu_s[0], s.mu_d[0] = mu_s, mu_d
    s.g[:] = [0.0, 0.0, 0.0]
    s.defineWorldBoundaries(L=[1.0, 1.0, 1.0])
    s.time_dt[0] = dt
    s.time_total[0] = 1.0
    return s


def _pair_geometry(s):
    """Normal of a (pointing from b to a), overlap and the tangential
    velocity of a's surface relative to b's, with the FP32 length rule."""
    xab = s.x[0, :3] - s.x[1, :3]
    ln = _f32len(xab)
    n = xab / ln
    delta = ln - 2 * s.radius[0]
    w = s.angvel[0]
    vt = (s.radius[0] + delta / 2.0) * numpy.cross(n, w)
    vt = vt - n * numpy.dot(n, vt)
    return n, delta, vt


def test_stick_slip_cycle_closed_form():
    """Model 2 (src/contactmodels.cuh:316-398): while sticking, the stored
    displacement is the sum of v_t dt and the force is -k_t dt_sum - gamma_t
    v_t; the first step whose trial force exceeds mu_s |f_n| slips, with
    |f_t| = mu_s |f_n| along the trial force, the displacement reset to
    -(mu_d |f_n| t + gamma_t v_t)/k_t and es_dot = |f_t| |v_t| / 2; from the
    reset the displacement grows by v_t dt again until the next slip."""
    s = fixed_spinning_pair()
    n, delta, vt = _pair_geometry(s)
    k_n, k_t, g_t = s.k_n[0], s.k_t[0], s.gamma_t[0]
    mu_s, mu_d, dt = s.mu_s[0], s.mu_d[0], s.time_dt[0]
    fn = -k_n * delta * n
    fn_len = numpy.linalg.norm(fn)
    rho_a = s.radius[0] + delta * 0.5
    o = oracle.Oracle(s.tobytes())
    disp = numpy.zeros(3)
    slips, stick_steps = [], 0
    for k in range(1, 400):
        o.step(1)
        trial_disp = disp + vt * dt
        trial = -k_t * trial_disp - g_t * vt
        tl = numpy.linalg.norm(trial)
        limit = mu_s * fn_len
        # Away from the threshold the branch is decided by the closed form.
        assert abs(tl - limit) > 1e-6 * limit
        F = o.get("force")[0]
        T = o.get("torque")[0]
        row = o.get("contacts")[0]
        slot = int(numpy.nonzero(row == 1)[0][0])
        D = o.get("delta_t")[0, slot]
        if tl <= limit:
            ft = trial
            disp = trial_disp
            stick_steps += 1
            assert o.get("es_dot")[0] == 0.0
        else:
            t_hat = trial / tl
            ft = limit * t_hat
            disp = -(mu_d * fn_len * t_hat + g_t * vt) / k_t
            slips.append(k)
            assert abs(o.get("es_dot")[0] - 0.5 * limit * numpy.linalg.norm(vt)) \
                < 1e-6 * limit * numpy.linalg.norm(vt)
        numpy.testing.assert_allclose(D, disp, rtol=1e-6, atol=1e-6 * numpy.abs(disp).max())
        numpy.testing.assert_allclose(F, fn + ft, rtol=1e-6, atol=1e-6 * fn_len)
        numpy.testing.assert_allclose(T, numpy.cross(-rho_a * n, ft), rtol=1e-6,
                                      atol=1e-6 * rho_a * limit)
        # The lower sphere receives the opposite force.
        numpy.testing.assert_allclose(o.get("force")[1], -(fn + ft), rtol=1e-6,
                                      atol=1e-6 * fn_len)
    # Stick, a slip, then the shorter stick/slip cycle from the reset.
    assert len(slips) >= 3 and stick_steps > 100
    gaps = numpy.diff(slips)
    assert slips[0] > gaps.max()
    assert gaps.max() - gaps.min() <= 1


@pytest.mark.parametrize("omega,gamma_t,regime", [(2.0, 5.0, "viscous"),
                                                  (2.0, 1.0e4, "dynamic"),
                                                  (0.02, 1.0e6, "static")])
def test_model1_friction_closed_form(omega, gamma_t, regime):
    """contactLinearViscous (src/contactmodels.cuh:122-240): f_n = k_n |delta|
    n at rest; f_t = -min(gamma_t |v_t|, mu |f_n|) v_t/|v_t| with mu = mu_d
    when |v_t| > 0.001f, else mu_s; no history, no energies."""
    s = fixed_spinning_pair(model=1, omega=omega, gamma_t=gamma_t, mu_s=0.6, mu_d=0.3)
    n, delta, vt = _pair_geometry(s)
    vtl = numpy.linalg.norm(vt)
    fn = -s.k_n[0] * delta * n
    mu = s.mu_d[0] if vtl > float(numpy.float32(0.001)) else s.mu_s[0]
    assert (regime == "static") == (vtl <= 0.001)
    limit = mu * numpy.linalg.norm(fn)
    visc = gamma_t * vtl
    assert (regime == "viscous") == (visc < limit)
    ft = -min(visc, limit) * vt / vtl
    o = oracle.Oracle(s.tobytes())
    o.step(3)
    numpy.testing.assert_allclose(o.get("force")[0], fn + ft, rtol=1e-6,
                                  atol=1e-6 * numpy.linalg.norm(fn))
    rho_a = s.radius[0] + delta / 2.0
    numpy.testing.assert_allclose(o.get("torque")[0], -rho_a * numpy.cross(n, ft), rtol=1e-6,
                                  atol=1e-6 * rho_a * numpy.linalg.norm(ft))
    assert o.get("es_dot")[0] == 0.0 and o.get("ev_dot")[0] == 0.0


@pytest.mark.parametrize("model", [1, 2, 3])
def test_youngs_modulus_stiffness_closed_form(model):
    """With E > 0.001 the stiffnesses are k_n = k_t = pi/2 E (r_a + r_b)/2
    (src/contactmodels.cuh:183,325-330,494): the normal force at a held
    overlap is k delta (k delta^3/2 for Hertz), and model 2's first stick
    force is k v_t dt + gamma_t v_t."""
    E = 3.0e6
    s = fixed_spinning_pair(model=model, E=E, k_n=1.0, k_t=1.0, gamma_t=0.0, mu_s=10.0,
                            mu_d=10.0)
    s.radius[1] = 0.015
    s.x[0, 2] = s.x[1, 2] + s.radius[0] + s.radius[1] - 1.0e-4
    n, _, _ = _pair_geometry(s)
    xab = s.x[0, :3] - s.x[1, :3]
    delta = _f32len(xab) - (s.radius[0] + s.radius[1])
    k = math.pi / 2.0 * E * (s.radius[0] + s.radius[1]) / 2.0
    o = oracle.Oracle(s.tobytes())
    o.step(1)
    F = o.get("force")[0]
    if model == 3:
        want_fn = k * abs(delta) ** 1.5
    else:
        want_fn = k * abs(delta)
    assert abs(numpy.dot(F, n) - want_fn) < 1e-6 * want_fn
    if model == 2:
        vt = (s.radius[0] + delta / 2.0) * numpy.cross(n, s.angvel[0])
        ft = -k * vt * s.time_dt[0]
        # n is along z: the horizontal components are the tangential force.
        numpy.testing.assert_allclose(F[:2], ft[:2], rtol=1e-6, atol=1e-6 * numpy.abs(ft).max())


def fixed_sphere_on_floor(omega=(0.0, 3.0, 0.0), gamma_wt=50.0, mu_wd=0.4, E=0.0,
                          overlap=1.0e-4, r=0.01, k_n=1.0e5):
    s = sim(np=1, sid="wallspin")
    s.radius[:] = r
    s.x[0, :] = [0.5, 0.5, r - overlap]
    s.fixvel[0] = -1.0
    s.angvel[0, :] = omega
    s.k_n[0] = k_n
    s.E[0] = E
    s.gamma_wn[0] = 0.0
    s.gamma_wt[0] = gamma_wt
    s.mu_ws[0] = 0.9  # the static wall coefficient does not enter the law
    s.mu_wd[0] = mu_wd
    s.g[:] = [0.0, 0.0, 0.0]
    s.defineWorldBoundaries(L=[1.0, 1.0, 1.0])
    s.time_dt[0] = 1.0e-5
    s.time_total[0] = 1.0
    return s


@pytest.mark.parametrize("gamma_wt,E", [(2.0, 0.0), (5.0e3, 0.0), (5.0e3, 2.0e7)])
def test_wall_friction_closed_form(gamma_wt, E):
    """contactLinear_wall (src/contactmodels.cuh:10-117) on the floor:
    f_n = k_n |delta| (k_n = pi/2 E r with E > 0.001); f_t = -gamma_wt v_t
    below mu_wd |f_n|, else -mu_wd |f_n| v_t/|v_t| with es_dot = |f_t||v_t|;
    the torque acts at the contact point r + delta/2 below the centre."""
    s = fixed_sphere_on_floor(gamma_wt=gamma_wt, E=E)
    r = s.radius[0]
    delta = s.x[0, 2] - r
    n = numpy.array([0.0, 0.0, 1.0])
    k = math.pi / 2.0 * E * r if E > 0.001 else s.k_n[0]
    fn = -k * delta * n
    vt = (r + delta / 2.0) * numpy.cross(n, s.angvel[0])
    vtl = numpy.linalg.norm(vt)
    limit = s.mu_wd[0] * numpy.linalg.norm(fn)
    if gamma_wt * vtl < limit:
        ft = -gamma_wt * vt
        es = 0.0
    else:
        ft = -limit * vt / vtl
        es = limit * vtl
    o = oracle.Oracle(s.tobytes())
    o.step(2)
    numpy.testing.assert_allclose(o.get("force")[0], fn + ft, rtol=1e-6,
                                  atol=1e-6 * numpy.linalg.norm(fn))
    numpy.testing.assert_allclose(o.get("torque")[0], numpy.cross(-(r + delta / 2.0) * n, ft),
                                  rtol=1e-6, atol=1e-9 * limit)
    assert abs(o.get("es_dot")[0] - es) <= 1e-6 * max(es, 1e-30)


def moving_wall_scene():
    """A fixed sphere pressed by a top wall that moves at 0.02 m/s."""
    s = sim(np=1, nw=1, sid="wallmove")
    r = 0.01
    s.radius[:] = r
    s.x[0, :] = [0.5, 0.5, 0.5]
    s.fixvel[0] = -1.0
    s.k_n[0] = 1.0e5
    s.gamma_wn[0] = 0.0
    s.gamma_wt[0] = 10.0
    s.mu_ws[0] = s.mu_wd[0] = 100.0
    s.g[:] = [0.0, 0.0, 0.0]
    s.defineWorldBoundaries(L=[1.0, 1.0, 1.0])
    s.wmode[0] = 2
    s.w_n[0] = [0.0, 0.0, -1.0]
    s.w_x[0] = 0.5 + r - 1.0e-4
    s.w_vel[0] = 0.02
    s.w_m[0] = 1.0
    s.time_dt[0] = 1.0e-6
    s.time_total[0] = 1.0
    return s


def test_moving_wall_adds_its_speed_to_every_component():
    """The wall velocity enters the contact velocity as a scalar added to all
    three components (Float3 + Float, src/vector_arithmetic.h; used at
    src/contactmodels.cuh:37): a fixed sphere under a top wall moving at w
    gets a tangential velocity (w, w, 0), so its friction points along
    -(1, 1, 0)."""
    s = moving_wall_scene()
    o = oracle.Oracle(s.tobytes())
    o.step(1)
    F = o.get("force")[0]
    assert abs(F[0] - F[1]) < 1e-12 * abs(F[0]) and F[0] < 0.0
    numpy.testing.assert_allclose(F[:2], [-10.0 * 0.02, -10.0 * 0.02], rtol=1e-9)
    assert F[2] < 0.0  # pushed down by the wall


# ---------------------------------------------------------------------------
# Per-contact invariants on a sheared bed (contact models 2 and 3)

def _sheared_log(steps_before=300, model=2):
    s = fixture("shear1k.bin")
    s.contactmodel[0] = model
    o = oracle.Oracle(s.tobytes())
    o.step(steps_before)
    oracle.set_contact_log(True)
    try:
        pre = {k: o.get(k) for k in ("x", "contacts", "distmod", "walls_nx")}
        o.step(1)
        log = o.contact_log()
    finally:
        oracle.set_contact_log(False)
    return s, o, pre, log


def test_coulomb_identities_on_a_sheared_bed():
    """For every sliding contact |f_t| = mu_s |f_n - f_c| and the reset
    displacement satisfies k_t dt = -(mu_d |f_n - f_c| t + gamma_t v_t); for
    every sticking one |f_t| <= mu_s |f_n - f_c|, the displacement is the
    previous one turned into the tangent plane plus v_t dt, and the force is
    -k_t dt - gamma_t v_t.  Normal forces are repulsive and along n."""
    s, o, pre, log = _sheared_log()
    mu_s, mu_d = s.mu_s[0], s.mu_d[0]
    g_t, dt = s.gamma_t[0], s.time_dt[0]
    act = log[..., CLOG["active"]] == 1.0
    slip = act & (log[..., CLOG["slip"]] == 1.0)
    stick = act & ~slip
    assert slip.sum() >= 20 and stick.sum() >= 200
    n = log[..., CLOG["n"]]
    fn, ft, fc = log[..., CLOG["f_n"]], log[..., CLOG["f_t"]], log[..., CLOG["f_c"]]
    vt, d, d0 = log[..., CLOG["vel_t"]], log[..., CLOG["delta_t"]], log[..., CLOG["delta_t0"]]
    kt = log[..., CLOG["k_t"]]
    nrm = numpy.linalg.norm
    fnl = nrm(fn, axis=-1)
    # Normal force: along n, repulsive.
    assert (numpy.einsum("...i,...i", fn, n)[act] >= 0.0).all()
    assert (nrm(numpy.cross(fn, n), axis=-1)[act] <= 1e-12 * fnl[act] + 1e-300).all()
    lim = nrm(fn - fc, axis=-1)
    ftl = nrm(ft, axis=-1)
    # Sliding: on the Coulomb cone, with Luding's reset.
    assert numpy.allclose(ftl[slip], mu_s * lim[slip], rtol=1e-6, atol=0)
    # A contact whose normal force the clamp zeroed (the dashpot pulls) has
    # a zero cone: no tangential force, and the reset leaves -gamma_t v_t.
    free = slip & (lim == 0.0)
    assert (ft[free] == 0.0).all()
    assert numpy.allclose(kt[free][:, None] * d[free], -g_t * vt[free], rtol=1e-12, atol=0)
    slip = slip & (lim > 0.0)
    assert slip.sum() >= 20
    t_hat = ft[slip] / ftl[slip][:, None]
    want = -(mu_d * lim[slip][:, None] * t_hat + g_t * vt[slip])
    assert numpy.allclose(kt[slip][:, None] * d[slip], want, rtol=1e-6,
                          atol=1e-6 * nrm(want, axis=-1).max())
    # Sticking: inside the cone; the spring law on the carried displacement.
    assert (ftl[stick] <= mu_s * lim[stick] * (1 + 1e-12)).all()
    assert numpy.allclose(d[stick], d0[stick] + vt[stick] * dt, rtol=0,
                          atol=1e-12 * nrm(d[stick], axis=-1).max())
    want = -kt[stick][:, None] * d[stick] - g_t * vt[stick]
    assert numpy.allclose(ft[stick], want, rtol=0, atol=1e-9 * nrm(want, axis=-1).max())
    # The carried displacement and the tangential velocity lie in the
    # tangent plane (to the 1e-7 by which |n| = x/len32 differs from 1).
    assert (numpy.abs(numpy.einsum("...i,...i", d0, n))[act]
            <= 1e-6 * nrm(d0, axis=-1)[act].max()).all()
    assert (numpy.abs(numpy.einsum("...i,...i", vt, n))[act]
            <= 1e-6 * nrm(vt, axis=-1)[act].max()).all()


def test_contact_forces_obey_newton_and_angular_momentum():
    """Each contact a-b seen from both sides: f_ab = -f_ba, the torques act
    at the common contact point (x_a - (r_a + delta/2) n = x_b + (r_b +
    delta/2) n), so that x_a x f_ab + T_a + x_b x f_ba + T_b = 0, and the
    particle's force and torque are the sums over its contacts (particles
    away from the walls)."""
    s, o, pre, log = _sheared_log(steps_before=200)
    rows = pre["contacts"]
    x = pre["x"]
    npart = rows.shape[0]
    act = log[..., CLOG["active"]] == 1.0
    n = log[..., CLOG["n"]]
    f = log[..., CLOG["f_n"]] + log[..., CLOG["f_t"]] + log[..., CLOG["f_c"]]
    delta = log[..., CLOG["delta"]]
    lever = x[:, None, 3] + delta / 2.0
    torque = numpy.cross(-lever[..., None] * n, log[..., CLOG["f_t"]])
    pairs = 0
    for a in range(npart):
        for sa in numpy.nonzero(act[a])[0]:
            b = rows[a, sa]
            sb = numpy.nonzero(rows[b] == a)[0]
            if len(sb) != 1 or not act[b, sb[0]]:
                continue
            sb = sb[0]
            fa, fb = f[a, sa], f[b, sb]
            scale = numpy.linalg.norm(fa)
            assert numpy.linalg.norm(fa + fb) <= 1e-6 * scale
            # Common contact point, with b's image shift.
            xa = x[a, :3]
            xb = x[b, :3] - pre["distmod"][a, sa, :3]
            ca = xa - lever[a, sa] * n[a, sa]
            cb = xb - lever[b, sb] * n[b, sb]
            assert numpy.linalg.norm(ca - cb) <= 1e-6 * x[a, 3]
            # Moments about b's centre (the remainder, x_b x (f_ab + f_ba),
            # vanishes by the third law checked above).
            L = numpy.cross(xa - xb, fa) + torque[a, sa] + torque[b, sb]
            assert numpy.linalg.norm(L) <= 1e-6 * scale * (x[a, 3] + x[b, 3])
            pairs += 1
    assert pairs >= 500
    # Particle sums away from the walls.
    F, T = o.get("force"), o.get("torque")
    wall_z = pre["walls_nx"][0, 3]
    free = (x[:, 2] - x[:, 3] > s.origo[2]) & (x[:, 2] + x[:, 3] < wall_z)
    assert free.sum() > 0.8 * npart
    Fsum = f.sum(axis=1)
    Tsum = torque.sum(axis=1)
    scale = numpy.abs(F[free]).max()
    assert numpy.abs(F[free] - Fsum[free]).max() <= 1e-9 * scale
    assert numpy.abs(T[free] - Tsum[free]).max() <= 1e-9 * numpy.abs(T[free]).max()


def test_hertz_contacts_on_a_sheared_bed():
    """Contact model 3 on the same bed: normal force along n and repulsive,
    the sliding contacts on the (velocity-dependent) Coulomb cone, the
    sticking ones inside it, and Newton's third law per contact."""
    s, o, pre, log = _sheared_log(steps_before=100, model=3)
    act = log[..., CLOG["active"]] == 1.0
    slip = act & (log[..., CLOG["slip"]] == 1.0)
    assert act.sum() > 500
    fn, ft, fc, vt = (log[..., CLOG[k]] for k in ("f_n", "f_t", "f_c", "vel_t"))
    nrm = numpy.linalg.norm
    lim = nrm(fn - fc, axis=-1)
    mu = numpy.where(nrm(vt, axis=-1) > float(numpy.float32(0.001)), s.mu_d[0], s.mu_s[0])
    ftl = nrm(ft, axis=-1)
    assert numpy.allclose(ftl[slip], (mu * lim)[slip], rtol=1e-6, atol=0)
    assert (ftl[act & ~slip] <= (mu * lim)[act & ~slip] * (1 + 1e-12)).all()
    rows = pre["contacts"]
    f = fn + ft + fc
    checked = 0
    for a, sa in zip(*numpy.nonzero(act)):
        b = rows[a, sa]
        sb = numpy.nonzero(rows[b] == a)[0]
        if len(sb) == 1 and act[b, sb[0]]:
            assert nrm(f[a, sa] + f[b, sb[0]]) <= 1e-6 * nrm(f[a, sa])
            checked += 1
    assert checked > 500


# ---------------------------------------------------------------------------
# Bonds (src/cohesion.cuh:8-233): the set-ups of the reference's disabled
# tests/bond_tests.py:109-188 (two bonded spheres, one perturbed)

def bonded_pair(mode, radius=0.5, d=0.1, k=1.16e9, dt_frac=1.0 / 400.0):
    """Two spheres a small gap d apart (so that no contact forms while the
    bond oscillates), bonded, one of them perturbed in `mode`."""
    s = sim(np=2, sid="bondtest")
    s.x[0, :] = [10.0, 10.0, 10.0]
    s.x[1, :] = [10.0 + 2.0 * radius + d, 10.0, 10.0]
    s.radius[:] = radius
    s.defineWorldBoundaries(L=[20.0, 20.0, 20.0])
    s.periodic[0] = 1
    s.g[:] = [0.0, 0.0, 0.0]
    # k_t != k_n, so that a stiffness used in the wrong mode shows.
    k_t = 0.7 * k
    s.k_n[0], s.k_t[0] = k, k_t
    s.gamma_n[0] = s.gamma_t[0] = 0.0
    s.bond(0, 1)
    m = _mass(radius)
    Ip = 0.4 * m * radius ** 2
    A = math.pi * radius ** 2
    I = 0.25 * math.pi * radius ** 4
    J = 2.0 * I
    if mode == "shear":
        s.vel[1, 2] = 1e-4
        omega = math.sqrt(2.0 * k_t * A / m)
    elif mode == "twist":
        s.angvel[1, 0] = 1e-4
        omega = math.sqrt(2.0 * k_t * J / Ip)
    elif mode == "bend":
        s.angvel[0, 1] = -1e-4
        s.angvel[1, 1] = 1e-4
        omega = math.sqrt(2.0 * k * I / Ip)
    else:  # axial
        s.vel[1, 0] = 1e-4
        omega = math.sqrt(2.0 * k * A / m)
    s.time_dt[0] = 2.0 * math.pi / omega * dt_frac
    s.time_total[0] = 1.0
    return s, omega


def _energy(a):
    return a.energy("kin") + a.energy("rot") + a.energy("bondpot")


def _bond_run(mode, dt_frac, periods=3.0):
    """Steps a bonded pair `periods` periods; returns the relative energy
    deviations, the relative coordinate's velocity history and the times."""
    s, omega = bonded_pair(mode, dt_frac=dt_frac)
    o = oracle.Oracle(s.tobytes())
    e0 = _energy(sim().frombytes(s.tobytes()))
    period = 2.0 * math.pi / omega
    steps = int(round(periods * period / s.time_dt[0]))
    rel, t, dev = [], [], []
    for k in range(steps):
        o.step(1)
        if mode in ("shear", "axial"):
            v = o.get("vel")
            c = 2 if mode == "shear" else 0
            rel.append(v[1, c] - v[0, c])
        else:
            w = o.get("angvel")
            c = 0 if mode == "twist" else 1
            rel.append(w[1, c] - w[0, c])
        t.append((k + 1) * s.time_dt[0])
        if k % 37 == 0:
            dev.append((_energy(sim().frombytes(o.tobytes())) - e0) / e0)
    return numpy.array(dev), numpy.array(rel), numpy.array(t), period, o


@pytest.mark.parametrize("mode", ["shear", "twist", "bend", "axial"])
def test_bond_modes_conserve_energy_and_oscillate_at_their_frequency(mode):
    """Energy: kinetic + rotational + bond potential (sim.energy, the
    reference's python/sphere.py:4311-4380) is conserved up to the
    first-order lag between the stored displacements and the velocities:
    the deviation stays bounded over three periods, below 2% at 2,000 steps
    per period (4% for the bend mode, whose two spheres both turn), and halves
    when the step halves.  Frequency: the relative
    coordinate of each mode is a harmonic oscillator, omega^2 = 2 k A / m
    (shear, axial), 2 k_t J / I_p (twist), 2 k_n I / I_p (bend) with I_p =
    2/5 m r^2; the measured half periods (sign changes of the relative
    velocity) match within 1%.  The other components stay zero, as the
    reference test (tests/bond_tests.py:109-188) asserts."""
    dev_a, _, _, _, _ = _bond_run(mode, 1.0 / 1000.0)
    dev_b, rel, t, period, o = _bond_run(mode, 1.0 / 2000.0)
    a_max, b_max = numpy.abs(dev_a).max(), numpy.abs(dev_b).max()
    assert b_max < 0.04
    assert 1.7 < a_max / b_max < 2.3
    crossings = numpy.nonzero(numpy.diff(numpy.sign(rel)) != 0)[0]
    assert len(crossings) >= 5
    half = numpy.diff(t[crossings]).mean()
    assert abs(half - period / 2.0) / (period / 2.0) < 0.01
    a = sim().frombytes(o.tobytes())
    if mode == "shear":
        assert (a.vel[:, 1] == 0.0).all() and (a.angvel == 0.0).all()
        assert (a.vel[:, 2] >= -1e-12).all()
    elif mode == "twist":
        assert (a.angvel[:, 1:] == 0.0).all()
        assert (a.angvel[:, 0] >= -1e-12).all()
    elif mode == "bend":
        assert (a.angvel[:, 1] != 0.0).all()
        assert (a.angvel[:, [0, 2]] == 0.0).all()


@pytest.mark.parametrize("which", ["sigma", "tau"])
def test_bond_breaks_at_the_first_step_reaching_the_strength(which):
    """The bond breaks in the step whose stresses sigma_max = |f_n|/A +
    |t_t| R/I or tau_max = |f_t|/A + |t_n| R/J first reach sigma_b or tau_b
    (>=), and from then on exerts no force (src/cohesion.cuh:193-206).  The
    stresses are those of the same bed stepped without a strength limit."""
    mode = "axial" if which == "sigma" else "shear"
    s, _ = bonded_pair(mode, k=1.0e8)
    s.vel[1, 0 if mode == "axial" else 2] = 0.05
    r = s.radius[0]
    A = math.pi * r ** 2
    k = s.k_n[0] if mode == "axial" else s.k_t[0]
    o = oracle.Oracle(s.tobytes())
    stress = []
    for _ in range(60):
        o.step(1)
        bd = o.get("bonds_delta")[0]
        if mode == "axial":
            stress.append(abs(k * A * bd[3]) / A)
        else:
            stress.append(numpy.linalg.norm(k * A * bd[:3]) / A)
    stress = numpy.array(stress)
    assert (numpy.diff(stress[:20]) > 0).all()
    strength = 0.5 * (stress[14] + stress[15])
    if which == "sigma":
        s.sigma_b[0], s.tau_b[0] = strength, 1e300
    else:
        s.sigma_b[0], s.tau_b[0] = 1e300, strength
    o = oracle.Oracle(s.tobytes())
    for k_step in range(1, 18):
        o.step(1)
        broken = int(o.get("bonds")[0, 0]) == s.nb0[0]
        assert broken == (k_step >= 16), k_step
        if k_step == 16:
            # The breaking step applies no bond force (the pair is apart).
            assert (o.get("force") == 0.0).all()
        elif k_step < 16:
            assert (o.get("force")[0] != 0.0).any()


@pytest.mark.parametrize("mode", ["axial", "bend"])
def test_bond_stiffness_option_sets_the_frequency(mode):
    """The engine's bond option SPH_S_BOND_K (here on the oracle): a bond
    stiffness k_b in N/m in place of k_n A, so the axial mode oscillates at
    omega^2 = 2 k_b / m and the bend mode at 2 (k_b / A) I / I_p, whatever
    the contact stiffness."""
    s, _ = bonded_pair(mode, k=1.16e9)
    k_b = 1.0e8
    r = s.radius[0]
    m = _mass(r)
    A = math.pi * r ** 2
    I = 0.25 * math.pi * r ** 4
    omega = (math.sqrt(2.0 * k_b / m) if mode == "axial"
             else math.sqrt(2.0 * (k_b / A) * I / (0.4 * m * r * r)))
    period = 2.0 * math.pi / omega
    s.time_dt[0] = period / 2000.0
    o = oracle.Oracle(s.tobytes())
    o.set_bond_model(bond_k=k_b)
    rel, t = [], []
    for k in range(int(3 * 2000)):
        o.step(1)
        if mode == "axial":
            v = o.get("vel")
            rel.append(v[1, 0] - v[0, 0])
        else:
            w = o.get("angvel")
            rel.append(w[1, 1] - w[0, 1])
        t.append((k + 1) * s.time_dt[0])
    rel, t = numpy.array(rel), numpy.array(t)
    crossings = numpy.nonzero(numpy.diff(numpy.sign(rel)) != 0)[0]
    half = numpy.diff(t[crossings]).mean()
    assert abs(half - period / 2.0) / (period / 2.0) < 0.01


def test_bond_extension_break_option():
    """SPH_S_BOND_BREAK_EXT: the bond breaks in the first step whose normal
    extension (-delta_n) reaches the fraction of the smaller radius, taken
    from the same pair stepped without the option; the stress strengths do
    not matter then."""
    s, _ = bonded_pair("axial", k=1.0e8)
    s.vel[1, 0] = 0.05
    s.radius[1] = 0.4
    s.x[1, 0] = s.x[0, 0] + 0.5 + 0.4 + 0.1
    rmin = 0.4
    o = oracle.Oracle(s.tobytes())
    ext = []
    for _ in range(60):
        o.step(1)
        ext.append(-o.get("bonds_delta")[0, 3])
    ext = numpy.array(ext)
    assert (numpy.diff(ext[:20]) > 0).all()
    frac = 0.5 * (ext[14] + ext[15]) / rmin
    s.sigma_b[0] = s.tau_b[0] = 0.0  # the stress criterion would break it at once
    o = oracle.Oracle(s.tobytes())
    o.set_bond_model(bond_break_ext=frac)
    for k_step in range(1, 18):
        o.step(1)
        broken = int(o.get("bonds")[0, 0]) == s.nb0[0]
        assert broken == (k_step >= 16), k_step


# ---------------------------------------------------------------------------
# Boundary conditions around the step

def wmode3_scene():
    """A top particle (fixvel 1) pushed in x by a free neighbour, and a
    fixvel particle below L_z/2, under a wmode-3 wall with tau_x = 3 Pa."""
    s = sim(np=3, nw=1, sid="wmode3")
    r = 0.01
    s.radius[:] = r
    s.x[0, :] = [0.5, 0.5, 0.8]             # top, fixvel 1
    s.x[1, :] = [0.5 - 2 * r + 2e-4, 0.5, 0.8]  # touching it in x, free
    s.x[2, :] = [0.3, 0.3, 0.2]             # below L_z/2, fixvel 1
    s.fixvel[0] = 1.0
    s.fixvel[2] = 1.0
    s.vel[0, 0] = 0.01
    s.vel[2, 0] = 0.02
    s.k_n[0] = s.k_t[0] = 1.0e5
    s.gamma_n[0] = s.gamma_t[0] = 0.0
    s.g[:] = [0.0, 0.0, -9.81]
    s.periodic[0] = 1
    s.defineWorldBoundaries(L=[1.0, 1.0, 1.0])
    s.wmode[0] = 3
    s.w_n[0] = [0.0, 0.0, -1.0]
    s.w_x[0] = 0.95
    s.w_m[0] = 2.0
    s.w_tau_x[0] = 3.0
    s.time_dt[0] = 1.0e-6
    s.time_total[0] = 1.0
    return s


def test_wmode3_shear_stress_closed_form():
    """wmode 3 (src/integration.cuh:64-68,116-129,457-484): top particles
    (fixvel > 0.0001, z > L_z/2) get a_x = (tau_x + tau_eff) Lx Ly / m_w with
    tau_eff = sum of their x forces / (Lx Ly), a_y = 0, gravity cancelled,
    no angular acceleration; particles below L_z/2 follow the fixvel rule."""
    s = wmode3_scene()
    r = s.radius[0]
    o = oracle.Oracle(s.tobytes())
    o.step(1)
    F = o.get("force")
    acc = o.get("acc")
    m0 = _mass(r)
    LxLy = 1.0 * 1.0
    tau_eff = F[0, 0] / LxLy
    assert F[0, 0] > 0.0  # pushed away from particle 1
    want = (3.0 + tau_eff) * LxLy / 2.0
    assert abs(acc[0, 0] - want) <= 1e-12 * abs(want)
    assert acc[0, 1] == 0.0
    assert abs(acc[0, 2] - F[0, 2] / m0) <= 1e-12 * max(abs(F[0, 2] / m0), 1e-300)
    assert (o.get("angacc")[0] == 0.0).all()
    # Below L_z/2: the ordinary fixvel rule, no shear-stress acceleration.
    assert acc[2, 0] == 0.0 and acc[2, 1] == 0.0
    # The free particle: Newton's law with gravity.
    assert abs(acc[1, 0] - F[1, 0] / m0) <= 1e-12 * abs(F[1, 0] / m0)
    assert abs(acc[1, 2] - (F[1, 2] / m0 - 9.81)) < 1e-12


def adaptive_scene():
    """A jittered lattice of spheres under a top wall that descends at
    200 m/s, with an adaptive grid."""
    rng = numpy.random.default_rng(5)
    g = numpy.arange(0.02, 0.28, 0.015)
    gz = numpy.arange(0.02, 0.19, 0.015)
    pts = numpy.array([(a, b, c) for c in gz for b in g for a in g])
    pts += rng.uniform(-0.002, 0.002, pts.shape)
    n = len(pts)
    s = sim(np=n, nw=1, sid="adaptive")
    s.radius[:] = rng.uniform(0.004, 0.0055, n)
    s.x[:, :] = pts
    s.periodic[0] = 1
    s.defineWorldBoundaries(L=[0.3, 0.3, 0.3])
    s.g[:] = [0.0, 0.0, 0.0]
    s.adaptive[0] = 1
    s.wmode[0] = 2
    s.w_n[0] = [0.0, 0.0, -1.0]
    s.w_x[0] = 0.3
    s.w_vel[0] = -200.0
    s.w_m[0] = 1.0
    s.time_dt[0] = 1.0e-5
    s.time_total[0] = 1.0
    return s


def test_adaptive_grid_follows_the_top_wall():
    """DEM::updateGridSize (src/device.cu:363-393,1050-1052): with an
    adaptive grid, each step first sets L_z to the top wall's position, so
    the cell keys of that step follow the floor formula with that L_z, and
    the written L_z is the last step's wall position."""
    s = adaptive_scene()
    o = oracle.Oracle(s.tobytes())
    num = s.num.astype(numpy.int64)
    for _ in range(40):
        wall_before = o.get("walls_nx")[0, 3]
        x_before = o.get("x")
        o.step(1)
        key = numpy.sort(o.get("cellid").astype(numpy.int64))
        L = [0.3, 0.3, wall_before]
        ijk = [numpy.floor((x_before[:, d] - 0.0) / (L[d] / num[d])).astype(numpy.int64)
               for d in range(3)]
        want = numpy.sort((ijk[2] * num[1]) * num[0] + ijk[1] * num[0] + ijk[0])
        assert numpy.array_equal(key, want)
    a = sim().frombytes(o.tobytes())
    assert a.L[2] == wall_before
    assert a.L[2] < 0.3 - 0.05


def rate_state_scene():
    """Three free spheres with fixvel 7 (switched), 1 and 12 (not), from
    t = 9 s in steps of 1/8 s, exact in binary, so that steps start exactly
    at t = 10 s and t = 15 s, where the >= comparisons decide."""
    s = sim(np=3, sid="ratestate")
    s.radius[:] = 0.01
    s.x[:, :] = [[0.2, 0.5, 0.5], [0.5, 0.5, 0.5], [0.8, 0.5, 0.5]]
    s.fixvel[:] = [7.0, 1.0, 12.0]
    s.vel[:, 0] = 0.001
    s.g[:] = [0.0, 0.0, 0.0]
    s.periodic[0] = 1
    s.defineWorldBoundaries(L=[1.0, 1.0, 1.0])
    s.time_dt[0] = 0.125
    s.time_current[0] = 9.0
    s.time_total[0] = 100.0
    return s


def test_rate_state_switch_closed_form():
    """src/device.cu:993-999,2454-2462 and src/integration.cuh:285-291: a
    step that starts at t >= 10 s (and < 15 s) multiplies vel.x of particles
    with 5 < fixvel < 10 by 10, once; the first step starting at t >= 15 s
    divides it by 10; other particles are untouched.  With fixvel in
    (0.0001, 10) the horizontal acceleration is zero, so vel.x is otherwise
    constant and x advances by vel.x dt."""
    s = rate_state_scene()
    dt = float(s.time_dt[0])
    o = oracle.Oracle(s.tobytes())
    t = 9.0
    v_expect = 0.001
    state = 1
    for _ in range(60):
        x0 = o.get("x")[0, 0]
        v_step = v_expect  # x advances with the velocity the step starts from
        if t >= 10.0 and t < 15.0 and state == 1:
            v_expect = v_expect * 10.0
            state = 2
        elif t >= 15.0 and state == 2:
            v_expect = v_expect / 10.0
            state = 1
        o.step(1)
        t += dt
        v = o.get("vel")
        assert v[0, 0] == v_expect
        assert v[1, 0] == 0.001 and v[2, 0] == 0.001
        dx = (o.get("x")[0, 0] - x0) % 1.0
        assert abs(dx - v_step * dt) < 1e-12
        if t == 10.125:
            assert v[0, 0] == 0.001 * 10.0  # the step starting at exactly 10 s
        if t == 15.125:
            assert v[0, 0] == 0.001 * 10.0 / 10.0
    assert t > 15.0 and v_expect == 0.001 * 10.0 / 10.0
