"""The pinned scenes on the GPU: the reference's oblique-collision and
Young's-modulus tests through sim.run(), and the closed-form friction, bond
and boundary-condition scenes through the C ABI.  Each run is compared bit
for bit with the oracle on the same image, and the reference's own
assertions are checked on the engine's output, so the engine inherits the
pins of test_oracle_friction_pins.py (the oracle against closed forms and
invariants) without going through the oracle.
"""

import numpy
import pytest

import oracle
from sphere_amd import engine
from sphere_amd.sim import sim

from . import test_oracle_friction_pins as pins
from .test_gpu_parity import EXACT_FIELDS, _compare

pytestmark = pytest.mark.gpu


def _oracle_last_output(s):
    """The oracle image that sim.run() writes last: the host loop's double
    additions of t and the file clock (src/device.cu:2536-2614), one output
    every file_dt, the last one read back by readlast()."""
    ref = oracle.Oracle(s.tobytes())
    t, ftc = float(s.time_current[0]), 0.0
    dt, fdt = float(s.time_dt[0]), float(s.time_file_dt[0])
    last = s.tobytes()
    while t <= float(s.time_total[0]):
        ref.step(1)
        t += dt
        ftc += dt
        if ftc >= fdt:
            last = ref.tobytes()
            ftc = 0.0
    return sim().frombytes(last)


def _run_and_read(s, tmp_path):
    assert s.run(verbose=False, root=str(tmp_path)) == 0
    after = sim(np=int(s.np[0]), sid=s.sid)
    after.readlast(verbose=False, root=str(tmp_path))
    want = _oracle_last_output(s)
    after.time_step_count[0] = want.time_step_count[0] = 0
    assert after.compare(want) == []
    return after


@pytest.mark.parametrize("E", [0.0, 7.0e9])
@pytest.mark.parametrize("case", ["no_slip", "slip", "slip_viscous"])
def test_reference_oblique_collisions_via_run(tmp_path, case, E):
    """tests/contactmodel.py:137-222 and tests/contactmodel_young.py:144-232
    through sim.run() and readlast()."""
    orig = pins.oblique_pair(case, E)
    after = _run_and_read(orig, tmp_path)
    pins.check_oblique(orig, after, case)


@pytest.mark.parametrize("r,z", [((1.0, 1.0), (2.0, 4.05)), ((2.0, 1.0), (2.0, 5.05)),
                                 ((1.0, 2.0), (2.0, 5.05))])
@pytest.mark.parametrize("gamma_n", [0.0, 1.0e6])
def test_reference_young_collisions_via_run(tmp_path, r, z, gamma_n):
    """tests/contactmodel_young.py:16-142."""
    orig = pins.young_pair(r, z, gamma_n)
    after = _run_and_read(orig, tmp_path)
    e0 = orig.totalKineticEnergy()
    if gamma_n == 0.0:
        assert abs((e0 - after.totalKineticEnergy()) / e0) < 1e-3
    else:
        e1 = after.totalKineticEnergy() + after.totalViscousEnergy()
        assert abs((e0 - e1) / e0) < 0.05


@pytest.mark.parametrize("case", ["elastic", "oblique", "viscous", "oblique_viscous"])
def test_reference_young_wall_collisions_via_run(tmp_path, case):
    """tests/contactmodel_wall_young.py."""
    orig = pins.young_wall(case)
    after = _run_and_read(orig, tmp_path)
    pins.check_young_wall(orig, after, case)


def _bit_exact(s, steps, chunks=4, fields=EXACT_FIELDS):
    image = s.tobytes()
    ref = oracle.Oracle(image)
    with engine.Context(image) as ctx:
        per = max(1, steps // chunks)
        done = 0
        while done < steps:
            k = min(per, steps - done)
            ref.step(k)
            ctx.step(k)
            done += k
            bad = _compare(ctx, ref, fields)
            assert bad == [], (done, bad)
        if int(s.nb0[0]) > 0:
            for f in ("bonds", "bonds_delta", "bonds_omega"):
                assert numpy.array_equal(ctx.get(f), ref.get(f)), f
        for f in ("walls_nx", "walls_mvfd"):
            assert numpy.array_equal(ctx.get(f), ref.get(f)), f


def test_stick_slip_pair_bit_exact():
    """The closed-form stick/slip cycle (three slips in 400 steps)."""
    _bit_exact(pins.fixed_spinning_pair(), 400, chunks=8)


@pytest.mark.parametrize("omega,gamma_t", [(2.0, 5.0), (2.0, 1.0e4), (0.02, 1.0e6)])
def test_model1_friction_pair_bit_exact(omega, gamma_t):
    s = pins.fixed_spinning_pair(model=1, omega=omega, gamma_t=gamma_t, mu_s=0.6, mu_d=0.3)
    _bit_exact(s, 20)


@pytest.mark.parametrize("model", [1, 2, 3])
def test_youngs_modulus_pair_bit_exact(model):
    s = pins.fixed_spinning_pair(model=model, E=3.0e6, k_n=1.0, k_t=1.0, gamma_t=0.0,
                                 mu_s=10.0, mu_d=10.0)
    s.radius[1] = 0.015
    s.x[0, 2] = s.x[1, 2] + s.radius[0] + s.radius[1] - 1.0e-4
    _bit_exact(s, 50)


@pytest.mark.parametrize("gamma_wt,E", [(2.0, 0.0), (5.0e3, 0.0), (5.0e3, 2.0e7)])
def test_wall_friction_bit_exact(gamma_wt, E):
    _bit_exact(pins.fixed_sphere_on_floor(gamma_wt=gamma_wt, E=E), 20)


@pytest.mark.parametrize("mode", ["shear", "twist", "bend", "axial"])
def test_bond_modes_bit_exact(mode):
    s, _ = pins.bonded_pair(mode, dt_frac=1.0 / 1000.0)
    _bit_exact(s, 1500, chunks=5)


@pytest.mark.parametrize("which", ["sigma", "tau"])
def test_bond_breakage_step_bit_exact(which):
    """The bond of the breakage pin, with its strength between the stresses
    of steps 15 and 16: broken in step 16 on both sides."""
    mode = "axial" if which == "sigma" else "shear"
    s, _ = pins.bonded_pair(mode, k=1.0e8)
    s.vel[1, 0 if mode == "axial" else 2] = 0.05
    o = oracle.Oracle(s.tobytes())
    stress = []
    A = numpy.pi * s.radius[0] ** 2
    k = s.k_n[0] if mode == "axial" else s.k_t[0]
    for _ in range(16):
        o.step(1)
        bd = o.get("bonds_delta")[0]
        stress.append(abs(k * A * bd[3]) / A if mode == "axial"
                      else numpy.linalg.norm(k * A * bd[:3]) / A)
    strength = 0.5 * (stress[14] + stress[15])
    s.sigma_b[0] = strength if which == "sigma" else 1e300
    s.tau_b[0] = strength if which == "tau" else 1e300
    image = s.tobytes()
    ref = oracle.Oracle(image)
    with engine.Context(image) as ctx:
        for k_step in range(1, 17):
            ref.step(1)
            ctx.step(1)
            b = ctx.get("bonds")
            assert numpy.array_equal(b, ref.get("bonds"))
            assert (int(b[0, 0]) == int(s.nb0[0])) == (k_step == 16)
            assert _compare(ctx, ref) == []


@pytest.mark.parametrize("scene,steps", [("moving_wall", 10), ("wmode3", 200),
                                         ("adaptive", 40), ("rate_state", 60)])
def test_boundary_scenes_bit_exact(scene, steps):
    """The closed-form scenes of the moving wall (the wall speed added to
    every component), the wmode-3 shear stress, the adaptive grid and the
    rate-state switch (steps starting exactly at 10 s and 15 s)."""
    s = getattr(pins, scene + "_scene")()
    _bit_exact(s, steps, chunks=min(steps, 20))
