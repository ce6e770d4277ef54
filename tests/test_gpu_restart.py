"""Restart sidecars (no counterpart in the reference, whose resumed runs
start with empty contact rows and acc0 = 0): a run resumed from an output
.bin and its sidecar is bit-identical to the uninterrupted run, and to the
oracle run uninterrupted from the same image."""

import os

import numpy
import pytest

import oracle
from sphere_amd import engine

from . import cases

FIELDS = ("x", "vel", "angvel", "angpos", "xyzsum", "acc", "angacc", "force", "torque", "es",
          "ev", "p", "contacts", "delta_t", "walls_nx", "walls_mvfd", "walls_acc")


def _state(ctx):
    return {f: ctx.get(f) for f in FIELDS}


def _same(a, b):
    return [f for f in FIELDS if not numpy.array_equal(a[f], b[f])]


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["shear1k", "settling"])
def test_resumed_run_is_bit_identical(tmp_path, case):
    if case == "shear1k":
        s, n1, n2 = cases.fixture("shear1k.bin"), 60, 60
    else:
        s, n1, n2 = cases.settling_case(1200, 12), 120, 80
    image = s.tobytes()
    whole = engine.Context(image)
    whole.step(n1 + n2)
    ref = _state(whole)
    whole.close()

    first = engine.Context(image)
    first.step(n1)
    binp = str(tmp_path / "run.output00001.bin")
    first.write_bin(binp)
    first.write_state(str(tmp_path / "run.output00001.state"))
    first.close()

    resumed = engine.Context(path=binp, sidecar=True)
    resumed.step(n2)
    got = _state(resumed)
    assert _same(got, ref) == []
    resumed.close()

    # The anchor: the oracle stepped n1 + n2 steps from the same image
    # without interruption.
    orc = oracle.Oracle(image)
    orc.step(n1 + n2)
    assert [f for f in FIELDS if not numpy.array_equal(got[f], orc.get(f))] == []

    # Without the sidecar the resumed run is the reference's: it differs.
    plain = engine.Context(path=binp)
    plain.step(n2)
    assert _same(_state(plain), ref) != []
    plain.close()


@pytest.mark.gpu
def test_sidecar_is_checked_against_the_simulation(tmp_path):
    s = cases.fixture("shear1k.bin")
    ctx = engine.Context(s.tobytes())
    ctx.step(5)
    st = str(tmp_path / "a.state")
    ctx.write_state(st)
    other = engine.Context(cases.settling_case(300, 11).tobytes())
    with pytest.raises(engine.EngineError) as e:
        other.read_state(st)
    assert "np" in str(e.value)
    raw = bytearray(open(st, "rb").read())
    raw[0] = ord("X")
    bad = str(tmp_path / "bad.state")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(engine.EngineError):
        ctx.read_state(bad)
    open(bad, "wb").write(bytes(open(st, "rb").read()[:-8]))
    with pytest.raises(engine.EngineError):
        ctx.read_state(bad)
    ctx.read_state(st)  # its own sidecar reads back
    other.close()
    ctx.close()


@pytest.mark.gpu
def test_run_writes_sidecars_next_to_outputs(tmp_path):
    s = cases.fixture("shear1k.bin")
    dt = float(s.time_dt[0])
    s.time_total[0] = float(s.time_current[0]) + 30.5 * dt
    s.time_file_dt[0] = 10.0 * dt
    binp = str(tmp_path / "in.bin")
    open(binp, "wb").write(s.tobytes())
    ctx = engine.Context(path=binp, sidecar=True)
    out = str(tmp_path / "out")
    ctx.run(out, "sc")
    ctx.close()
    bins = sorted(f for f in os.listdir(out) if f.endswith(".bin"))
    states = sorted(f for f in os.listdir(out) if f.endswith(".state"))
    assert len(bins) >= 2 and [b[:-4] for b in bins] == [t[:-6] for t in states]


@pytest.mark.gpu
def test_sim_run_resumes_from_output_and_sidecar(tmp_path):
    """The sim API: run with sidecars, read an output back, resume it with
    its sidecar; the resumed output equals the uninterrupted run's."""
    from sphere_amd.sim import sim
    s = cases.fixture("shear1k.bin")
    s.sid = "resume"
    dt = float(s.time_dt[0])
    t0 = float(s.time_current[0])
    s.time_total[0] = t0 + 40.5 * dt
    s.time_file_dt[0] = 20.0 * dt
    root = str(tmp_path)
    assert s.run(verbose=False, root=root, sidecar=True) == 0
    out = os.path.join(root, "output")
    whole = sim()
    whole.readbin(os.path.join(out, "resume.output00002.bin"), verbose=False)
    part = sim()
    part.readbin(os.path.join(out, "resume.output00001.bin"), verbose=False)
    part.sid = "resume2"
    assert part.run(verbose=False, root=root,
                    restart_state=os.path.join(out, "resume.output00001.state")) == 0
    # The resumed run keeps the output numbering (step_count) of its input.
    again = sim()
    again.readbin(os.path.join(out, "resume2.output00002.bin"), verbose=False)
    assert float(again.time_current[0]) == float(whole.time_current[0])
    for f in ("x", "vel", "angvel", "angpos", "xyzsum", "es", "ev", "w_x", "w_vel"):
        assert numpy.array_equal(numpy.asarray(getattr(again, f)),
                                 numpy.asarray(getattr(whole, f))), f
    # And the oracle, uninterrupted from the input over as many steps as the
    # host loop takes to write file 2 (its own double additions of t and of
    # the file clock).
    t, ftc, nfile, steps = t0, 0.0, 0, 0
    while nfile < 2:
        t += dt
        ftc += dt
        steps += 1
        if ftc >= float(s.time_file_dt[0]):
            nfile += 1
            ftc = 0.0
    orc = oracle.Oracle(s.tobytes())
    orc.step(steps)
    want = sim().frombytes(orc.tobytes())
    again.time_step_count[0] = want.time_step_count[0]
    assert again.compare(want) == []
