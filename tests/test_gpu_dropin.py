"""The drop-in boundary on the GPU: sim.run() / the `sphere` CLI write the
reference's output files, and the reference's own contact-model tests
(tests/contactmodel.py, tests/contactmodel_wall.py) pass through them."""

import os
import subprocess

import numpy
import pytest

import oracle
from sphere_amd import engine, scenes
from sphere_amd.sim import sim, status

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rel(a, b):
    return abs((a - b) / a)


def _pair_sim(r, z, v0=1.0, gamma_n=0.0):
    s = sim(np=2, sid="contactmodeltest")
    s.radius[:] = r
    s.x[0, :] = [5.0, 5.0, z[0]]
    s.x[1, :] = [5.0, 5.0, z[1]]
    s.vel[0, 2] = v0
    s.defineWorldBoundaries(L=[10, 10, 10])
    s.initTemporal(total=0.1, file_dt=0.01)
    s.gamma_n[0] = gamma_n
    return s


@pytest.mark.parametrize("r,z", [((1.0, 1.0), (2.0, 4.05)), ((2.0, 1.0), (2.0, 5.05)),
                                 ((1.0, 2.0), (2.0, 5.05))])
def test_reference_elastic_collisions_via_run(tmp_path, r, z):
    """tests/contactmodel.py:16-70 through sim.run() and readlast()."""
    orig = _pair_sim(r, z)
    assert orig.run(verbose=False, root=str(tmp_path)) == 0
    after = sim(np=2, sid="contactmodeltest")
    after.readlast(verbose=False, root=str(tmp_path))
    assert _rel(orig.totalKineticEnergy(), after.totalKineticEnergy()) < 1e-3
    if r[0] == r[1]:
        assert _rel(orig.vel[0, 2], after.vel[1, 2]) < 1e-3


def test_reference_viscous_collision_via_run(tmp_path):
    """tests/contactmodel.py:72-95."""
    orig = _pair_sim((1.0, 1.0), (2.0, 4.05), gamma_n=1.0e6)
    assert orig.run(verbose=False, root=str(tmp_path)) == 0
    after = sim(np=2, sid="contactmodeltest")
    after.readlast(verbose=False, root=str(tmp_path))
    assert orig.vel[0, 2] > after.vel[1, 2]
    assert _rel(orig.totalKineticEnergy(),
                after.totalKineticEnergy() + after.totalViscousEnergy()) < 0.05


def test_reference_wall_rebound_via_run(tmp_path):
    """tests/contactmodel_wall.py: elastic rebound from the bottom wall."""
    s = sim(np=1, sid="wallbounce")
    s.radius[:] = 1.0
    s.x[0, :] = [5.0, 5.0, 1.05]
    s.vel[0, 2] = -1.0
    s.gamma_wn[0] = 0.0
    s.defineWorldBoundaries(L=[10, 10, 10])
    s.initTemporal(total=0.1, file_dt=0.01)
    assert s.run(verbose=False, root=str(tmp_path)) == 0
    a = sim(np=1, sid="wallbounce")
    a.readlast(verbose=False, root=str(tmp_path))
    assert _rel(-s.vel[0, 2], a.vel[0, 2]) < 1e-3


def test_run_outputs_match_oracle(tmp_path):
    """sph_run writes output00000 (the input state) and one file per file_dt,
    each identical to the oracle stepped the same number of steps, and
    status.dat in the reference format."""
    s = scenes.settling_box(n=500, seed=5, box=(0.12, 0.12, 0.3))
    dt = float(s.time_dt[0])
    s.time_total[0] = 90.5 * dt
    s.time_file_dt[0] = 30.0 * dt
    s.sid = "runcheck"
    assert s.run(verbose=False, root=str(tmp_path)) == 0
    out = tmp_path / "output"
    last = status("runcheck", root=str(tmp_path))
    line = (out / "runcheck.status.dat").read_text().split()
    assert len(line) == 3 and int(line[2]) == last
    first = (out / "runcheck.output00000.bin").read_bytes()
    ref = oracle.Oracle(s.tobytes())
    assert oracle.Oracle(first).tobytes() == first
    # Replay the host loop's step counts with the same double additions.
    t, ftc, k = 0.0, 0.0, 0
    files = {}
    nfile = 0
    while t <= float(s.time_total[0]):
        ref.step(1)
        t += dt
        ftc += dt
        if ftc >= float(s.time_file_dt[0]):
            nfile += 1
            files[nfile] = ref.tobytes()
            ftc = 0.0
    # The replay counts files with the host loop's own double additions.
    assert len(files) == last >= 2
    for i, img in files.items():
        got = (out / ("runcheck.output%05d.bin" % i)).read_bytes()
        a = sim().frombytes(got)
        b = sim().frombytes(img)
        a.time_step_count[0] = b.time_step_count[0] = 0
        assert a.compare(b) == [], i


def test_cli_runs_from_project_root(tmp_path):
    s = scenes.settling_box(n=200, seed=6, box=(0.1, 0.1, 0.2))
    dt = float(s.time_dt[0])
    s.time_total[0] = 20.5 * dt
    s.time_file_dt[0] = 10.0 * dt
    s.sid = "clirun"
    s.writebin(str(tmp_path / "input"), verbose=False)
    cli = os.path.join(ROOT, "sphere_amd", "bin", "sphere")
    r = subprocess.run([cli, "-q", "input/clirun.bin"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "output" / "clirun.output00002.bin").exists()
    assert status("clirun", root=str(tmp_path)) == 2
