"""The C ABI library and the host logic around it, without a GPU.

Input validation happens on the host before any device call, so the
reference's error behaviour (checkValues, src/sphere.cpp:145-345; the
version check, src/file_io.cpp:47-54) is testable here.
"""

import os
import re
import subprocess

import numpy
import pytest

from sphere_amd import binfmt, engine
from sphere_amd.sim import sim

from . import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sphere_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sph_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = engine.load()
    syms = declared_symbols()
    assert len(syms) >= 18
    missing = [s for s in syms if not hasattr(L, s)]
    assert missing == []
    assert set(syms) == set(engine.SYMBOLS)


def _small():
    s = sim(np=3)
    s.radius[:] = 0.01
    s.x[:] = [[0.1, 0.1, 0.1], [0.2, 0.1, 0.1], [0.1, 0.2, 0.1]]
    s.defineWorldBoundaries(L=[0.3, 0.3, 0.3])
    s.time_dt[0] = 1e-6
    s.time_total[0] = 1.0
    return s


def _open_error(image):
    with pytest.raises(engine.EngineError) as ei:
        engine.Context(image)
    return ei.value.code, str(ei.value)


def test_rejects_wrong_version():
    img = bytearray(_small().tobytes())
    img[0:8] = numpy.array([2.14]).tobytes()
    code, msg = _open_error(bytes(img))
    assert code == -2 and "version" in msg


@pytest.mark.parametrize("mutate,needle", [
    (lambda s: s.x.__setitem__((1, 0), numpy.nan), "non-real position"),
    (lambda s: s.x.__setitem__((1, 0), 0.31), "outside of the computational grid"),
    (lambda s: s.radius.__setitem__(2, 0.0), "radius"),
    (lambda s: s.origo.__setitem__(0, -0.1), "Negative values in grid.origo"),
    (lambda s: s.k_n.__setitem__(0, 0.0), "k_n"),
    (lambda s: s.rho.__setitem__(0, -1.0), "rho"),
    (lambda s: s.time_current.__setitem__(0, 2.0), "time.current"),
    (lambda s: (s.periodic.__setitem__(0, 1), s.num.__setitem__(0, 2)), "at least 3 cells"),
])
def test_check_values_errors(mutate, needle):
    s = _small()
    mutate(s)
    code, msg = _open_error(s.tobytes())
    assert code == -3 and needle in msg


def test_particle_above_top_wall_rejected():
    s = _small()
    s.nw = numpy.ones(1, dtype=numpy.uint32)
    s.wmode = numpy.array([1], dtype=numpy.int32)
    s.w_n = numpy.array([[0.0, 0.0, -1.0]])
    s.w_x = numpy.array([0.05])
    s.w_m = numpy.ones(1)
    s.w_vel = numpy.zeros(1)
    s.w_force = numpy.zeros(1)
    s.w_sigma0 = numpy.zeros(1)
    code, msg = _open_error(s.tobytes())
    assert code == -3 and "upper, dynamic wall" in msg


def test_truncated_file_rejected():
    img = _small().tobytes()
    code, _ = _open_error(img[:200])
    assert code == -1


def test_bond_index_out_of_range_rejected():
    s = _small()
    s.bond(0, 7)
    code, msg = _open_error(s.tobytes())
    assert code == -3 and "bond" in msg


def test_cli_help_version_and_missing_input():
    cli = os.path.join(ROOT, "sphere_amd", "bin", "sphere")
    assert subprocess.run([cli, "--help"], capture_output=True).returncode == 0
    assert subprocess.run([cli, "-V"], capture_output=True).returncode == 0
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 1 and "missing input binary file" in r.stderr
    r = subprocess.run([cli, "-q", "/nonexistent/x.bin"], capture_output=True, text=True)
    assert r.returncode == 1 and "Could not read input binary file" in r.stderr


def test_forcechains_help_version_and_missing_input():
    """src/forcechains.cpp:16-103: help and version exit 0, no input file
    exits 1 with the reference's message."""
    exe = os.path.join(ROOT, "sphere_amd", "bin", "forcechains")
    r = subprocess.run([exe, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "force chain visualizer" in r.stdout
    r = subprocess.run([exe, "-V"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout == "Force chain calculator, sphere version 2.15\n"
    r = subprocess.run([exe, "-f", "txt"], capture_output=True, text=True)
    assert r.returncode == 1 and "missing input binary file" in r.stderr
    r = subprocess.run([exe, "-lc"], capture_output=True, text=True)
    assert r.returncode == 1 and "needs a value" in r.stderr


def test_reference_fixture_decodes_and_upgrades():
    """output/cube-init.output00029.bin (v2.11): header values and the
    upgrade to a v2.15 image of the expected size."""
    s = cases.fixture("cube-init.output00029.bin")
    assert int(s.np[0]) == 1000 and list(s.num) == [3, 3, 112] and int(s.periodic[0]) == 1
    assert float(s.time_current[0]) == 5.800000296761987
    assert float(s.k_n[0]) == 1.0e8 and int(s.nw[0]) == 0
    img = s.tobytes()
    assert len(img) == binfmt.image_size(1000)
    t = sim().frombytes(img)
    assert t.compare(s) == []


def test_python_codec_round_trip_with_walls_and_bonds():
    s = cases.fixture("shear1k.bin")
    s.bond(3, 4)
    s.bonds_delta_n[-1] = 1.5e-7
    s.bonds_omega_t[-1] = [1.0, -2.0, 3.0]
    img = s.tobytes()
    t = sim().frombytes(img)
    assert t.compare(s) == []
    assert t.tobytes() == img
