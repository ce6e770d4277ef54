"""Output-side readers on the GPU against the oracle's restatement of the
reference: DEM::findOverlaps (`sphere --contacts`) and DEM::porosity
(`sphere --porosity`), bit for bit, on the current state of a context."""

import numpy
import pytest

import oracle
from sphere_amd import engine

from . import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["shear1k", "settling", "crowded"])
def test_find_overlaps_matches_oracle(case):
    if case == "shear1k":
        s = cases.fixture("shear1k.bin")
    elif case == "settling":
        s = cases.settling_case(1200, 12)
    else:
        s = cases.crowded_case()
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    ref.step(30)
    ctx.step(30)
    pa, da = ctx.find_overlaps()
    pb, db = ref.find_overlaps()
    assert len(pb) > 0
    assert numpy.array_equal(pa, pb) and numpy.array_equal(da, db)
    ctx.close()


def test_find_overlaps_in_a_batch_and_with_a_short_buffer():
    a, b = cases.fixture("shear1k.bin"), cases.settling_case(300, 11)
    ctx = engine.Context(images=[a.tobytes(), b.tobytes()])
    for k, s in enumerate((a, b)):
        pa, da = ctx.find_overlaps(sim=k)
        pb, db = oracle.Oracle(s.tobytes()).find_overlaps()
        assert numpy.array_equal(pa, pb) and numpy.array_equal(da, db)
    # a buffer shorter than the list: the first pairs, and the full count
    import ctypes
    L = engine.load()
    n = ctypes.c_uint64()
    pairs = numpy.zeros((5, 2), dtype=numpy.uint32)
    delta = numpy.zeros(5)
    assert L.sph_find_overlaps(ctx._h, 0, pairs.ctypes.data, delta.ctypes.data, 5,
                               ctypes.byref(n)) == 0
    pb, db = oracle.Oracle(a.tobytes()).find_overlaps()
    assert n.value == len(pb)
    assert numpy.array_equal(pairs, pb[:5]) and numpy.array_equal(delta, db[:5])
    ctx.close()


@pytest.mark.parametrize("case,slices", [("shear1k", 8), ("settling", 5)])
def test_porosity_matches_oracle(case, slices):
    s = cases.fixture("shear1k.bin") if case == "shear1k" else cases.settling_case(1200, 12)
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ctx = engine.Context(image)
    ref.step(20)
    ctx.step(20)
    za, pa = ctx.porosity(slices)
    zb, pb = ref.porosity(slices)
    assert numpy.array_equal(za, zb) and numpy.array_equal(pa, pb)
    with pytest.raises(engine.EngineError):
        ctx.porosity(100000)
    ctx.close()


def test_sim_find_overlaps_and_contact_consumers():
    """sim.findOverlaps fills .pairs (2 x k) and .overlaps like the reference
    (python/sphere.py:4868-4883); the consumers count and price them."""
    s = cases.fixture("shear1k.bin")
    pb, db = oracle.Oracle(s.tobytes()).find_overlaps()
    assert len(pb) > 1000
    s.findOverlaps()
    assert s.pairs.dtype == numpy.int32 and s.pairs.shape == (2, len(pb))
    assert numpy.array_equal(s.pairs.T, pb) and numpy.array_equal(s.overlaps, db)
    cn = s.findCoordinationNumber()
    assert cn.sum() == 2 * len(pb)
    assert s.findMeanCoordinationNumber() == 2.0 * len(pb) / int(s.np[0])
    s.findContactStresses("average")
    assert numpy.allclose(s.f_n_magn, s.k_n[0] * numpy.abs(db))
    # lens area of two equal spheres: pi (r^2 - (d/2)^2)
    r = 1.0
    t = type(s)(np=2)
    t.radius[:] = r
    a = t.contactSurfaceArea(0, 1, -0.2)
    assert abs(a - numpy.pi * (r * r - 0.9 ** 2)) < 1e-12
    loaded = s.findLoadedContacts(0.0)
    assert len(loaded[0]) == len(pb)
    por, depth = s.porosity(5)
    zb, pb2 = oracle.Oracle(s.tobytes()).porosity(5)
    assert numpy.array_equal(depth, zb) and numpy.array_equal(por, pb2)


def test_cli_contacts_prints_the_overlaps(tmp_path):
    """`sphere --contacts input/<sid>.bin` prints "i<TAB>j<TAB>delta" per
    overlapping pair and exits 0 (src/main.cpp, src/sphere.cpp:904-913)."""
    import os
    import subprocess
    s = cases.fixture("shear1k.bin")
    s.sid = "contactlist"
    s.writebin(str(tmp_path / "input"), verbose=False)
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "sphere_amd", "bin", "sphere")
    # -q: without it the listing follows the "processing input file" line,
    # as the reference's does (src/main.cpp:169-177)
    r = subprocess.run([cli, "-q", "--contacts", "input/contactlist.bin"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = numpy.loadtxt(r.stdout.splitlines())
    pb, db = oracle.Oracle(s.tobytes()).find_overlaps()
    assert data.shape == (len(pb), 3)
    assert numpy.array_equal(data[:, :2].astype(numpy.int64), pb.astype(numpy.int64))
    # std::ostream's default six significant digits, as the reference prints
    assert numpy.allclose(data[:, 2], db, rtol=1e-5, atol=0)


def _forcechains_expected(s, pairs, delta, fmt, threedim, lower, upper, sid):
    """DEM::forcechains' printout (src/sphere.cpp:727-901), formatted as
    std::ostream's defaults (printf %g) from the oracle's overlap list."""
    g = lambda v: "%g" % v
    x = numpy.column_stack([s.x, s.radius])
    lo = numpy.minimum(0.0, s.x.min(axis=0))
    hi = numpy.maximum(0.0, s.x.max(axis=0))
    k_n = float(s.k_n[0])
    f_max = -k_n * delta.min()
    out = []
    if fmt == "txt":
        h = "x_1, [m]\t" + ("y_1, [m]\t" if threedim else "") + "z_1, [m]\t" + "x_2, [m]\t"
        h += ("y_2, [m]\t" if threedim else "") + "z_2, [m]\t" + "||f_n||, [N]" + "i\tj\n"
        out.append(h)
    else:
        out.append("#!/usr/bin/env gnuplot\n"
                   "# This Gnuplot script is automatically generated using\n"
                   "# the forcechain utility in sphere. For more information,\n"
                   "# see https://github.com/anders-dc/sphere\n"
                   "set size ratio -1\n")
        if fmt == "png":
            out.append("set term pngcairo size 30 cm,20 cm\nset out '%s-fc.png'\n" % sid)
        out.append("set xlabel '\\sffamily $x_1$, [m]'\n")
        if threedim:
            out.append("set ylabel '\\sffamily $x_2$, [m]'\n"
                       "set zlabel '\\sffamily $x_3$, [m]' offset 2\n")
        else:
            out.append("set ylabel '\\sffamily $x_3$, [m]' offset 2\n")
        out.append("set cblabel '\\sffamily $||\\boldsymbol{f}_n||$, [N]'\n"
                   "set xyplane at %s\nset pm3d\nset view 90.0,0.0\n"
                   "set palette defined ( 1 '#d3d3d3', 2 '#000000')\n"
                   "set cbrange [%s:%s]\n\n" % (g(lo[2]), g(lower), g(upper)))
    for k, ((i, j), d) in enumerate(zip(pairs, delta)):
        f = -k_n * d
        if f < lower:
            continue
        a, b = x[i], x[j]
        if fmt == "txt":
            row = g(a[0]) + "\t" + (g(a[1]) + "\t" if threedim else "") + g(a[2]) + "\t"
            row += g(b[0]) + "\t" + (g(b[1]) + "\t" if threedim else "") + g(b[2]) + "\t"
            out.append(row + g(f) + "%d\t%d\n" % (i, j))
        elif f > lower and s.fixvel[i] + s.fixvel[j] == 0.0:
            row = "set arrow %d from " % (k + 1) + g(a[0]) + ","
            row += (g(a[1]) + "," if threedim else "") + g(a[2]) + " to " + g(b[0]) + ","
            row += (g(b[1]) + "," if threedim else "") + g(b[2])
            out.append(row + " nohead lw " + g(f / f_max * 16.0) + " lc palette cb " + g(f) + "\n")
    if fmt != "txt":
        foot = ("splot " if threedim else "plot ") + "[%s:%s] " % (g(lo[0]), g(hi[0]))
        if threedim:
            foot += "[%s:%s] " % (g(lo[1]), g(hi[1]))
        out.append(foot + "[%s:%s] NaN notitle\n" % (g(lo[2]), g(hi[2])))
    return "".join(out)


@pytest.mark.parametrize("fmt,twod,lower", [("txt", False, 0.0), ("txt", True, 0.3),
                                             ("png", False, 0.0), ("interactive", True, 0.5)])
def test_forcechains_tool(tmp_path, fmt, twod, lower):
    """`forcechains` (src/forcechains.cpp) prints the same script / table as
    the reference would from the oracle's overlap list, byte for byte."""
    import os
    import subprocess
    s = cases.fixture("shear1k.bin")
    s.sid = "fc.test"
    s.writebin(str(tmp_path), verbose=False)
    pb, db = oracle.Oracle(s.tobytes()).find_overlaps()
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "sphere_amd", "bin", "forcechains")
    args = [exe, "-f", fmt, "-lc", str(lower), "-uc", "2"] + (["-2d"] if twod else [])
    r = subprocess.run(args + [str(tmp_path / "fc.test.bin")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    want = _forcechains_expected(s, pb.astype(numpy.int64), db, fmt, not twod, lower, 2.0,
                                 "fc-test")
    assert r.stdout == want


def _overlap_case(kind):
    """Inputs that stress the overlap grid: a dense pile with coincident
    particles (long per-particle runs), a bed with one particle far away
    (a coarsened grid), zero radii, and positions that are not finite (the
    all-pairs fallback)."""
    from sphere_amd.sim import sim
    rng = numpy.random.default_rng(17)
    if kind == "coincident":
        n = 600
        s = sim(np=n)
        s.x[:] = rng.integers(0, 6, (n, 3)) * 0.5
        s.radius[:] = rng.uniform(0.2, 0.6, n)
    elif kind == "far_outlier":
        n = 3000
        s = sim(np=n)
        s.x[:] = rng.uniform(0.0, 1.0, (n, 3))
        s.radius[:] = rng.uniform(0.01, 0.03, n)
        s.x[n // 2] = [1.0e7, -3.0e6, 2.0e5]
    elif kind == "zero_radii":
        n = 500
        s = sim(np=n)
        s.x[:] = rng.integers(0, 4, (n, 3)) * 0.25
        s.radius[:] = 0.0
        s.radius[::7] = 0.2
    else:
        n = 800
        s = sim(np=n)
        s.x[:] = rng.uniform(0.0, 1.0, (n, 3))
        s.radius[:] = rng.uniform(0.02, 0.05, n)
        s.x[5, 0] = numpy.inf
        s.x[9, 1] = numpy.nan
    s.fixvel[::11] = 1.0
    s.vel[:, 0] = rng.normal(0.0, 1.0, n)
    s.defineWorldBoundaries(L=[3.0, 3.0, 3.0])
    with numpy.errstate(divide="ignore"):  # zero radii: the dt rule divides by 0
        s.initTemporal(total=1.0, file_dt=0.1)
    return s


@pytest.mark.parametrize("kind", ["coincident", "far_outlier", "zero_radii", "not_finite"])
def test_find_overlaps_edge_inputs(kind):
    s = _overlap_case(kind)
    image = s.tobytes()
    pb, db = oracle.Oracle(image).find_overlaps()
    with engine.Context(image, check=False) as ctx:
        pa, da = ctx.find_overlaps()
    if kind == "coincident":
        assert len(pb) > 10 * len(s.radius)
    assert numpy.array_equal(pa, pb) and numpy.array_equal(da, db)


def test_find_overlaps_large_cloud():
    """60 000 polydisperse particles at random in a box (about four overlaps
    each): the grid search against the oracle's all-pairs loop."""
    from sphere_amd.sim import sim
    rng = numpy.random.default_rng(23)
    n = 60000
    s = sim(np=n)
    s.x[:] = rng.uniform(0.0, 1.0, (n, 3)) * [2.0, 1.0, 0.5]
    s.radius[:] = rng.uniform(0.012, 0.02, n)
    s.fixvel[::13] = 1.0
    s.defineWorldBoundaries(L=[2.0, 1.0, 0.5])
    s.initTemporal(total=1.0, file_dt=0.1)
    image = s.tobytes()
    pb, db = oracle.Oracle(image).find_overlaps()
    with engine.Context(image, check=False) as ctx:
        pa, da = ctx.find_overlaps()
    assert len(pb) > 2 * n
    assert numpy.array_equal(pa, pb) and numpy.array_equal(da, db)
