"""Every scene the GPU tests use is built (and stepped on the oracle) here."""

import numpy
import pytest

import oracle
from sphere_amd import scenes

from . import cases


@pytest.mark.parametrize("n,seed", [(300, 11), (1200, 12)])
def test_settling_case_builds(n, seed):
    s = cases.settling_case(n, seed)
    o = oracle.Oracle(s.tobytes())
    o.step(2)


@pytest.mark.parametrize("n,seed,shrink", [(500, 21, 0.97), (2000, 22, 0.95), (800, 31, 0.96),
                                           (800, 32, 0.96), (400, 41, 0.97), (600, 7, 0.97)])
def test_shear_case_builds(n, seed, shrink):
    s = scenes.shear_box(n=n, seed=seed, phi=0.3, shrink=shrink)
    o = oracle.Oracle(s.tobytes())
    o.step(2)
    assert (o.get("contacts") != n).sum() > 0


@pytest.mark.parametrize("name", ["shear1k.bin", "cube-init.output00029.bin", "init-1e4-slab.bin"])
def test_fixture_loads_and_is_contact_rich(name):
    s = cases.fixture(name)
    n = int(s.np[0])
    o = oracle.Oracle(s.tobytes())
    o.step(1)
    assert (o.get("contacts") != n).sum() > n


def test_bonded_cases_build():
    from .test_gpu_bonds import _bonded
    import numpy
    for frac, seed, sb in ((0.2, 3, numpy.inf), (0.3, 4, 300.0)):
        s = _bonded(frac=frac, seed=seed, sigma_b=sb)
        assert int(s.nb0[0]) > 10
        o = oracle.Oracle(s.tobytes())
        o.step(2)


def test_bond_breakage_case_breaks_on_oracle():
    from .test_gpu_bonds import _bonded
    s = _bonded(frac=0.3, seed=4, sigma_b=300.0)
    o = oracle.Oracle(s.tobytes())
    o.step(120)
    assert (o.get("bonds")[:, 0] == int(s.nb0[0])).any()


def test_wmode3_adaptive_rate_state_cases_build():
    s = cases.fixture("shear1k.bin")
    s.shear(shear_stress=2.0e3)
    assert int(s.wmode[0]) == 3
    oracle.Oracle(s.tobytes()).step(2)
    s = cases.fixture("shear1k.bin")
    s.adaptiveGrid()
    oracle.Oracle(s.tobytes()).step(2)
    s = cases.fixture("shear1k.bin")
    top = s.fixvel > 0
    assert top.any()


def test_crowded_case_overflows_contact_row():
    s = cases.crowded_case()
    n = int(s.np[0])
    o = oracle.Oracle(s.tobytes())
    o.step(3)
    row = o.get("contacts")[0]
    assert (row != n).all()  # all NC slots of the large sphere are taken
    d = numpy.linalg.norm(s.x[1:] - s.x[0], axis=1) - (s.radius[1:] + s.radius[0])
    assert (d < 0).sum() > 32


def test_fcc_case_crowds_every_wavefront():
    """The dense packing: interior spheres with 12 contacts, more than 64
    contacts in every run of 16 sorted particles, a partial last run, and
    a packing that stays bounded over 200 steps."""
    s = cases.fcc_case()
    n = int(s.np[0])
    assert n % 16 != 0
    o = oracle.Oracle(s.tobytes())
    o.step(1)
    z = (o.get("contacts") != n).sum(1)
    assert z.max() == 12
    zs = z[o.get("perm").astype(numpy.int64)]
    assert numpy.add.reduceat(zs, numpy.arange(0, n, 16)).min() > 64
    o.step(200)
    assert numpy.isfinite(o.get("x")).all()


def test_capillary_case_builds_and_bridges():
    s = cases.fixture("shear1k.bin")
    s.contactmodel[0] = 1
    s.kappa[0] = 0.5
    s.db[0] = 2.0e-4
    s.V_b[0] = 1.0e-9
    o = oracle.Oracle(s.tobytes())
    o.step(2)
    f = o.get("force")
    assert numpy.isfinite(f).all()


def test_tiled_shear_cell_moves_like_its_copies():
    """A periodic cell tiled 3 x 2 is the same bed: after a few oracle steps
    every copy's particles are where the single cell's are (to the rounding
    of the shifted coordinates), and the contact count scales by six."""
    s = cases.fixture("shear1k.bin")
    n = int(s.np[0])
    t = scenes.tile_xy(s, 3, 2)
    assert int(t.np[0]) == 6 * n
    a, b = oracle.Oracle(s.tobytes()), oracle.Oracle(t.tobytes())
    a.step(5)
    b.step(5)
    xa, xb = a.get("x")[:, :3], b.get("x")[:, :3]
    L = numpy.asarray(s.L, dtype=numpy.float64)
    for k in range(6):
        i, j = k % 3, k // 3
        d = xb[k * n:(k + 1) * n] - (xa + [i * L[0], j * L[1], 0.0])
        assert numpy.abs(d).max() < 1e-9
    va, vb = a.get("vel")[:, :3], b.get("vel")[:, :3]
    assert numpy.abs(vb - numpy.tile(va, (6, 1))).max() < 1e-6 * numpy.abs(va).max()
    za = int((a.get("contacts") != n).sum())
    zb = int((b.get("contacts") != 6 * n).sum())
    assert zb == 6 * za


def test_batch_shear_cells_sizes_and_parameters():
    s = cases.fixture("shear1k.bin")
    sizes = [1000, 3000, 12000, 20000]
    cells = scenes.batch_shear_cells(sizes, [s], seed=3)
    assert [int(c.np[0]) for c in cells] == sizes
    mus = [float(c.mu_s[0]) for c in cells]
    assert len(set(mus)) == len(mus) and all(0.3 <= m <= 0.7 for m in mus)
    assert all(5e3 <= float(c.w_sigma0[0]) <= 1.6e5 for c in cells)
    assert float(s.w_sigma0[0]) == 1e4  # the base cell is untouched


def test_batch_shear_cells_with_own_stiffness_stay_stable():
    """The stiffest and the softest draw of the batch's stiffness range, each
    with its own dt, step stably on the oracle from the tiled bed."""
    s = cases.fixture("shear1k.bin")
    for kn in (1.0e6, 1.0e8):
        c = scenes.batch_shear_cells([2000], [s], seed=5, kn_range=(kn, kn),
                                     stress_range=(1.6e5, 1.6e5))[0]
        assert float(c.k_n[0]) == pytest.approx(kn)
        assert float(c.time_dt[0]) == pytest.approx(scenes.dt_rule(float(c.radius.min()),
                                                                   2600.0, kn))
        o = oracle.Oracle(c.tobytes())
        o.step(200)
        x = o.get("x")[:, :3]
        assert numpy.isfinite(x).all()
        assert (x[:, 2] > 0.0).all() and (x[:, 2] < float(c.L[2])).all()


def test_lattice_bed_is_a_sheared_bonded_bed():
    """The full-size bed builder at a small size: every bond joins two
    touching lattice neighbours, radii stay below 0.55 of the spacing, the
    top wall is stress-controlled and the top layer moves."""
    s = cases.lattice_bed(6000, seed=3, median=0.005, sigma=0.25, k_n=1.0e8,
                          normal_stress=20.0e3, shear_vel=5.0e-4, bond_frac=0.1)
    n = int(s.np[0])
    assert n == 6000 and int(s.periodic[0]) == 1 and list(s.wmode) == [1]
    assert float(s.w_sigma0[0]) == 20.0e3
    assert s.radius.max() <= 0.55 * 0.01
    nb = int(s.nb0[0])
    assert nb == len(s.bonds) > 50
    a, b = s.bonds[:, 0].astype(int), s.bonds[:, 1].astype(int)
    dist = numpy.linalg.norm(s.x[b] - s.x[a], axis=1)
    assert numpy.all(dist < s.radius[a] + s.radius[b])
    assert numpy.all(numpy.isin(b - a, [1, 23, 23 * 23]))  # lattice neighbours (23 per row)
    top = s.fixvel > 0
    assert numpy.any(s.vel[top, 0] > 0.0) and numpy.all(s.vel[~top] == 0.0)
    assert s.sigma_b[0] == s.tau_b[0] == 1.0e8 * 0.05 * 0.005
