"""Scenes shared by the CPU and GPU tests (built on the CPU, seeded)."""

import os

import numpy

from sphere_amd import scenes
from sphere_amd.sim import sim

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    """A committed .bin fixture, read (and upgraded to v2.15 when older)."""
    s = sim()
    s.readbin(os.path.join(GOLDEN, name), verbose=False)
    return s


def settling_case(n, seed):
    """A closed box with downward and sideways initial velocities, so that
    particles reach the floor and the side walls within a short test."""
    box = (0.12, 0.12, 0.3) if n <= 400 else (0.16, 0.16, 0.45)
    s = scenes.settling_box(n=n, seed=seed, box=box)
    rng = numpy.random.default_rng(seed)
    s.vel[:, 2] = -rng.uniform(0.5, 2.0, n)
    s.vel[:, 0] = rng.uniform(-0.5, 0.5, n)
    return s


def crowded_case(n_small=44, seed=9):
    """One large sphere overlapped by more small spheres than a contact row
    holds (NC = 32), so that the slot policy's overflow rules (slot nc is
    overwritten, the rest of the cell is skipped) decide the contact rows."""
    from sphere_amd.sim import sim
    rng = numpy.random.default_rng(seed)
    s = sim(np=1 + n_small, sid="crowded")
    s.radius[0] = 1.0
    s.radius[1:] = rng.uniform(0.25, 0.3, n_small)
    s.x[0, :] = [3.0, 3.0, 3.0]
    # Fibonacci sphere directions, each small sphere 0.05 inside contact.
    k = numpy.arange(n_small) + 0.5
    phi = numpy.arccos(1.0 - 2.0 * k / n_small)
    th = numpy.pi * (1.0 + 5.0 ** 0.5) * k
    d = numpy.stack([numpy.cos(th) * numpy.sin(phi), numpy.sin(th) * numpy.sin(phi),
                     numpy.cos(phi)], axis=1)
    s.x[1:, :] = s.x[0, :] + d * (1.0 + s.radius[1:, None] - 0.05)
    s.vel[1:, :] = rng.normal(0.0, 0.01, (n_small, 3))
    s.defineWorldBoundaries(L=[6.0, 6.0, 6.0])
    s.initTemporal(total=1.0, file_dt=0.1)
    return s


def lattice_bed(n, seed, median, sigma, k_n=1.16e7, normal_stress=10.0e3, shear_vel=1.0e-3,
                bond_frac=0.0, aspect=2.0, rho=2600.0, mu=0.5):
    """A direct-shear bed of any size, built in O(n): lognormal radii
    (median, sigma), clipped at 0.55 of the lattice spacing 2 median, on a
    jittered cubic lattice periodic in x and y, so that a particle touches at
    most its six lattice neighbours (about half of them do); the top wall at
    `normal_stress` (consolidate) and the top layer sheared at `shear_vel`
    (shear).  `bond_frac` of the touching lattice pairs are bonded, breaking
    at a stress of k_n * 0.05 * median (a bond stretched by 5% of a median
    radius)."""
    import math
    rng = numpy.random.default_rng(seed)
    d = 2.0 * median
    r = numpy.minimum(scenes.lognormal_radii(n, median, sigma, rng), 0.55 * d)
    nxy = int(math.ceil((n * aspect) ** (1.0 / 3.0)))
    nz = int(math.ceil(n / float(nxy * nxy)))
    i = numpy.arange(n)
    ix, iy, iz = i % nxy, (i // nxy) % nxy, i // (nxy * nxy)
    x = numpy.stack([ix, iy, iz], axis=1) * d + 0.5 * d + 0.02 * d * rng.uniform(-1.0, 1.0,
                                                                                (n, 3))
    s = sim(np=n, sid="bed-%d" % n)
    s.radius = r
    s.x = x
    s.defaultParams(mu_s=mu, mu_d=mu, rho=rho, k_n=k_n, k_t=k_n)
    scenes._set_contact_params(s, k_n, 0.5, mu, float(r.min()))
    s.g[:] = [0.0, 0.0, -9.81]
    s.periodicBoundariesXY()
    s.defineWorldBoundaries(L=[nxy * d, nxy * d, (nz + 1) * d])
    dt = scenes.dt_rule(float(r.min()), rho, k_n)
    s.time_dt[0] = dt
    s.time_current[0] = 0.0
    s.time_total[0] = 1.0e3
    s.time_file_dt[0] = 1.0e3
    s.consolidate(normal_stress)
    z_min = float(numpy.min(s.x[:, 2] - s.radius))
    z_max = float(numpy.max(s.x[:, 2] + s.radius))
    s.shear(shear_strain_rate=shear_vel / (z_max - z_min))
    if bond_frac > 0.0:
        pairs = []
        for step, ok in ((1, ix < nxy - 1), (nxy, iy < nxy - 1), (nxy * nxy, iz < nz - 1)):
            a = i[ok & (i + step < n)]
            b = a + step
            dist = numpy.linalg.norm(s.x[b] - s.x[a], axis=1)
            pairs.append(numpy.stack([a, b], axis=1)[dist < r[a] + r[b]])
        pairs = numpy.concatenate(pairs)
        pick = numpy.sort(rng.choice(len(pairs), size=max(1, int(bond_frac * len(pairs))),
                                     replace=False))
        nb = len(pick)
        s.bonds = pairs[pick].astype(numpy.uint32)
        s.nb0[0] = nb
        s.lambda_bar[0] = 1.0
        s.bonds_delta_n = numpy.zeros(nb)
        s.bonds_delta_t = numpy.zeros((nb, 3))
        s.bonds_omega_n = numpy.zeros(nb)
        s.bonds_omega_t = numpy.zeros((nb, 3))
        s.sigma_b[0] = k_n * 0.05 * median
        s.tau_b[0] = k_n * 0.05 * median
    return s


def fcc_case(cells=(4, 4, 6), drop=5, seed=3, overlap=0.004, k_n=1.16e7):
    """Equal spheres (r = 0.005 m) on a face-centred cubic lattice, closer
    than touching by `overlap` of the diameter, so that an interior sphere
    starts with 12 contacts: a wavefront of 16 particles holds far more than
    64 contacts.  The last `drop` lattice sites are left empty (np is not a
    multiple of 16), small random velocities, gravity, walls on every side."""
    rng = numpy.random.default_rng(seed)
    r = 0.005
    a = 2.0 * r * (1.0 - overlap) * 2.0 ** 0.5  # nearest neighbours 2 r (1 - overlap)
    basis = numpy.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    nx, ny, nz = cells
    pts = [(numpy.array([i, j, k]) + b) * a for k in range(nz) for j in range(ny)
           for i in range(nx) for b in basis]
    pts = numpy.array(pts)[: len(pts) - drop] + 1.5 * r
    n = len(pts)
    s = sim(np=n, sid="fcc-%d" % n)
    s.radius[:] = r
    s.x = pts
    s.vel[:, :3] = rng.normal(0.0, 0.01, (n, 3))
    s.defaultParams(mu_s=0.5, mu_d=0.5, rho=2600.0, k_n=k_n, k_t=k_n)
    scenes._set_contact_params(s, k_n, 0.5, 0.5, r)
    s.g[:] = [0.0, 0.0, -9.81]
    s.normalBoundariesXY()
    s.defineWorldBoundaries(L=[nx * a + 3.0 * r, ny * a + 3.0 * r, nz * a + 3.0 * r])
    dt = scenes.dt_rule(r, 2600.0, k_n)
    s.time_dt[0] = dt
    s.time_current[0] = 0.0
    s.time_total[0] = dt * 1000
    s.time_file_dt[0] = dt * 1000
    return s
