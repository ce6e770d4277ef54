"""Seeded synthetic inputs for the DEM workloads.

The reference's own setup helpers draw from numpy's unseeded global RNG
(python/sphere.py:2358-2846); these builders use explicit seeds so that the
HIP engine and the CPU oracle consume identical .bin images.  Parameter
rules follow python/sphere.py:3762-3873 (defaultParams) and 3307-3376
(initTemporal), with the configuration texts' choices:

* k_t = k_n, gamma_n = gamma_t = frac * 2 sqrt(m_min k_n) with the true
  smallest mass m_min = rho 4/3 pi r_min^3;
* dt = 0.075 sqrt(m_min / k_n)  (initTemporal(epsilon=0.075));
* the sorting cell is 2.1 r_max (python/sphere.py:2632,2642-2644).
"""

import math

import numpy

from .sim import sim


def critical_damping(r_min, rho, k_n):
    m_min = rho * 4.0 / 3.0 * math.pi * r_min ** 3
    return 2.0 * math.sqrt(m_min * k_n)


def rsa_positions(radii, lo, hi, rng, periodic=(False, False, False), max_tries=20000,
                  cap=16):
    """Random sequential insertion without overlap, largest radius first.

    Centres lie in [lo + r, hi - r] on bounded axes and in [lo, hi) on
    periodic axes, where overlaps are also checked through the periodic
    image.  A uniform grid of cell size 2 r_max holds up to `cap` particles
    per cell; each attempt is checked against the 27 neighbouring cells with
    one vectorised distance test.
    """
    radii = numpy.asarray(radii, dtype=numpy.float64)
    n = radii.size
    lo = numpy.asarray(lo, dtype=numpy.float64)
    hi = numpy.asarray(hi, dtype=numpy.float64)
    per = numpy.asarray(periodic, dtype=bool)
    span = hi - lo
    cell = 2.0 * radii.max()
    dims = numpy.maximum(1, numpy.floor(span / cell).astype(numpy.int64))
    csz = span / dims
    ncell = int(numpy.prod(dims))
    slots = numpy.full((ncell, cap), -1, dtype=numpy.int64)
    count = numpy.zeros(ncell, dtype=numpy.int64)
    x = numpy.zeros((n, 3))
    offs = numpy.array([(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1)
                        for dx in (-1, 0, 1)], dtype=numpy.int64)
    # Largest first: late large particles would find no gap.
    for i in numpy.argsort(-radii, kind="stable"):
        r = radii[i]
        a = lo + numpy.where(per, 0.0, r)
        b = hi - numpy.where(per, 0.0, r)
        placed = False
        for _ in range(max_tries):
            c = a + (b - a) * rng.random(3)
            k = numpy.minimum((numpy.floor((c - lo) / csz)).astype(numpy.int64), dims - 1)
            nb = k[None, :] + offs
            ok_cell = numpy.ones(27, dtype=bool)
            for d in range(3):
                if per[d]:
                    nb[:, d] %= dims[d]
                else:
                    ok_cell &= (nb[:, d] >= 0) & (nb[:, d] < dims[d])
            nb = nb[ok_cell]
            cid = numpy.unique(nb[:, 0] + dims[0] * (nb[:, 1] + dims[1] * nb[:, 2]))
            js = slots[cid].ravel()
            js = js[js >= 0]
            if js.size:
                dv = c[None, :] - x[js]
                for d in range(3):
                    if per[d]:
                        dv[:, d] -= span[d] * numpy.round(dv[:, d] / span[d])
                dd = numpy.einsum("ij,ij->i", dv, dv)
                if numpy.any(dd < (r + radii[js]) ** 2):
                    continue
            ci = int(k[0] + dims[0] * (k[1] + dims[1] * k[2]))
            if count[ci] >= cap:
                continue
            slots[ci, count[ci]] = i
            count[ci] += 1
            x[i] = c
            placed = True
            break
        if not placed:
            raise RuntimeError("random sequential insertion failed at particle %d" % i)
    return x


def _set_contact_params(s, k_n, damping_frac, mu, r_min):
    rho = float(s.rho[0])
    s.k_n[0] = k_n
    s.k_t[0] = k_n
    gam = damping_frac * critical_damping(r_min, rho, k_n)
    s.gamma_n[0] = gam
    s.gamma_t[0] = gam
    s.mu_s[0] = mu
    s.mu_d[0] = mu
    s.mu_ws[0] = mu
    s.mu_wd[0] = mu


def dt_rule(r_min, rho, k_n, factor=0.075):
    m_min = rho * 4.0 / 3.0 * math.pi * r_min ** 3
    return factor * math.sqrt(m_min / k_n)


def settling_box(n=4096, seed=1001, box=(0.2, 0.2, 0.6), r_range=(0.005, 0.01), k_n=1.16e7,
                 damping_frac=0.5, mu=0.5, steps=2000, rho=2600.0):
    """Gravity settling in a closed box with frictional walls (the 4,096-sphere
    settling workload).  Side walls at 0 and L (periodic = 0), bottom at
    z = 0, top wall at L.z (nw = 0)."""
    rng = numpy.random.default_rng(seed)
    s = sim(np=n, sid="settling-%d" % n)
    s.radius = rng.uniform(r_range[0], r_range[1], n)
    s.x = rsa_positions(s.radius, [0.0, 0.0, 0.0], list(box), rng)
    s.defaultParams(mu_s=mu, mu_d=mu, rho=rho, k_n=k_n, k_t=k_n, gamma_wn=1.0e4, gamma_wt=1.0e4)
    _set_contact_params(s, k_n, damping_frac, mu, s.radius.min())
    s.g[:] = [0.0, 0.0, -9.81]
    s.normalBoundariesXY()
    s.defineWorldBoundaries(L=list(box))
    dt = dt_rule(s.radius.min(), rho, k_n)
    s.time_dt[0] = dt
    s.time_current[0] = 0.0
    s.time_total[0] = dt * steps
    s.time_file_dt[0] = dt * steps
    return s


def lognormal_radii(n, median, sigma, rng):
    return median * numpy.exp(sigma * rng.standard_normal(n))


def shear_box(n=2000, seed=2001, median=0.01, sigma=0.2, phi=0.25, lxy=None, k_n=1.16e7,
              damping_frac=0.5, mu=0.5, normal_stress=10.0e3, shear_vel=1.0e-3, rho=2600.0,
              shrink=1.0):
    """A periodic-x/y shear cell from a random loose packing.

    Particles are inserted without overlap at packing fraction `phi` in a
    box periodic in x and y; with shrink < 1 the insertion uses radii scaled
    by `shrink` and the true radii are restored afterwards, which leaves
    small overlaps (a contact-rich start).  Then the reference's consolidate()
    and shear() set the top wall (wmode 1, sigma0) and the fixed-velocity
    bottom and top layers (python/sphere.py:3019-3048,3097-3165).
    """
    rng = numpy.random.default_rng(seed)
    r = lognormal_radii(n, median, sigma, rng)
    vol = float(numpy.sum(4.0 / 3.0 * math.pi * r ** 3))
    if lxy is None:
        lxy = max(6 * 2.1 * r.max(), (vol / phi / 2.0) ** (1.0 / 3.0))
    h = vol / phi / (lxy * lxy)
    s = sim(np=n, sid="shear-%d" % n)
    s.radius = r
    s.x = rsa_positions(r * shrink, [0.0, 0.0, 0.0], [lxy, lxy, h], rng,
                        periodic=(True, True, False))
    s.defaultParams(mu_s=mu, mu_d=mu, rho=rho, k_n=k_n, k_t=k_n)
    _set_contact_params(s, k_n, damping_frac, mu, r.min())
    s.g[:] = [0.0, 0.0, -9.81]
    s.periodicBoundariesXY()
    z_top = float(numpy.max(s.x[:, 2] + r))
    s.defineWorldBoundaries(L=[lxy, lxy, z_top * 1.2])
    s.consolidate(normal_stress)
    s.shear(shear_strain_rate=shear_vel / (float(numpy.max(s.x[:, 2] + r)) -
                                           float(numpy.min(s.x[:, 2] - r))))
    dt = dt_rule(r.min(), rho, k_n)
    s.time_dt[0] = dt
    s.time_current[0] = 0.0
    s.time_total[0] = 1.0
    s.time_file_dt[0] = 1.0
    return s


def grow_pack(s, open_stepper, phi_rsa=0.28, stages=40, steps_per_stage=200, settle_steps=3000,
              seed=0, friction_during_growth=0.0, log=None):
    """Densify a periodic-x/y packing by growing the radii in place.

    `s` carries the true radii, the box (origo, L with L[2] the fixed top
    during growth) and the final contact parameters.  Particles are first
    inserted without overlap with radii scaled by s0 = (phi_rsa / phi)^(1/3)
    (phi the final packing fraction of the box), then the radii grow back to
    their true values in `stages` increments, each followed by
    `steps_per_stage` DEM steps without gravity and without friction.  Then
    gravity and friction are restored and the bed settles for `settle_steps`.

    `open_stepper(image)` returns an object with step(n), get(name),
    set(name, value) and tobytes() -- a GPU engine context or the CPU
    oracle.  Returns the settled sim.
    """
    rng = numpy.random.default_rng(seed)
    r = numpy.asarray(s.radius, dtype=numpy.float64).copy()
    n = r.size
    vol = float(numpy.sum(4.0 / 3.0 * math.pi * r ** 3))
    box = float((s.L[0] - s.origo[0]) * (s.L[1] - s.origo[1]) * (s.L[2] - s.origo[2]))
    phi = vol / box
    s0 = (phi_rsa / phi) ** (1.0 / 3.0)
    if s0 >= 1.0:
        s0 = 1.0
    s.x = rsa_positions(r * s0, list(s.origo), list(s.L), rng,
                        periodic=(s.periodic[0] in (1, 2), s.periodic[0] == 1, False))
    g = s.g.copy()
    mu = (float(s.mu_s[0]), float(s.mu_d[0]), float(s.mu_ws[0]), float(s.mu_wd[0]))
    s.g[:] = 0.0
    s.mu_s[0] = s.mu_d[0] = friction_during_growth
    s.mu_ws[0] = s.mu_wd[0] = friction_during_growth
    s.radius = r * s0
    st = open_stepper(s.tobytes())
    for k in range(1, stages + 1):
        sk = s0 + (1.0 - s0) * k / stages
        x = st.get("x").copy()
        x[:, 3] = r * sk
        st.set("x", x)
        st.step(steps_per_stage)
        if log:
            log("grow stage %d/%d scale %.4f" % (k, stages, sk))
    s.frombytes(st.tobytes())
    s.g[:] = g
    s.mu_s[0], s.mu_d[0], s.mu_ws[0], s.mu_wd[0] = mu
    s.vel[:] = 0.0
    s.angvel[:] = 0.0
    s.time_current[0] = 0.0
    if settle_steps:
        st = open_stepper(s.tobytes())
        st.step(settle_steps)
        s.frombytes(st.tobytes())
        s.time_current[0] = 0.0
    return s


def shear_cell(n, seed, median=0.01, sigma=0.2, lxy=None, phi=0.58, k_n=1.16e7,
               damping_frac=0.5, mu=0.5, normal_stress=10.0e3, shear_vel=1.0e-3, rho=2600.0,
               open_stepper=None, stages=40, steps_per_stage=200, settle_steps=3000,
               consolidate_steps=2000, log=None):
    """The direct-shear workload: lognormal radii, a packing grown to
    fraction `phi` in a box periodic in x and y, settled under gravity, then
    consolidate(normal_stress) (wmode 1) and shear() with the top layer at
    `shear_vel` (python/sphere.py:3019-3048,3097-3165).  With `open_stepper`
    (see grow_pack) the packing is relaxed by DEM steps; the final
    consolidation also runs `consolidate_steps` before the shear setup."""
    rng = numpy.random.default_rng(seed)
    r = lognormal_radii(n, median, sigma, rng)
    vol = float(numpy.sum(4.0 / 3.0 * math.pi * r ** 3))
    if lxy is None:
        lxy = max(6 * 2.1 * r.max(), (vol / phi / 1.5) ** (1.0 / 3.0))
    h = vol / phi / (lxy * lxy)
    s = sim(np=n, sid="shear-%d" % n)
    s.radius = r
    s.defaultParams(mu_s=mu, mu_d=mu, rho=rho, k_n=k_n, k_t=k_n)
    _set_contact_params(s, k_n, damping_frac, mu, r.min())
    s.g[:] = [0.0, 0.0, -9.81]
    s.periodicBoundariesXY()
    s.defineWorldBoundaries(L=[lxy, lxy, h])
    dt = dt_rule(r.min(), rho, k_n)
    s.time_dt[0] = dt
    s.time_current[0] = 0.0
    s.time_total[0] = 1.0e3
    s.time_file_dt[0] = 1.0e3
    grow_pack(s, open_stepper, stages=stages, steps_per_stage=steps_per_stage,
              settle_steps=settle_steps, seed=seed + 1, log=log)
    s.consolidate(normal_stress)
    if consolidate_steps:
        st = open_stepper(s.tobytes())
        st.step(consolidate_steps)
        s.frombytes(st.tobytes())
        s.time_current[0] = 0.0
    z_min = float(numpy.min(s.x[:, 2] - s.radius))
    z_max = float(numpy.max(s.x[:, 2] + s.radius))
    s.shear(shear_strain_rate=shear_vel / (z_max - z_min))
    return s
