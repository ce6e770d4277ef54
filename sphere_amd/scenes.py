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

import copy
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
    """k_t = k_n; gamma_n = gamma_t = damping_frac * critical (true m_min);
    the wall viscosities take the same value.  (The reference's default wall
    viscosity of 1e4 N s/m, python/sphere.py:3773-3774, makes the explicit
    update unstable for the smallest particles at dt = 0.075 sqrt(m_min/k_n):
    gamma dt / m_min exceeds 2.)"""
    rho = float(s.rho[0])
    s.k_n[0] = k_n
    s.k_t[0] = k_n
    gam = damping_frac * critical_damping(r_min, rho, k_n)
    s.gamma_n[0] = gam
    s.gamma_t[0] = gam
    s.gamma_wn[0] = gam
    s.gamma_wt[0] = gam
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
    s.defaultParams(mu_s=mu, mu_d=mu, rho=rho, k_n=k_n, k_t=k_n)
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


def lattice_box(n=5000, seed=1, median=0.005, sigma=0.2, aspect=2.0, k_n=1.16e7,
                damping_frac=0.5, mu=0.5, rho=2600.0, v0=0.2):
    """A granular gas collapsing under gravity in a closed box: lognormal
    radii on a jittered cubic lattice (spacing 2.2 r_max, so no initial
    overlap), random velocities of about v0, side walls at 0 and L.  Built
    in O(n) (no insertion search), for batches of many simulations."""
    rng = numpy.random.default_rng(seed)
    r = lognormal_radii(n, median, sigma, rng)
    d = 2.2 * r.max()
    nxy = max(3, int(math.ceil((n / aspect) ** (1.0 / 3.0))))
    nz = max(3, int(math.ceil(n / (nxy * nxy))))
    ix = numpy.arange(n) % nxy
    iy = (numpy.arange(n) // nxy) % nxy
    iz = numpy.arange(n) // (nxy * nxy)
    jit = (d / 2.0 - r.max()) * rng.uniform(-0.9, 0.9, (n, 3))
    s = sim(np=n, sid="lattice-%d" % n)
    s.radius = r
    s.x = numpy.stack([(ix + 0.5) * d, (iy + 0.5) * d, (iz + 0.5) * d], axis=1) + jit
    s.vel[:, :3] = rng.normal(0.0, v0 / math.sqrt(3.0), (n, 3))
    s.defaultParams(mu_s=mu, mu_d=mu, rho=rho, k_n=k_n, k_t=k_n)
    _set_contact_params(s, k_n, damping_frac, mu, r.min())
    s.g[:] = [0.0, 0.0, -9.81]
    s.normalBoundariesXY()
    s.defineWorldBoundaries(L=[nxy * d, nxy * d, nz * d * 1.05])
    dt = dt_rule(r.min(), rho, k_n)
    s.time_dt[0] = dt
    s.time_current[0] = 0.0
    s.time_total[0] = dt * 1.0e6
    s.time_file_dt[0] = dt * 1.0e6
    return s


def collapse_column(n=2000, seed=3001, r_range=(0.008, 0.012), footprint=0.3, k_n=1.0e7,
                    damping_frac=0.2, mu=0.5, rho=2600.0, v_std=0.01, open_stepper=None,
                    settle_steps=100000, settle_damping=1.0, spread=3.0):
    """A granular column released onto a frictional floor: n spheres with
    radii uniform in `r_range` on a jittered cubic lattice (spacing 2.02
    r_max, no overlaps) in a footprint x footprint box with frictional side
    walls, settled under gravity for `settle_steps` steps with
    `open_stepper` (at `settle_damping` times critical damping) and brought
    to rest, then placed at the centre of a floor `spread` times wider (the
    side walls move out: the column is released).  Initial velocities
    N(0, v_std) per component, seeded.  The contact parameters of the result
    are k_n = k_t, gamma_n = gamma_t = damping_frac times the critical
    damping of the smallest pair, mu_s = mu_d = mu.  (2,000 spheres of these
    radii fill a 0.3 m footprint to about 0.3 m once settled.)"""
    rng = numpy.random.default_rng(seed)
    r = rng.uniform(r_range[0], r_range[1], n)
    d = 2.02 * r.max()
    m = max(2, int(footprint // d))
    nz = int(math.ceil(n / float(m * m)))
    i = numpy.arange(n)
    ix, iy, iz = i % m, (i // m) % m, i // (m * m)
    x0 = 0.5 * (footprint - m * d)
    jit = (d / 2.0 - r.max()) * rng.uniform(-0.9, 0.9, (n, 3))
    s = sim(np=n, sid="column-%d" % n)
    s.radius = r
    s.x = numpy.stack([x0 + (ix + 0.5) * d, x0 + (iy + 0.5) * d, (iz + 0.5) * d], axis=1) + jit
    s.defaultParams(mu_s=mu, mu_d=mu, rho=rho, k_n=k_n, k_t=k_n)
    s.g[:] = [0.0, 0.0, -9.81]
    s.normalBoundariesXY()
    height = (nz + 1) * d
    s.defineWorldBoundaries(L=[footprint, footprint, height])
    dt = dt_rule(float(r.min()), rho, k_n)
    s.time_dt[0] = dt
    s.time_current[0] = 0.0
    s.time_total[0] = 1.0e3
    s.time_file_dt[0] = 1.0e3
    if open_stepper is not None and settle_steps > 0:
        _set_contact_params(s, k_n, settle_damping, mu, float(r.min()))
        st = open_stepper(s.tobytes())
        _relax(st, settle_steps, 0)
        _quench(st)
        s.frombytes(st.tobytes())
    _set_contact_params(s, k_n, damping_frac, mu, float(r.min()))
    wide = spread * footprint
    s.x[:, 0] += 0.5 * (wide - footprint)
    s.x[:, 1] += 0.5 * (wide - footprint)
    s.defineWorldBoundaries(L=[wide, wide, height])
    s.time_current[0] = 0.0
    s.vel[:, :3] = rng.normal(0.0, v_std, (n, 3))
    s.angvel[:] = 0.0
    return s


def batch_sizes(nsims=256, lo=500, hi=20000, seed=7):
    """Log-uniform particle counts in [lo, hi] for a batch of simulations."""
    rng = numpy.random.default_rng(seed)
    return numpy.exp(rng.uniform(math.log(lo), math.log(hi), nsims)).astype(int)


def consolidation_batch(nsims, seed, open_stepper, cell_settle_steps=20000,
                        cell_consolidate_steps=5000, ncells=4):
    """The batched consolidation sweep: `nsims` beds of log-uniform size in
    [500, 20000] particles (batch_sizes), each a settled 500-particle cell
    (radii uniform in [0.005, 0.01] m, one of `ncells` seeds, periodic in x
    and y, under a stress-controlled top wall, not sheared) tiled in x and y
    to its size, with its own normal stress (uniform in [5, 160] kPa),
    stiffness (log-uniform in [1e6, 1e8] N/m, with its own critical damping
    and dt) and friction (uniform in [0.3, 0.7]) (batch_shear_cells)."""
    sizes = batch_sizes(nsims, seed=seed)
    bases = [settled_shear_cell(500, seed + k, open_stepper=open_stepper,
                                settle_steps=cell_settle_steps,
                                consolidate_steps=cell_consolidate_steps,
                                r_uniform=(0.005, 0.01), shear=False)
             for k in range(ncells)]
    return batch_shear_cells(sizes, bases, seed=seed, kn_range=(1.0e6, 1.0e8))


_PER_PARTICLE = ("x", "radius", "xyzsum", "vel", "fixvel", "force", "angpos", "angvel",
                 "torque", "es_dot", "es", "ev_dot", "ev", "p", "color")


def tile_xy(s, mx, my):
    """mx x my copies of a cell periodic in x and y, side by side: a cell
    of mx L_x by my L_y with the same bed (the periodic images of the
    original become real neighbours).  The grid keeps its cell size, and
    the walls' masses scale with the area, so that a stress-controlled wall
    moves as before.  Contact histories are not carried (the .bin holds
    none)."""
    if int(s.periodic[0]) != 1:
        raise ValueError("tile_xy needs a cell periodic in x and y")
    if int(s.nb0[0]):
        raise ValueError("tile_xy does not copy bonds")
    n = int(s.np[0])
    L0, L1 = float(s.L[0]), float(s.L[1])
    t = copy.deepcopy(s)
    shift = numpy.array([[i * L0, j * L1, 0.0] for j in range(my) for i in range(mx)])
    for name in _PER_PARTICLE:
        a = numpy.asarray(getattr(s, name))
        setattr(t, name, numpy.concatenate([a] * (mx * my), axis=0))
    t.x = t.x + numpy.repeat(shift, n, axis=0)
    t.np[0] = n * mx * my
    t.L[0], t.L[1] = L0 * mx, L1 * my
    t.num[0], t.num[1] = int(s.num[0]) * mx, int(s.num[1]) * my
    t.w_m = numpy.asarray(s.w_m, dtype=numpy.float64) * (mx * my)
    t.w_force = numpy.asarray(s.w_force, dtype=numpy.float64) * (mx * my)
    t.sid = "%s-x%dy%d" % (s.sid, mx, my)
    return t


def near_square(m):
    """The factor pair (a, b) of m with a >= b and a - b smallest."""
    b = int(math.isqrt(m))
    while m % b:
        b -= 1
    return m // b, b


def batch_shear_cells(sizes, bases, seed=7, mu_range=(0.3, 0.7),
                      stress_range=(5.0e3, 1.6e5), kn_range=None, damping_frac=0.5):
    """A batch of sheared beds for many-simulation runs: simulation k is
    base cell k % len(bases) (settled direct-shear cells, periodic in x and
    y) tiled to the nearest multiple of its size in `sizes`, with its own
    Coulomb friction and normal stress drawn uniformly from the ranges and,
    with `kn_range`, its own stiffness drawn log-uniformly (k_t = k_n,
    damping_frac of critical damping, and its own dt = 0.075
    sqrt(m_min / k_n))."""
    rng = numpy.random.default_rng(seed)
    out = []
    for k, n in enumerate(sizes):
        b = bases[k % len(bases)]
        nb = int(b.np[0])
        m = max(1, int(round(float(n) / nb)))
        mx, my = near_square(m)
        t = tile_xy(b, mx, my)
        mu = rng.uniform(*mu_range)
        if kn_range is not None:
            k_n = math.exp(rng.uniform(math.log(kn_range[0]), math.log(kn_range[1])))
            r_min = float(numpy.min(t.radius))
            _set_contact_params(t, k_n, damping_frac, mu, r_min)
            t.time_dt[0] = dt_rule(r_min, float(t.rho[0]), k_n)
        t.mu_s[0] = mu
        t.mu_d[0] = mu
        t.w_sigma0 = numpy.asarray(t.w_sigma0, dtype=numpy.float64).copy()
        t.w_sigma0[0] = rng.uniform(*stress_range)
        out.append(t)
    return out


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


def _quench(st):
    """Remove all kinetic energy, of the particles (keeping fixvel in vel.w)
    and of the dynamic walls."""
    v = st.get("vel").copy()
    v[:, :3] = 0.0
    st.set("vel", v)
    st.set("angvel", numpy.zeros((v.shape[0], 3)))
    w = st.get("walls_mvfd")
    if w.shape[0]:
        w = w.copy()
        w[:, 1] = 0.0
        st.set("walls_mvfd", w)


def _relax(st, steps, quench_every):
    done = 0
    while done < steps:
        k = min(quench_every, steps - done) if quench_every else steps - done
        if quench_every:
            _quench(st)
        st.step(k)
        done += k


def grow_pack(s, open_stepper, phi_rsa=0.30, stages=30, steps_per_stage=100, seed=0,
              quench_every=100, log=None):
    """Grow radii in place inside the box of `s` (top at L[2], nw == 0).

    Particles are inserted without overlap with radii r * s0,
    s0 = (phi_rsa / phi)^(1/3) for the fraction phi the true radii fill in
    the box, then grow back to r in `stages` equal increments, each followed
    by `steps_per_stage` DEM steps without gravity or friction; kinetic
    energy is removed every `quench_every` steps.  Meant for fractions well
    below jamming, where the increments leave no lasting overlap.
    `open_stepper(image)` returns an engine context or oracle."""
    rng = numpy.random.default_rng(seed)
    r = numpy.asarray(s.radius, dtype=numpy.float64).copy()
    vol = float(numpy.sum(4.0 / 3.0 * math.pi * r ** 3))
    box = float((s.L[0] - s.origo[0]) * (s.L[1] - s.origo[1]) * (s.L[2] - s.origo[2]))
    s0 = min(1.0, (phi_rsa * box / vol) ** (1.0 / 3.0))
    s.x = rsa_positions(r * s0, list(s.origo), list(s.L), rng,
                        periodic=(s.periodic[0] in (1, 2), s.periodic[0] == 1, False))
    g = s.g.copy()
    mu = (float(s.mu_s[0]), float(s.mu_d[0]), float(s.mu_ws[0]), float(s.mu_wd[0]))
    s.g[:] = 0.0
    s.mu_s[0] = s.mu_d[0] = s.mu_ws[0] = s.mu_wd[0] = 0.0
    s.radius = r * s0
    st = open_stepper(s.tobytes())
    for k in range(1, stages + 1):
        x = st.get("x").copy()
        x[:, 3] = r * (s0 + (1.0 - s0) * k / stages)
        st.set("x", x)
        _relax(st, steps_per_stage, quench_every)
        if log:
            log("grow stage %d/%d" % (k, stages))
    s.frombytes(st.tobytes())
    s.radius = r.copy()
    s.g[:] = g
    s.mu_s[0], s.mu_d[0], s.mu_ws[0], s.mu_wd[0] = mu
    s.vel[:] = 0.0
    s.angvel[:] = 0.0
    s.time_current[0] = 0.0
    return s


def _median_pressure(st):
    """Median of the per-particle pressure over all particles: it becomes
    non-zero only when most particles carry load (a jammed packing)."""
    p = st.get("p")
    return float(numpy.median(p)) if p.size else 0.0


def jam_pack(s, open_stepper, p_target, phi_rsa=0.30, d_scale=0.02, max_stages=400,
             steps_per_stage=100, quench_every=50, seed=0, log=None):
    """Jam a packing at a target particle pressure, then rescale to the
    true radii.

    Particles are inserted without overlap into the box of `s` with radii
    r * s0 (s0 from phi_rsa).  The radii then grow (without gravity and
    friction, kinetic energy quenched every `quench_every` steps) by
    factors of (1 + d_scale); whenever the median particle pressure p
    (the reference's per-particle p field) exceeds 2 p_target the last
    increment is undone and the increment halved, until p lies within
    [p_target / 2, 2 p_target].  Finally every length (positions, box, grid
    origin) is divided by the final radius scale, which restores the true
    radii without changing the contact geometry.  `open_stepper(image)`
    returns an engine context or oracle (step/get/set/tobytes).
    """
    rng = numpy.random.default_rng(seed)
    r = numpy.asarray(s.radius, dtype=numpy.float64).copy()
    vol = float(numpy.sum(4.0 / 3.0 * math.pi * r ** 3))
    box = float((s.L[0] - s.origo[0]) * (s.L[1] - s.origo[1]) * (s.L[2] - s.origo[2]))
    sc = min(1.0, (phi_rsa * box / vol) ** (1.0 / 3.0))
    s.x = rsa_positions(r * sc, list(s.origo), list(s.L), rng,
                        periodic=(s.periodic[0] in (1, 2), s.periodic[0] == 1, False))
    g = s.g.copy()
    mu = (float(s.mu_s[0]), float(s.mu_d[0]), float(s.mu_ws[0]), float(s.mu_wd[0]))
    s.g[:] = 0.0
    s.mu_s[0] = s.mu_d[0] = s.mu_ws[0] = s.mu_wd[0] = 0.0
    s.radius = r * sc
    st = open_stepper(s.tobytes())
    d = d_scale
    for stage in range(max_stages):
        x_prev = st.get("x").copy()
        trial = sc * (1.0 + d)
        x = x_prev.copy()
        x[:, 3] = r * trial
        st.set("x", x)
        _relax(st, steps_per_stage, quench_every)
        pm = _median_pressure(st)
        if log:
            log("jam stage %d scale %.6f step %.2e median p %.3e" % (stage, trial, d, pm))
        if pm > 2.0 * p_target and d > 1e-7:
            st.set("x", x_prev)
            _relax(st, steps_per_stage, quench_every)
            d *= 0.5
            continue
        sc = trial
        if pm >= 0.5 * p_target:
            break
        if pm > 0.0:
            d = min(d, 0.5 * d_scale)
    s.frombytes(st.tobytes())
    inv = 1.0 / sc
    s.x = s.x * inv
    s.radius = r.copy()
    s.L = s.L * inv
    s.origo = s.origo * inv
    s.g[:] = g
    s.mu_s[0], s.mu_d[0], s.mu_ws[0], s.mu_wd[0] = mu
    s.vel[:] = 0.0
    s.angvel[:] = 0.0
    s.time_current[0] = 0.0
    # Cells of at least 2.1 r_max after the rescale.
    for dd in range(3):
        s.num[dd] = max(3, int(numpy.floor((s.L[dd] - s.origo[dd]) / (2.1 * r.max()))))
    return sc


def contact_number(st, n):
    """Mean number of occupied contact slots per particle."""
    return float((st.get("contacts") != n).sum()) / max(n, 1)


def shear_cell(n, seed, median=0.01, sigma=0.2, lxy=None, k_n=1.16e7, damping_frac=0.5,
               mu=0.5, normal_stress=10.0e3, shear_vel=1.0e-3, rho=2600.0, open_stepper=None,
               steps_per_stage=100, consolidate_steps=2000, quench_every=50, relax_steps=1000,
               log=None):
    """The direct-shear workload: lognormal radii in a box periodic in x and
    y, jammed at the target normal stress (jam_pack), then consolidate
    (normal_stress on the top wall, wmode 1) under gravity, then shear() with
    the top layer moving at `shear_vel` (python/sphere.py:3019-3048,
    3097-3165).  `open_stepper` runs the DEM relaxations: consolidation runs
    `consolidate_steps` with kinetic-energy quenching, then `relax_steps`
    without."""
    rng = numpy.random.default_rng(seed)
    r = lognormal_radii(n, median, sigma, rng)
    vol = float(numpy.sum(4.0 / 3.0 * math.pi * r ** 3))
    phi0 = 0.58
    if lxy is None:
        lxy = max(6 * 2.1 * r.max(), (vol / phi0 / 1.5) ** (1.0 / 3.0))
    h = vol / phi0 / (lxy * lxy)
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
    jam_pack(s, open_stepper, p_target=normal_stress, steps_per_stage=steps_per_stage,
             quench_every=quench_every, seed=seed + 1, log=log)
    s.consolidate(normal_stress)
    st = open_stepper(s.tobytes())
    _relax(st, consolidate_steps, quench_every)
    _relax(st, relax_steps, 0)
    if log:
        log("consolidated: mean contacts per particle %.2f" % contact_number(st, n))
    s.frombytes(st.tobytes())
    s.time_current[0] = 0.0
    z_min = float(numpy.min(s.x[:, 2] - s.radius))
    z_max = float(numpy.max(s.x[:, 2] + s.radius))
    s.shear(shear_strain_rate=shear_vel / (z_max - z_min))
    return s


def settled_shear_cell(n, seed, median=0.01, sigma=0.2, lxy=None, depth=0.35, phi_rsa=0.33,
                       k_n=1.16e7, damping_frac=0.5, mu=0.5, normal_stress=10.0e3,
                       shear_vel=1.0e-3, rho=2600.0, open_stepper=None, gravity_factor=30.0,
                       settle_steps=40000, ramp_stages=8, ramp_steps=2000, quench_every=500,
                       consolidate_steps=10000, free_steps=2000, chunk=2000,
                       grow_to_phi=0.5, grow_stages=30, grow_steps=100, log=None,
                       r_uniform=None, shear=True):
    """The direct-shear workload, prepacked by seeded gravity settling.

    Lognormal radii (median, sigma) are inserted without overlap at fraction
    `phi_rsa` into a box periodic in x and y (side `lxy`; by default the side
    that gives a settled bed about `depth` deep, 1.6 m for 1e5 particles);
    the bed settles under
    `gravity_factor` x 9.81 m/s^2 for `settle_steps` (an accelerated
    pluviation: the fall time shrinks by sqrt(gravity_factor)), then gravity
    returns to 9.81 m/s^2, consolidate(normal_stress) puts the stress-
    controlled top wall on the bed (python/sphere.py:3019-3048) and it
    relaxes for `consolidate_steps` (gravity is lowered in `ramp_stages`
    geometric stages of `ramp_steps` each, with the kinetic energy removed
    every `quench_every` steps, before the wall is placed); finally shear()
    fixes the bottom layer
    and drives the top layer at `shear_vel` (python/sphere.py:3097-3165).
    With `grow_to_phi`, the particles are first inserted at `phi_rsa` into a
    box of that fraction and grown to their true radii (grow_pack) before
    the box top is lifted and the bed settles.
    `open_stepper(image)` runs the DEM steps (GPU context or oracle).
    With `r_uniform` = (lo, hi) the radii are uniform in [lo, hi] instead;
    with shear=False the bed is left consolidated under the wall (no
    sheared layers)."""
    rng = numpy.random.default_rng(seed)
    if r_uniform is not None:
        r = rng.uniform(r_uniform[0], r_uniform[1], n)
    else:
        r = lognormal_radii(n, median, sigma, rng)
    vol = float(numpy.sum(4.0 / 3.0 * math.pi * r ** 3))
    if lxy is None:
        # A wide, shallow cell: `depth` of settled bed at a fraction near 0.56.
        lxy = max(6 * 2.1 * r.max(), math.sqrt(vol / (0.56 * depth)))
    h = vol / phi_rsa / (lxy * lxy)
    s = sim(np=n, sid="shear-%d" % n)
    s.radius = r
    s.defaultParams(mu_s=mu, mu_d=mu, rho=rho, k_n=k_n, k_t=k_n)
    _set_contact_params(s, k_n, damping_frac, mu, r.min())
    s.periodicBoundariesXY()
    s.defineWorldBoundaries(L=[lxy, lxy, h])
    dt = dt_rule(r.min(), rho, k_n)
    s.time_dt[0] = dt
    s.time_current[0] = 0.0
    s.time_total[0] = 1.0e3
    s.time_file_dt[0] = 1.0e3
    if grow_to_phi:
        # Shrink the box to `grow_to_phi` and grow the radii into it: the bed
        # starts dense but unjammed, so the fall that follows is short.
        s.L[2] = vol / grow_to_phi / (lxy * lxy)
        grow_pack(s, open_stepper, phi_rsa=phi_rsa, stages=grow_stages,
                  steps_per_stage=grow_steps, seed=seed + 1, quench_every=grow_steps)
        s.L[2] = h
    else:
        s.x = rsa_positions(r, [0.0, 0.0, 0.0], [lxy, lxy, h], rng,
                            periodic=(True, True, False))
    s.g[:] = [0.0, 0.0, -9.81 * gravity_factor]
    st = open_stepper(s.tobytes())
    done = 0
    while done < settle_steps:
        k = min(chunk, settle_steps - done)
        st.step(k)
        done += k
        if log:
            log("settle %d/%d steps, contacts per particle %.2f" % (done, settle_steps,
                                                                   contact_number(st, n)))
    # Lower gravity back to 9.81 m/s^2 in geometric stages, removing the
    # kinetic energy the bed gains as its elastic compression relaxes.
    if gravity_factor > 1.0:
        for gk in numpy.geomspace(gravity_factor, 1.0, ramp_stages + 1)[1:]:
            s.frombytes(st.tobytes())
            s.g[:] = [0.0, 0.0, -9.81 * gk]
            s.vel[:] = 0.0
            s.angvel[:] = 0.0
            s.time_current[0] = 0.0
            st = open_stepper(s.tobytes())
            _relax(st, ramp_steps, quench_every)
            if log:
                log("gravity x%.2f: contacts per particle %.2f" % (gk, contact_number(st, n)))
    s.frombytes(st.tobytes())
    s.g[:] = [0.0, 0.0, -9.81]
    s.time_current[0] = 0.0
    s.consolidate(normal_stress)
    st = open_stepper(s.tobytes())
    done = 0
    while done < consolidate_steps:
        k = min(chunk, consolidate_steps - done)
        # The load is applied at once; quenching damps the wall-bed
        # oscillation, except over the last `free_steps`.
        q = quench_every if done + k <= consolidate_steps - free_steps else 0
        _relax(st, k, q)
        done += k
        if log:
            log("consolidate %d/%d steps, contacts per particle %.2f" % (
                done, consolidate_steps, contact_number(st, n)))
    s.frombytes(st.tobytes())
    s.time_current[0] = 0.0
    if not shear:
        return s
    z_min = float(numpy.min(s.x[:, 2] - s.radius))
    z_max = float(numpy.max(s.x[:, 2] + s.radius))
    s.shear(shear_strain_rate=shear_vel / (z_max - z_min))
    return s
