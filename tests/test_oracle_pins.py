"""Pins of the CPU oracle against closed forms, brute force and invariants.

Each test checks the oracle against something it does not share code with:
numpy restatements of a closed-form rule (hash, sort order, cell ranges,
FP32 length, reduction tree), an O(n^2) brute-force contact search, the
analytic two-body collision, restitution and duration of the linear and
Hertzian springs, the 5/7 rolling law, the TY3 constant-acceleration
trajectory, and the reference's own invariant tests
(tests/contactmodel.py, tests/contactmodel_wall.py; tolerances 1e-3 / 5%).
"""

import math

import numpy
import pytest

import oracle
from sphere_amd import scenes
from sphere_amd.sim import sim

NC = 32


# ---------------------------------------------------------------------------
# Arithmetic rules

def test_sqrt32_is_fp32_root_of_fp64_value():
    """src/vector_arithmetic.h:317-324: sqrtf receives a double (narrowed)."""
    rng = numpy.random.default_rng(1)
    v = numpy.concatenate([rng.uniform(0, 1e-3, 2000), rng.uniform(0, 10, 2000),
                           [2.0000000001, 1e-300, 0.0, 4.0, 1e-8]])
    want = numpy.sqrt(v.astype(numpy.float32)).astype(numpy.float64)
    got = numpy.array([oracle.sqrt32(x) for x in v])
    assert numpy.array_equal(got, want)
    # The P1 example: differs from the FP64 root at 1.7e-8 relative.
    x = 2.0000000001
    assert abs(oracle.sqrt32(x) - math.sqrt(x)) / math.sqrt(x) > 1e-8


def _tree_sum_numpy(v):
    n = v.size
    nb = (n + 255) // 256
    pad = numpy.zeros(nb * 256)
    pad[:n] = v
    parts = []
    for b in range(nb):
        c = 0.0 + pad[256 * b:256 * (b + 1)]
        i = 128
        while i:
            c[:i] = c[:i] + c[i:2 * i]
            i //= 2
        parts.append(c[0])
    s = 0.0
    for p in parts:
        s += p
    return s


def test_tree_sum_follows_block_tree_then_serial():
    """summation + integrateWalls (src/integration.cuh:309-341,375-378)."""
    rng = numpy.random.default_rng(2)
    for n in (1, 255, 256, 257, 1000, 4097):
        v = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
        assert oracle.tree_sum(v) == _tree_sum_numpy(v)
    # An input where the order matters: the tree pairs 1e16 with -1e16
    # before adding the ones, a serial sum loses them.
    v = numpy.zeros(512)
    v[0], v[128] = 1e16, -1e16
    v[1:128] = 1.0
    assert oracle.tree_sum(v) == _tree_sum_numpy(v) == 127.0
    serial = 0.0
    for x in v:
        serial += x
    assert serial != 127.0


# ---------------------------------------------------------------------------
# Hash, sort, cells

def _random_box(n, seed, periodic=1, L=(0.3, 0.25, 0.4)):
    rng = numpy.random.default_rng(seed)
    s = sim(np=n)
    s.radius = rng.uniform(0.004, 0.01, n)
    s.x = numpy.column_stack([rng.uniform(0, L[d], n) for d in range(3)])
    s.periodic[0] = periodic
    s.defineWorldBoundaries(L=list(L))
    s.time_dt[0] = 1e-6
    s.time_total[0] = 1.0
    return s


def test_hash_matches_floor_formula():
    """calcCellID: floor((x - origo) / (L / num)), z-major linear index."""
    s = _random_box(3000, 3)
    o = oracle.Oracle(s.tobytes())
    o.stage("hash")
    key = o.get("cellid")
    L, num, org = s.L, s.num.astype(numpy.int64), s.origo
    ijk = [numpy.floor((s.x[:, d] - org[d]) / (L[d] / num[d])).astype(numpy.int64)
           for d in range(3)]
    want = (ijk[2] * num[1]) * num[0] + ijk[1] * num[0] + ijk[0]
    assert numpy.array_equal(key.astype(numpy.int64), want)
    assert numpy.array_equal(o.get("perm"), numpy.arange(3000, dtype=numpy.uint32))


def test_hash_at_upper_boundary_p17():
    """A particle at exactly x == L (kept by the `> L` wrap) is neither
    clamped into the last cell of its row nor wrapped to the first: its key
    is that of the first cell of the next row (ix == num)."""
    s = _random_box(10, 4)
    s.x[3, 0] = s.L[0]
    o = oracle.Oracle(s.tobytes())
    o.stage("hash")
    key = int(o.get("cellid")[3])
    nx, ny = int(s.num[0]), int(s.num[1])
    iy = int(s.x[3, 1] // (s.L[1] / ny))
    iz = int(s.x[3, 2] // (s.L[2] / int(s.num[2])))
    row_first = (iz * ny + iy) * nx
    assert key == row_first + nx
    assert key != row_first + nx - 1  # not clamped


def test_sort_is_stable_lexicographic():
    """thrust::sort_by_key on (cellID, identity) == lexsort by (key, index)."""
    s = _random_box(5000, 5)
    o = oracle.Oracle(s.tobytes())
    o.stage("hash")
    key = o.get("cellid").copy()
    o.stage("sort")
    perm = o.get("perm")
    want = numpy.lexsort((numpy.arange(5000), key))
    assert numpy.array_equal(perm, want.astype(numpy.uint32))
    assert numpy.array_equal(o.get("cellid"), key[want])


def test_cell_ranges():
    """cellStart / cellEnd from the sorted keys (src/sorting.cuh:95-118)."""
    s = _random_box(4000, 6)
    o = oracle.Oracle(s.tobytes())
    for st in ("hash", "sort", "reorder"):
        o.stage(st)
    skey = o.get("cellid").astype(numpy.int64)
    cs, ce = o.get("cellstart"), o.get("cellend")
    ncells = int(numpy.prod(s.num))
    for c in range(ncells):
        idx = numpy.nonzero(skey == c)[0]
        if idx.size == 0:
            assert cs[c] == 0xFFFFFFFF
        else:
            assert cs[c] == idx[0] and ce[c] == idx[-1] + 1


# ---------------------------------------------------------------------------
# Contact search

def _brute_force_contacts(s):
    """All overlapping pairs by the FP32 length rule, periodic images in x/y."""
    x, r = s.x, s.radius
    n = x.shape[0]
    L = s.L
    per = int(s.periodic[0])
    out = [set() for _ in range(n)]
    for a in range(n):
        d = x[a] - x
        if per in (1, 2):
            d[:, 0] -= L[0] * numpy.round(d[:, 0] / L[0])
        if per == 1:
            d[:, 1] -= L[1] * numpy.round(d[:, 1] / L[1])
        dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        ln = numpy.sqrt(dd.astype(numpy.float32)).astype(numpy.float64)
        hit = numpy.nonzero(ln - (r[a] + r) < 0.0)[0]
        out[a] = set(int(b) for b in hit if b != a)
    return out


@pytest.mark.parametrize("periodic", [0, 1, 2])
def test_topology_matches_brute_force(periodic):
    """Every overlapping pair is registered in both rows, nothing else."""
    rng = numpy.random.default_rng(7 + periodic)
    n = 600
    s = sim(np=n)
    s.radius = rng.uniform(0.008, 0.012, n)
    L = [0.2, 0.2, 0.2]
    s.x = numpy.column_stack([rng.uniform(0.012, L[d] - 0.012, n) for d in range(3)])
    if periodic:
        s.x[:, 0] = rng.uniform(0, L[0], n)
    if periodic == 1:
        s.x[:, 1] = rng.uniform(0, L[1], n)
    s.periodic[0] = periodic
    s.defineWorldBoundaries(L=L)
    s.time_dt[0] = 1e-7
    s.time_total[0] = 1.0
    o = oracle.Oracle(s.tobytes())
    for st in ("hash", "sort", "reorder", "topology"):
        o.stage(st)
    rows = o.get("contacts")
    want = _brute_force_contacts(s)
    assert sum(len(w) for w in want) > n  # dense enough to matter
    for a in range(n):
        got = set(int(b) for b in rows[a] if b != n)
        if len(want[a]) <= NC:
            assert got == want[a], a
    # Periodic shifts: x_ab + distmod is the minimum image.
    dm = o.get("distmod")
    for a in range(0, n, 7):
        for k in range(NC):
            b = int(rows[a, k])
            if b == n:
                continue
            d = s.x[a] - s.x[b] + dm[a, k, :3]
            assert numpy.linalg.norm(d) < s.radius[a] + s.radius[b] + 1e-9
            assert dm[a, k, 3] == s.radius[b]


def _slot_policy(row, n, finds):
    """findContactsInCell's slot choice (src/contactsearch.cuh:267-299) for a
    sequence of overlapping partners `finds` [(cell, partner)], written
    independently of the oracle."""
    row = list(row)
    nc = 0
    done_cells = set()
    for cell, b in finds:
        if cell in done_cells:
            continue
        cpos = nc
        for i in range(NC):
            if row[i] == n:
                cpos = i
            elif row[i] == b:
                cpos = i
                break
        if cpos < NC:
            row[cpos] = b
        nc += 1
        if nc > NC:
            done_cells.add(cell)
    return row


def test_slot_policy_with_history_and_overflow():
    """A large sphere among many small ones: > NC overlaps, pre-filled rows
    (live contacts, a reusable empty slot, stale partners)."""
    rng = numpy.random.default_rng(11)
    n_small = 60
    n = n_small + 1
    s = sim(np=n)
    s.radius[:] = 0.003
    s.radius[0] = 0.02
    c = numpy.array([0.1, 0.1, 0.1])
    s.x[0] = c
    v = rng.standard_normal((n_small, 3))
    v /= numpy.linalg.norm(v, axis=1)[:, None]
    s.x[1:] = c + v * (0.02 + 0.003 - 0.0005)
    s.periodic[0] = 0
    s.defineWorldBoundaries(L=[0.2, 0.2, 0.2], dx=0.0421)
    s.time_dt[0] = 1e-7
    s.time_total[0] = 1.0
    o = oracle.Oracle(s.tobytes())
    rows = o.get("contacts").copy()
    pre = rows[0].copy()
    pre[:] = n
    pre[0:10] = numpy.arange(1, 11)     # live contacts found again
    pre[10] = 55                        # live, found again later
    pre[12] = 40                        # stale position for a real partner
    pre[31] = n                         # empty slots remain at 11, 13..31
    rows[0] = pre
    o.set("contacts", rows)
    for st in ("hash", "sort", "reorder"):
        o.stage(st)
    # Candidate order of particle 0: 27 cells z-major, ascending sorted index.
    perm = o.get("perm")
    skey = o.get("cellid").astype(numpy.int64)
    num = s.num.astype(numpy.int64)
    g = numpy.floor(s.x[0] / (s.L / num)).astype(numpy.int64)
    finds = []
    ia = int(numpy.nonzero(perm == 0)[0][0])
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                t = g + numpy.array([dx, dy, dz])
                if numpy.any(t < 0) or numpy.any(t >= num):
                    continue
                cell = (t[2] * num[1]) * num[0] + t[1] * num[0] + t[0]
                for ib in numpy.nonzero(skey == cell)[0]:
                    if ib == ia:
                        continue
                    b = int(perm[ib])
                    d = s.x[0] - s.x[b]
                    ln = float(numpy.sqrt(numpy.float32((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])))
                    if ln - (s.radius[0] + s.radius[b]) < 0.0:
                        finds.append((int(cell), b))
    assert len(finds) > NC
    o.stage("topology")
    got = o.get("contacts")[0]
    assert list(got) == _slot_policy(pre, n, finds)


# ---------------------------------------------------------------------------
# Two-body and wall closed forms (tests/contactmodel.py:16-222 set-ups)

def _pair(r=(1.0, 1.0), z=(2.0, 4.05), v0=1.0, model=2, gamma_n=0.0, k_n=1.16e9, mu=0.5):
    s = sim(np=2, sid="pair")
    s.radius[:] = r
    s.x[0, :] = [5.0, 5.0, z[0]]
    s.x[1, :] = [5.0, 5.0, z[1]]
    s.vel[0, 2] = v0
    s.contactmodel[0] = model
    s.k_n[0] = s.k_t[0] = k_n
    s.gamma_n[0] = gamma_n
    s.mu_s[0] = s.mu_d[0] = mu
    s.defineWorldBoundaries(L=[10, 10, 10])
    s.initTemporal(total=0.1, file_dt=0.01)
    return s


def _kin(s):
    m = s.rho[0] * 4.0 / 3.0 * math.pi * s.radius ** 3
    return m, 0.5 * m * numpy.sum(s.vel ** 2, axis=1)


@pytest.mark.parametrize("model", [1, 2])
@pytest.mark.parametrize("r", [(1.0, 1.0), (2.0, 1.0), (1.0, 2.0)])
def test_elastic_collision_conserves_momentum_and_energy(model, r):
    s = _pair(r=r, z=(2.0, 2.0 + r[0] + r[1] + 0.05), model=model)
    m, ke0 = _kin(s)
    p0 = numpy.sum(m * s.vel[:, 2])
    o = oracle.Oracle(s.tobytes())
    o.step(int(0.1 / s.time_dt[0]) + 1)
    a = sim().frombytes(o.tobytes())
    _, ke1 = _kin(a)
    assert abs(ke1.sum() - ke0.sum()) / ke0.sum() < 1e-3
    assert abs(numpy.sum(m * a.vel[:, 2]) - p0) / abs(p0) < 1e-6
    # Closed form of the 1-D elastic collision.
    u = (m[0] - m[1]) / (m[0] + m[1])
    w = 2 * m[0] / (m[0] + m[1])
    assert abs(a.vel[0, 2] - u) < 1e-3 and abs(a.vel[1, 2] - w) < 1e-3


def test_linear_spring_contact_duration():
    """Half period of the linear spring: t_c = pi sqrt(m_eff / k_n)."""
    s = _pair(z=(2.0, 4.0 - 1e-9), k_n=1.16e7)
    m, _ = _kin(s)
    meff = m[0] * m[1] / (m[0] + m[1])
    tc = math.pi * math.sqrt(meff / 1.16e7)
    s.time_dt[0] = tc / 4000.0
    o = oracle.Oracle(s.tobytes())
    steps, touching = 0, 0
    while steps < 12000:
        o.step(1)
        steps += 1
        if (o.get("contacts")[0] != 2).any():
            touching += 1
        elif touching:
            break
    assert abs(touching * s.time_dt[0] - tc) / tc < 2e-3


def _clamped_dashpot_restitution(zeta):
    """Restitution of m d'' + gamma d' + k d = 0 when the contact force is
    clamped at zero (f_n = max(0, k d + gamma d'), src/contactmodels.cuh:333):
    the force vanishes at t* where w0^2 d + 2 zeta w0 d' = 0, after which the
    spheres separate at the velocity d'(t*)."""
    w0 = 1.0
    wd = w0 * math.sqrt(1 - zeta * zeta)

    def d(t):
        return math.exp(-zeta * w0 * t) * math.sin(wd * t) / wd

    def dv(t):
        return math.exp(-zeta * w0 * t) * (math.cos(wd * t) - zeta * w0 / wd * math.sin(wd * t))

    def f(t):
        return w0 * w0 * d(t) + 2 * zeta * w0 * dv(t)

    lo, hi = 0.5 * math.pi / wd, math.pi / wd
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            lo = mid
        else:
            hi = mid
    return -dv(0.5 * (lo + hi))


def test_viscous_restitution_closed_form():
    """Linear spring-dashpot restitution with the reference's fmax(0, .)
    clamp, against the closed-form solution of the damped oscillator."""
    k = 1.16e7
    s = _pair(z=(2.0, 4.0 - 1e-9), k_n=k)
    m, _ = _kin(s)
    meff = m[0] * m[1] / (m[0] + m[1])
    b = 0.2
    s.gamma_n[0] = 2 * b * math.sqrt(meff * k)
    s.time_dt[0] = math.pi * math.sqrt(meff / k) / 2000.0
    o = oracle.Oracle(s.tobytes())
    o.step(6000)
    a = sim().frombytes(o.tobytes())
    e = (a.vel[1, 2] - a.vel[0, 2]) / 1.0
    e_theory = _clamped_dashpot_restitution(b)
    assert abs(e - e_theory) / e_theory < 2e-3
    # Energy balance incl. the viscous dissipation ev (tests/contactmodel.py:90).
    _, ke0 = _kin(s)
    _, ke1 = _kin(a)
    assert abs((ke1.sum() + a.ev.sum()) - ke0.sum()) / ke0.sum() < 0.05


def test_hertz_duration_scaling():
    """Hertz F = k delta^(3/2): t_c = 3.2145 (m_eff^2 / k^2 v)^(1/5) for
    impact speed v; elastic (restitution 1)."""
    k = 1.0e9
    res = []
    for v0 in (1.0, 4.0):
        s = _pair(z=(2.0, 4.0 - 1e-12), model=3, k_n=k, v0=v0)
        m, _ = _kin(s)
        meff = m[0] * m[1] / (m[0] + m[1])
        tc = 3.2145 * (meff ** 2 / (k ** 2 * v0)) ** 0.2
        s.time_dt[0] = tc / 3000.0
        o = oracle.Oracle(s.tobytes())
        steps, touching = 0, 0
        while steps < 9000:
            o.step(1)
            steps += 1
            if (o.get("contacts")[0] != 2).any():
                touching += 1
            elif touching:
                break
        a = sim().frombytes(o.tobytes())
        res.append((touching * s.time_dt[0], tc, a.vel[1, 2] - a.vel[0, 2], v0))
    for t_meas, tc, vrel, v0 in res:
        assert abs(t_meas - tc) / tc < 0.01
        assert abs(vrel - v0) / v0 < 1e-3


def test_wall_elastic_rebound():
    """tests/contactmodel_wall.py: a sphere bounces off the bottom wall with
    v -> -v (gamma_wn = 0)."""
    s = sim(np=1, sid="wall")
    s.radius[:] = 1.0
    s.x[0, :] = [5.0, 5.0, 1.05]
    s.vel[0, 2] = -1.0
    s.gamma_wn[0] = 0.0
    s.k_n[0] = 1.16e9
    s.defineWorldBoundaries(L=[10, 10, 10])
    s.initTemporal(total=0.1, file_dt=0.01)
    o = oracle.Oracle(s.tobytes())
    o.step(int(0.1 / s.time_dt[0]) + 1)
    a = sim().frombytes(o.tobytes())
    assert abs(a.vel[0, 2] - 1.0) < 1e-3


def test_sliding_sphere_rolls_at_five_sevenths():
    """A sphere sliding on the frictional bottom wall under gravity ends up
    rolling at 5/7 of its initial speed (Coulomb sliding phase, then the
    tangential viscous term holds v_t ~ 0)."""
    s = sim(np=1, sid="roll")
    r = 0.01
    s.radius[:] = r
    m = 2600.0 * 4.0 / 3.0 * math.pi * r ** 3
    k = 1.16e7
    s.k_n[0] = s.k_t[0] = k
    s.g[:] = [0.0, 0.0, -9.81]
    s.gamma_wn[0] = 0.5 * 2 * math.sqrt(m * k)
    s.gamma_wt[0] = 1.0e4
    s.mu_ws[0] = s.mu_wd[0] = 0.3
    s.x[0, :] = [0.5, 0.5, r - m * 9.81 / k]
    s.vel[0, 0] = 0.5
    s.defineWorldBoundaries(L=[1.0, 1.0, 0.2])
    s.time_dt[0] = 0.05 * math.sqrt(m / k)
    s.time_total[0] = 1.0
    o = oracle.Oracle(s.tobytes())
    t_slip = 2.0 / 7.0 * 0.5 / (0.3 * 9.81)
    o.step(int(1.5 * t_slip / s.time_dt[0]))
    a = sim().frombytes(o.tobytes())
    assert abs(a.vel[0, 0] - 5.0 / 7.0 * 0.5) / (5.0 / 7.0 * 0.5) < 0.01
    # Rolling: angular speed matches v / r with the sign of the rolling sense.
    assert abs(a.angvel[0, 1] * r - a.vel[0, 0]) / a.vel[0, 0] < 0.02


# ---------------------------------------------------------------------------
# Integration and walls

def test_ty3_constant_acceleration_trajectory():
    """TY3 with acc0 = 0 on the first step (src/integration.cuh:74-76,226-263):
    v_N = v0 + g dt (N + 1/2); x follows by summing the exact update."""
    s = sim(np=1)
    s.radius[:] = 0.01
    s.x[0] = [0.5, 0.5, 50.0]
    s.vel[0] = [0.1, 0.0, 0.0]
    s.g[:] = [0.0, 0.0, -9.81]
    s.defineWorldBoundaries(L=[1.0, 1.0, 100.0])
    dt = 1e-3
    s.time_dt[0] = dt
    s.time_total[0] = 10.0
    o = oracle.Oracle(s.tobytes())
    N = 500
    o.step(N)
    v = o.get("vel")[0]
    x = o.get("x")[0]
    g = -9.81
    assert abs(v[2] - g * dt * (N + 0.5)) < 1e-12 * abs(g * dt * N)
    from fractions import Fraction as Fr
    G, D = Fr(g), Fr(dt)
    z = Fr(50.0) + Fr(2, 3) * G * D * D
    for k in range(1, N):
        z += (G * D * (k + Fr(1, 2))) * D + Fr(1, 2) * G * D * D
    assert abs(x[2] - float(z)) < 1e-11 * 50.0
    assert abs(x[0] - (0.5 + 0.1 * dt * N)) < 1e-12


@pytest.mark.parametrize("fixvel,expect", [(1.0, "vertical only, no gravity"),
                                           (-1.0, "fixed"), (12.0, "2-D")])
def test_fixvel_modes(fixvel, expect):
    """src/integration.cuh:131-160."""
    s = sim(np=1)
    s.radius[:] = 0.01
    s.x[0] = [0.5, 0.5, 0.5]
    s.vel[0] = [0.2, 0.3, 0.4]
    s.fixvel[0] = fixvel
    s.g[:] = [1.0, 2.0, -9.81]
    s.defineWorldBoundaries(L=[1.0, 1.0, 1.0])
    s.time_dt[0] = 1e-4
    s.time_total[0] = 1.0
    o = oracle.Oracle(s.tobytes())
    o.step(10)
    v = o.get("vel")[0]
    if fixvel == 1.0:
        assert v[0] == 0.2 and v[1] == 0.3 and abs(v[2] - 0.4) < 1e-15
    elif fixvel == -1.0:
        assert list(v[:3]) == [0.2, 0.3, 0.4]
    else:
        assert v[1] == 0.3 and v[0] != 0.2 and v[2] != 0.4


def test_periodic_wrap_rule():
    """x < origo: += L; x > L: -= L; x == L survives (P17)."""
    s = sim(np=2)
    s.radius[:] = 0.01
    s.x[0] = [0.999999, 0.5, 0.5]
    s.vel[0] = [1.0, 0.0, 0.0]
    s.x[1] = [0.3, 0.000001, 0.5]
    s.vel[1] = [0.0, -1.0, 0.0]
    s.periodic[0] = 1
    s.defineWorldBoundaries(L=[1.0, 1.0, 1.0])
    s.time_dt[0] = 1e-5
    s.time_total[0] = 1.0
    o = oracle.Oracle(s.tobytes())
    o.step(1)
    x = o.get("x")
    assert 0.0 <= x[0, 0] < 0.01 and 0.99 < x[1, 1] <= 1.0


def test_stress_controlled_wall_trajectory():
    """integrateWalls (wmode 1) with no particle touching: acceleration
    -sigma0 Lx Ly / m, TY3 with acc0 = 0 on the first step."""
    s = sim(np=1, nw=1)
    s.radius[:] = 0.01
    s.x[0] = [0.5, 0.5, 0.05]
    s.defineWorldBoundaries(L=[1.0, 1.0, 1.0])
    s.wmode[0] = 1
    s.w_x[0] = 0.9
    s.w_m[0] = 3.0
    s.w_sigma0[0] = 2.0
    dt = 1e-3
    s.time_dt[0] = dt
    s.time_total[0] = 10.0
    o = oracle.Oracle(s.tobytes())
    N = 100
    o.step(N)
    acc = -2.0 * 1.0 * 1.0 / 3.0
    w = o.get("walls_mvfd")[0]
    assert abs(w[1] - acc * dt * (N + 0.5)) < 1e-12
    assert o.get("walls_acc")[0] == acc


def test_bond_energy_conservation():
    """Two bonded spheres oscillating along their axis without damping: the
    bond potential (sim.energy('bondpot'), python/sphere.py:4360-4378) plus
    the kinetic energy stays constant (the disabled tests/bond_tests.py)."""
    s = sim(np=2)
    s.radius[:] = 0.01
    s.x[0] = [0.5, 0.5, 0.5]
    s.x[1] = [0.52, 0.5, 0.5]
    s.vel[0, 0] = -0.01
    s.vel[1, 0] = 0.01
    s.k_n[0] = s.k_t[0] = 1.0e8
    s.bond(0, 1)
    s.defineWorldBoundaries(L=[1.0, 1.0, 1.0])
    m = 2600.0 * 4.0 / 3.0 * math.pi * 0.01 ** 3
    A = math.pi * 0.01 ** 2
    period = 2 * math.pi * math.sqrt(0.5 * m / (1.0e8 * A))
    s.time_dt[0] = period / 2000.0
    s.time_total[0] = 1.0
    o = oracle.Oracle(s.tobytes())
    e0 = None
    for _ in range(8):
        o.step(250)
        a = sim().frombytes(o.tobytes())
        _, ke = _kin(a)
        e = ke.sum() + a.energy("bondpot")
        e0 = e if e0 is None else e0
        assert abs(e - e0) / e0 < 5e-3
    f = o.get("force")
    assert numpy.allclose(f[0], -f[1], rtol=0, atol=1e-12 * numpy.abs(f).max())


# ---------------------------------------------------------------------------
# Whole-step invariants

def test_step_is_deterministic_and_thread_independent():
    s = scenes.settling_box(n=800, seed=3, box=(0.12, 0.12, 0.3))
    img = s.tobytes()
    oracle.set_threads(1)
    a = oracle.Oracle(img)
    a.step(50)
    oracle.set_threads(4)
    b = oracle.Oracle(img)
    b.step(50)
    oracle.set_threads(0)
    assert a.tobytes() == b.tobytes()


def test_codec_round_trip_through_oracle():
    """Python writer -> oracle reader -> oracle writer is the identity."""
    s = scenes.shear_box(n=300, seed=5, phi=0.25)
    s.bond(0, 1)
    s.bond(5, 9)
    img = s.tobytes()
    assert oracle.Oracle(img).tobytes() == img


def _wall1_only_case():
    """Two resting spheres in a triaxial box of five dynamic walls; sphere 0
    presses on side wall 1 (normal -x) and touches nothing else."""
    from sphere_amd.sim import sim
    s = sim(np=2, sid="wall1")
    s.radius[:] = 0.01
    s.x[0, :] = [0.0897, 0.05, 0.05]
    s.x[1, :] = [0.05, 0.05, 0.05]
    s.defineWorldBoundaries(L=[0.1, 0.1, 0.1])
    s.initTemporal(total=1.0, file_dt=0.1)
    s.g[:] = 0.0
    s.triaxial(wvel=0.0, normal_stress=1.0e3)
    # the box walls back at the domain faces, wall 1 just inside sphere 0
    s.w_x[:] = [0.1, 0.0995, 0.0, 0.1, 0.0]
    s.w_m[:] = 1.0
    return s


def test_per_wall_sums_integrate_each_walls_own_force():
    """Reference behaviour (P8): every dynamic wall integrates wall 0's sum,
    here zero, so side wall 1 feels nothing.  With per-wall sums, wall 1's
    force is the sum of its own per-particle forces: one nonzero leaf, so
    exactly that particle's wall-1 force; the other walls stay at zero."""
    s = _wall1_only_case()
    img = s.tobytes()
    try:
        oracle.set_wall_sums_per_wall(False)
        ref = oracle.Oracle(img)
        ref.step(1)
        oracle.set_wall_sums_per_wall(True)
        fix = oracle.Oracle(img)
        fix.step(1)
    finally:
        oracle.set_wall_sums_per_wall(False)
    n = 2
    fpp = fix.get("walls_force_pp").reshape(5, n)
    assert fpp[1, 0] != 0.0 and fpp[1, 1] == 0.0
    assert (fpp[[0, 2, 3, 4]] == 0.0).all()
    mref = ref.get("walls_mvfd").reshape(-1, 4)
    mfix = fix.get("walls_mvfd").reshape(-1, 4)
    assert (mref[:, 2] == 0.0).all()
    assert mfix[1, 2] == fpp[1, 0]
    assert mfix[0, 2] == 0.0 and (mfix[2:, 2] == 0.0).all()
    # wall 1 now moves under its own force (it did not before)
    assert fix.get("walls_acc")[1] != ref.get("walls_acc")[1]


def _three_spheres(fix2=False):
    from sphere_amd.sim import sim
    s = sim(np=3)
    s.radius[:] = [0.01, 0.01, 0.02]
    s.x[0] = [0.05, 0.05, 0.05]
    s.x[1] = [0.068, 0.05, 0.05]
    s.x[2] = [0.05, 0.05, 0.079]
    if fix2:
        s.fixvel[2] = 1.0
    s.defineWorldBoundaries(L=[0.1, 0.1, 0.2])
    s.initTemporal(total=1.0, file_dt=0.1)
    return s


def test_find_overlaps_pairs_against_kd_tree_and_exclusions():
    """DEM::findOverlaps: the pair set equals an independent k-d tree search
    (pairs closer than r_i + r_j), the overlaps equal the distances less the
    radius sums, and pairs with a fixed particle are left out."""
    from scipy.spatial import cKDTree
    pairs, delta = oracle.Oracle(_three_spheres().tobytes()).find_overlaps()
    assert pairs.tolist() == [[0, 1], [0, 2]]
    assert delta == pytest.approx([-0.002, -0.001], rel=1e-12)
    pairs, _ = oracle.Oracle(_three_spheres(fix2=True).tobytes()).find_overlaps()
    assert pairs.tolist() == [[0, 1]]
    from . import cases
    s = cases.settling_case(300, 11)
    x = numpy.asarray(s.x)[:, :3]
    r = numpy.asarray(s.radius)
    s.fixvel[:] = 0.0
    pairs, delta = oracle.Oracle(s.tobytes()).find_overlaps()
    cand = cKDTree(x).query_pairs(2.0 * r.max(), output_type="ndarray")
    d = numpy.linalg.norm(x[cand[:, 1]] - x[cand[:, 0]], axis=1) - (r[cand[:, 0]] + r[cand[:, 1]])
    want = cand[d < 0]
    want = want[numpy.lexsort((want[:, 1], want[:, 0]))]
    assert numpy.array_equal(pairs, want)
    dd = d[d < 0][numpy.lexsort((cand[d < 0][:, 1], cand[d < 0][:, 0]))]
    assert delta == pytest.approx(dd, rel=1e-9, abs=1e-15)


def test_porosity_closed_forms_and_the_references_cap_rule():
    """DEM::porosity: a sphere inside a slice removes its volume; a sphere
    centred on a slice face removes a half (the cap of height r) from each
    side; a sphere whose centre lies inside a slice and crosses its lower
    face has the reference's rule V_sphere + cap(z_low - (z_c - r))
    subtracted (src/sphere.cpp:565-570), kept as it is."""
    from sphere_amd.sim import sim
    o = oracle.Oracle(_three_spheres().tobytes())
    z, por = o.porosity(4)
    V = lambda r: 4.0 / 3.0 * math.pi * r ** 3
    Vs = 0.05 * 0.1 * 0.1
    assert z.tolist() == pytest.approx([0.025, 0.075, 0.125, 0.175])
    # spheres 0 and 1 centred on the face z = 0.05: half each side
    assert por[0] == pytest.approx(1.0 - V(0.01) / Vs, rel=1e-13)
    assert por[1] == pytest.approx(1.0 - (V(0.01) + V(0.02)) / Vs, rel=1e-13)
    assert por[2] == 1.0 and por[3] == 1.0
    s = sim(np=1)
    s.radius[:] = 0.01
    s.x[0] = [0.05, 0.05, 0.055]
    s.defineWorldBoundaries(L=[0.1, 0.1, 0.2])
    s.initTemporal(total=1.0, file_dt=0.1)
    z, por = oracle.Oracle(s.tobytes()).porosity(4)
    cap = lambda h, r: math.pi * h * h / 3.0 * (3.0 * r - h)
    assert por[1] == pytest.approx(1.0 - (V(0.01) + cap(0.005, 0.01)) / Vs, rel=1e-13)
    assert por[0] == pytest.approx(1.0 - cap(0.005, 0.01) / Vs, rel=1e-13)
    with pytest.raises(ValueError):
        oracle.Oracle(s.tobytes()).porosity(11)  # slices of 0.018 m < 2 r


def test_tangent_oracle_values_are_the_oracle_s():
    """The dual-number build computes the same values bit for bit."""
    from . import cases
    s = cases.settling_case(300, 11)
    image = s.tobytes()
    a, b = oracle.Oracle(image), oracle.TangentOracle(image)
    b.seed_param("k_n", 1.0)
    a.step(40)
    b.step(40)
    for f in ("x", "vel", "angvel", "force", "torque", "contacts", "delta_t", "es", "ev", "p",
              "walls_force_pp"):
        assert numpy.array_equal(a.get(f), b.get(f)), f


def test_tangent_oracle_against_central_differences():
    """Directional derivatives of the dual-number build against central
    differences of the plain oracle, both with FP64 lengths, over 40 steps
    of a box where contacts form, to 1e-6 relative: three parameters and an
    initial-velocity direction.  (Steps of 1e-3 of a parameter: smaller
    steps are dominated by the rounding of the two runs' difference.)"""
    from . import cases
    s = cases.settling_case(300, 11)
    image = s.tobytes()
    c = numpy.random.default_rng(5).standard_normal((300, 3))
    d = numpy.random.default_rng(6).standard_normal((300, 3))
    lib_dual = oracle._lib(True)
    oracle.set_exact_length(True)
    lib_dual.orc_set_exact_length(1)
    try:
        for kind in ("k_n", "gamma_n", "mu_s", "vel"):
            t = oracle.TangentOracle(image)
            if kind == "vel":
                t.seed_field("vel", d)
            else:
                t.seed_param(kind, 1.0)
            t.step(40)
            tan = float(numpy.sum(c * t.tangent("x")))
            vals = []
            for sign in (1.0, -1.0):
                s2 = sim().frombytes(image)
                if kind == "vel":
                    h = 1e-4
                    s2.vel[:, :3] += sign * h * d
                else:
                    h = 1e-3 * float(getattr(s2, kind)[0])
                    getattr(s2, kind)[0] += sign * h
                o = oracle.Oracle(s2.tobytes())
                o.step(40)
                vals.append(float(numpy.sum(c * o.get("x")[:, :3])))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(tan - fd) <= 1e-6 * abs(fd), (kind, tan, fd)
    finally:
        oracle.set_exact_length(False)
        lib_dual.orc_set_exact_length(0)
