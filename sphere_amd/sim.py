"""The `sim` object: the reference's Python front end for a DEM run.

Attribute names, method names, argument meaning and error behaviour mirror
python/sphere.py (class sim, :33-7449), restricted to the fluid-free DEM
path.  What differs by design:

* run() does not shell out to a CUDA binary: it writes the same
  <root>/input/<sid>.bin and calls the HIP library in-process through its
  C ABI (libsphere_hip.so), which writes <root>/output/<sid>.output%05d.bin
  and <root>/output/<sid>.status.dat exactly where the reference's `sphere`
  executable writes them.  <root> is ".." (the reference's convention,
  python/sphere.py:4508,4533) unless SPHERE_ROOT or the `root` argument
  says otherwise.
* The random setup helpers accept an optional numpy Generator; without one
  they draw from numpy's global RNG as the reference does.
"""

import math
import os
import shutil

import numpy

from . import binfmt

VERSION = binfmt.VERSION


def _root(root=None):
    if root is not None:
        return root
    return os.environ.get("SPHERE_ROOT", "..")


def V_sphere(r):
    """Volume of a sphere of radius r (python/sphere.py:7760-7767)."""
    return 4.0 / 3.0 * math.pi * r ** 3.0


class sim:
    """A DEM simulation: particles, walls, bonds, parameters and time."""

    def __init__(self, sid="unnamed", np=0, nd=3, nw=0, fluid=False, cfd_solver=0):
        if fluid:
            raise NotImplementedError(
                "fluid coupling (Navier-Stokes / Darcy) is not part of this engine")
        # python/sphere.py:57-455
        self.version = numpy.ones(1, dtype=numpy.float64) * VERSION
        self.nd = numpy.ones(1, dtype=numpy.int32) * int(nd)
        self.np = numpy.ones(1, dtype=numpy.uint32) * int(np)
        self.sid = sid
        self.time_dt = numpy.zeros(1, dtype=numpy.float64)
        self.time_current = numpy.zeros(1, dtype=numpy.float64)
        self.time_total = numpy.zeros(1, dtype=numpy.float64)
        self.time_file_dt = numpy.zeros(1, dtype=numpy.float64)
        self.time_step_count = numpy.zeros(1, dtype=numpy.uint32)
        self.origo = numpy.zeros(self.nd[0], dtype=numpy.float64)
        self.L = numpy.zeros(self.nd[0], dtype=numpy.float64)
        self.num = numpy.zeros(self.nd[0], dtype=numpy.uint32)
        self.periodic = numpy.zeros(1, dtype=numpy.uint32)
        self.adaptive = numpy.zeros(1, dtype=numpy.uint32)
        n = int(self.np[0])
        self.x = numpy.zeros((n, self.nd[0]), dtype=numpy.float64)
        self.radius = numpy.ones(n, dtype=numpy.float64)
        self.xyzsum = numpy.zeros((n, 3), dtype=numpy.float64)
        self.vel = numpy.zeros((n, self.nd[0]), dtype=numpy.float64)
        self.fixvel = numpy.zeros(n, dtype=numpy.float64)
        self.force = numpy.zeros((n, self.nd[0]), dtype=numpy.float64)
        self.angpos = numpy.zeros((n, self.nd[0]), dtype=numpy.float64)
        self.angvel = numpy.zeros((n, self.nd[0]), dtype=numpy.float64)
        self.torque = numpy.zeros((n, self.nd[0]), dtype=numpy.float64)
        self.es_dot = numpy.zeros(n, dtype=numpy.float64)
        self.es = numpy.zeros(n, dtype=numpy.float64)
        self.ev_dot = numpy.zeros(n, dtype=numpy.float64)
        self.ev = numpy.zeros(n, dtype=numpy.float64)
        self.p = numpy.zeros(n, dtype=numpy.float64)
        self.g = numpy.array([0.0, 0.0, 0.0], dtype=numpy.float64)
        self.k_n = numpy.ones(1, dtype=numpy.float64) * 1.16e9
        self.k_t = numpy.ones(1, dtype=numpy.float64) * 1.16e9
        self.k_r = numpy.zeros(1, dtype=numpy.float64)
        self.E = numpy.zeros(1, dtype=numpy.float64)
        self.gamma_n = numpy.zeros(1, dtype=numpy.float64)
        self.gamma_t = numpy.zeros(1, dtype=numpy.float64)
        self.gamma_r = numpy.zeros(1, dtype=numpy.float64)
        self.mu_s = numpy.ones(1, dtype=numpy.float64) * 0.5
        self.mu_d = numpy.ones(1, dtype=numpy.float64) * 0.5
        self.mu_r = numpy.zeros(1, dtype=numpy.float64)
        self.gamma_wn = numpy.ones(1, dtype=numpy.float64) * 1.0e3
        self.gamma_wt = numpy.zeros(1, dtype=numpy.float64)
        self.mu_ws = numpy.ones(1, dtype=numpy.float64) * 0.5
        self.mu_wd = numpy.ones(1, dtype=numpy.float64) * 0.5
        self.rho = numpy.ones(1, dtype=numpy.float64) * 2600.0
        self.contactmodel = numpy.ones(1, dtype=numpy.uint32) * 2
        self.kappa = numpy.zeros(1, dtype=numpy.float64)
        self.db = numpy.zeros(1, dtype=numpy.float64)
        self.V_b = numpy.zeros(1, dtype=numpy.float64)
        self.nw = numpy.ones(1, dtype=numpy.uint32) * int(nw)
        nwi = int(self.nw[0])
        self.wmode = numpy.zeros(nwi, dtype=numpy.int32)
        self.w_n = numpy.zeros((nwi, self.nd[0]), dtype=numpy.float64)
        if nwi >= 1:
            self.w_n[0, 2] = -1.0
        if nwi >= 2:
            self.w_n[1, 0] = -1.0
        if nwi >= 3:
            self.w_n[2, 0] = 1.0
        if nwi >= 4:
            self.w_n[3, 1] = -1.0
        if nwi >= 5:
            self.w_n[4, 1] = 1.0
        self.w_x = numpy.ones(nwi, dtype=numpy.float64)
        self.w_m = numpy.zeros(nwi, dtype=numpy.float64)
        self.w_vel = numpy.zeros(nwi, dtype=numpy.float64)
        self.w_force = numpy.zeros(nwi, dtype=numpy.float64)
        self.w_sigma0 = numpy.zeros(nwi, dtype=numpy.float64)
        self.w_sigma0_A = numpy.zeros(1, dtype=numpy.float64)
        self.w_sigma0_f = numpy.zeros(1, dtype=numpy.float64)
        self.w_tau_x = numpy.zeros(1, dtype=numpy.float64)
        self.lambda_bar = numpy.ones(1, dtype=numpy.float64)
        self.nb0 = numpy.zeros(1, dtype=numpy.uint32)
        self.sigma_b = numpy.ones(1, dtype=numpy.float64) * numpy.inf
        self.tau_b = numpy.ones(1, dtype=numpy.float64) * numpy.inf
        self.bonds = numpy.zeros((0, 2), dtype=numpy.uint32)
        self.bonds_delta_n = numpy.zeros(0, dtype=numpy.float64)
        self.bonds_delta_t = numpy.zeros((0, self.nd[0]), dtype=numpy.float64)
        self.bonds_omega_n = numpy.zeros(0, dtype=numpy.float64)
        self.bonds_omega_t = numpy.zeros((0, self.nd[0]), dtype=numpy.float64)
        self.fluid = False
        self.color = numpy.zeros(n, dtype=numpy.int32)

    # -- identity ---------------------------------------------------------
    def id(self, sid=""):
        if sid == "":
            return self.sid
        self.sid = sid

    def idAppend(self, string):
        self.sid += string

    # -- comparison (python/sphere.py:456-793) ----------------------------
    _CMP_FIELDS = (
        "version", "nd", "np", "time_dt", "time_current", "time_total",
        "time_file_dt", "time_step_count", "origo", "L", "num", "periodic",
        "adaptive", "x", "radius", "xyzsum", "vel", "fixvel", "force", "angpos",
        "angvel", "torque", "es_dot", "es", "ev_dot", "ev", "p", "g", "k_n",
        "k_t", "k_r", "E", "gamma_n", "gamma_t", "gamma_r", "mu_s", "mu_d",
        "mu_r", "gamma_wn", "gamma_wt", "mu_ws", "mu_wd", "rho", "contactmodel",
        "kappa", "db", "V_b", "nw", "wmode", "w_n", "w_x", "w_m", "w_vel",
        "w_force", "w_sigma0", "w_sigma0_A", "w_sigma0_f", "w_tau_x",
        "lambda_bar", "nb0", "sigma_b", "tau_b", "bonds", "bonds_delta_n",
        "bonds_delta_t", "bonds_omega_n", "bonds_omega_t", "color")

    def compare(self, other, ignore=()):
        """Field-by-field exact comparison.  Returns the list of differing
        field names (empty when equal), like sim.__cmp__'s report."""
        diff = []
        for f in self._CMP_FIELDS:
            if f in ignore:
                continue
            a = numpy.asarray(getattr(self, f))
            b = numpy.asarray(getattr(other, f))
            if a.shape != b.shape and a.size == b.size:
                a = a.reshape(b.shape)
            if a.shape != b.shape or not numpy.array_equal(a, b):
                diff.append(f)
        return diff

    def __eq__(self, other):
        return isinstance(other, sim) and not self.compare(other)

    # -- particles ----------------------------------------------------------
    def addParticle(self, x, radius, xyzsum=numpy.zeros(3), vel=numpy.zeros(3),
                    fixvel=numpy.zeros(1), force=numpy.zeros(3), angpos=numpy.zeros(3),
                    angvel=numpy.zeros(3), torque=numpy.zeros(3), es_dot=numpy.zeros(1),
                    es=numpy.zeros(1), ev_dot=numpy.zeros(1), ev=numpy.zeros(1),
                    p=numpy.zeros(1), color=0):
        """python/sphere.py:821-890."""
        self.np = self.np + 1
        self.x = numpy.append(self.x, [x], axis=0)
        self.radius = numpy.append(self.radius, radius)
        self.vel = numpy.append(self.vel, [vel], axis=0)
        self.xyzsum = numpy.append(self.xyzsum, [xyzsum], axis=0)
        self.fixvel = numpy.append(self.fixvel, fixvel)
        self.force = numpy.append(self.force, [force], axis=0)
        self.angpos = numpy.append(self.angpos, [angpos], axis=0)
        self.angvel = numpy.append(self.angvel, [angvel], axis=0)
        self.torque = numpy.append(self.torque, [torque], axis=0)
        self.es_dot = numpy.append(self.es_dot, es_dot)
        self.es = numpy.append(self.es, es)
        self.ev_dot = numpy.append(self.ev_dot, ev_dot)
        self.ev = numpy.append(self.ev, ev)
        self.p = numpy.append(self.p, p)
        self.color = numpy.append(self.color, color)

    def deleteParticle(self, i):
        """python/sphere.py:892-938."""
        if type(i) == tuple:
            raise Exception("Cannot parse tuples as index value. "
                            "Valid types are int, list and numpy.ndarray")
        if isinstance(i, numpy.ndarray):
            self.np -= i.size
        elif isinstance(i, list):
            self.np -= len(i)
        else:
            self.np -= 1
        for f in ("x", "vel", "xyzsum", "force", "angpos", "angvel", "torque"):
            setattr(self, f, numpy.delete(getattr(self, f), i, axis=0))
        for f in ("radius", "fixvel", "es_dot", "es", "ev_dot", "ev", "p", "color"):
            setattr(self, f, numpy.delete(getattr(self, f), i))

    def deleteAllParticles(self):
        """python/sphere.py:940-963."""
        self.np[0] = 0
        nd = self.nd[0]
        self.x = numpy.zeros((0, nd))
        self.radius = numpy.ones(0)
        self.xyzsum = numpy.zeros((0, 3))
        self.vel = numpy.zeros((0, nd))
        self.fixvel = numpy.zeros(0)
        self.force = numpy.zeros((0, nd))
        self.angpos = numpy.zeros((0, nd))
        self.angvel = numpy.zeros((0, nd))
        self.torque = numpy.zeros((0, nd))
        self.es_dot = numpy.zeros(0)
        self.es = numpy.zeros(0)
        self.ev_dot = numpy.zeros(0)
        self.ev = numpy.zeros(0)
        self.p = numpy.zeros(0)
        self.color = numpy.zeros(0, dtype=numpy.int32)

    # -- binary I/O --------------------------------------------------------
    def readbin(self, targetbin, verbose=True, bonds=True, sigma0mod=True,
                esysparticle=False):
        """Read a .bin file (python/sphere.py:965-1353)."""
        if verbose:
            print("Input file: {0}".format(targetbin))
        with open(targetbin, "rb") as fh:
            data = fh.read()
        binfmt.decode(data, self, bonds=bonds, sigma0mod=sigma0mod)

    def frombytes(self, data):
        binfmt.decode(data, self)
        return self

    def tobytes(self):
        return binfmt.encode(self)

    def writebin(self, folder=None, verbose=True):
        """Write <folder>/<sid>.bin (python/sphere.py:1355-1553).  The default
        folder is <root>/input/."""
        if folder is None:
            folder = os.path.join(_root(), "input")
        os.makedirs(folder, exist_ok=True)
        fn = os.path.join(folder, self.sid + ".bin")
        if verbose:
            print("Output file: {0}".format(fn))
        with open(fn, "wb") as fh:
            fh.write(binfmt.encode(self))
        return fn

    def readfirst(self, verbose=True, root=None):
        self.readbin(os.path.join(_root(root), "output", self.sid + ".output00000.bin"), verbose)

    def readsecond(self, verbose=True, root=None):
        self.readbin(os.path.join(_root(root), "output", self.sid + ".output00001.bin"), verbose)

    def readstep(self, step, verbose=True, root=None):
        fn = os.path.join(_root(root), "output", "{0}.output{1:0=5}.bin".format(self.sid, step))
        self.readbin(fn, verbose)

    def readlast(self, verbose=True, root=None):
        """python/sphere.py:2314-2327."""
        lastfile = status(self.sid, root=root)
        self.readstep(lastfile, verbose, root=root)

    def readTime(self, time, verbose=True, root=None):
        """python/sphere.py:2329-2356."""
        self.readfirst(verbose=False, root=root)
        t_first = self.currentTime()
        n_first = self.time_step_count[0]
        self.readlast(verbose=False, root=root)
        t_last = self.currentTime()
        n_last = self.time_step_count[0]
        if time < t_first or time > t_last:
            raise Exception("Error: The specified time {} s is outside the range of "
                            "output files [{}; {}] s.".format(time, t_first, t_last))
        dt_dn = (t_last - t_first) / (n_last - n_first)
        step = int((time - t_first) / dt_dn) + n_first + 1
        self.readstep(step, verbose=verbose, root=root)

    # -- setup helpers -----------------------------------------------------
    def generateRadii(self, psd="logn", mean=440e-6, variance=8.8e-9, histogram=False,
                      rng=None):
        """python/sphere.py:2358-2416."""
        rs = rng if rng is not None else numpy.random
        n = int(self.np[0])
        if psd == "logn":
            mu = math.log((mean ** 2) / math.sqrt(variance + mean ** 2))
            sigma = math.sqrt(math.log(variance / (mean ** 2) + 1))
            self.radius = rs.lognormal(mu, sigma, n)
        elif psd == "uni":
            self.radius = rs.uniform(mean - variance, mean + variance, n)
        else:
            raise Exception("Particle size distribution type not understood (" + str(psd)
                            + "). Valid values are 'uni' or 'logn'")

    def contactModel(self, contactmodel):
        self.contactmodel[0] = contactmodel

    def wall0iz(self):
        if self.nw[0] > 0:
            return int(self.w_x[0] / (self.L[2] / self.num[2]))
        raise Exception("No dynamic top wall present!")

    def normalBoundariesXY(self):
        self.periodic[0] = 0

    def periodicBoundariesXY(self):
        self.periodic[0] = 1

    def periodicBoundariesX(self):
        self.periodic[0] = 2

    def adaptiveGrid(self):
        self.adaptive[0] = 1

    def staticGrid(self):
        self.adaptive[0] = 0

    def defineWorldBoundaries(self, L, origo=[0.0, 0.0, 0.0], dx=-1):
        """python/sphere.py:2607-2653."""
        if dx > 0.0:
            cellsize_min = dx
        else:
            if self.np[0] < 1:
                raise Exception("Error: You need to define dx in defineWorldBoundaries "
                                "if there are no particles in the simulation.")
            cellsize_min = 2.1 * numpy.amax(self.radius)
        self.origo[:] = origo[:]
        self.L[:] = L[:]
        for d in range(3):
            self.num[d] = numpy.ceil((self.L[d] - self.origo[d]) / cellsize_min)
        if self.num[0] < 3 or self.num[1] < 3 or self.num[2] < 3:
            raise Exception("Error: The grid must be at least 3 cells in each direction\n"
                            "Grid: x={}, y={}, z={}\n".format(*self.num)
                            + "Please increase the world size.")

    def initGrid(self, dx=-1):
        """python/sphere.py:2655-2685."""
        cellsize_min = dx if dx > 0.0 else 2.1 * numpy.amax(self.radius)
        for d in range(3):
            self.num[d] = numpy.ceil((self.L[d] - self.origo[d]) / cellsize_min)
        if self.num[0] < 4 or self.num[1] < 4 or self.num[2] < 4:
            raise Exception("Error: The grid must be at least 3 cells in each direction\n"
                            "Grid: x={}, y={}, z={}".format(*self.num))
        if self.nw > 0:
            self.w_x[0] = self.L[0]

    def initGridAndWorldsize(self, margin=2.0):
        """python/sphere.py:2687-2723."""
        r_max = numpy.amax(self.radius)
        self.origo = numpy.array([numpy.amin(self.x[:, d] - self.radius[:]) for d in range(3)]) \
            - margin * r_max
        self.L = numpy.array([numpy.amax(self.x[:, d] + self.radius[:]) for d in range(3)]) \
            + margin * r_max
        cellsize_min = 2.1 * r_max
        for d in range(3):
            self.num[d] = numpy.ceil((self.L[d] - self.origo[d]) / cellsize_min)
        if self.num[0] < 4 or self.num[1] < 4 or self.num[2] < 4:
            raise Exception("Error: The grid must be at least 3 cells in each direction, "
                            "num = " + str(self.num))
        if self.nw > 0:
            self.w_x[0] = self.L[0]

    def initGridPos(self, gridnum=numpy.array([12, 12, 36])):
        """python/sphere.py:2725-2785: loose cubic packing."""
        self.num = numpy.asarray(gridnum).astype(numpy.uint32)
        r_max = numpy.amax(self.radius)
        cellsize = 2.1 * r_max
        self.L = self.num * cellsize
        n = int(self.np[0])
        if (int(self.num[0]) * int(self.num[1]) * int(self.num[2]) - 2 ** 3) < n:
            raise NameError("Error! The grid is not sufficiently large.")
        if self.periodic[0] == 1:
            self.num[0] -= 1
            self.num[1] -= 1
        if (int(self.num[0]) * int(self.num[1]) * int(self.num[2]) - 2 * 3 * 3) < n:
            raise NameError("Error! The grid is not sufficiently large.")
        nx, ny = int(self.num[0]), int(self.num[1])
        for i in range(n):
            gp = numpy.array([i % nx, (i // nx) % nx, i // (nx * ny)])
            self.x[i, :] = gp * cellsize + 0.5 * cellsize
            if self.periodic[0] == 1 and gp[2] % 2:
                self.x[i, 0] += 0.5 * cellsize
                self.x[i, 1] += 0.5 * cellsize
        if self.periodic[0] == 1:
            self.num[0] += 1
            self.num[1] += 1

    def initRandomGridPos(self, gridnum=numpy.array([12, 12, 32]), padding=2.1, rng=None):
        """python/sphere.py:2787-2846: loose cubic packing with random offsets."""
        rs = rng if rng is not None else numpy.random
        coarsegrid = numpy.floor(numpy.asarray(gridnum) / 2)
        r_max = numpy.amax(self.radius)
        cellsize = padding * r_max * 2
        n = int(self.np[0])
        if ((coarsegrid[0] - 1) * (coarsegrid[1] - 1) * (coarsegrid[2] - 1)) < n:
            raise NameError("Error! The grid is not sufficiently large.")
        i = numpy.arange(n)
        gp = numpy.empty((n, 3))
        gp[:, 0] = i % coarsegrid[0]
        gp[:, 1] = numpy.floor(i / coarsegrid[0]) % coarsegrid[1]
        gp[:, 2] = numpy.floor(i / (coarsegrid[0] * coarsegrid[1]))
        r = self.radius * 1.05
        for d in range(3):
            u = rs.random_sample(n) if hasattr(rs, "random_sample") else rs.random(n)
            self.x[:, d] = gp[:, d] * cellsize + ((cellsize - r) - r) * u + r
        x_max = numpy.max(self.x[:, 0] + self.radius)
        y_max = numpy.max(self.x[:, 1] + self.radius)
        z_max = numpy.max(self.x[:, 2] + self.radius)
        self.num[0] = numpy.ceil(x_max / cellsize)
        self.num[1] = numpy.ceil(y_max / cellsize)
        self.num[2] = numpy.ceil(z_max / cellsize)
        self.L = self.num * cellsize

    # -- bonds --------------------------------------------------------------
    def bond(self, i, j):
        """Register a bond between particles i and j (python/sphere.py:4075-4119)."""
        self.lambda_bar[0] = 1.0
        self.nb0[0] += 1
        if self.nb0[0] == 1:
            self.bonds = numpy.zeros((1, 2), dtype=numpy.uint32)
            self.bonds[0] = [i, j]
            self.bonds_delta_n = numpy.zeros(1)
            self.bonds_delta_t = numpy.zeros((1, self.nd[0]))
            self.bonds_omega_n = numpy.zeros(1)
            self.bonds_omega_t = numpy.zeros((1, self.nd[0]))
        else:
            self.bonds = numpy.vstack((self.bonds, [i, j])).astype(numpy.uint32)
            self.bonds_delta_n = numpy.append(self.bonds_delta_n, [0.0])
            self.bonds_delta_t = numpy.append(self.bonds_delta_t, [[0.0, 0.0, 0.0]], axis=0)
            self.bonds_omega_n = numpy.append(self.bonds_omega_n, [0.0])
            self.bonds_omega_t = numpy.append(self.bonds_omega_t, [[0.0, 0.0, 0.0]], axis=0)

    # -- boundary conditions -------------------------------------------------
    def zeroKinematics(self):
        """python/sphere.py:2930-2947."""
        n, nd = int(self.np[0]), int(self.nd[0])
        self.force = numpy.zeros((n, nd))
        self.torque = numpy.zeros((n, nd))
        self.vel = numpy.zeros((n, nd))
        self.angvel = numpy.zeros((n, nd))
        self.angpos = numpy.zeros((n, nd))
        self.es = numpy.zeros(n)
        self.ev = numpy.zeros(n)
        self.xyzsum = numpy.zeros((n, 3))

    def adjustUpperWall(self, z_adjust=1.1):
        """python/sphere.py:2949-2970."""
        self.nw = numpy.ones(1, dtype=numpy.uint32)
        self.wmode = numpy.zeros(1, dtype=numpy.int32)
        self.w_n = numpy.zeros((1, self.nd[0]))
        self.w_n[0, 2] = -1.0
        self.w_vel = numpy.zeros(1)
        self.w_force = numpy.zeros(1)
        self.w_sigma0 = numpy.zeros(1)
        self.w_x = numpy.zeros(1)
        self.w_m = numpy.zeros(1)
        self.adjustWall(idx=0, adjust=z_adjust)

    def adjustWall(self, idx, adjust=1.1, wall_mass_factor=1.0):
        """python/sphere.py:2972-3017."""
        if idx == 0:
            dim = 2
        elif idx in (1, 2):
            dim = 0
        elif idx in (3, 4):
            dim = 1
        else:
            raise Exception("adjustWall: idx value not understood")
        xmin = numpy.min(self.x[:, dim] - self.radius)
        xmax = numpy.max(self.x[:, dim] + self.radius)
        cellsize = self.L[0] / self.num[0]
        self.num[dim] = numpy.ceil(((xmax - xmin) * adjust + xmin) / cellsize)
        self.L[dim] = (xmax - xmin) * adjust + xmin
        if idx in (0, 1, 3):
            self.w_x[idx] = xmax
        else:
            self.w_x[idx] = xmin
        self.w_m[idx] = self.totalMass()

    def consolidate(self, normal_stress=10e3):
        """python/sphere.py:3019-3048."""
        self.nw[0] = 1
        if normal_stress <= 0.0:
            raise Exception("consolidate() error: The normal stress should be a positive "
                            "value, but is " + str(normal_stress) + " Pa")
        self.zeroKinematics()
        self.adjustUpperWall()
        self.wmode = numpy.array([1], dtype=numpy.int32)
        self.w_sigma0 = numpy.ones(1) * normal_stress
        self.w_m[0] = self.totalMass()

    def uniaxialStrainRate(self, wvel=-0.001):
        """python/sphere.py:3050-3066."""
        self.zeroKinematics()
        self.adjustUpperWall()
        self.wmode = numpy.array([2], dtype=numpy.int32)
        self.w_vel = numpy.array([wvel])

    def triaxial(self, wvel=-0.001, normal_stress=10.0e3):
        """python/sphere.py:3068-3095."""
        self.zeroKinematics()
        self.nw[0] = 5
        self.wmode = numpy.array([2, 1, 1, 1, 1], dtype=numpy.int32)
        self.w_vel = numpy.array([1, 0, 0, 0, 0]) * wvel
        self.w_sigma0 = numpy.array([0, 1, 1, 1, 1]) * normal_stress
        self.w_n = numpy.array(([0, 0, -1], [-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0]),
                               dtype=numpy.float64)
        self.w_x = numpy.zeros(5)
        self.w_m = numpy.zeros(5)
        self.w_force = numpy.zeros(5)
        for i in range(5):
            self.adjustWall(idx=i)

    def shear(self, shear_strain_rate=1.0, shear_stress=False):
        """python/sphere.py:3097-3165."""
        self.nw[0] = 1
        z_min = numpy.min(self.x[:, 2] - self.radius)
        z_max = numpy.max(self.x[:, 2] + self.radius)
        cellsize = self.L[0] / self.num[0]
        self.num[2] += 1
        self.L[2] = self.num[2] * cellsize
        self.zeroKinematics()
        self.wmode = numpy.array([1], dtype=numpy.int32)
        d_max_below = numpy.max(self.radius[numpy.nonzero(self.x[:, 2] <
                                                          (z_max - z_min) * 0.3)]) * 2.0
        I = numpy.nonzero(self.x[:, 2] < (z_min + d_max_below))
        self.fixvel[I] = 1
        self.angvel[I, :] = 0.0
        self.vel[I, 0] = 0.0
        self.vel[I, 1] = 0.0
        self.color[I] = -1
        d_max_top = numpy.max(self.radius[numpy.nonzero(self.x[:, 2] >
                                                        (z_max - z_min) * 0.7)]) * 2.0
        I = numpy.nonzero(self.x[:, 2] > (z_max - d_max_top))
        self.fixvel[I] = 1
        self.angvel[I, :] = 0.0
        if shear_stress is False:
            self.vel[I, 0] = (z_max - z_min) * shear_strain_rate
        else:
            self.vel[I, 0] = 0.0
            self.wmode[0] = 3
            self.w_tau_x[0] = float(shear_stress)
        self.vel[I, 1] = 0.0
        self.color[I] = -1
        self.gamma_wt[0] = 0.0
        self.mu_ws[0] = 0.0
        self.mu_wd[0] = 0.0

    def initTemporal(self, total, current=0.0, file_dt=0.05, step_count=0, dt=-1,
                     epsilon=0.01):
        """python/sphere.py:3307-3376.  dt = epsilon/sqrt(k_max/m_min) unless given."""
        if dt > 0.0:
            self.time_dt[0] = dt
            if self.np[0] > 0:
                print("Warning: Manually specifying the time step length when "
                      "simulating particles may produce instabilities.")
        elif self.np[0] > 0:
            r_min = numpy.min(self.radius)
            m_min = self.rho * 4.0 / 3.0 * numpy.pi * r_min ** 3
            if self.E > 0.001:
                k_max = numpy.max(numpy.pi / 2.0 * self.E * self.radius)
            else:
                k_max = numpy.max([self.k_n[:], self.k_t[:]])
            self.time_dt[0] = (epsilon / (numpy.sqrt(k_max / m_min))).ravel()[0]
        else:
            raise Exception("Error: Could not automatically set a time step.")
        self.time_current[0] = current
        self.time_total[0] = total
        self.time_file_dt[0] = file_dt
        self.time_step_count[0] = 0

    def defaultParams(self, mu_s=0.5, mu_d=0.5, mu_r=0.0, rho=2600, k_n=1.16e9,
                      k_t=1.16e9, k_r=0, gamma_n=0.0, gamma_t=0.0, gamma_r=0.0,
                      gamma_wn=1.0e4, gamma_wt=1.0e4, capillaryCohesion=0):
        """python/sphere.py:3762-3873."""
        self.rho = numpy.ones(1) * rho
        self.k_n = numpy.ones(1) * k_n
        self.k_t = numpy.ones(1) * k_t
        self.k_r = numpy.ones(1) * k_r
        self.gamma_n = numpy.ones(1) * gamma_n
        self.gamma_t = numpy.ones(1) * gamma_t
        self.gamma_r = numpy.ones(1) * gamma_r
        self.mu_s = numpy.ones(1) * mu_s
        self.mu_d = numpy.ones(1) * mu_d
        self.mu_r = numpy.ones(1) * mu_r
        self.gamma_wn[0] = gamma_wn
        self.gamma_wt[0] = gamma_wt
        self.mu_ws = self.mu_s
        self.mu_wd = self.mu_d
        theta = 0.0
        if capillaryCohesion == 1:
            self.kappa[0] = 2.0 * math.pi * gamma_t * numpy.cos(theta)
            self.V_b[0] = 1e-12
        else:
            self.kappa[0] = 0.0
            self.V_b[0] = 0.0
        self.db[0] = (1.0 + theta / 2.0) * self.V_b[0] ** (1.0 / 3.0)

    def setStiffnessNormal(self, k_n):
        self.k_n[0] = k_n

    def setStiffnessTangential(self, k_t):
        self.k_t[0] = k_t

    def setYoungsModulus(self, E):
        self.E[0] = E

    def setDampingNormal(self, gamma, over_damping=False):
        """python/sphere.py:3910-3945 (critical damping from smallestMass())."""
        self.gamma_n[0] = gamma
        critical_gamma = 2.0 * numpy.sqrt(self.smallestMass() * self.k_n[0])
        damping_ratio = gamma / critical_gamma
        if damping_ratio > 1.0 and not over_damping:
            raise Exception("Warning: The system is over-dampened (ratio = "
                            + str(damping_ratio) + ") in the normal component.\n"
                            "Call this function once more with `over_damping=True` if this "
                            "is what you want. \nCritical damping = " + str(critical_gamma)
                            + ".")

    def setDampingTangential(self, gamma, over_damping=False):
        """python/sphere.py:3947-3977."""
        self.gamma_t[0] = gamma
        damping_ratio = gamma / (2.0 * numpy.sqrt(self.smallestMass() * self.k_t[0]))
        if damping_ratio > 1.0 and not over_damping:
            raise Exception("Warning: The system is over-dampened (ratio = "
                            + str(damping_ratio) + ") in the tangential component.\n"
                            "Call this function once more with `over_damping=True` if this "
                            "is what you want.")

    def setStaticFriction(self, mu_s):
        self.mu_s[0] = mu_s

    def setDynamicFriction(self, mu_d):
        self.mu_d[0] = mu_d

    def setTopWallNormalStressModulation(self, A, f, plot=False):
        self.w_sigma0_A[0] = A
        self.w_sigma0_f[0] = f

    def disableTopWallNormalStressModulation(self):
        self.w_sigma0_A[0] = 0.0
        self.w_sigma0_f[0] = 0.0

    # -- derived quantities (python/sphere.py:4121-4440) -----------------
    def currentTime(self, value=-1):
        if value != -1:
            self.time_current[0] = value
        return self.time_current[0]

    def currentNormalStress(self, type="defined"):
        if type == "defined":
            return self.w_sigma0[0] + self.w_sigma0_A[0] * numpy.sin(
                2.0 * numpy.pi * self.w_sigma0_f[0] * self.time_current[0])
        elif type == "effective":
            return self.w_force[0] / (self.L[0] * self.L[1])
        raise Exception("Normal stress type " + type + " not understood")

    def surfaceArea(self, idx):
        return 4.0 * numpy.pi * self.radius[idx] ** 2

    def volume(self, idx):
        return V_sphere(self.radius[idx])

    def mass(self, idx):
        return self.rho[0] * self.volume(idx)

    def totalMass(self):
        """Sequential sum of rho*V over particles, in index order, as the
        reference's loop (python/sphere.py:4175-4184) accumulates it."""
        m = self.rho[0] * (4.0 / 3.0 * math.pi * numpy.asarray(self.radius) ** 3.0)
        return float(numpy.cumsum(numpy.concatenate(([0.0], m)))[-1])

    def smallestMass(self):
        """The reference returns the smallest *volume* here (no density),
        python/sphere.py:4186-4195; kept for drop-in behaviour."""
        return V_sphere(numpy.min(self.radius))

    def largestMass(self):
        return V_sphere(numpy.max(self.radius))

    def momentOfInertia(self, idx):
        return 2.0 / 5.0 * self.mass(idx) * self.radius[idx] ** 2

    def kineticEnergy(self, idx):
        return 0.5 * self.mass(idx) * numpy.sqrt(numpy.dot(self.vel[idx, :], self.vel[idx, :])) ** 2

    def totalKineticEnergy(self):
        return sum(self.kineticEnergy(i) for i in range(int(self.np[0])))

    def rotationalEnergy(self, idx):
        return 0.5 * self.momentOfInertia(idx) * numpy.sqrt(
            numpy.dot(self.angvel[idx, :], self.angvel[idx, :])) ** 2

    def totalRotationalEnergy(self):
        return sum(self.rotationalEnergy(i) for i in range(int(self.np[0])))

    def viscousEnergy(self, idx):
        return self.ev[idx]

    def totalViscousEnergy(self):
        return float(numpy.sum(self.ev))

    def frictionalEnergy(self, idx):
        return self.es[idx]

    def totalFrictionalEnergy(self):
        return float(numpy.sum(self.es))

    def momentum(self, idx):
        return self.mass(idx) * self.vel[idx, :]

    def totalMomentum(self):
        m = self.rho[0] * 4.0 / 3.0 * math.pi * self.radius ** 3.0
        return numpy.sum(m[:, None] * self.vel, axis=0)

    def energy(self, method):
        """python/sphere.py:4311-4380 (fluid-free methods)."""
        if method == "pot":
            m = self.rho[0] * 4.0 / 3.0 * math.pi * self.radius ** 3.0
            return float(numpy.sum(m * math.sqrt(numpy.dot(self.g, self.g)) * self.x[:, 2]))
        elif method == "kin":
            return self.totalKineticEnergy()
        elif method == "rot":
            return self.totalRotationalEnergy()
        elif method == "shear":
            return float(numpy.sum(self.es))
        elif method == "shearrate":
            return float(numpy.sum(self.es_dot))
        elif method == "visc_n":
            return float(numpy.sum(self.ev))
        elif method == "visc_n_rate":
            return float(numpy.sum(self.ev_dot))
        elif method == "bondpot":
            if self.nb0[0] > 0:
                R_bar = self.lambda_bar[0] * numpy.minimum(
                    self.radius[self.bonds[:, 0]], self.radius[self.bonds[:, 1]])
                A = numpy.pi * R_bar ** 2
                I = 0.25 * numpy.pi * R_bar ** 4
                J = I * 2.0
                bondpot_fn = numpy.sum(0.5 * A * self.k_n[0] * numpy.abs(self.bonds_delta_n) ** 2)
                bondpot_ft = numpy.sum(0.5 * A * self.k_t[0] *
                                       numpy.linalg.norm(self.bonds_delta_t, axis=1) ** 2)
                bondpot_tn = numpy.sum(0.5 * J * self.k_t[0] * numpy.abs(self.bonds_omega_n) ** 2)
                bondpot_tt = numpy.sum(0.5 * I * self.k_n[0] *
                                       numpy.linalg.norm(self.bonds_omega_t, axis=1) ** 2)
                return float(bondpot_fn + bondpot_ft + bondpot_tn + bondpot_tt)
            return 0.0
        raise Exception("Unknownw energy() method '" + method + "'")

    def voidRatio(self):
        """python/sphere.py:4382-4400."""
        if self.nw[0] == 0:
            V_total = self.L[0] * self.L[1] * self.L[2]
        else:
            V_total = self.L[0] * self.L[1] * self.w_x[0]
        V_solid = numpy.sum(V_sphere(self.radius))
        return (V_total - V_solid) / V_solid

    def bulkPorosity(self, trim=True):
        """python/sphere.py:4402-4435."""
        V_total = 0.0
        if trim:
            min_z = numpy.min(self.x[:, 2] - self.radius)
            max_z = numpy.max(self.x[:, 2] + self.radius)
            V_total = self.L[0] * self.L[1] * (max_z - min_z)
        else:
            if self.nw[0] == 0:
                V_total = self.L[0] * self.L[1] * self.L[2]
            else:
                V_total = self.L[0] * self.L[1] * self.w_x[0]
        V_solid = numpy.sum(V_sphere(self.radius))
        return (V_total - V_solid) / V_total

    def shearDisplacement(self):
        """python/sphere.py:4777-4792."""
        fixvel = numpy.nonzero(self.fixvel > 0.0)
        return numpy.max(self.xyzsum[fixvel, 0])

    def shearVelocity(self):
        fixvel = numpy.nonzero(self.fixvel > 0.0)
        return numpy.max(self.vel[fixvel, 0])

    def shearVel(self):
        return self.shearVelocity()

    def shearStrain(self):
        """python/sphere.py:4816-4835."""
        h0 = self.w_x[0]
        return self.shearDisplacement() / h0

    def shearStrainRate(self):
        return self.shearVelocity() / self.w_x[0]

    def shearStress(self, type="effective"):
        """python/sphere.py:6407-6436 (fluid-free part)."""
        if type == "defined":
            return self.w_tau_x[0]
        elif type == "effective":
            fixvel = numpy.nonzero(self.fixvel > 0.0)
            force = numpy.zeros(3)
            for i in fixvel[0]:
                if self.x[i, 2] > numpy.max(self.x[:, 2] - self.radius) * 0.5:
                    force += -self.force[i, :]
            return force[0] / (self.L[0] * self.L[1])
        raise Exception("Shear stress type " + type + " not understood")

    def findOverlaps(self, device=-1):
        """Pairwise overlaps (python/sphere.py:4868-4883): DEM::findOverlaps on
        the GPU (sph_find_overlaps; the reference runs `sphere --contacts`
        and reads its printout back).  Sets ``self.pairs`` (2 x k, int32,
        i < j) and ``self.overlaps`` (k), as the reference does."""
        from . import engine
        with engine.Context(self.tobytes(), device=device, check=False) as ctx:
            pairs, overlaps = ctx.find_overlaps()
        self.pairs = numpy.ascontiguousarray(pairs.T.astype(numpy.int32))
        self.overlaps = overlaps

    def porosity(self, slices=10, verbose=False, device=-1):
        """Porosity in horizontal slabs (python/sphere.py:4437-4480):
        DEM::porosity on the GPU (sph_porosity) where the reference runs
        its `porosity` program.  Returns (porosities, slab depths)."""
        from . import engine
        with engine.Context(self.tobytes(), device=device, check=False) as ctx:
            depth, por = ctx.porosity(int(slices))
        return por, depth

    # Consumers of findOverlaps' pair list (python/sphere.py:4885-5017).
    def findCoordinationNumber(self):
        """Contacts per particle from self.pairs."""
        self.coordinationnumber = numpy.bincount(
            self.pairs.reshape(-1).astype(numpy.int64), minlength=int(self.np[0]))
        return self.coordinationnumber

    def findMeanCoordinationNumber(self):
        return numpy.mean(self.coordinationnumber)

    def findNormalForces(self):
        """Elastic normal force magnitude k_n |delta| per overlapping pair."""
        self.findOverlaps()
        self.f_n_magn = self.k_n * numpy.abs(self.overlaps)

    def contactSurfaceArea(self, i, j, overlap):
        """Area of the lens where spheres i and j (centre distance
        r_i + r_j + overlap) intersect: pi a^2, with a the radius of the
        intersection circle from the triangle of sides d, r_i, r_j."""
        ri = self.radius[i]
        rj = self.radius[j]
        d = ri + rj + overlap
        a = numpy.sqrt((-d + ri - rj) * (-d - ri + rj) * (-d + ri + rj) * (d + ri + rj)) / (2.0 * d)
        return numpy.pi * a ** 2

    def contactParticleArea(self, i, j):
        """Cross-section of the pair's mean radius."""
        return numpy.pi * ((self.radius[i] + self.radius[j]) * 0.5) ** 2

    def findAllContactSurfaceAreas(self):
        return self.contactSurfaceArea(self.pairs[0, :], self.pairs[1, :], self.overlaps)

    def findAllAverageParticlePairAreas(self):
        return self.contactParticleArea(self.pairs[0, :], self.pairs[1, :])

    def findContactStresses(self, area="average"):
        """Normal stress per pair: k_n |delta| over the mean-radius section
        ('average') or the lens area ('contact')."""
        self.findNormalForces()
        if area == "average":
            areas = self.findAllAverageParticlePairAreas()
        elif area == "contact":
            areas = self.findAllContactSurfaceAreas()
        else:
            raise Exception('Contact area type "' + area + '" not understood')
        self.sigma_contacts = self.f_n_magn / areas

    def findLoadedContacts(self, threshold):
        self.findContactStresses()
        return numpy.nonzero(self.sigma_contacts >= threshold)

    def dry(self):
        self.fluid = False

    # -- running -------------------------------------------------------------
    def run(self, verbose=True, hideinputfile=False, dry=False, valgrind=False,
            cudamemcheck=False, device=-1, root=None, sidecar=False, restart_state=None,
            bond_k=0.0, bond_break_ext=0.0):
        """Run the simulation (python/sphere.py:4482-4540).

        Writes <root>/input/<sid>.bin, then runs the DEM loop in-process on the
        GPU and writes <root>/output/<sid>.output%05d.bin + status.dat.
        Returns 0 on success; on failure prints the reference's warning and
        returns the (negative) library status.  Not in the reference:
        `sidecar` also writes a restart sidecar (.state) next to every
        output, and `restart_state` (the sidecar of the output this sim was
        read from) makes the run continue bit for bit where that one left
        off.  `bond_k` (N/m) and `bond_break_ext` (fraction of the smaller
        radius) select the engine's bond options (SPH_S_BOND_K,
        SPH_S_BOND_BREAK_EXT; 0 keeps the reference's bond law), which the
        .bin layout cannot carry."""
        from . import engine
        rt = _root(root)
        fn = self.writebin(os.path.join(rt, "input"), verbose=False)
        state_in = fn[:-4] + ".state"
        if restart_state is not None:
            shutil.copyfile(restart_state, state_in)
            sidecar = True
        elif os.path.exists(state_in):
            os.remove(state_in)  # a sidecar left by an earlier resumed run
        if valgrind or cudamemcheck:
            print("Warning: valgrind/cuda-memcheck wrappers are not available; "
                  "running normally")
        opts = {}
        if bond_k:
            opts["bond_k"] = float(bond_k)
        if bond_break_ext:
            opts["bond_break_ext"] = float(bond_break_ext)
        status_ = engine.run_file(fn, os.path.join(rt, "output"), self.sid,
                                  verbose=verbose, dry=dry, device=device, sidecar=sidecar,
                                  scalars=opts)
        if status_ != 0:
            print("Warning: the sphere run returned with status " + str(status_))
        return status_

    def cleanup(self, root=None):
        cleanup(self, root=root)


def status(project, root=None):
    """Last output file number from <root>/output/<sid>.status.dat
    (python/sphere.py:7721-7741)."""
    fn = os.path.join(_root(root), "output", "{0}.status.dat".format(project))
    with open(fn) as fh:
        data = fh.read()
    return int(data.split()[2])


def cleanup(s, root=None):
    """Remove the input/output files of a sim (python/sphere.py:7743-7758)."""
    import glob
    rt = _root(root)
    pats = [os.path.join(rt, "input", s.sid + ".bin"),
            os.path.join(rt, "output", s.sid + ".*.bin"),
            os.path.join(rt, "output", s.sid + ".status.dat")]
    for pat in pats:
        for f in glob.glob(pat):
            try:
                os.remove(f)
            except OSError:
                pass


def run(binary, verbose=True, hideinputfile=False, root=None):
    """Run the engine on an input file (python/sphere.py:7617-7636)."""
    from . import engine
    rt = _root(root)
    sid = os.path.splitext(os.path.basename(binary))[0]
    path = binary if os.path.isabs(binary) else os.path.join(rt, binary)
    return engine.run_file(path, os.path.join(rt, "output"), sid, verbose=verbose)
