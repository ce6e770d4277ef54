"""sphere .bin codec (fluid-free part), written for this package.

Layout follows src/file_io.cpp:420-731 (writer) and python/sphere.py:965-1553
(reader, which upgrades files written by versions >= 1.02 and stamps them
2.15).  All integers are little-endian; there is no padding.
"""

import numpy

VERSION = 2.15

# Field groups written per particle, in file order (src/file_io.cpp:455-507).
_F64 = numpy.dtype("<f8")
_U32 = numpy.dtype("<u4")
_I32 = numpy.dtype("<i4")


class BinFormatError(Exception):
    pass


class _Reader:
    def __init__(self, data):
        self.b = memoryview(data)
        self.off = 0

    def take(self, dtype, count):
        dtype = numpy.dtype(dtype)
        n = dtype.itemsize * int(count)
        if self.off + n > len(self.b):
            raise BinFormatError("short read at byte %d (need %d more)" % (self.off, n))
        a = numpy.frombuffer(self.b[self.off:self.off + n], dtype=dtype).copy()
        self.off += n
        return a

    def f64(self, count=1):
        return self.take(_F64, count)

    def u32(self, count=1):
        return self.take(_U32, count)

    def i32(self, count=1):
        return self.take(_I32, count)


def decode(data, sim, bonds=True, sigma0mod=True):
    """Fill the attributes of `sim` from a .bin image (python/sphere.py:965-1353).

    Versions >= 1.02 are accepted; the fields missing in older versions get
    the reference reader's defaults and the version is stamped 2.15.
    Fluid blocks are not read (the DEM path carries none).
    """
    r = _Reader(data)
    version = r.f64()
    v = float(version[0])
    sim.version = version
    sim.nd = r.i32()
    sim.np = r.u32()
    nd, np_ = int(sim.nd[0]), int(sim.np[0])
    if nd != 3:
        raise BinFormatError("nd = %d, only 3-D files are supported" % nd)
    sim.time_dt = r.f64()
    sim.time_current = r.f64()
    sim.time_total = r.f64()
    sim.time_file_dt = r.f64()
    sim.time_step_count = r.u32()
    sim.origo = r.f64(nd)
    sim.L = r.f64(nd)
    sim.num = r.u32(nd)
    sim.periodic = r.i32()
    sim.adaptive = r.i32() if v >= 2.14 else numpy.zeros(1, dtype=numpy.int32)
    xr = r.f64(np_ * 4).reshape(np_, 4)
    sim.x = xr[:, :3].copy()
    sim.radius = xr[:, 3].copy()
    if v >= 1.03:
        sim.xyzsum = r.f64(np_ * 3).reshape(np_, 3)
    else:
        xy = r.f64(np_ * 2).reshape(np_, 2)
        sim.xyzsum = numpy.zeros((np_, 3))
        sim.xyzsum[:, :2] = xy
    vr = r.f64(np_ * 4).reshape(np_, 4)
    sim.vel = vr[:, :3].copy()
    sim.fixvel = vr[:, 3].copy()
    sim.force = r.f64(np_ * 3).reshape(np_, 3)
    sim.angpos = r.f64(np_ * 3).reshape(np_, 3)
    sim.angvel = r.f64(np_ * 3).reshape(np_, 3)
    sim.torque = r.f64(np_ * 3).reshape(np_, 3)
    sim.es_dot = r.f64(np_)
    sim.es = r.f64(np_)
    sim.ev_dot = r.f64(np_)
    sim.ev = r.f64(np_)
    sim.p = r.f64(np_)
    sim.g = r.f64(nd)
    sim.k_n = r.f64()
    sim.k_t = r.f64()
    sim.k_r = r.f64()
    sim.E = r.f64() if v >= 2.13 else numpy.zeros(1)
    sim.gamma_n = r.f64()
    sim.gamma_t = r.f64()
    sim.gamma_r = r.f64()
    sim.mu_s = r.f64()
    sim.mu_d = r.f64()
    sim.mu_r = r.f64()
    sim.gamma_wn = r.f64()
    sim.gamma_wt = r.f64()
    sim.mu_ws = r.f64()
    sim.mu_wd = r.f64()
    sim.rho = r.f64()
    sim.contactmodel = r.u32()
    sim.kappa = r.f64()
    sim.db = r.f64()
    sim.V_b = r.f64()
    sim.nw = r.u32()
    nw = int(sim.nw[0])
    sim.wmode = r.i32(nw)
    wn = r.f64(nw * 4).reshape(nw, 4)
    sim.w_n = wn[:, :3].copy()
    sim.w_x = wn[:, 3].copy()
    wm = r.f64(nw * 4).reshape(nw, 4)
    sim.w_m = wm[:, 0].copy()
    sim.w_vel = wm[:, 1].copy()
    sim.w_force = wm[:, 2].copy()
    sim.w_sigma0 = wm[:, 3].copy()
    if sigma0mod:
        sim.w_sigma0_A = r.f64()
        sim.w_sigma0_f = r.f64()
    sim.w_tau_x = r.f64() if v >= 2.1 else numpy.zeros(1)
    if bonds:
        sim.lambda_bar = r.f64()
        sim.nb0 = r.u32()
        sim.sigma_b = r.f64()
        sim.tau_b = r.f64()
        nb = int(sim.nb0[0])
        sim.bonds = r.u32(nb * 2).reshape(nb, 2)
        sim.bonds_delta_n = r.f64(nb)
        sim.bonds_delta_t = r.f64(nb * 3).reshape(nb, 3)
        sim.bonds_omega_n = r.f64(nb)
        sim.bonds_omega_t = r.f64(nb * 3).reshape(nb, 3)
    else:
        sim.nb0 = numpy.zeros(1, dtype=numpy.uint32)
    if v >= 1.02:
        sim.color = r.i32(np_)
    else:
        sim.color = numpy.zeros(np_, dtype=numpy.int32)
    sim.version = numpy.array([VERSION])
    return r.off


def _a(x, dtype, n=None):
    a = numpy.ascontiguousarray(numpy.asarray(x, dtype=dtype)).ravel()
    if n is not None and a.size != n:
        raise BinFormatError("expected %d values, got %d" % (n, a.size))
    return a.tobytes()


def encode(sim):
    """Serialize `sim` as a v2.15 .bin image (src/file_io.cpp:420-731)."""
    np_ = int(sim.np[0])
    nw = int(sim.nw[0])
    nb = int(sim.nb0[0])
    parts = [
        _a([VERSION], _F64, 1),
        _a(sim.nd, _I32, 1),
        _a(sim.np, _U32, 1),
        _a(sim.time_dt, _F64, 1),
        _a(sim.time_current, _F64, 1),
        _a(sim.time_total, _F64, 1),
        _a(sim.time_file_dt, _F64, 1),
        _a(sim.time_step_count, _U32, 1),
        _a(sim.origo, _F64, 3),
        _a(sim.L, _F64, 3),
        _a(sim.num, _U32, 3),
        _a(sim.periodic, _I32, 1),
        _a(sim.adaptive, _I32, 1),
    ]
    xr = numpy.empty((np_, 4))
    xr[:, :3] = numpy.asarray(sim.x, dtype=numpy.float64).reshape(np_, 3)
    xr[:, 3] = numpy.asarray(sim.radius, dtype=numpy.float64).reshape(np_)
    parts.append(_a(xr, _F64, np_ * 4))
    parts.append(_a(sim.xyzsum, _F64, np_ * 3))
    vr = numpy.empty((np_, 4))
    vr[:, :3] = numpy.asarray(sim.vel, dtype=numpy.float64).reshape(np_, 3)
    vr[:, 3] = numpy.asarray(sim.fixvel, dtype=numpy.float64).reshape(np_)
    parts.append(_a(vr, _F64, np_ * 4))
    for name in ("force", "angpos", "angvel", "torque"):
        parts.append(_a(getattr(sim, name), _F64, np_ * 3))
    for name in ("es_dot", "es", "ev_dot", "ev", "p"):
        parts.append(_a(getattr(sim, name), _F64, np_))
    parts.append(_a(sim.g, _F64, 3))
    for name in ("k_n", "k_t", "k_r", "E", "gamma_n", "gamma_t", "gamma_r", "mu_s",
                 "mu_d", "mu_r", "gamma_wn", "gamma_wt", "mu_ws", "mu_wd", "rho"):
        parts.append(_a(numpy.asarray(getattr(sim, name)).ravel()[:1], _F64, 1))
    parts.append(_a(numpy.asarray(sim.contactmodel).ravel()[:1], _U32, 1))
    for name in ("kappa", "db", "V_b"):
        parts.append(_a(numpy.asarray(getattr(sim, name)).ravel()[:1], _F64, 1))
    parts.append(_a(sim.nw, _U32, 1))
    parts.append(_a(numpy.asarray(sim.wmode).ravel()[:nw], _I32, nw))
    wn = numpy.empty((nw, 4))
    if nw:
        wn[:, :3] = numpy.asarray(sim.w_n, dtype=numpy.float64).reshape(-1, 3)[:nw]
        wn[:, 3] = numpy.asarray(sim.w_x, dtype=numpy.float64).ravel()[:nw]
    parts.append(_a(wn, _F64, nw * 4))
    wm = numpy.empty((nw, 4))
    if nw:
        wm[:, 0] = numpy.asarray(sim.w_m, dtype=numpy.float64).ravel()[:nw]
        wm[:, 1] = numpy.asarray(sim.w_vel, dtype=numpy.float64).ravel()[:nw]
        wm[:, 2] = numpy.asarray(sim.w_force, dtype=numpy.float64).ravel()[:nw]
        wm[:, 3] = numpy.asarray(sim.w_sigma0, dtype=numpy.float64).ravel()[:nw]
    parts.append(_a(wm, _F64, nw * 4))
    parts.append(_a(numpy.asarray(sim.w_sigma0_A).ravel()[:1], _F64, 1))
    parts.append(_a(numpy.asarray(sim.w_sigma0_f).ravel()[:1], _F64, 1))
    parts.append(_a(numpy.asarray(sim.w_tau_x).ravel()[:1], _F64, 1))
    parts.append(_a(numpy.asarray(sim.lambda_bar).ravel()[:1], _F64, 1))
    parts.append(_a(sim.nb0, _U32, 1))
    parts.append(_a(numpy.asarray(sim.sigma_b).ravel()[:1], _F64, 1))
    parts.append(_a(numpy.asarray(sim.tau_b).ravel()[:1], _F64, 1))
    parts.append(_a(numpy.asarray(sim.bonds).reshape(-1)[:2 * nb], _U32, 2 * nb))
    parts.append(_a(numpy.asarray(sim.bonds_delta_n).ravel()[:nb], _F64, nb))
    parts.append(_a(numpy.asarray(sim.bonds_delta_t).reshape(-1)[:3 * nb], _F64, 3 * nb))
    parts.append(_a(numpy.asarray(sim.bonds_omega_n).ravel()[:nb], _F64, nb))
    parts.append(_a(numpy.asarray(sim.bonds_omega_t).reshape(-1)[:3 * nb], _F64, 3 * nb))
    parts.append(_a(sim.color, _I32, np_))
    return b"".join(parts)


def image_size(np_, nw=0, nb0=0):
    """Byte size of a fluid-free v2.15 image."""
    return 348 + 228 * np_ + 68 * nw + 72 * nb0
