"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  It is the checker for the HIP product path and is
never the thing measured or shipped.
"""

import ctypes
import os

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_DUAL = None

NC = 32

# Field ids and layouts, shared with include/sphere_hip.h.
FIELDS = {
    "x": (0, numpy.float64, (-1, 4)),
    "vel": (1, numpy.float64, (-1, 4)),
    "angvel": (2, numpy.float64, (-1, 3)),
    "angpos": (3, numpy.float64, (-1, 3)),
    "force": (4, numpy.float64, (-1, 3)),
    "torque": (5, numpy.float64, (-1, 3)),
    "xyzsum": (6, numpy.float64, (-1, 3)),
    "acc": (7, numpy.float64, (-1, 3)),
    "angacc": (8, numpy.float64, (-1, 3)),
    "es_dot": (9, numpy.float64, (-1,)),
    "es": (10, numpy.float64, (-1,)),
    "ev_dot": (11, numpy.float64, (-1,)),
    "ev": (12, numpy.float64, (-1,)),
    "p": (13, numpy.float64, (-1,)),
    "contacts": (14, numpy.uint32, (-1, NC)),
    "delta_t": (15, numpy.float64, (-1, NC, 3)),
    "distmod": (16, numpy.float64, (-1, NC, 4)),
    "cellid": (17, numpy.uint32, (-1,)),
    "perm": (18, numpy.uint32, (-1,)),
    "cellstart": (19, numpy.uint32, (-1,)),
    "cellend": (20, numpy.uint32, (-1,)),
    "walls_nx": (21, numpy.float64, (-1, 4)),
    "walls_mvfd": (22, numpy.float64, (-1, 4)),
    "walls_acc": (23, numpy.float64, (-1,)),
    "walls_force_pp": (24, numpy.float64, (-1,)),
    "bonds": (25, numpy.uint32, (-1, 2)),
    "bonds_delta": (26, numpy.float64, (-1, 4)),
    "bonds_omega": (27, numpy.float64, (-1, 4)),
    "color": (28, numpy.int32, (-1,)),
}

STAGES = {"hash": 0, "sort": 1, "reorder": 2, "topology": 3, "interact": 4,
          "bonds": 5, "integrate": 6, "walls": 7, "advance": 8}


def lib_path(dual=False):
    return os.path.join(_HERE, "liboracle_dual.so" if dual else "liboracle.so")


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    if force or not os.path.exists(lib_path()) or not os.path.exists(lib_path(True)):
        subprocess.check_call(["make", "-s", "-B" if force else "all"], cwd=_HERE)


def _lib(dual=False):
    global _LIB, _DUAL
    if (_DUAL if dual else _LIB) is None:
        if not os.path.exists(lib_path(dual)):
            build()
        L = ctypes.CDLL(lib_path(dual))
        L.orc_open_mem.restype = ctypes.c_void_p
        L.orc_open_mem.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p,
                                   ctypes.c_size_t]
        L.orc_close.argtypes = [ctypes.c_void_p]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.orc_stage.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_size_t]
        L.orc_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_size_t]
        L.orc_field_bytes.restype = ctypes.c_int64
        L.orc_field_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_write_mem.restype = ctypes.c_int64
        L.orc_write_mem.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.orc_np.restype = ctypes.c_uint32
        L.orc_np.argtypes = [ctypes.c_void_p]
        L.orc_time.restype = ctypes.c_double
        L.orc_time.argtypes = [ctypes.c_void_p]
        L.orc_iter.restype = ctypes.c_int64
        L.orc_iter.argtypes = [ctypes.c_void_p]
        L.orc_sqrt32.restype = ctypes.c_double
        L.orc_sqrt32.argtypes = [ctypes.c_double]
        L.orc_tree_sum.restype = ctypes.c_double
        L.orc_tree_sum.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_exact_length.argtypes = [ctypes.c_int]
        L.orc_set_wall_sums_per_wall.argtypes = [ctypes.c_int]
        L.orc_find_overlaps.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        L.orc_find_overlaps.restype = None
        L.orc_porosity.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_void_p]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_set_contact_log.argtypes = [ctypes.c_int]
        L.orc_set_bond_model.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_set_bond_model.restype = None
        L.orc_contact_log_width.restype = ctypes.c_int
        L.orc_get_contact_log.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        if dual:
            L.orc_dual_seed_param.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
            L.orc_dual_seed_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
            L.orc_dual_get_tangent.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
            _DUAL = L
        else:
            _LIB = L
    return _DUAL if dual else _LIB


def set_threads(n):
    _lib().orc_set_threads(int(n))


def get_threads():
    return _lib().orc_get_threads()


def set_exact_length(on):
    """Gradient-check aid: FP64 instead of FP32 lengths in every step that
    follows (process-wide; the reference's FP32 lengths are the default)."""
    _lib().orc_set_exact_length(1 if on else 0)


def set_contact_log(on):
    """Test aid: record every evaluated contact of the following interact
    stages (contact models 2 and 3; process-wide, off by default)."""
    _lib().orc_set_contact_log(1 if on else 0)


# Columns of one contact record (Oracle.contact_log).
CLOG = {"active": 0, "slip": 1, "delta": 2, "n": slice(3, 6), "f_n": slice(6, 9),
        "f_t": slice(9, 12), "f_c": slice(12, 15), "vel_t": slice(15, 18),
        "delta_t": slice(18, 21), "delta_t0": slice(21, 24), "k_t": 24}


def set_wall_sums_per_wall(on):
    """Each dynamic wall integrates its own force sum (the engine's
    SPH_S_WALL_SUMS = 1) instead of wall 0's (the reference)."""
    _lib().orc_set_wall_sums_per_wall(1 if on else 0)


class Oracle:
    """One oracle simulation, created from a v2.15 .bin image."""

    DUAL = False

    def __init__(self, image):
        L = self._L = _lib(self.DUAL)
        buf = ctypes.create_string_buffer(bytes(image), len(image))
        err = ctypes.create_string_buffer(256)
        h = L.orc_open_mem(buf, len(image), err, 256)
        if not h:
            raise ValueError("oracle: " + err.value.decode())
        self._h = ctypes.c_void_p(h)

    def close(self):
        if self._h:
            self._L.orc_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def np(self):
        return int(self._L.orc_np(self._h))

    @property
    def time(self):
        return float(self._L.orc_time(self._h))

    @property
    def iter(self):
        return int(self._L.orc_iter(self._h))

    def step(self, n=1):
        rc = self._L.orc_step(self._h, int(n))
        if rc != 0:
            raise RuntimeError("orc_step failed: %d" % rc)

    def stage(self, name):
        rc = self._L.orc_stage(self._h, STAGES[name])
        if rc != 0:
            raise RuntimeError("orc_stage failed: %d" % rc)

    def get(self, name):
        fid, dt, shape = FIELDS[name]
        nbytes = self._L.orc_field_bytes(self._h, fid)
        out = numpy.empty(nbytes // numpy.dtype(dt).itemsize, dtype=dt)
        rc = self._L.orc_get_field(self._h, fid, out.ctypes.data, nbytes)
        if rc != 0:
            raise RuntimeError("orc_get_field(%s) failed: %d" % (name, rc))
        return out.reshape(shape)

    def set(self, name, value):
        fid, dt, shape = FIELDS[name]
        a = numpy.ascontiguousarray(value, dtype=dt)
        rc = self._L.orc_set_field(self._h, fid, a.ctypes.data, a.nbytes)
        if rc != 0:
            raise RuntimeError("orc_set_field(%s) failed: %d" % (name, rc))

    def set_bond_model(self, bond_k=0.0, bond_break_ext=0.0):
        """The engine's bond options for this simulation: a bond stiffness in
        N/m and a break at a normal extension of bond_break_ext times the
        smaller radius (0 = the reference's law)."""
        self._L.orc_set_bond_model(self._h, float(bond_k), float(bond_break_ext))

    def contact_log(self):
        """The contact records of the last interact stage (np x NC x width,
        original-index rows; columns in CLOG), when set_contact_log is on."""
        w = self._L.orc_contact_log_width()
        out = numpy.zeros(self.np * NC * w)
        if self._L.orc_get_contact_log(self._h, out.ctypes.data, out.size) != 0:
            raise RuntimeError("no contact log (set_contact_log(True) before stepping)")
        return out.reshape(self.np, NC, w)

    def find_overlaps(self):
        """DEM::findOverlaps: (pairs (k, 2) i < j, overlaps (k,))."""
        L = self._L
        n = ctypes.c_uint64()
        L.orc_find_overlaps(self._h, None, None, 0, ctypes.byref(n))
        pairs = numpy.zeros((n.value, 2), dtype=numpy.uint32)
        delta = numpy.zeros(n.value)
        L.orc_find_overlaps(self._h, pairs.ctypes.data, delta.ctypes.data, n.value,
                            ctypes.byref(n))
        return pairs, delta

    def porosity(self, z_slices=10):
        """DEM::porosity: (z positions, porosities) of z_slices slices."""
        z = numpy.zeros(z_slices)
        por = numpy.zeros(z_slices)
        rc = self._L.orc_porosity(self._h, int(z_slices), z.ctypes.data, por.ctypes.data)
        if rc:
            raise ValueError("The number of z-slices is too high.")
        return z, por

    def tobytes(self):
        n = self._L.orc_write_mem(self._h, None, 0)
        buf = ctypes.create_string_buffer(n)
        self._L.orc_write_mem(self._h, buf, n)
        return buf.raw


class TangentOracle(Oracle):
    """The oracle's dual-number build (liboracle_dual.so): the same values
    bit for bit, each carrying its derivative along one seeded direction of
    the parameters and the initial state.  Reversing the engine's adjoint
    gives the same directional derivatives (forward mode vs reverse mode)."""

    DUAL = True

    def seed_param(self, name, d=1.0):
        """Direction component of a parameter (ADJOINT_PARAMS names)."""
        i = ("k_n", "k_t", "gamma_n", "gamma_t", "mu_s", "mu_d", "gamma_wn", "gamma_wt",
             "mu_wd").index(name)
        if self._L.orc_dual_seed_param(self._h, i, float(d)) != 0:
            raise RuntimeError("orc_dual_seed_param failed")

    def seed_field(self, name, direction):
        """Direction component of the state "x", "vel" or "angvel" (np x 3)."""
        a = numpy.ascontiguousarray(direction, dtype=numpy.float64)
        if self._L.orc_dual_seed_field(self._h, FIELDS[name][0], a.ctypes.data) != 0:
            raise RuntimeError("orc_dual_seed_field failed")

    def tangent(self, name):
        """Directional derivative of "x", "vel" or "angvel" (np x 3)."""
        out = numpy.zeros((self.np, 3))
        if self._L.orc_dual_get_tangent(self._h, FIELDS[name][0], out.ctypes.data) != 0:
            raise RuntimeError("orc_dual_get_tangent failed")
        return out


def sqrt32(v):
    """The reference's length() rule on one value: FP32 root of v, widened."""
    return float(_lib().orc_sqrt32(float(v)))


def tree_sum(values):
    """The reference's wall-force reduction of `values` (block trees of 256,
    then a serial sum of the block results)."""
    a = numpy.ascontiguousarray(values, dtype=numpy.float64)
    return float(_lib().orc_tree_sum(a.ctypes.data, a.size))
