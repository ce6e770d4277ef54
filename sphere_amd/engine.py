"""ctypes binding of libsphere_hip.so, the HIP engine's C ABI.

The product path: every call here runs the HIP kernels of
sphere_amd/csrc.  There is no CPU fallback; a missing or unloadable
library raises EngineUnavailable.
"""

import ctypes
import os

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphere_hip.so")
NC = 32

_LIB = None


class EngineUnavailable(RuntimeError):
    pass


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sphere_hip error %d: %s" % (code, msg))
        self.code = code


# Field ids and shapes (include/sphere_hip.h, enum sph_field).
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

SCALARS = {"time": 0, "time_total": 1, "dt": 2, "file_dt": 3, "iter": 4, "step_count": 5,
           "np": 6, "nw": 7, "nb0": 8, "ncells": 9, "vl_skin": 10, "sorts": 11, "vl_builds": 12,
           "nsims": 13, "contact_lanes": 14, "sort_mode": 15, "resort_merge_max": 16,
           "graph_steps": 17, "xcd_order": 18, "adj_checkpoint": 19, "fused": 20,
           "wall_sums": 21, "k_n": 22, "integrate_lanes": 23, "bond_k": 24,
           "bond_break_ext": 25, "contact_wg": 26, "vl_cap": 27, "radix_bits": 28}

# Exported symbols of the C ABI (include/sphere_hip.h).
SYMBOLS = ("sph_open", "sph_open_mem", "sph_step", "sph_sync", "sph_run", "sph_write_bin",
           "sph_write_mem", "sph_get_field", "sph_set_field", "sph_field_bytes",
           "sph_get_scalar", "sph_set_scalar", "sph_get_contacts", "sph_get_cells",
           "sph_profile", "sph_profile_report", "sph_close", "sph_last_error",
           "sph_open_error", "sph_version_string", "sph_open_batch", "sph_open_batch_mem",
           "sph_num_sims", "sph_write_bin_sim", "sph_write_mem_sim", "sph_get_field_sim",
           "sph_set_field_sim", "sph_field_bytes_sim", "sph_get_scalar_sim",
           "sph_set_scalar_sim", "sph_adjoint", "sph_adjoint_samples", "sph_adjoint_batch",
           "sph_adjoint_samples_batch", "sph_adjoint_record", "sph_adjoint_reverse",
           "sph_selftest_sqrt32",
           "sph_write_state", "sph_read_state", "sph_write_state_sim", "sph_read_state_sim",
           "sph_find_overlaps", "sph_porosity", "sph_autotune", "sph_porosity_noalloc")

# Parameters of sph_adjoint's g_params, in order.
ADJOINT_PARAMS = ("k_n", "k_t", "gamma_n", "gamma_t", "mu_s", "mu_d", "gamma_wn", "gamma_wt",
                  "mu_wd")

QUIET = 0x1
NOCHECK = 0x2
SIDECAR = 0x4


def load():
    """Load libsphere_hip.so (raises EngineUnavailable if absent)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise EngineUnavailable(
            "libsphere_hip.so is not built (%s); run `python -m sphere_amd.build` "
            "or __graft_entry__.build()" % LIB_PATH)
    try:
        L = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise EngineUnavailable("cannot load %s: %s" % (LIB_PATH, e))
    vp = ctypes.c_void_p
    L.sph_open.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint, ctypes.POINTER(vp)]
    L.sph_open_mem.argtypes = [vp, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint,
                               ctypes.POINTER(vp)]
    L.sph_step.argtypes = [vp, ctypes.c_uint64]
    L.sph_sync.argtypes = [vp]
    L.sph_run.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p]
    L.sph_write_bin.argtypes = [vp, ctypes.c_char_p]
    L.sph_write_mem.argtypes = [vp, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.sph_get_field.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t]
    L.sph_set_field.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t]
    L.sph_field_bytes.argtypes = [vp, ctypes.c_int]
    L.sph_field_bytes.restype = ctypes.c_int64
    L.sph_get_scalar.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.sph_set_scalar.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    L.sph_get_contacts.argtypes = [vp, vp, vp, vp]
    L.sph_get_cells.argtypes = [vp, vp, vp, vp, vp]
    L.sph_profile.argtypes = [vp, ctypes.c_int]
    L.sph_profile_report.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    L.sph_close.argtypes = [vp]
    L.sph_close.restype = None
    L.sph_last_error.argtypes = [vp]
    L.sph_last_error.restype = ctypes.c_char_p
    L.sph_open_error.restype = ctypes.c_char_p
    L.sph_version_string.restype = ctypes.c_char_p
    L.sph_open_batch.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_int,
                                 ctypes.c_uint, ctypes.POINTER(vp)]
    L.sph_open_batch_mem.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t),
                                     ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                     ctypes.POINTER(vp)]
    L.sph_num_sims.argtypes = [vp]
    L.sph_write_bin_sim.argtypes = [vp, ctypes.c_int, ctypes.c_char_p]
    L.sph_find_overlaps.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_uint64,
                                    ctypes.POINTER(ctypes.c_uint64)]
    L.sph_porosity.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
    L.sph_porosity_noalloc.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
    L.sph_write_state.argtypes = [vp, ctypes.c_char_p]
    L.sph_read_state.argtypes = [vp, ctypes.c_char_p]
    L.sph_write_state_sim.argtypes = [vp, ctypes.c_int, ctypes.c_char_p]
    L.sph_read_state_sim.argtypes = [vp, ctypes.c_int, ctypes.c_char_p]
    L.sph_write_mem_sim.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t,
                                    ctypes.POINTER(ctypes.c_size_t)]
    L.sph_get_field_sim.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_size_t]
    L.sph_set_field_sim.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_size_t]
    L.sph_field_bytes_sim.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.sph_field_bytes_sim.restype = ctypes.c_int64
    L.sph_get_scalar_sim.argtypes = [vp, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double)]
    L.sph_set_scalar_sim.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.sph_adjoint.argtypes = [vp, ctypes.c_uint64, vp, vp, vp, vp, vp, vp, vp]
    L.sph_adjoint_batch.argtypes = [vp, ctypes.c_uint64, vp, vp, vp, vp, vp, vp, vp]
    L.sph_adjoint_samples_batch.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, vp, vp, vp,
                                            vp, vp, vp]
    L.sph_autotune.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    L.sph_adjoint_samples.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, vp, vp, vp, vp, vp,
                                      vp]
    L.sph_adjoint_record.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, vp]
    L.sph_adjoint_reverse.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, vp, vp, vp, vp, vp,
                                      vp]
    L.sph_selftest_sqrt32.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_uint64)]
    _LIB = L
    return L


class Context:
    """Simulations resident on the GPU (a sph_ctx): one .bin image (or path),
    or a batch of images stepped together.  Field and scalar accessors take
    the simulation index `sim` (default 0)."""

    def __init__(self, image=None, path=None, device=-1, quiet=True, check=True, images=None,
                 paths=None, sidecar=False):
        L = load()
        h = ctypes.c_void_p()
        flags = (QUIET if quiet else 0) | (0 if check else NOCHECK) | (SIDECAR if sidecar else 0)
        if paths is not None:
            arr = (ctypes.c_char_p * len(paths))(*[p.encode() for p in paths])
            rc = L.sph_open_batch(arr, len(paths), device, flags, ctypes.byref(h))
        elif images is not None:
            bufs = [ctypes.create_string_buffer(bytes(im), len(im)) for im in images]
            ptrs = (ctypes.c_void_p * len(bufs))(*[ctypes.addressof(b) for b in bufs])
            sizes = (ctypes.c_size_t * len(bufs))(*[len(im) for im in images])
            rc = L.sph_open_batch_mem(ptrs, sizes, len(bufs), device, flags, ctypes.byref(h))
        elif path is not None:
            rc = L.sph_open(path.encode(), device, flags, ctypes.byref(h))
        else:
            buf = ctypes.create_string_buffer(bytes(image), len(image))
            rc = L.sph_open_mem(buf, len(image), device, flags, ctypes.byref(h))
        if rc != 0:
            raise EngineError(rc, L.sph_open_error().decode())
        self._h = h
        self._L = L
        self.nsims = L.sph_num_sims(h)

    def _chk(self, rc):
        if rc != 0:
            raise EngineError(rc, self._L.sph_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.sph_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, n=1, sync=True):
        self._chk(self._L.sph_step(self._h, int(n)))
        if sync:
            self.sync()

    def sync(self):
        self._chk(self._L.sph_sync(self._h))

    def get(self, name, sim=0):
        fid, dt, shape = FIELDS[name]
        nbytes = self._L.sph_field_bytes_sim(self._h, sim, fid)
        if nbytes < 0:
            raise EngineError(int(nbytes), "bad simulation index or field")
        out = numpy.empty(nbytes // numpy.dtype(dt).itemsize, dtype=dt)
        self._chk(self._L.sph_get_field_sim(self._h, sim, fid, out.ctypes.data, nbytes))
        return out.reshape(shape)

    def set(self, name, value, sim=0):
        fid, dt, shape = FIELDS[name]
        a = numpy.ascontiguousarray(value, dtype=dt)
        self._chk(self._L.sph_set_field_sim(self._h, sim, fid, a.ctypes.data, a.nbytes))

    def scalar(self, name, sim=0):
        v = ctypes.c_double()
        self._chk(self._L.sph_get_scalar_sim(self._h, sim, SCALARS[name], ctypes.byref(v)))
        return v.value

    def set_scalar(self, name, value, sim=0):
        self._chk(self._L.sph_set_scalar_sim(self._h, sim, SCALARS[name], float(value)))

    def tobytes(self, sim=0):
        n = ctypes.c_size_t()
        self._chk(self._L.sph_write_mem_sim(self._h, sim, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._chk(self._L.sph_write_mem_sim(self._h, sim, buf, n.value, ctypes.byref(n)))
        return buf.raw

    def write_bin(self, path, sim=0):
        self._chk(self._L.sph_write_bin_sim(self._h, sim, path.encode()))

    def find_overlaps(self, sim=0):
        """DEM::findOverlaps on the GPU (sph_find_overlaps): (pairs (k, 2),
        i < j, ascending; overlaps (k,), negative)."""
        n = ctypes.c_uint64()
        self._chk(self._L.sph_find_overlaps(self._h, sim, None, None, 0, ctypes.byref(n)))
        pairs = numpy.zeros((n.value, 2), dtype=numpy.uint32)
        delta = numpy.zeros(n.value)
        self._chk(self._L.sph_find_overlaps(self._h, sim, pairs.ctypes.data, delta.ctypes.data,
                                            n.value, ctypes.byref(n)))
        return pairs, delta

    def porosity(self, z_slices=10, sim=0):
        """DEM::porosity on the GPU (sph_porosity): (z positions, porosities)."""
        z = numpy.zeros(int(z_slices))
        por = numpy.zeros(int(z_slices))
        self._chk(self._L.sph_porosity(self._h, sim, int(z_slices), z.ctypes.data,
                                       por.ctypes.data))
        return z, por

    def write_state(self, path, sim=0):
        """Restart sidecar (sph_write_state): the contact rows, displacements,
        accelerations, wall accelerations and loop counter a .bin lacks."""
        self._chk(self._L.sph_write_state_sim(self._h, sim, path.encode()))

    def read_state(self, path, sim=0):
        self._chk(self._L.sph_read_state_sim(self._h, sim, path.encode()))

    def run(self, out_dir, sid):
        self._chk(self._L.sph_run(self._h, out_dir.encode(), sid.encode()))

    def rows(self):
        """Particle rows of the host arrays the adjoint takes and returns:
        every simulation's particles, in simulation order."""
        return int(sum(self.scalar("np", sim=k) for k in range(self.nsims)))

    def _adjoint_result(self, outs, gp):
        res = {"x": outs[0], "v": outs[1], "w": outs[2]}
        if self.nsims == 1:
            res.update({name: float(v) for name, v in zip(ADJOINT_PARAMS, gp[0])})
        else:
            res.update({name: gp[:, k].copy() for k, name in enumerate(ADJOINT_PARAMS)})
        return res

    def adjoint(self, nsteps, dL_dx=None, dL_dv=None, dL_dw=None):
        """Gradient of L = sum dL_dx.x + dL_dv.v + dL_dw.w after `nsteps` steps
        with respect to the state before them and to the contact parameters
        (sph_adjoint; sph_adjoint_batch for a batch, whose particle arrays
        concatenate the simulations).  Advances the simulation by nsteps.
        Returns a dict with "x", "v", "w" (rows x 3) and one entry per
        ADJOINT_PARAMS name (a float, or one value per simulation)."""
        n = self.rows()

        def seed(a):
            if a is None:
                return None
            return numpy.ascontiguousarray(a, dtype=numpy.float64).reshape(n, 3)

        seeds = [seed(a) for a in (dL_dx, dL_dv, dL_dw)]
        outs = [numpy.zeros((n, 3)) for _ in range(3)]
        gp = numpy.zeros((self.nsims, len(ADJOINT_PARAMS)))
        ptr = [None if a is None else a.ctypes.data for a in seeds]
        fn = self._L.sph_adjoint if self.nsims == 1 else self._L.sph_adjoint_batch
        self._chk(fn(self._h, int(nsteps), ptr[0], ptr[1], ptr[2], outs[0].ctypes.data,
                     outs[1].ctypes.data, outs[2].ctypes.data, gp.ctypes.data))
        return self._adjoint_result(outs, gp)

    def adjoint_samples(self, every, nsamples, dL_dx=None, dL_dv=None):
        """Gradient of L = sum_k dL_dx[k] . x(k) + dL_dv[k] . v(k), x(k) and
        v(k) the state after (k + 1) * every steps (sph_adjoint_samples, or
        sph_adjoint_samples_batch for a batch), with respect to the state
        now and to the contact parameters.  Advances the simulation by
        every * nsamples steps; returns the same dict as adjoint()."""
        n = self.rows()

        def seeds(a):
            if a is None:
                return None
            return numpy.ascontiguousarray(a, dtype=numpy.float64).reshape(int(nsamples), n, 3)

        sx, sv = seeds(dL_dx), seeds(dL_dv)
        outs = [numpy.zeros((n, 3)) for _ in range(3)]
        gp = numpy.zeros((self.nsims, len(ADJOINT_PARAMS)))
        fn = (self._L.sph_adjoint_samples if self.nsims == 1
              else self._L.sph_adjoint_samples_batch)
        self._chk(fn(self._h, int(every), int(nsamples), None if sx is None else sx.ctypes.data,
                     None if sv is None else sv.ctypes.data, outs[0].ctypes.data,
                     outs[1].ctypes.data, outs[2].ctypes.data, gp.ctypes.data))
        return self._adjoint_result(outs, gp)

    def adjoint_record(self, every, nsamples):
        """The forward half of adjoint_samples (sph_adjoint_record): steps
        every * nsamples steps recording the tape and returns the positions
        after every, 2 every, ... steps, shape (nsamples, rows, 3)."""
        out = numpy.zeros((int(nsamples), self.rows(), 3))
        self._chk(self._L.sph_adjoint_record(self._h, int(every), int(nsamples),
                                             out.ctypes.data))
        return out

    def adjoint_reverse(self, every, nsamples, dL_dx=None, dL_dv=None):
        """The reverse half (sph_adjoint_reverse) of the run adjoint_record
        recorded: the same dict as adjoint_samples, bit for bit."""
        n = self.rows()

        def seeds(a):
            if a is None:
                return None
            return numpy.ascontiguousarray(a, dtype=numpy.float64).reshape(int(nsamples), n, 3)

        sx, sv = seeds(dL_dx), seeds(dL_dv)
        outs = [numpy.zeros((n, 3)) for _ in range(3)]
        gp = numpy.zeros((self.nsims, len(ADJOINT_PARAMS)))
        self._chk(self._L.sph_adjoint_reverse(
            self._h, int(every), int(nsamples), None if sx is None else sx.ctypes.data,
            None if sv is None else sv.ctypes.data, outs[0].ctypes.data, outs[1].ctypes.data,
            outs[2].ctypes.data, gp.ctypes.data))
        return self._adjoint_result(outs, gp)

    def autotune(self, steps_per_trial=50):
        """sph_autotune: choose the contact lanes, the contact workgroup
        size, the candidate-list capacity, the radix digit width and the
        workgroup order by timing the contact kernel over steps of this
        simulation, and the radix digit width by timing sorts (identical
        results for every setting; the simulation advances).  Returns the
        contact kernel's best mean time (us); the scalars then read the
        chosen settings."""
        us = ctypes.c_double()
        self._chk(self._L.sph_autotune(self._h, int(steps_per_trial), ctypes.byref(us)))
        return us.value

    def profile(self, enable=True):
        self._chk(self._L.sph_profile(self._h, 1 if enable else 0))

    def profile_report(self):
        buf = ctypes.create_string_buffer(1 << 16)
        self._L.sph_profile_report(self._h, buf, len(buf))
        rep = {}
        for line in buf.value.decode().splitlines():
            name, us, n = line.split()
            rep[name] = (float(us), int(n))
        return rep


def run_file(path, out_dir, sid, verbose=True, dry=False, device=-1, sidecar=False,
             scalars=None):
    """Open a .bin and run DEM::startTime's loop; returns 0 or a negative code.
    With `sidecar`, a restart sidecar next to the input is read and one is
    written next to every output.  `scalars` ({name: value}, SCALARS names)
    are set after opening (options the .bin cannot carry)."""
    try:
        ctx = Context(path=path, device=device, quiet=not verbose, sidecar=sidecar)
    except EngineError as e:
        print(str(e))
        return e.code
    try:
        for name, value in (scalars or {}).items():
            ctx.set_scalar(name, value)
        if dry:
            return 0
        ctx.run(out_dir, sid)
        return 0
    except EngineError as e:
        print(str(e))
        return e.code
    finally:
        ctx.close()


def selftest_sqrt32(lo=0, hi=1 << 32, device=0):
    """Mismatches of the engine's FP32 square root against the correctly
    rounded one over the FP32 bit patterns [lo, hi)."""
    L = load()
    n = ctypes.c_uint64()
    rc = L.sph_selftest_sqrt32(device, lo, hi, ctypes.byref(n))
    if rc:
        raise EngineError(rc, "sph_selftest_sqrt32 failed (%d)" % rc)
    return n.value
