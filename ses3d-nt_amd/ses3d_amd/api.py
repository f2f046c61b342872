"""Python host interface over the C ABI in include/ses3d_hip.h.

This is the Python mirror of the reference's Fortran module API for the hot
path (evolution_mod, source/receiver channels, src/ses3d-nt.F90:240-270):
one `Simulation` per mesh, loaded with a model, sources and receivers, then
stepped. It loads the in-tree ses3d-nt_amd/libses3d_hip.so and fails loudly
when that library is missing: there is no CPU fallback.

6-D arrays are float64 numpy arrays of shape (p+1, p+1, p+1, nz, ny, nx) in C
order, byte-identical to the reference's column-major
f(0:nx-1, 0:ny-1, 0:nz-1, 0:p, 0:p, 0:p).
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# SES3D_HIP_LIB selects another build of the same library (e.g. a tuning
# variant); the default is the in-tree build.
LIB_PATH = os.environ.get("SES3D_HIP_LIB", os.path.join(_PKG, "libses3d_hip.so"))
_LIB = None

FP64, FP32 = 0, 1
RHOINV, LAMBDA, MU, A, B, C_, Q = range(7)
VX, VY, VZ = 0, 1, 2
SXX, SXY, SXZ, SYY, SYZ, SZZ = 3, 4, 5, 6, 7, 8
MM, TAPER = 18, 19
GRAD_RHO, GRAD_CS, GRAD_CP, GRAD_CSH, GRAD_CSV = 20, 21, 22, 23, 24
# symmetric displacement gradient (after enable_displacement_gradient)
EXX, EXY, EXZ, EYY, EYZ, EZZ = 26, 27, 28, 29, 30, 31
# volume curl fields (after capture_curl_strain and one step)
ROT_VX, ROT_VY, ROT_VZ = 50, 51, 52
MXX = 32
CH = {"vx": 0, "vy": 1, "vz": 2, "n": 3, "e": 4, "z": 5, "pressure": 6,
      "rot_vx": 7, "rot_vy": 8, "rot_vz": 9,
      "mxx": 10, "mxy": 11, "mxz": 12, "myy": 13, "myz": 14, "mzz": 15,
      "fx": 16, "fy": 17, "fz": 18}
WAVELET = {"RICKER": 0, "GAUSS": 1, "SIN**3": 2, "BRUNE": 3, "DELTA": 4}

EXPORTS = [
    "s3d_last_error", "s3d_device_count", "s3d_create", "s3d_destroy", "s3d_create_shared",
    "s3d_set_model_slab", "s3d_set_model", "s3d_init", "s3d_cfl", "s3d_gll",
    "s3d_autotune", "s3d_locate", "s3d_wavelet", "s3d_add_source_mt", "s3d_add_source_sf",
    "s3d_add_receiver", "s3d_step", "s3d_current_step", "s3d_get_trace",
    "s3d_get_field", "s3d_set_field", "s3d_vrange", "s3d_kinetic_energy",
    "s3d_synchronize", "s3d_timing_enable", "s3d_kernel_times",
    "s3d_device_bytes", "s3d_batch_create", "s3d_batch_step", "s3d_set_tile",
    "s3d_batch_synchronize", "s3d_batch_destroy", "s3d_grad_add", "s3d_vrange_noalloc",
    "s3d_set_lds_layout", "s3d_get_tuning", "s3d_enable_displacement_gradient",
    "s3d_capture_curl_strain", "s3d_set_store_order", "s3d_get_store_order",
]


class S3dGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lpd", C.c_int),
                ("lat_min", C.c_double), ("lat_max", C.c_double),
                ("lon_min", C.c_double), ("lon_max", C.c_double),
                ("rad_min", C.c_double), ("rad_max", C.c_double),
                ("taper_width", C.c_int), ("taper_slope", C.c_double),
                ("dt", C.c_double), ("nt", C.c_int)]


class S3dOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("kahan_v", C.c_int),
                ("kahan_sigma", C.c_int), ("device", C.c_int),
                ("xcd_remap", C.c_int), ("bf16_static", C.c_int),
                ("state_buffers", C.c_int), ("memory_vars", C.c_int)]


class S3dError(RuntimeError):
    pass


def lib():
    """The HIP library; raises when it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise S3dError("libses3d_hip.so not built: run python -c "
                           "'import __graft_entry__ as g; g.build()'")
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        L.s3d_last_error.restype = C.c_char_p
        L.s3d_device_count.argtypes = [C.POINTER(C.c_int)]
        L.s3d_create.argtypes = [C.POINTER(S3dGrid), C.POINTER(S3dOptions), C.POINTER(vp)]
        L.s3d_destroy.argtypes = [vp]
        L.s3d_create_shared.argtypes = [vp, C.POINTER(vp)]
        L.s3d_set_model_slab.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
        L.s3d_set_model.argtypes = [vp, C.c_int, dp]
        L.s3d_init.argtypes = [vp]
        L.s3d_cfl.argtypes = [vp, C.POINTER(C.c_double)]
        L.s3d_autotune.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.s3d_set_tile.argtypes = [vp, C.c_int]
        L.s3d_set_lds_layout.argtypes = [vp, C.c_int]
        L.s3d_get_tuning.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.s3d_gll.argtypes = [C.c_int, dp, dp, dp]
        L.s3d_locate.argtypes = [vp, C.c_double, C.c_double, C.c_double, ip, dp]
        L.s3d_wavelet.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, dp]
        L.s3d_add_source_mt.argtypes = [vp, ip, dp, dp]
        L.s3d_add_source_sf.argtypes = [vp, ip, dp, dp]
        L.s3d_add_receiver.argtypes = [vp, ip, dp, C.c_int, ip, C.POINTER(C.c_int)]
        L.s3d_step.argtypes = [vp, C.c_int]
        L.s3d_current_step.argtypes = [vp, C.POINTER(C.c_int)]
        L.s3d_get_trace.argtypes = [vp, C.c_int, C.c_int, dp]
        L.s3d_get_field.argtypes = [vp, C.c_int, dp]
        L.s3d_set_field.argtypes = [vp, C.c_int, dp]
        L.s3d_grad_add.argtypes = [vp, dp, dp, dp, C.c_int]
        L.s3d_enable_displacement_gradient.argtypes = [vp]
        L.s3d_capture_curl_strain.argtypes = [vp]
        L.s3d_set_store_order.argtypes = [vp, C.c_int]
        L.s3d_get_store_order.argtypes = [vp, C.POINTER(C.c_int)]
        L.s3d_vrange.argtypes = [vp, dp]
        L.s3d_vrange_noalloc.argtypes = [vp, dp]
        L.s3d_kinetic_energy.argtypes = [vp, C.POINTER(C.c_double)]
        L.s3d_synchronize.argtypes = [vp]
        L.s3d_timing_enable.argtypes = [vp, C.c_int]
        L.s3d_kernel_times.argtypes = [vp, dp, ip]
        L.s3d_device_bytes.argtypes = [vp, C.POINTER(C.c_double)]
        L.s3d_batch_create.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(vp)]
        L.s3d_batch_step.argtypes = [vp, C.c_int]
        L.s3d_batch_synchronize.argtypes = [vp]
        L.s3d_batch_destroy.argtypes = [vp]
        _LIB = L
    return _LIB


def _chk(rc):
    if rc != 0:
        raise S3dError(lib().s3d_last_error().decode())
    return rc


def device_count():
    n = C.c_int(0)
    lib().s3d_device_count(C.byref(n))
    return n.value


def gll(p):
    k, w, d = np.zeros(p + 1), np.zeros(p + 1), np.zeros((p + 1) ** 2)
    _chk(lib().s3d_gll(p, k, w, d))
    return k, w, d.reshape(p + 1, p + 1)


def wavelet(name, onset, width, dt, nt):
    out = np.zeros(nt)
    _chk(lib().s3d_wavelet(WAVELET[name], onset, width, dt, nt, out))
    return out


class Simulation:
    """Device state of one simulation (one mesh, one model)."""

    def __init__(self, grid, precision="fp32", kahan_v=True, kahan_sigma=False,
                 device=0, xcd_remap=True, bf16_static=False, state_buffers=0,
                 memory_vars=0):
        """kahan_v, kahan_sigma: False/0 off, True/1 FP32 compensation
        words, 2 or "bf16" BF16 compensation words (velocity and stress). bf16_static: store the moduli,
        tau and dt/M in BF16 (FP32 only). state_buffers: 2 (or 0, the
        default) read the state from one block and write it to another, 1
        update one block in place (half the state memory). memory_vars: 6
        (or 0, the default) memory variables per relaxation mechanism, or 5,
        their deviatoric part (the zz field is then -(xx + yy), read only)."""
        g = S3dGrid(grid.nx, grid.ny, grid.nz, grid.lpd, grid.lat_min,
                    grid.lat_max, grid.lon_min, grid.lon_max, grid.rad_min,
                    grid.rad_max, grid.taper_width, grid.taper_slope, grid.dt,
                    grid.nt)
        ks = 2 if kahan_sigma == "bf16" else int(kahan_sigma)
        kv = 2 if kahan_v == "bf16" else int(kahan_v)
        o = S3dOptions(FP32 if precision == "fp32" else FP64, kv,
                       ks, int(device), int(xcd_remap), int(bf16_static), int(state_buffers),
                       int(memory_vars))
        h = C.c_void_p()
        _chk(lib().s3d_create(C.byref(g), C.byref(o), C.byref(h)))
        self._h = h
        self.grid = grid
        self.p = grid.lpd
        self.precision = precision
        self.shape = (grid.lpd + 1,) * 3 + (grid.nz, grid.ny, grid.nx)

    @classmethod
    def shared(cls, model):
        """A further simulation on `model`'s mesh and Earth model
        (s3d_create_shared): own state, sources and receivers; the static
        fields are `model`'s. Ready to step."""
        self = cls.__new__(cls)
        h = C.c_void_p()
        _chk(lib().s3d_create_shared(model._h, C.byref(h)))
        self._h = h
        self._model = model          # the model's simulation must outlive this one
        self.grid = model.grid
        self.p = model.p
        self.precision = model.precision
        self.shape = model.shape
        return self

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _chk(lib().s3d_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- setup -------------------------------------------------------------
    def set_model(self, param, field):
        _chk(lib().s3d_set_model(self._h, param, np.ascontiguousarray(field, np.float64).reshape(-1)))

    def set_model_slab(self, param, ez0, field):
        f = np.ascontiguousarray(field, np.float64)
        nez = f.shape[3]
        _chk(lib().s3d_set_model_slab(self._h, param, ez0, nez, f.reshape(-1)))

    def init(self):
        _chk(lib().s3d_init(self._h))

    def autotune(self):
        """(elements per workgroup, element kernel ms) of the fastest tile."""
        t, ms = C.c_int(), C.c_double()
        _chk(lib().s3d_autotune(self._h, C.byref(t), C.byref(ms)))
        return t.value, ms.value

    def set_tile(self, tile):
        """Element tile without timing (s3d_set_tile): 0 (default), 1, 2 or 4
        elements per workgroup."""
        _chk(lib().s3d_set_tile(self._h, int(tile)))

    def set_lds_layout(self, layout):
        """LDS staging layout of the FP32 element kernel (s3d_set_lds_layout):
        0 planar, 1 node-interleaved."""
        _chk(lib().s3d_set_lds_layout(self._h, int(layout)))

    def set_store_order(self, early):
        """Store order of the FP32 element kernel (s3d_set_store_order): 0
        the new state after the weak divergence, 1 before it."""
        _chk(lib().s3d_set_store_order(self._h, int(early)))

    def store_order(self):
        x = C.c_int()
        _chk(lib().s3d_get_store_order(self._h, C.byref(x)))
        return x.value

    def tuning(self):
        """(elements per workgroup, XCD-aware block order, LDS layout) in use."""
        t, r, l = C.c_int(), C.c_int(), C.c_int()
        _chk(lib().s3d_get_tuning(self._h, C.byref(t), C.byref(r), C.byref(l)))
        return t.value, r.value, l.value

    def cfl(self):
        x = C.c_double()
        _chk(lib().s3d_cfl(self._h, C.byref(x)))
        return x.value

    def locate(self, lat, lon, depth):
        e = np.zeros(3, np.int32)
        vm = np.zeros((self.p + 1) ** 3)
        rc = lib().s3d_locate(self._h, lat, lon, depth, e, vm)
        if rc == 1:
            return None
        _chk(rc)
        return e, vm

    def add_source_mt(self, exyz, vm, stf6):
        _chk(lib().s3d_add_source_mt(self._h, np.asarray(exyz, np.int32),
                                     np.ascontiguousarray(vm, np.float64),
                                     np.ascontiguousarray(stf6, np.float64).reshape(-1)))

    def add_source_sf(self, exyz, vm, stf3):
        _chk(lib().s3d_add_source_sf(self._h, np.asarray(exyz, np.int32),
                                     np.ascontiguousarray(vm, np.float64),
                                     np.ascontiguousarray(stf3, np.float64).reshape(-1)))

    def add_receiver(self, exyz, vm, channels):
        ch = np.array([CH[c] for c in channels], np.int32)
        rid = C.c_int()
        _chk(lib().s3d_add_receiver(self._h, np.asarray(exyz, np.int32),
                                    np.ascontiguousarray(vm, np.float64), len(ch), ch,
                                    C.byref(rid)))
        return rid.value

    # -- stepping and outputs ------------------------------------------------
    def step(self, n=1):
        _chk(lib().s3d_step(self._h, n))

    @property
    def it(self):
        x = C.c_int()
        _chk(lib().s3d_current_step(self._h, C.byref(x)))
        return x.value

    def trace(self, rec, ch):
        out = np.zeros(self.grid.nt)
        _chk(lib().s3d_get_trace(self._h, rec, ch, out))
        return out

    def get(self, field):
        out = np.zeros(self.shape)
        _chk(lib().s3d_get_field(self._h, field, out.reshape(-1)))
        return out

    def set(self, field, arr):
        _chk(lib().s3d_set_field(self._h, field, np.ascontiguousarray(arr, np.float64).reshape(-1)))

    def enable_displacement_gradient(self):
        """Integrate the displacement gradient from the next step on
        (s3d_enable_displacement_gradient); read it with get(EXX .. EZZ)."""
        _chk(lib().s3d_enable_displacement_gradient(self._h))

    def capture_curl_strain(self):
        """Store the strain-rate parts of the current velocity for the volume
        curl fields; after one more step get(ROT_VX .. ROT_VZ) returns them
        (s3d_capture_curl_strain)."""
        _chk(lib().s3d_capture_curl_strain(self._h))

    def grad_add(self, vx_fw, vy_fw, vz_fw, increment):
        """grad_rho += increment*dt*(v_fw . v)/J with a forward velocity
        snapshot in the 6-D layout (reference ses3d_grad,
        src/SOURCE/ses3d_grad.F90:212-213)."""
        f = [np.ascontiguousarray(x, np.float64).reshape(-1) for x in (vx_fw, vy_fw, vz_fw)]
        _chk(lib().s3d_grad_add(self._h, f[0], f[1], f[2], increment))

    def vrange(self):
        out = np.zeros(2)
        _chk(lib().s3d_vrange(self._h, out))
        return out

    def vrange_noalloc(self):
        """s3d_vrange without any allocation (fixed per-device scratch)."""
        out = np.zeros(2)
        _chk(lib().s3d_vrange_noalloc(self._h, out))
        return out

    def kinetic_energy(self):
        x = C.c_double()
        _chk(lib().s3d_kinetic_energy(self._h, C.byref(x)))
        return x.value

    def synchronize(self):
        _chk(lib().s3d_synchronize(self._h))

    def timing(self, enable=True):
        _chk(lib().s3d_timing_enable(self._h, int(enable)))

    def kernel_times(self):
        ms = np.zeros(3)
        cnt = np.zeros(3, np.int32)
        _chk(lib().s3d_kernel_times(self._h, ms, cnt))
        return ms, cnt

    def device_bytes(self):
        x = C.c_double()
        _chk(lib().s3d_device_bytes(self._h, C.byref(x)))
        return x.value


class Batch:
    """Several Simulations on one device advanced together (one launch per
    kernel class and GLL order per step, chunks replayed as one hipGraph)."""

    def __init__(self, sims):
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s._h.value for s in self.sims])
        h = C.c_void_p()
        _chk(lib().s3d_batch_create(len(self.sims), arr, C.byref(h)))
        self._h = h

    def step(self, n=1):
        _chk(lib().s3d_batch_step(self._h, n))

    def synchronize(self):
        _chk(lib().s3d_batch_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().s3d_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_case(case, precision="fp32", kahan_v=True, kahan_sigma=False,
               slab=None, xcd_remap=True, autotune=False, bf16_static=False, state_buffers=0,
               memory_vars=0):
    """Simulation loaded with a synth.Case (model, sources, receivers).
    The model is uploaded `slab` element layers at a time (all at once when
    None). Returns (simulation, list of receivers that were placed)."""
    g = case.grid
    sim = Simulation(g, precision, kahan_v, kahan_sigma, xcd_remap=xcd_remap,
                     bf16_static=bf16_static, state_buffers=state_buffers,
                     memory_vars=memory_vars)
    nz = g.nz
    step = slab or nz
    for ez0 in range(0, nz, step):
        nez = min(step, nz - ez0)
        m = case.model_slab(ez0, nez)
        sim.set_model_slab(RHOINV, ez0, m["rhoinv"])
        sim.set_model_slab(LAMBDA, ez0, m["lambda"])
        sim.set_model_slab(MU, ez0, m["mu"])
        if "q" in m:
            sim.set_model_slab(Q, ez0, m["q"])
        del m
    sim.init()
    if autotune:
        sim.autotune()
    return sim, _add_channels(sim, case, case.sources)


def _add_channels(sim, case, sources):
    """Locate and add `sources` and the case's receivers; returns the
    receivers that were placed."""
    g = case.grid
    for s in sources:
        loc = sim.locate(s.lat, s.lon, s.depth)
        if loc is None:
            continue
        f = wavelet(s.wavelet, s.onset, s.width, g.dt, g.nt)
        sim.add_source_mt(loc[0], loc[1], np.outer(np.asarray(s.moment_tensor), f))
    placed = []
    for r in case.receivers:
        loc = sim.locate(r.lat, r.lon, r.depth)
        if loc is None:
            continue
        sim.add_receiver(loc[0], loc[1], r.channels)
        placed.append(r)
    return placed


def build_shared_sources(case, precision="fp32", kahan_v=True, kahan_sigma=False,
                         slab=None, autotune=False, state_buffers=0, memory_vars=0):
    """One simulation per source of `case`, all on one copy of the case's
    model (the first owns it, the others are Simulation.shared): the
    many-sources-one-model workload. Returns (simulations, placed receivers)."""
    from dataclasses import replace
    first, placed = build_case(replace(case, sources=case.sources[:1]), precision, kahan_v,
                               kahan_sigma, slab=slab, autotune=autotune,
                               state_buffers=state_buffers, memory_vars=memory_vars)
    sims = [first]
    for s in case.sources[1:]:
        sim = Simulation.shared(first)
        _add_channels(sim, case, [s])
        sims.append(sim)
    return sims, placed
