"""ctypes front end of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference the HIP path is compared with. The product
never imports it.

All 6-D arrays are numpy float64 arrays of shape (p+1, p+1, p+1, nz, ny, nx)
in C order, which is byte-identical to the reference's column-major
f(0:nx-1, 0:ny-1, 0:nz-1, 0:p, 0:p, 0:p) (src/elastic_vars.F90:112).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# field ids (s3d_oracle.h)
VX, VY, VZ = 0, 1, 2
SXX, SXY, SXZ, SYY, SYZ, SZZ = 3, 4, 5, 6, 7, 8
EXX, EYY, EZZ, EXY, EXZ, EYZ, EYX, EZX, EZY = range(9, 18)
MM, TAPER, THETA, PHI, R = 18, 19, 20, 21, 22
SX, SY, SZ = 23, 24, 25
MXX = 32
# displacement gradient d_a u_b at DXUX + 3*a + b (dxux dxuy dxuz dyux ... dzuz)
DXUX = 50
RHOINV, LAMBDA, MU, A, B, C_, Q = range(7)
CH = {"vx": 0, "vy": 1, "vz": 2, "n": 3, "e": 4, "z": 5, "pressure": 6,
      "rot_vx": 7, "rot_vy": 8, "rot_vz": 9,
      "mxx": 10, "mxy": 11, "mxz": 12, "myy": 13, "myz": 14, "mzz": 15,
      "fx": 16, "fy": 17, "fz": 18}
WAVELET = {"RICKER": 0, "GAUSS": 1, "SIN**3": 2, "BRUNE": 3, "DELTA": 4}


class OraConfig(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lpd", C.c_int),
                ("lat_min", C.c_double), ("lat_max", C.c_double),
                ("lon_min", C.c_double), ("lon_max", C.c_double),
                ("rad_min", C.c_double), ("rad_max", C.c_double),
                ("taper_width", C.c_int), ("taper_slope", C.c_double),
                ("dt", C.c_double), ("nt", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.ora_create.restype = C.c_void_p
        L.ora_create.argtypes = [C.POINTER(OraConfig)]
        L.ora_destroy.argtypes = [C.c_void_p]
        L.ora_set_threads.argtypes = [C.c_int]
        L.ora_gll.argtypes = [C.c_int, dp, dp, dp]
        L.ora_set_model.argtypes = [C.c_void_p, C.c_int, dp]
        L.ora_init.argtypes = [C.c_void_p]
        L.ora_cfl.restype = C.c_double
        L.ora_cfl.argtypes = [C.c_void_p]
        L.ora_locate.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, ip, dp]
        L.ora_lagrange.argtypes = [C.c_int, C.c_double, dp]
        L.ora_stf.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, dp]
        L.ora_add_source_mt.argtypes = [C.c_void_p, ip, dp, dp]
        L.ora_add_source_sf.argtypes = [C.c_void_p, ip, dp, dp]
        L.ora_add_receiver.argtypes = [C.c_void_p, ip, dp, C.c_int, ip]
        L.ora_step.argtypes = [C.c_void_p, C.c_int]
        L.ora_evolution.argtypes = [C.c_void_p]
        L.ora_get_field.argtypes = [C.c_void_p, C.c_int, dp]
        L.ora_set_field.argtypes = [C.c_void_p, C.c_int, dp]
        L.ora_get_trace.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.ora_current_it.argtypes = [C.c_void_p]
        L.ora_vrange.argtypes = [C.c_void_p, dp]
        L.ora_kinetic_energy.restype = C.c_double
        L.ora_kinetic_energy.argtypes = [C.c_void_p]
        L.ora_scalars.argtypes = [C.c_void_p, dp]
        L.ora_coords_1d.argtypes = [C.c_void_p, dp, dp, dp]
        L.ora_profiles_1d.argtypes = [C.c_void_p, dp, dp, dp]
        L.ora_visco_enabled.argtypes = [C.c_void_p]
        L.ora_enable_displacement_gradient.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def gll(p):
    k = np.zeros(p + 1)
    w = np.zeros(p + 1)
    d = np.zeros((p + 1) * (p + 1))
    if lib().ora_gll(p, k, w, d):
        raise ValueError("unsupported lpd %d" % p)
    # d[i + (p+1) j] = l_i'(x_j)  ->  D[j, i]
    return k, w, d.reshape(p + 1, p + 1)


def lagrange(p, x):
    out = np.zeros(p + 1)
    if lib().ora_lagrange(p, x, out):
        raise ValueError(p)
    return out


def stf(wavelet, onset, width, dt, nt):
    out = np.zeros(nt)
    if lib().ora_stf(WAVELET[wavelet], onset, width, dt, nt, out):
        raise ValueError(wavelet)
    return out


class Oracle:
    """One FP64 simulation with the reference's data layout and step order."""

    def __init__(self, grid, threads=None):
        cfg = OraConfig(grid.nx, grid.ny, grid.nz, grid.lpd,
                        grid.lat_min, grid.lat_max, grid.lon_min, grid.lon_max,
                        grid.rad_min, grid.rad_max, grid.taper_width,
                        grid.taper_slope, grid.dt, grid.nt)
        if threads:
            lib().ora_set_threads(int(threads))
        self._h = lib().ora_create(C.byref(cfg))
        if not self._h:
            raise ValueError("ora_create failed")
        self.grid = grid
        self.p = grid.lpd
        self.shape = (grid.lpd + 1,) * 3 + (grid.nz, grid.ny, grid.nx)

    def close(self):
        if self._h:
            lib().ora_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_model(self, param, field):
        f = np.ascontiguousarray(field, dtype=np.float64).reshape(-1)
        lib().ora_set_model(self._h, param, f)

    def init(self):
        lib().ora_init(self._h)

    def cfl(self):
        return lib().ora_cfl(self._h)

    def locate(self, lat, lon, depth):
        e = np.zeros(3, dtype=np.int32)
        vm = np.zeros((self.p + 1) ** 3)
        if lib().ora_locate(self._h, lat, lon, depth, e, vm):
            return None
        return e, vm

    def add_source_mt(self, exyz, vm, stf6):
        lib().ora_add_source_mt(self._h, np.asarray(exyz, np.int32), np.ascontiguousarray(vm, np.float64),
                                np.ascontiguousarray(stf6, np.float64).reshape(-1))

    def add_source_sf(self, exyz, vm, stf3):
        lib().ora_add_source_sf(self._h, np.asarray(exyz, np.int32), np.ascontiguousarray(vm, np.float64),
                                np.ascontiguousarray(stf3, np.float64).reshape(-1))

    def add_receiver(self, exyz, vm, channels):
        ch = np.array([CH[c] for c in channels], dtype=np.int32)
        return lib().ora_add_receiver(self._h, np.asarray(exyz, np.int32),
                                      np.ascontiguousarray(vm, np.float64), len(ch), ch)

    def step(self, n=1):
        return lib().ora_step(self._h, n)

    def evolution(self):
        lib().ora_evolution(self._h)

    def enable_displacement_gradient(self):
        """Integrate the displacement gradient (K3) from the next step on."""
        if lib().ora_enable_displacement_gradient(self._h):
            raise MemoryError("displacement gradient")

    def get(self, field):
        out = np.zeros(self.shape)
        if lib().ora_get_field(self._h, field, out.reshape(-1)):
            raise ValueError(field)
        return out

    def set(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        if lib().ora_set_field(self._h, field, a):
            raise ValueError(field)

    def trace(self, rec, ch):
        out = np.zeros(self.grid.nt)
        lib().ora_get_trace(self._h, rec, ch, out)
        return out

    def vrange(self):
        out = np.zeros(2)
        lib().ora_vrange(self._h, out)
        return out

    def kinetic_energy(self):
        return lib().ora_kinetic_energy(self._h)

    def scalars(self):
        out = np.zeros(10)
        lib().ora_scalars(self._h, out)
        keys = ["dx", "dy", "dz", "jac", "x_loc_min", "x_loc_max", "y_loc_min",
                "y_loc_max", "z_loc_min", "z_loc_max"]
        return dict(zip(keys, out))

    def coords_1d(self):
        g = self.grid
        P1 = g.lpd + 1
        th = np.zeros(g.nx * P1)
        ph = np.zeros(g.ny * P1)
        r = np.zeros(g.nz * P1)
        lib().ora_coords_1d(self._h, th, ph, r)
        return th.reshape(P1, g.nx), ph.reshape(P1, g.ny), r.reshape(P1, g.nz)

    def profiles_1d(self):
        g = self.grid
        P1 = g.lpd + 1
        a = np.zeros(g.nx * P1)
        b = np.zeros(g.ny * P1)
        c = np.zeros(g.nz * P1)
        lib().ora_profiles_1d(self._h, a, b, c)
        return a.reshape(P1, g.nx), b.reshape(P1, g.ny), c.reshape(P1, g.nz)

    @property
    def visco(self):
        return bool(lib().ora_visco_enabled(self._h))

    @property
    def it(self):
        return lib().ora_current_it(self._h)


def build_case(case, threads=None):
    """Oracle loaded with a synth.Case: model, sources, receivers, ready to
    step. Returns (oracle, receiver list with channel names)."""
    g = case.grid
    o = Oracle(g, threads=threads)
    m = case.model_full()
    o.set_model(RHOINV, m["rhoinv"])
    o.set_model(LAMBDA, m["lambda"])
    o.set_model(MU, m["mu"])
    if "q" in m:
        o.set_model(Q, m["q"])
    o.init()
    for s in case.sources:
        loc = o.locate(s.lat, s.lon, s.depth)
        if loc is None:
            continue
        f = stf(s.wavelet, s.onset, s.width, g.dt, g.nt)
        o.add_source_mt(loc[0], loc[1], np.outer(np.asarray(s.moment_tensor), f))
    recs = []
    for r in case.receivers:
        loc = o.locate(r.lat, r.lon, r.depth)
        if loc is None:
            continue
        o.add_receiver(loc[0], loc[1], r.channels)
        recs.append(r)
    return o, recs
