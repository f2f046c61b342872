"""Seeded synthetic workloads: meshes, Earth models, sources and receivers.

Every case is described in the reference's own terms: the &grid / &model /
&time NAMELIST values (src/parser/parser.f90:148-260), 6-D model parameters
rhoinv, lambda, mu, q per element-local node (src/model_paras.F90:39-46), and
moment-tensor sources and receivers given by lat / lon / depth
(src/parser/parser.f90:265-420).

Model arrays are float64 numpy arrays of shape (p+1, p+1, p+1, nz, ny, nx) in
C order, i.e. byte-identical to the reference's column-major
f(0:nx-1, 0:ny-1, 0:nz-1, 0:p, 0:p, 0:p).

Workloads (all seeded with numpy.random.default_rng):
  homogeneous   12x12x8 elements, p = 4, 10 x 10 deg x 500 km, vp 6000,
                vs 3500, rho 2700, elastic, 1 source, 16 receivers
  layered       96x96x32 (or any size), p = 4, 40 x 40 deg x 1400 km,
                piecewise-linear 1-D model with a 35 km crust and 410 / 660 km
                discontinuities, Q_mu 300 (crust) / 600 (mantle), +-5 % smooth
                Gaussian random vs perturbation with 200 km correlation length
"""
from dataclasses import dataclass, field, replace
import math

import numpy as np

EARTH_RADIUS = 6371000.0   # src/parameters.F90:64


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    lpd: int
    lat_min: float
    lat_max: float
    lon_min: float
    lon_max: float
    rad_min: float
    rad_max: float
    taper_width: int
    taper_slope: float
    dt: float
    nt: int

    @property
    def n_elem(self):
        return self.nx * self.ny * self.nz

    @property
    def nodes(self):
        """element-local GLL nodes per step: n_elem * (p+1)^3"""
        return self.n_elem * (self.lpd + 1) ** 3

    @property
    def unique_nodes(self):
        p = self.lpd
        return (self.nx * p + 1) * (self.ny * p + 1) * (self.nz * p + 1)


@dataclass
class Source:
    lat: float
    lon: float
    depth: float
    moment_tensor: tuple          # M_xx M_xy M_xz M_yy M_yz M_zz
    wavelet: str = "RICKER"
    onset: float = 40.0
    width: float = 20.0


@dataclass
class Receiver:
    lat: float
    lon: float
    depth: float
    station: str = ""
    channels: tuple = ("vx", "vy", "vz")


@dataclass
class Case:
    name: str
    grid: Grid
    model: str                    # "homogeneous" | "layered"
    seed: int
    visco: bool
    sources: list = field(default_factory=list)
    receivers: list = field(default_factory=list)
    params: dict = field(default_factory=dict)
    _grf: object = None

    # -- model -------------------------------------------------------------
    def model_slab(self, ez0, nez):
        """rhoinv, lambda, mu, q for element layers ez0 .. ez0+nez-1.
        For the layered model the returned arrays are work buffers that the
        next call of the same slab size overwrites: copy them to keep them."""
        return model_slab(self, ez0, nez)

    def model_full(self):
        return model_slab(self, 0, self.grid.nz)


# ---------------------------------------------------------------------------
# GLL points and the reference coordinate tables
# ---------------------------------------------------------------------------

def gll_knots(p):
    """GLL points: -1, 1 and the roots of P_p' (src/gll_mod.f90:61-165)."""
    if p == 1:
        return np.array([-1.0, 1.0])
    c = np.zeros(p + 1)
    c[p] = 1.0
    inner = np.polynomial.legendre.legroots(np.polynomial.legendre.legder(c))
    return np.concatenate([[-1.0], np.sort(inner.real), [1.0]])


def coords_1d(p, elms, lo, hi):
    """coordinates_1d (src/coordinates_mod.f90:114-141): shape (p+1, elms)."""
    kn = 0.5 * (1.0 + gll_knots(p))
    delta = (hi - lo) / elms
    out = lo + (np.arange(elms)[None, :] + kn[:, None]) * delta
    out[p, elms - 1] = hi
    return out


def axes(grid):
    """theta (p+1, nx), phi (p+1, ny), r (p+1, nz) node coordinates."""
    d2r = math.pi / 180.0
    x0 = (90.0 - grid.lat_max) * d2r
    x1 = (90.0 - grid.lat_min) * d2r
    y0 = grid.lon_min * d2r
    y1 = grid.lon_max * d2r
    th = coords_1d(grid.lpd, grid.nx, x0, x0 + (x1 - x0))
    ph = coords_1d(grid.lpd, grid.ny, y0, y0 + (y1 - y0))
    r = coords_1d(grid.lpd, grid.nz, grid.rad_min + (grid.rad_max - grid.rad_min), grid.rad_min)
    return th, ph, r


def cfl_dt(grid, vp_of_r, courant=0.3):
    """dt such that the reference's cfl_loc (src/model_paras.F90:529-561)
    equals `courant`, for a model whose fastest speed at radius r is vp_of_r."""
    th, ph, r = axes(grid)
    vmax = vp_of_r(r)                              # (p+1, nz)
    q = 0.0
    dth = np.diff(th, axis=0)                      # (p, nx)
    dph = np.diff(ph, axis=0)
    dr = -np.diff(r, axis=0)                       # (p, nz) positive
    # theta: v(r_k) / (r_k * dtheta_i)
    q = max(q, np.max(vmax[:, :, None] / (r[:, :, None] * dth.min())))
    # phi: v / (r sin(theta) dphi); smallest sin(theta) on the section
    smin = np.sin(th).min()
    q = max(q, np.max(vmax / (r * smin * dph.min())))
    q = max(q, np.max(vmax[1:] / dr))
    return courant / q


# ---------------------------------------------------------------------------
# 1-D layered model
# ---------------------------------------------------------------------------

# depth ranges in km; (vp, vs, rho) at top and bottom of each layer, km/s, g/cm3
_LAYERS = [
    (0.0, 35.0, (5.80, 3.20, 2.60), (6.50, 3.70, 2.90), 300.0),
    (35.0, 410.0, (8.04, 4.48, 3.38), (8.91, 4.77, 3.54), 600.0),
    (410.0, 660.0, (9.13, 4.93, 3.72), (10.20, 5.52, 3.99), 600.0),
    (660.0, 3000.0, (10.75, 5.94, 4.38), (13.00, 7.00, 5.40), 600.0),
]


def layered_1d(r):
    """vp, vs, rho (SI) and Q_mu of the piecewise-linear layered model."""
    d = (EARTH_RADIUS - np.asarray(r, dtype=np.float64)) / 1000.0
    vp = np.empty_like(d)
    vs = np.empty_like(d)
    rho = np.empty_like(d)
    q = np.empty_like(d)
    for top, bot, a, b, qm in _LAYERS:
        m = (d >= top) & (d < bot) if top > 0 else (d < bot)
        f = (d[m] - top) / (bot - top)
        vp[m] = 1000.0 * (a[0] + f * (b[0] - a[0]))
        vs[m] = 1000.0 * (a[1] + f * (b[1] - a[1]))
        rho[m] = 1000.0 * (a[2] + f * (b[2] - a[2]))
        q[m] = qm
    return vp, vs, rho, q


class _GRF:
    """Smooth Gaussian random field on a coarse (theta, phi, r) lattice with a
    Gaussian covariance of correlation length `corr` metres (arc length at the
    surface for the angular axes), scaled so that max |field| = 1. Values at
    GLL nodes come from separable linear interpolation."""

    def __init__(self, grid, seed, corr=200e3, spacing=25e3):
        th, ph, r = axes(grid)
        self.t0, self.t1 = th.min(), th.max()
        self.p0, self.p1 = ph.min(), ph.max()
        self.r0, self.r1 = r.min(), r.max()
        Lt = (self.t1 - self.t0) * EARTH_RADIUS
        Lp = (self.p1 - self.p0) * EARTH_RADIUS
        Lr = self.r1 - self.r0
        n = [max(4, int(math.ceil(L / spacing)) + 1) for L in (Lt, Lp, Lr)]
        pad = int(math.ceil(2 * corr / spacing))
        rng = np.random.default_rng(seed + 7919)
        shp = [m + 2 * pad for m in n]
        w = rng.standard_normal(shp)
        k = [np.fft.fftfreq(m, d=spacing) * 2 * np.pi for m in shp]
        kk = (k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2)
        filt = np.exp(-kk * corr * corr / 8.0)
        g = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
        g = g[pad:pad + n[0], pad:pad + n[1], pad:pad + n[2]]
        g = g - g.mean()
        g = g / np.abs(g).max()
        self.g = g
        self.n = n

    @staticmethod
    def _weights(x, lo, hi, n):
        t = (np.asarray(x, dtype=np.float64).ravel() - lo) / (hi - lo) * (n - 1)
        t = np.clip(t, 0.0, n - 1 - 1e-9)
        i0 = np.floor(t).astype(np.int64)
        return i0, t - i0

    def sample(self, th, ph, r, scratch=None):
        """field on the tensor lattice th (a), ph (b), r (c) -> (c, b, a).
        `scratch` (a dict) keeps the work arrays between calls of one shape:
        reusing memory that is already mapped is much faster than fresh
        allocations for arrays of this size."""
        it, ft = self._weights(th, self.t0, self.t1, self.n[0])
        ip, fp = self._weights(ph, self.p0, self.p1, self.n[1])
        ir, fr = self._weights(r, self.r0, self.r1, self.n[2])
        sc = {} if scratch is None else scratch
        na, nb, nc = it.size, ip.size, ir.size
        key = (na, nb, nc)
        if sc.get("key") != key:
            sc.clear()
            sc["key"] = key
            sc["h1"] = np.empty((self.n[0], self.n[1], nc))
            sc["h2"] = np.empty((self.n[0], nb, nc))
            sc["h3"] = np.empty((nc, nb, na))
            sc["t"] = np.empty((nc, nb, na))
        g = self.g                                            # (t, p, r)
        h1 = sc["h1"]
        np.multiply(g[:, :, ir], (1.0 - fr), out=h1)
        h1 += g[:, :, ir + 1] * fr                            # (t, p, c)
        h2 = sc["h2"]
        np.multiply(h1[:, ip, :], (1.0 - fp)[None, :, None], out=h2)
        h2 += h1[:, ip + 1, :] * fp[None, :, None]            # (t, b, c)
        h2t = np.ascontiguousarray(h2.transpose(2, 1, 0))    # (c, b, t)
        h3, tmp = sc["h3"], sc["t"]
        np.take(h2t, it, axis=2, out=h3)
        h3 *= (1.0 - ft)[None, None, :]
        np.take(h2t, it + 1, axis=2, out=tmp)
        tmp *= ft[None, None, :]
        h3 += tmp
        return h3                                             # (c, b, a)


def model_slab(case, ez0, nez):
    g = case.grid
    P1 = g.lpd + 1
    shp = (P1, P1, P1, nez, g.ny, g.nx)
    if case.model == "homogeneous":
        vp, vs, rho = case.params["vp"], case.params["vs"], case.params["rho"]
        mu = rho * vs * vs
        lam = rho * vp * vp - 2.0 * mu
        out = {"rhoinv": np.full(shp, 1.0 / rho), "lambda": np.full(shp, lam),
               "mu": np.full(shp, mu)}
        if case.visco:
            out["q"] = np.full(shp, case.params.get("q", 600.0))
        return out
    if case.model != "layered":
        raise ValueError(case.model)
    th, ph, r = axes(g)
    rs = r[:, ez0:ez0 + nez]                       # (P1, nez)
    vp1, vs1, rho1, q1 = layered_1d(rs)            # (P1[k], nez)
    if case._grf is None:
        case._grf = _GRF(g, case.seed)
    sc = case.params.setdefault("_scratch", {})
    if sc.get("shape") != shp:
        sc.clear()
        sc["shape"] = shp
        for name in ("rhoinv", "lambda", "mu", "q"):
            sc[name] = np.zeros(shp)
        sc["grf"] = {}
    # lattice orders (k, ez), (j, ey), (i, ex) keep the ex axis contiguous
    d = case._grf.sample(th.ravel(), ph.ravel(), rs.ravel(), sc["grf"])
    d = d.reshape(P1, nez, P1, g.ny, P1, g.nx).transpose(0, 2, 4, 1, 3, 5)
    amp = case.params.get("dvs", 0.05)
    bshape = (P1, 1, 1, nez, 1, 1)
    mu, lam, rhoinv = sc["mu"], sc["lambda"], sc["rhoinv"]
    np.multiply(d, amp, out=mu)
    mu += 1.0
    mu *= vs1.reshape(bshape)                      # vs
    mu *= mu                                       # vs^2
    mu *= rho1.reshape(bshape)                     # rho vs^2
    np.multiply(mu, -2.0, out=lam)
    lam += (rho1 * vp1 * vp1).reshape(bshape)      # rho vp^2 - 2 mu
    rhoinv[...] = (1.0 / rho1).reshape(bshape)
    out = {"rhoinv": rhoinv, "lambda": lam, "mu": mu}
    if case.visco:
        sc["q"][...] = q1.reshape(bshape)
        out["q"] = sc["q"]
    return out


# ---------------------------------------------------------------------------
# sources and receivers
# ---------------------------------------------------------------------------

def random_moment_tensor(rng, scale=1e19):
    m = rng.standard_normal(6)
    return tuple(float(x) for x in scale * m / np.linalg.norm(m))


def receivers_uniform(rng, grid, n, channels=("vx", "vy", "vz"), margin=None):
    if margin is None:
        margin = grid.taper_width + 0.5
    # keep at least one element's width to draw from on small meshes
    mx = min(margin, 0.5 * (grid.nx - 1))
    my = min(margin, 0.5 * (grid.ny - 1))
    dlat = (grid.lat_max - grid.lat_min) / grid.nx
    dlon = (grid.lon_max - grid.lon_min) / grid.ny
    lat = rng.uniform(grid.lat_min + mx * dlat, grid.lat_max - mx * dlat, n)
    lon = rng.uniform(grid.lon_min + my * dlon, grid.lon_max - my * dlon, n)
    return [Receiver(float(a), float(b), 0.0, "S%03d" % i, tuple(channels))
            for i, (a, b) in enumerate(zip(lat, lon))]


# ---------------------------------------------------------------------------
# the workloads
# ---------------------------------------------------------------------------

def homogeneous(seed=0, nt=1000, nx=12, ny=12, nz=8, lpd=4, n_rec=16,
                visco=False, courant=0.3):
    """Homogeneous elastic section, 10 x 10 deg x 500 km (12x12x8, p = 4)."""
    g = Grid(nx, ny, nz, lpd, -5.0, 5.0, -5.0, 5.0, 5871000.0, 6371000.0,
             2, 1.2, 0.0, nt)
    vp = 6000.0
    g.dt = cfl_dt(g, lambda r: np.full_like(r, vp), courant)
    rng = np.random.default_rng(seed)
    src = Source(float(rng.uniform(-1.0, 1.0)), float(rng.uniform(-1.0, 1.0)), 50000.0,
                 random_moment_tensor(rng))
    recs = receivers_uniform(rng, g, n_rec)
    return Case("homogeneous", g, "homogeneous", seed, visco, [src], recs,
                {"vp": vp, "vs": 3500.0, "rho": 2700.0, "q": 600.0})


def layered(seed=1, nt=4000, nx=96, ny=96, nz=32, lpd=4, n_rec=256,
            visco=True, n_src=1, courant=0.3, lat=20.0, lon=20.0, depth_km=1400.0):
    """Layered visco-elastic section 40 x 40 deg x 1400 km."""
    g = Grid(nx, ny, nz, lpd, -lat, lat, -lon, lon, EARTH_RADIUS - depth_km * 1e3,
             EARTH_RADIUS, 3, 1.2, 0.0, nt)
    # fastest speed: vp of the layered model (the vs perturbation never makes
    # vs exceed vp)
    g.dt = cfl_dt(g, lambda r: layered_1d(r)[0], courant)
    rng = np.random.default_rng(seed)
    srcs = []
    for _ in range(n_src):
        srcs.append(Source(float(rng.uniform(-0.25 * lat, 0.25 * lat)),
                           float(rng.uniform(-0.25 * lon, 0.25 * lon)),
                           float(rng.uniform(20e3, 80e3)), random_moment_tensor(rng)))
    recs = receivers_uniform(rng, g, n_rec)
    return Case("layered", g, "layered", seed, visco, srcs, recs, {"dvs": 0.05})


def scaled(case, **kw):
    """Same workload with a different grid (sizes, nt); dt recomputed."""
    g = replace(case.grid, **kw)
    c = replace(case, grid=g, _grf=None)
    if case.model == "homogeneous":
        g.dt = cfl_dt(g, lambda r: np.full_like(r, case.params["vp"]))
    else:
        g.dt = cfl_dt(g, lambda r: layered_1d(r)[0])
    return c


def multi_source(seed=5, n_src=32, nt=4000, nx=128, ny=128, nz=48, lpd=4, n_rec=256):
    """Many seeded moment-tensor sources on one layered visco-elastic
    model (40 x 40 deg x 1400 km): each source is its own simulation sharing
    the model."""
    return layered(seed=seed, nt=nt, nx=nx, ny=ny, nz=nz, lpd=lpd, n_rec=n_rec, n_src=n_src)


def uneven_batch(seed=4, n_sims=64, nt=2000, n_rec=32, min_e=(8, 8, 8), max_e=(48, 48, 24),
                 orders=(4, 6)):
    """Independent regional simulations of uneven size: element counts drawn
    between min_e and max_e, GLL order from `orders`, each with its own seeded
    perturbed layered model, source and receivers. Element size is kept near
    that of the layered workload (0.42 deg x 0.42 deg x 43.75 km)."""
    rng = np.random.default_rng(seed)
    cases = []
    for s in range(n_sims):
        nx = int(rng.integers(min_e[0], max_e[0] + 1))
        ny = int(rng.integers(min_e[1], max_e[1] + 1))
        nz = int(rng.integers(min_e[2], max_e[2] + 1))
        p = int(rng.choice(orders))
        half_lat = 0.2083 * nx
        half_lon = 0.2083 * ny
        c = layered(seed=1000 + s, nt=nt, nx=nx, ny=ny, nz=nz, lpd=p, n_rec=n_rec,
                    lat=half_lat, lon=half_lon, depth_km=43.75 * nz)
        cases.append(c)
    return cases
