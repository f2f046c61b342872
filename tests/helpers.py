"""Shared test helpers: seeded random states and error norms."""
import numpy as np

from ses3d_amd import synth


def small_grid(nx=3, ny=4, nz=2, lpd=4, nt=50, taper_width=1, dt=0.2):
    return synth.Grid(nx, ny, nz, lpd, -3.0, 4.0, 10.0, 18.0, 6071000.0,
                      6371000.0, taper_width, 1.2, dt, nt)


def unique_to_local(u, grid):
    """Global unique-node array u[gz, gy, gx] -> 6-D element-local copies."""
    p = grid.lpd
    P1 = p + 1
    i = np.arange(P1)
    gx = (np.arange(grid.nx)[None, :] * p + i[:, None])      # (P1, nx)
    gy = (np.arange(grid.ny)[None, :] * p + i[:, None])
    gz = (np.arange(grid.nz)[None, :] * p + i[:, None])
    # out[k, j, i, ez, ey, ex] = u[gz[k,ez], gy[j,ey], gx[i,ex]]
    return u[gz[:, None, None, :, None, None], gy[None, :, None, None, :, None],
             gx[None, None, :, None, None, :]]


def random_unique(grid, rng, scale=1.0):
    p = grid.lpd
    shp = (grid.nz * p + 1, grid.ny * p + 1, grid.nx * p + 1)
    return scale * rng.standard_normal(shp)


def random_model(grid, rng, visco=False, aniso=False):
    P1 = grid.lpd + 1
    shp = (P1, P1, P1, grid.nz, grid.ny, grid.nx)
    vp = rng.uniform(5500.0, 8000.0, shp)
    vs = vp / rng.uniform(1.65, 1.85, shp)
    rho = rng.uniform(2600.0, 3400.0, shp)
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    m = {"rhoinv": 1.0 / rho, "lambda": lam, "mu": mu}
    if visco:
        m["q"] = rng.uniform(80.0, 600.0, shp)
    if aniso:
        m["a"] = 0.02 * lam * rng.standard_normal(shp)
        m["b"] = 0.02 * mu * rng.standard_normal(shp)
        m["c"] = 0.02 * lam * rng.standard_normal(shp)
    return m


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


def local_to_unique(f, grid):
    """6-D element-local copies -> unique-node array u[gz, gy, gx], taking
    for every node the copy of the highest-index element that holds it."""
    p = grid.lpd
    P1 = p + 1
    out = np.zeros((grid.nz * p + 1, grid.ny * p + 1, grid.nx * p + 1))
    for k in range(P1):
        for j in range(P1):
            for i in range(P1):
                out[k:k + grid.nz * p:p, j:j + grid.ny * p:p, i:i + grid.nx * p:p] = f[k, j, i]
    return out


def node_coords(o):
    """(theta, phi, r) at every element-local node of an initialised oracle."""
    import oracle
    return o.get(oracle.THETA), o.get(oracle.PHI), o.get(oracle.R)


def write_config(case, d, precision="fp32", nt=None, extra="", model=None, workflow="forward"):
    """Write the reference's NAMELIST config and raw float64 model files for a
    synth.Case into directory d; returns the config path. With `model` (a
    dict of &model entries such as model_type and output file names) no
    model files are written and those entries take their place."""
    import os
    g = case.grid
    files = {}
    if model is None:
        m = case.model_full()
        for k in ("rhoinv", "lambda", "mu", "q"):
            if k in m:
                p = os.path.join(d, k)
                np.ascontiguousarray(m[k], dtype="<f8").tofile(p)
                files[k] = "'%s'" % p
    else:
        files = {k: (v if isinstance(v, str) and v.startswith(".") else repr(v))
                 for k, v in model.items()}
    lines = ["&grid", " nx = %d" % g.nx, " ny = %d" % g.ny, " nz = %d" % g.nz,
             " lpd = %d" % g.lpd, " taper_width = %d" % g.taper_width,
             " taper_slope = %r" % g.taper_slope, "/", "",
             "&model", " lat_min = %r, lat_max = %r," % (g.lat_min, g.lat_max),
             " lon_min = %r, lon_max = %r," % (g.lon_min, g.lon_max),
             " rad_min = %r, rad_max = %r," % (g.rad_min, g.rad_max)]
    for k, v in files.items():
        lines.append(" %s = %s" % (k, v))
    lines += ["/", "", "&time", " dt = %r" % float(g.dt), " nt = %d" % (nt or g.nt),
              " date_time = 2009, 8, 20, 0, 6, 35, 0, 0", "/", "",
              "&general", " workflow = '%s'" % workflow, " event_name = 'synthetic'",
              " log_file_dir = './'", "/", "",
              "&gpu", " precision = '%s'" % precision, "/", ""]
    for s in case.sources:
        lines += ["&source_mt", " lat = %r" % s.lat, " lon = %r" % s.lon,
                  " depth = %r" % s.depth, " wavelet = '%s'" % s.wavelet,
                  " onset = %r" % s.onset, " width = %r" % s.width,
                  " moment_tensor = " + " ".join("%r" % x for x in s.moment_tensor), "/", ""]
    for r in case.receivers:
        lines += ["&receiver", " lat = %r" % r.lat, " lon = %r" % r.lon,
                  " depth = %r" % r.depth, " network = 'XX'", " station = '%s'" % r.station,
                  " location = '00'",
                  " attributes = " + " ".join("'%s'" % c for c in r.channels),
                  " override = .TRUE.", " prefix = '%s/'" % d, "/", ""]
    path = os.path.join(d, "config")
    with open(path, "w") as f:
        f.write("\n".join(lines) + extra)
    return path


def read_sac(path):
    """(header floats, header ints, data) of a binary SAC file."""
    raw = open(path, "rb").read()
    f = np.frombuffer(raw[:280], dtype="<f4")
    i = np.frombuffer(raw[280:420], dtype="<i4")
    data = np.frombuffer(raw[632:], dtype="<f4")
    return f, i, data
