"""Point location and every receiver channel kind of the HIP path against the
FP64 oracle, through the C ABI (reference src/config/point.f90:150-254,
src/config/channel_mod.f90:270-336,505-1095)."""
import math

import numpy as np
import pytest

import oracle
from helpers import rel_l2, small_grid
from ses3d_amd import api, synth

pytestmark = pytest.mark.gpu

EARTH_RADIUS = 6371000.0


def _model(shape, rho=2700.0, vp=6000.0, vs=3500.0):
    mu = rho * vs * vs
    return {"rhoinv": np.full(shape, 1.0 / rho), "lambda": np.full(shape, rho * vp * vp - 2 * mu),
            "mu": np.full(shape, mu)}


def _pair(grid, precision="fp64", **kw):
    o = oracle.Oracle(grid)
    s = api.Simulation(grid, precision, **kw)
    for name, arr in _model(o.shape).items():
        o.set_model({"rhoinv": oracle.RHOINV, "lambda": oracle.LAMBDA, "mu": oracle.MU}[name], arr)
        s.set_model({"rhoinv": api.RHOINV, "lambda": api.LAMBDA, "mu": api.MU}[name], arr)
    o.init()
    s.init()
    return o, s


@pytest.mark.parametrize("p", [2, 4, 7])
def test_locate_matches_oracle(p):
    """s3d_locate against ora_locate: random points and points at GLL nodes
    (element faces, the free surface), the same element and weights."""
    grid = small_grid(nx=5, ny=4, nz=3, lpd=p, nt=4, taper_width=0)
    o, s = _pair(grid)
    rng = np.random.default_rng(p)
    pts = list(zip(rng.uniform(grid.lat_min, grid.lat_max, 80), rng.uniform(grid.lon_min, grid.lon_max, 80),
                   rng.uniform(0.0, grid.rad_max - grid.rad_min, 80)))
    th, ph, r = o.coords_1d()
    for q in range(40):
        ex, ey, ez = rng.integers(0, grid.nx), rng.integers(0, grid.ny), rng.integers(0, grid.nz)
        i, j, k = rng.integers(0, p + 1, 3)
        pts.append((90.0 - math.degrees(th[i, ex]), math.degrees(ph[j, ey]), EARTH_RADIUS - r[k, ez]))
    pts.append((0.5 * (grid.lat_min + grid.lat_max), 0.5 * (grid.lon_min + grid.lon_max), 0.0))
    n_in = 0
    for lat, lon, depth in pts:
        a, b = o.locate(lat, lon, depth), s.locate(lat, lon, depth)
        assert (a is None) == (b is None), (lat, lon, depth)
        if a is None:
            continue
        n_in += 1
        assert np.array_equal(a[0], b[0]), (lat, lon, depth)
        assert np.abs(a[1] - b[1]).max() < 1e-14, (lat, lon, depth)
    assert n_in > 100


def _node_point(o, e, ijk):
    th, ph, r = o.coords_1d()
    return (90.0 - math.degrees(th[ijk[0], e[0]]), math.degrees(ph[ijk[1], e[1]]),
            EARTH_RADIUS - r[ijk[2], e[2]])


ECHO = ["mxx", "mxy", "mxz", "myy", "myz", "mzz", "fx", "fy", "fz"]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-4)])
def test_all_channel_kinds_match_oracle(precision, tol):
    """Moment-tensor and single-force sources (two moment tensors in one
    element: the later overwrites, its zero components are dropped), and
    receivers recording vx vy vz n e z pressure and the nine source echoes,
    one of them with 16 channels; GPU traces against the oracle. The echo
    channels are formed in FP64 from the FP64 source tables in both
    precisions, so they match to rounding."""
    nt = 150
    case = synth.homogeneous(nt=nt, nx=8, ny=8, nz=5, n_rec=4)
    g = case.grid
    o, s = _pair(g, precision)
    e_src = (3, 4, 1)
    rng = np.random.default_rng(11)
    stf = [oracle.stf("RICKER", 12.0 + 4 * q, 8.0, g.dt, nt) for q in range(3)]
    M1 = np.array([1.0, -0.4, 0.7, 0.2, -0.9, 0.5]) * 1e18
    M2 = np.array([0.0, 0.8, 0.0, -0.6, 0.0, 0.0]) * 1e18
    F = np.array([3.0, -1.0, 2.0]) * 1e12
    p1 = _node_point(o, e_src, (1, 2, 2))
    p2 = (p1[0] + 0.05, p1[1] - 0.05, p1[2] + 3000.0)
    srcs = [("mt", p1, np.outer(M1, stf[0])), ("sf", p2, np.outer(F, stf[1])),
            ("mt", p2, np.outer(M2, stf[2]))]
    for kind, pt, tab in srcs:
        a, b = o.locate(*pt), s.locate(*pt)
        assert tuple(a[0]) == e_src and tuple(b[0]) == e_src
        if kind == "mt":
            o.add_source_mt(a[0], a[1], tab)
            s.add_source_mt(b[0], b[1], tab)
        else:
            o.add_source_sf(a[0], a[1], tab)
            s.add_source_sf(b[0], b[1], tab)
    recs = [(p1, ECHO + ["vx", "n", "e", "z", "pressure", "vz", "mxy"]),
            (_node_point(o, e_src, (3, 1, 2)), ["mzz", "fy", "vy"])]
    for r in case.receivers:
        recs.append(((r.lat, r.lon, r.depth), ["vx", "vy", "vz", "n", "e", "z", "pressure", "mxx"]))
    for pt, chans in recs:
        a, b = o.locate(*pt), s.locate(*pt)
        o.add_receiver(a[0], a[1], chans)
        s.add_receiver(b[0], b[1], chans)
    o.step(nt)
    s.step(nt)
    for ri, (_, chans) in enumerate(recs):
        for ci, ch in enumerate(chans):
            ref, got = o.trace(ri, ci), s.trace(ri, ci)
            if ch in ECHO:
                if np.abs(ref).max() == 0:
                    assert not got.any(), (ri, ch)
                else:
                    assert rel_l2(got, ref) < 1e-13, (ri, ch)
            else:
                assert np.linalg.norm(ref) > 0, (ri, ch)
                assert rel_l2(got, ref) < tol, (ri, ch, rel_l2(got, ref))
    # the channels that alias each other
    assert np.array_equal(s.trace(0, 10), -s.trace(0, 9))       # n = -vx
    assert np.array_equal(s.trace(0, 14), s.trace(0, 12))       # z = vz
    assert np.array_equal(s.trace(0, 15), s.trace(0, 1))        # mxy twice


def test_echo_channels_in_a_batch_match_solo_runs():
    """Batched receiver launches read each member's own source tables."""
    case = synth.homogeneous(seed=4, nt=40, nx=6, ny=6, nz=4, n_rec=2)
    from dataclasses import replace
    src = case.sources[0]
    case = replace(case, receivers=case.receivers + [
        synth.Receiver(src.lat, src.lon, src.depth, "SRC", ("mxx", "myz", "vz"))])
    solo, recs = api.build_case(case, precision="fp32")
    solo.step(40)
    m, _ = api.build_case(case, precision="fp32")
    other, _ = api.build_case(synth.homogeneous(seed=5, nt=40, nx=5, ny=4, nz=4, n_rec=1), precision="fp32")
    b = api.Batch([m, other])
    b.step(40)
    b.synchronize()
    for r in range(len(recs)):
        for ch in range(len(recs[r].channels)):
            assert np.array_equal(m.trace(r, ch), solo.trace(r, ch)), (r, ch)
    assert np.abs(m.trace(len(recs) - 1, 0)).max() > 0
    b.close()


def test_batch_members_refuse_sources_receivers_and_solo_steps():
    """A batch caches its members' source and receiver buffers: while it
    exists, members refuse what would reallocate them, and solo steps."""
    case = synth.homogeneous(seed=3, nt=10, nx=4, ny=4, nz=4, n_rec=1)
    s, recs = api.build_case(case, precision="fp32")
    b = api.Batch([s])
    b.step(2)
    loc = s.locate(case.receivers[0].lat, case.receivers[0].lon, 0.0)
    with pytest.raises(api.S3dError, match="member of 1 batch"):
        s.add_receiver(loc[0], loc[1], ["vx"])
    with pytest.raises(api.S3dError, match="member of 1 batch"):
        s.add_source_mt(loc[0], loc[1], np.ones((6, case.grid.nt)))
    with pytest.raises(api.S3dError, match="member of 1 batch"):
        s.step(1)
    b.step(2)
    b.close()
    s.step(1)
    assert s.it == 5
    s.add_receiver(loc[0], loc[1], ["vx"])
    s.step(1)
    s.close()


def test_receiver_channel_limits():
    case = synth.homogeneous(seed=3, nt=10, nx=4, ny=4, nz=4, n_rec=1)
    s, _ = api.build_case(case, precision="fp32")
    loc = s.locate(case.receivers[0].lat, case.receivers[0].lon, 0.0)
    s.add_receiver(loc[0], loc[1], ["vx"] * 32)
    with pytest.raises(api.S3dError, match="1..32 channels"):
        s.add_receiver(loc[0], loc[1], ["vx"] * 33)
    s.close()
