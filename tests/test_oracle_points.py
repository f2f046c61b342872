"""Pins of the oracle's point location and of the receiver channels that
read sources (reference src/config/point.f90:150-254,553-657,826-889;
src/interpolation_mod.f90:83-205; src/config/channel_mod.f90:284-292,
505-545,859-1095).

What fixes them, independently of the oracle's own formulas:
  * a point placed at a GLL node of the mesh lies in an element holding that
    node, and its Lagrange weights are one-hot at that node; the global node
    index recovered from (element, local index) must be the node's own, so a
    mirrored reference coordinate or an element index counted from the wrong
    end fails;
  * Lagrange interpolation on p+1 GLL points is exact for polynomials of
    degree <= p, and the mapping from an element's reference cube to
    (theta, phi, r) is affine per axis: interpolating the node coordinates,
    or any tensor-product polynomial of degree <= p in them, at a located
    point returns the point's own value;
  * a source and a receiver at one interior GLL node: the echo channels must
    give back the injected glut, -M_ij stf(it) / w, with w the node's
    quadrature weight J W r^2 sin(theta) = -M/rho (the mass matrix, pinned by
    the volume integral elsewhere), or the injected force -F stf(it), at the
    step's own source-time sample; a receiver in another element reads zero;
  * n, e, z are -vx, vy, vz.
"""
import math

import numpy as np
import pytest

import oracle
from helpers import small_grid

EARTH_RADIUS = 6371000.0


def _oracle(grid, rho=2700.0, vp=6000.0, vs=3500.0):
    o = oracle.Oracle(grid)
    mu = rho * vs * vs
    o.set_model(oracle.RHOINV, np.full(o.shape, 1.0 / rho))
    o.set_model(oracle.LAMBDA, np.full(o.shape, rho * vp * vp - 2 * mu))
    o.set_model(oracle.MU, np.full(o.shape, mu))
    o.init()
    return o


def _grid(p, nt=10):
    # asymmetric element counts so that a transposed axis cannot pass
    return small_grid(nx=5, ny=4, nz=3, lpd=p, nt=nt, taper_width=0)


def _to_geo(theta, phi, r):
    return 90.0 - math.degrees(theta), math.degrees(phi), EARTH_RADIUS - r


def _global_nodes(o):
    """Global node coordinates along each axis: theta[gx], phi[gy], r[gz]."""
    th, ph, r = o.coords_1d()              # (p+1, n) each
    p = o.p

    def glob(a):
        n = a.shape[1]
        out = np.empty(n * p + 1)
        for e in range(n):
            out[e * p:e * p + p + 1] = a[:, e]
        return out
    return glob(th), glob(ph), glob(r)


@pytest.mark.parametrize("p", [2, 3, 4, 5, 6, 7])
def test_locate_at_gll_nodes_is_one_hot(p):
    g = _grid(p)
    o = _oracle(g)
    TH, PH, R = _global_nodes(o)
    P1 = p + 1
    # every theta and phi index strictly inside the domain (the outer faces are
    # outside the open interval the reference accepts); every r index from the
    # free surface (gz = 0, depth 0, accepted) down to the face above the
    # bottom (the bottom face itself is outside)
    gxs = range(1, g.nx * p)
    gys = range(1, g.ny * p)
    gzs = range(0, g.nz * p)
    rng = np.random.default_rng(p)
    picks = [(gx, gy, gz) for gx in gxs for gy in gys for gz in gzs]
    sel = rng.choice(len(picks), size=min(400, len(picks)), replace=False)
    # always include element faces in each direction and the surface
    extra = [(p, p, 0), (p, 1, p), (1, p, p), (2 * p, 2 * p, 2 * p), (1, 1, 0)]
    for q in list(sel) + extra:
        gx, gy, gz = picks[q] if not isinstance(q, tuple) else q
        lat, lon, depth = _to_geo(TH[gx], PH[gy], R[gz])
        loc = o.locate(lat, lon, depth)
        assert loc is not None, (gx, gy, gz)
        e, vm = loc
        vm = vm.reshape(P1, P1, P1)          # [k][j][i]
        kk, jj, ii = np.unravel_index(np.argmax(vm), vm.shape)
        one_hot = np.zeros_like(vm)
        one_hot[kk, jj, ii] = 1.0
        assert np.abs(vm - one_hot).max() < 1e-9, (gx, gy, gz, np.abs(vm - one_hot).max())
        # the node the weights select is the node the point was placed at
        assert (e[0] * p + ii, e[1] * p + jj, e[2] * p + kk) == (gx, gy, gz)


def test_points_outside_the_domain_are_rejected():
    g = _grid(4)
    o = _oracle(g)
    TH, PH, R = _global_nodes(o)
    mid = (TH[len(TH) // 2], PH[len(PH) // 2], R[len(R) // 2])
    for th, ph, r in ((TH[0], mid[1], mid[2]), (TH[-1], mid[1], mid[2]),
                      (mid[0], PH[0], mid[2]), (mid[0], PH[-1], mid[2]),
                      (mid[0], mid[1], R[-1]), (mid[0], mid[1], R[0] + 1.0)):
        assert o.locate(*_to_geo(th, ph, r)) is None
    assert o.locate(*_to_geo(*mid)) is not None


def _random_points(g, rng, n):
    lat = rng.uniform(g.lat_min, g.lat_max, n)
    lon = rng.uniform(g.lon_min, g.lon_max, n)
    depth = rng.uniform(0.0, g.rad_max - g.rad_min, n)
    return zip(lat, lon, depth)


@pytest.mark.parametrize("p", [2, 3, 4, 5, 6, 7])
def test_interpolation_reproduces_coordinates_and_polynomials(p):
    g = _grid(p)
    o = _oracle(g)
    th, ph, r = o.coords_1d()                 # (p+1, n)
    P1 = p + 1
    rng = np.random.default_rng(10 + p)
    # tensor-product polynomial of degree p in each normalised coordinate
    c = rng.standard_normal((P1, P1, P1))
    t0, t1 = th.min(), th.max()
    f0, f1 = ph.min(), ph.max()
    r0, r1 = r.min(), r.max()

    def poly(T, F, Rr):
        a = (T - t0) / (t1 - t0)
        b = (F - f0) / (f1 - f0)
        z = (Rr - r0) / (r1 - r0)
        return np.einsum("abc,a...,b...,c...->...", c,
                         np.stack([a ** q for q in range(P1)]),
                         np.stack([b ** q for q in range(P1)]),
                         np.stack([z ** q for q in range(P1)]))

    for lat, lon, depth in _random_points(g, rng, 60):
        loc = o.locate(lat, lon, depth)
        assert loc is not None
        e, vm = loc
        vm = vm.reshape(P1, P1, P1)           # [k][j][i]
        T = th[:, e[0]][None, None, :]
        F = ph[:, e[1]][None, :, None]
        Rr = r[:, e[2]][:, None, None]
        theta = math.radians(90.0 - lat)
        phi = math.radians(lon)
        rad = EARTH_RADIUS - depth
        assert abs((vm * T).sum() - theta) < 1e-13
        assert abs((vm * F).sum() - phi) < 1e-13
        assert abs((vm * Rr).sum() - rad) < 1e-13 * rad
        want = poly(np.array(theta), np.array(phi), np.array(rad))
        got = (vm * poly(T, F, Rr)).sum()
        assert abs(got - want) < 1e-10 * (1.0 + abs(want)), (got, want)


def _interior_node(o, g, e, ijk):
    """(lat, lon, depth) of local node ijk of element e."""
    th, ph, r = o.coords_1d()
    return _to_geo(th[ijk[0], e[0]], ph[ijk[1], e[1]], r[ijk[2], e[2]])


def test_echo_channels_give_back_the_injected_glut_and_force():
    p = 4
    nt = 12
    g = _grid(p, nt=nt)
    rho = 2700.0
    o = _oracle(g, rho=rho)
    e_src, node = (2, 1, 1), (1, 2, 3)        # an interior node of one element
    loc = o.locate(*_interior_node(o, g, e_src, node))
    assert loc is not None and tuple(loc[0]) == e_src
    vm = loc[1]
    assert abs(vm.max() - 1.0) < 1e-12
    # quadrature weight of the node: J W r^2 sin(theta) = -M / rho, M the mass
    # matrix (the node is inside one element, so M is that element's entry)
    P1 = p + 1
    mm = o.get(oracle.MM)[node[2], node[1], node[0], e_src[2], e_src[1], e_src[0]]
    w = -mm / rho
    rng = np.random.default_rng(5)
    s = rng.standard_normal(nt)
    M = np.array([1.0, -2.0, 3.0, 0.5, -1.5, 2.5]) * 1e17
    F = np.array([4.0, -5.0, 6.0]) * 1e10
    o.add_source_mt(loc[0], vm, np.outer(M, s))
    o.add_source_sf(loc[0], vm, np.outer(F, s))
    names = ["mxx", "mxy", "mxz", "myy", "myz", "mzz", "fx", "fy", "fz"]
    o.add_receiver(loc[0], vm, names)
    # a receiver in another element records zeros
    other = o.locate(*_interior_node(o, g, (0, 3, 2), (2, 2, 2)))
    o.add_receiver(other[0], other[1], ["mxx", "fz"])
    o.step(nt)
    for q in range(6):
        want = -M[q] * s / w
        assert np.abs(o.trace(0, q) - want).max() < 1e-12 * np.abs(want).max(), names[q]
    for q in range(3):
        want = -F[q] * s
        assert np.abs(o.trace(0, 6 + q) - want).max() < 1e-12 * np.abs(want).max(), names[6 + q]
    assert not o.trace(1, 0).any() and not o.trace(1, 1).any()


def test_later_source_overwrites_components_in_one_element():
    """add_m** assigns the element's glut (src/config/channel_mod.f90:866):
    with two moment tensors in one element, the later one's nonzero
    components replace the earlier one's; its all-zero components are dropped
    (src/config/source.f90:320-323) and leave the earlier values."""
    p = 4
    nt = 8
    g = _grid(p, nt=nt)
    o = _oracle(g)
    e_src, node = (1, 2, 1), (2, 1, 2)
    loc = o.locate(*_interior_node(o, g, e_src, node))
    vm = loc[1]
    s1 = np.linspace(1.0, 2.0, nt)
    s2 = np.linspace(-3.0, 1.0, nt)
    M1 = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0]) * 1e16
    M2 = np.array([0.0, -7.0, 0.0, 0.0, 9.0, 0.0]) * 1e16
    o.add_source_mt(loc[0], vm, np.outer(M1, s1))
    o.add_source_mt(loc[0], vm, np.outer(M2, s2))
    o.add_receiver(loc[0], vm, ["mxx", "mxy", "mxz", "myy", "myz", "mzz"])
    o.step(nt)
    ratio = [o.trace(0, q) / (-s1 if M2[q] == 0 else -s2) for q in range(6)]
    scale = ratio[0][0] / M1[0]               # 1 / w, the same for every component
    for q in range(6):
        want = (M1[q] if M2[q] == 0 else M2[q]) * scale
        assert np.allclose(ratio[q], want, rtol=1e-12, atol=0), q


def test_n_e_z_channels_are_minus_vx_vy_vz():
    p = 4
    nt = 40
    g = _grid(p, nt=nt)
    g.dt = 0.05
    o = _oracle(g)
    src = o.locate(*_interior_node(o, g, (2, 2, 1), (2, 2, 2)))
    s = np.sin(np.linspace(0.0, 3.0, nt)) ** 2
    o.add_source_mt(src[0], src[1], np.outer([1e18, 0.3e18, -0.2e18, 0.5e18, 0.1e18, -1e18], s))
    rec = o.locate(*_interior_node(o, g, (2, 1, 1), (1, 3, 2)))
    o.add_receiver(rec[0], rec[1], ["vx", "vy", "vz", "n", "e", "z"])
    o.step(nt)
    vx, vy, vz = o.trace(0, 0), o.trace(0, 1), o.trace(0, 2)
    assert np.abs(vx).max() > 0 and np.abs(vy).max() > 0 and np.abs(vz).max() > 0
    assert np.array_equal(o.trace(0, 3), -vx)
    assert np.array_equal(o.trace(0, 4), vy)
    assert np.array_equal(o.trace(0, 5), vz)
