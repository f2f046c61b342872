"""Edge cases of the mesh and of the run against the FP64 oracle: a single
element, meshes one element thick along one or two axes (every node on a
domain face), tiles that overhang the last element, the plane-pair tile
order with an odd number of planes, and a run with neither sources nor
receivers."""
import numpy as np
import pytest

from helpers import random_model, rel_l2, small_grid
from ses3d_amd import api
from test_gpu_parity import FIELDS, _pair, _random_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 1, 1), (1, 3, 2), (3, 2, 1), (5, 1, 3)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32", 1e-5)])
def test_tiny_and_thin_meshes_match_oracle(shape, precision, tol):
    nx, ny, nz = shape
    grid = small_grid(nx=nx, ny=ny, nz=nz, lpd=4, taper_width=0)
    rng = np.random.default_rng(40 + nx + 3 * ny + 7 * nz)
    o, s = _pair(grid, random_model(grid, rng, visco=True), precision=precision)
    _random_state(o, s, grid, rng, visco=True)
    o.step(5)
    s.step(5)
    for f in FIELDS:
        assert rel_l2(s.get(f), o.get(f)) < tol, (shape, f)


@pytest.mark.parametrize("tile", [1, 2, 4])
def test_tiles_overhanging_the_last_element(tile):
    """7 x 3 x 1 elements: 21 elements, so tiles of 2 and 4 leave idle slots
    in the last workgroup; the results equal the oracle's."""
    grid = small_grid(nx=7, ny=3, nz=1, lpd=4, taper_width=0)
    rng = np.random.default_rng(77)
    o, s = _pair(grid, random_model(grid, rng, visco=True), precision="fp32")
    s.set_tile(tile)
    _random_state(o, s, grid, rng, visco=True)
    o.step(4)
    s.step(4)
    for f in FIELDS:
        assert rel_l2(s.get(f), o.get(f)) < 1e-5, (tile, f)


@pytest.mark.parametrize("tile", [1, 2, 4])
@pytest.mark.parametrize("nz", [3, 4])
def test_plane_pair_order_matches_oracle(tile, nz):
    """4 x 4 planes hold whole tiles of 1, 2 and 4 elements, so the element
    kernel runs the tiles of planes 2m and 2m+1 alternately; with 3 planes
    the last one runs alone. Every element is still advanced once: the
    fields equal the oracle's, FP64 and FP32."""
    grid = small_grid(nx=4, ny=4, nz=nz, lpd=4, taper_width=0)
    for precision, tol in (("fp64", 1e-11), ("fp32", 1e-5)):
        rng = np.random.default_rng(90 + nz)
        o, s = _pair(grid, random_model(grid, rng, visco=True), precision=precision)
        s.set_tile(tile)
        _random_state(o, s, grid, rng, visco=True)
        o.step(4)
        s.step(4)
        for f in FIELDS:
            assert rel_l2(s.get(f), o.get(f)) < tol, (precision, tile, f)
        s.close()


def test_run_without_sources_or_receivers():
    grid = small_grid(nx=3, ny=2, nz=2, nt=20)
    rng = np.random.default_rng(5)
    o, s = _pair(grid, random_model(grid, rng), precision="fp32")
    s.step(20)   # one graph chunk and four direct steps
    assert np.array_equal(s.vrange(), [0.0, 0.0])
    assert s.it == 20
    for f in FIELDS:
        assert not np.any(s.get(f)), f
    with pytest.raises(api.S3dError):
        s.trace(0, 0)
    with pytest.raises(api.S3dError, match="past nt"):
        s.step(1)
