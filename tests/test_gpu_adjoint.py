"""Adjoint sensitivity accumulation (reference ses3d_grad,
src/SOURCE/ses3d_grad.F90:159-213) through the C ABI."""
import math

import numpy as np
import pytest

from helpers import random_model, random_unique, rel_l2, small_grid, unique_to_local
from ses3d_amd import api

pytestmark = pytest.mark.gpu

PARAMS = {"rhoinv": api.RHOINV, "lambda": api.LAMBDA, "mu": api.MU, "a": api.A, "b": api.B,
          "c": api.C_, "q": api.Q}


def _sim(grid, rng, precision, **kw):
    s = api.Simulation(grid, precision, **kw)
    for k, arr in random_model(grid, rng).items():
        s.set_model(PARAMS[k], arr)
    s.init()
    return s


def _jacobian(grid):
    """J = dx dy dz / 8 with dx, dy the element widths in radians (colatitude,
    longitude) and dz in metres (reference src/geometric_paras.f90:129)."""
    dx = math.radians(grid.lat_max - grid.lat_min) / grid.nx
    dy = math.radians(grid.lon_max - grid.lon_min) / grid.ny
    dz = (grid.rad_max - grid.rad_min) / grid.nz
    return dx * dy * dz / 8.0


def _set_velocity(s, grid, rng):
    for f in (api.VX, api.VY, api.VZ):
        s.set(f, unique_to_local(random_unique(grid, rng, 1e-3), grid))


@pytest.mark.parametrize("lpd", [4, 3])
def test_grad_rho_accumulates_the_correlation(lpd):
    """grad_rho = sum over calls of increment*dt*(v_fw . v)/J at every
    element-local node; a dropped component, a transposed snapshot layout or
    a wrong J or time factor fails."""
    grid = small_grid(lpd=lpd, nx=3, ny=2, nz=2)
    rng = np.random.default_rng(40 + lpd)
    s = _sim(grid, rng, "fp64")
    assert not s.get(api.GRAD_RHO).any()
    J = _jacobian(grid)
    want = np.zeros(s.shape)
    for inc in (3, 1):
        _set_velocity(s, grid, rng)
        v = [s.get(f) for f in (api.VX, api.VY, api.VZ)]
        fw = [rng.standard_normal(s.shape) for _ in range(3)]
        s.grad_add(fw[0], fw[1], fw[2], inc)
        want += inc * grid.dt * (fw[0] * v[0] + fw[1] * v[1] + fw[2] * v[2]) / J
    got = s.get(api.GRAD_RHO)
    assert rel_l2(got, want) < 1e-14
    # the strain-rate densities are zero: the reference never loads the
    # forward strain rates they need
    for f in (api.GRAD_CS, api.GRAD_CP, api.GRAD_CSH, api.GRAD_CSV):
        assert not s.get(f).any()
    s.close()


def test_grad_rho_of_the_field_with_itself_is_its_energy_density():
    """With v_fw = v the density is increment*dt*|v|^2/J: non-negative, and
    for a uniform velocity the same at every node."""
    grid = small_grid(nx=2, ny=2, nz=2)
    rng = np.random.default_rng(5)
    s = _sim(grid, rng, "fp64")
    const = np.full(s.shape, 2.5e-3)
    for f, scale in ((api.VX, 1.0), (api.VY, -2.0), (api.VZ, 0.5)):
        s.set(f, scale * const)
    v = [s.get(f) for f in (api.VX, api.VY, api.VZ)]
    s.grad_add(v[0], v[1], v[2], 2)
    g = s.get(api.GRAD_RHO)
    expect = 2 * grid.dt * (2.5e-3 ** 2) * (1.0 + 4.0 + 0.25) / _jacobian(grid)
    assert np.allclose(g, expect, rtol=1e-14, atol=0)
    s.close()


def test_grad_rho_fp32_uses_the_compensated_velocity():
    """In FP32 with compensated velocity the density follows the FP64 one to
    FP32 rounding of the velocity."""
    grid = small_grid(nx=3, ny=2, nz=2)
    rng = np.random.default_rng(9)
    model = random_model(grid, rng)
    sims = []
    for prec in ("fp64", "fp32"):
        s = api.Simulation(grid, prec, kahan_v=True)
        for k, arr in model.items():
            s.set_model(PARAMS[k], arr)
        s.init()
        sims.append(s)
    vr = np.random.default_rng(10)
    vel = [unique_to_local(random_unique(grid, vr, 1e-3), grid) for _ in range(3)]
    fw = [vr.standard_normal(sims[0].shape) for _ in range(3)]
    for s in sims:
        for f, x in zip((api.VX, api.VY, api.VZ), vel):
            s.set(f, x)
        s.grad_add(fw[0], fw[1], fw[2], 4)
    assert rel_l2(sims[1].get(api.GRAD_RHO), sims[0].get(api.GRAD_RHO)) < 1e-6
    for s in sims:
        s.close()


def test_grad_add_checks_its_arguments():
    grid = small_grid(nx=2, ny=2, nz=2)
    s = _sim(grid, np.random.default_rng(1), "fp64")
    z = np.zeros(s.shape)
    with pytest.raises(api.S3dError):
        s.grad_add(z, z, z, 0)
    s.close()
