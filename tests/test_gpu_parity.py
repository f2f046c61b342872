"""HIP path vs the FP64 oracle, through the C ABI (libses3d_hip.so)."""
import numpy as np
import pytest

import oracle
from helpers import random_model, random_unique, rel_l2, small_grid, unique_to_local
from ses3d_amd import api, synth

pytestmark = pytest.mark.gpu

PARAMS = {"rhoinv": (oracle.RHOINV, api.RHOINV), "lambda": (oracle.LAMBDA, api.LAMBDA),
          "mu": (oracle.MU, api.MU), "a": (oracle.A, api.A), "b": (oracle.B, api.B),
          "c": (oracle.C_, api.C_), "q": (oracle.Q, api.Q)}


def _pair(grid, model, precision="fp64", kahan_v=True, kahan_sigma=False):
    o = oracle.Oracle(grid)
    s = api.Simulation(grid, precision, kahan_v=kahan_v, kahan_sigma=kahan_sigma)
    for k, arr in model.items():
        o.set_model(PARAMS[k][0], arr)
        s.set_model(PARAMS[k][1], arr)
    o.init()
    s.init()
    return o, s


def _random_state(o, s, grid, rng, visco=False):
    for f in (api.VX, api.VY, api.VZ):
        v = unique_to_local(random_unique(grid, rng, 1e-3), grid)
        o.set(f, v)
        s.set(f, v)
    P1 = grid.lpd + 1
    shp = (P1, P1, P1, grid.nz, grid.ny, grid.nx)
    for f in range(api.SXX, api.SZZ + 1):
        sg = 1e5 * rng.standard_normal(shp)
        o.set(f, sg)
        s.set(f, sg)
    if visco:
        for f in range(api.MXX, api.MXX + 18):
            m = 1e-9 * rng.standard_normal(shp)
            o.set(f, m)
            s.set(f, m)


FIELDS = [api.VX, api.VY, api.VZ] + list(range(api.SXX, api.SZZ + 1))


@pytest.mark.parametrize("lpd", [4, 2, 3, 5, 6, 7])
def test_one_step_fp64_random_state(lpd):
    grid = small_grid(lpd=lpd, nx=3, ny=2, nz=2)
    rng = np.random.default_rng(100 + lpd)
    o, s = _pair(grid, random_model(grid, rng))
    _random_state(o, s, grid, rng)
    o.step(1)
    s.step(1)
    for f in FIELDS:
        assert rel_l2(s.get(f), o.get(f)) < 1e-12, f


@pytest.mark.parametrize("aniso", [False, True])
def test_one_step_fp64_visco(aniso):
    grid = small_grid(nx=3, ny=3, nz=2)
    rng = np.random.default_rng(7)
    o, s = _pair(grid, random_model(grid, rng, visco=True, aniso=aniso))
    assert o.visco
    _random_state(o, s, grid, rng, visco=True)
    for _ in range(3):
        o.step(1)
        s.step(1)
    for f in FIELDS + list(range(api.MXX, api.MXX + 18)):
        assert rel_l2(s.get(f), o.get(f)) < 1e-12, f


def test_mass_matrix_and_taper_fp64():
    grid = small_grid(nx=4, ny=3, nz=3, taper_width=1)
    rng = np.random.default_rng(3)
    o, s = _pair(grid, random_model(grid, rng))
    assert rel_l2(s.get(api.MM), o.get(oracle.MM)) < 1e-14
    assert rel_l2(s.get(api.TAPER), o.get(oracle.TAPER)) < 1e-15
    assert abs(s.cfl() - o.cfl()) < 1e-14 * o.cfl()


def _run_case(case, precision, steps, **kw):
    sim, recs = api.build_case(case, precision=precision, **kw)
    sim.step(steps)
    return sim, recs


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32", 1e-4)])
def test_homogeneous_traces(precision, tol):
    case = synth.homogeneous(nt=300, nx=8, ny=8, nz=5, n_rec=6)
    o, orecs = oracle.build_case(case)
    o.step(300)
    sim, recs = _run_case(case, precision, 300)
    assert len(recs) == len(orecs) > 0
    errs = []
    for r in range(len(recs)):
        for ch in range(3):
            errs.append(rel_l2(sim.trace(r, ch), o.trace(r, ch)))
    print(precision, "max relL2", max(errs), "median", np.median(errs))
    assert max(errs) < tol


def test_graph_replay_matches_eager_steps():
    """Chunks of 16 steps replay one captured hipGraph; single steps launch
    eagerly. Every replay is checked against the eager run, with eager work
    (field reads, v range) and synchronisation between replays."""
    case = synth.homogeneous(nt=64, nx=6, ny=6, nz=4, n_rec=4)
    a, recs = api.build_case(case, precision="fp32")
    b, _ = api.build_case(case, precision="fp32")
    for chunk in range(4):
        for _ in range(16):
            a.step(1)
        b.step(16)
        assert a.it == b.it == 16 * (chunk + 1)
        for r in range(len(recs)):
            for ch in range(3):
                assert np.array_equal(a.trace(r, ch), b.trace(r, ch)), (chunk, r, ch)
        assert np.array_equal(a.get(api.VX), b.get(api.VX))
        assert np.array_equal(a.vrange(), b.vrange())
    with pytest.raises(api.S3dError, match="past nt"):
        b.step(1)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_matches_individual_runs(precision):
    """Simulations of different sizes and GLL orders advanced as one batch
    (batched launches, graph-replayed chunks) give the same traces and
    wavefields as the same simulations stepped one by one."""
    cases = [synth.homogeneous(seed=1, nt=40, nx=5, ny=4, nz=4, lpd=4, n_rec=3),
             synth.homogeneous(seed=2, nt=40, nx=4, ny=5, nz=4, lpd=6, n_rec=2),
             synth.layered(seed=3, nt=40, nx=6, ny=6, nz=4, lpd=4, n_rec=3),
             synth.homogeneous(seed=4, nt=40, nx=4, ny=4, nz=5, lpd=5, n_rec=2)]
    solo = [api.build_case(c, precision=precision) for c in cases]
    bat = [api.build_case(c, precision=precision) for c in cases]
    for s, _ in solo:
        s.step(35)
    b = api.Batch([s for s, _ in bat])
    b.step(3)
    b.step(16)
    b.step(16)
    b.synchronize()
    for (s1, recs), (s2, _) in zip(solo, bat):
        assert s1.it == s2.it == 35
        for r in range(len(recs)):
            for ch in range(3):
                assert rel_l2(s2.trace(r, ch), s1.trace(r, ch)) < 1e-12
        assert rel_l2(s2.get(api.VZ), s1.get(api.VZ)) < 1e-12
    b.close()


def test_autotuned_tiles_give_identical_results():
    """Every element tile computes the same per-node arithmetic, so the
    autotuned run is bit-identical to the default one; the tuner refuses a
    state that is not at rest."""
    case = synth.layered(seed=5, nt=30, nx=7, ny=5, nz=4, lpd=4, n_rec=3)
    a, recs = api.build_case(case, precision="fp32")
    b, _ = api.build_case(case, precision="fp32", autotune=True)
    a.step(30)
    b.step(30)
    for r in range(len(recs)):
        for ch in range(3):
            assert np.array_equal(a.trace(r, ch), b.trace(r, ch))
    with pytest.raises(api.S3dError, match="at rest"):
        b.autotune()


def _traces(sim, recs):
    return np.array([[sim.trace(r, ch) for ch in range(3)] for r in range(len(recs))])


def test_precision_modes_against_fp64():
    """FP32 modes against the FP64 path on a small layered visco-elastic
    mesh: the compensated modes stay within 1e-5 per trace; BF16-stored
    moduli / 1/M perturb the model itself (2^-9 relative), which must show
    as a misfit far above the FP32 ones (so the BF16 storage is in use)."""
    case = synth.layered(seed=7, nt=400, nx=10, ny=10, nz=5, lpd=4, n_rec=8)
    ref, recs = api.build_case(case, precision="fp64")
    ref.step(400)
    t64 = _traces(ref, recs)
    norms = np.linalg.norm(t64, axis=2)
    keep = norms > 1e-3 * norms.max()
    out = {}
    for name, kw in [("kv", dict(kahan_v=True)),
                     ("kv+ks", dict(kahan_v=True, kahan_sigma=True)),
                     ("kv+ks_bf16", dict(kahan_v=True, kahan_sigma="bf16")),
                     ("kv_bf16+ks_bf16", dict(kahan_v="bf16", kahan_sigma="bf16")),
                     ("kv+bf16_static", dict(kahan_v=True, bf16_static=True))]:
        sim, _ = api.build_case(case, precision="fp32", **kw)
        sim.step(400)
        t = _traces(sim, recs)
        err = np.linalg.norm(t - t64, axis=2)[keep] / norms[keep]
        out[name] = (float(np.median(err)), float(err.max()))
    print(out)
    for name in ("kv", "kv+ks", "kv+ks_bf16", "kv_bf16+ks_bf16"):
        assert out[name][1] < 1e-5, (name, out[name])
    assert out["kv+ks_bf16"][0] < 2 * out["kv+ks"][0] + 1e-8
    assert out["kv_bf16+ks_bf16"][0] < 2 * out["kv+ks_bf16"][0] + 1e-8
    assert 1e-5 < out["kv+bf16_static"][0] < 5e-2


@pytest.mark.parametrize("ks", [False, True, "bf16"])
def test_fp32_state_block_fields(ks):
    """Stress and memory variables share one element-major block whose
    component count depends on the stress compensation (24, 30 or 27
    words): every field reads back as set (FP32 rounding) and a few steps
    follow the FP64 oracle. A wrong block stride or component offset gives
    O(1) errors."""
    grid = small_grid(nx=3, ny=3, nz=2)
    rng = np.random.default_rng(11)
    o, s = _pair(grid, random_model(grid, rng, visco=True), precision="fp32", kahan_v=True,
                 kahan_sigma=ks)
    assert o.visco
    _random_state(o, s, grid, rng, visco=True)
    state = list(range(api.SXX, api.SZZ + 1)) + list(range(api.MXX, api.MXX + 18))
    for f in state:
        assert rel_l2(s.get(f), o.get(f)) < 1e-7, f
    for _ in range(3):
        o.step(1)
        s.step(1)
    for f in FIELDS + state[6:]:
        assert rel_l2(s.get(f), o.get(f)) < 1e-4, f


def test_bf16_stress_compensation_field_io():
    """With BF16 compensation words a stress field reads back as the FP32
    value that was set (the compensation is cleared on set), and a run
    keeps hi - lo readable."""
    grid = small_grid(lpd=4, nx=3, ny=3, nz=2)
    rng = np.random.default_rng(3)
    model = random_model(grid, rng)
    s = api.Simulation(grid, "fp32", kahan_v=True, kahan_sigma="bf16")
    for k, arr in model.items():
        s.set_model(PARAMS[k][1], arr)
    s.init()
    P1 = grid.lpd + 1
    sg = 1e5 * rng.standard_normal((P1, P1, P1, grid.nz, grid.ny, grid.nx))
    s.set(api.SYY, sg)
    assert np.array_equal(s.get(api.SYY), sg.astype(np.float32).astype(np.float64))
    s.step(3)
    assert np.all(np.isfinite(s.get(api.SYY)))


def test_bf16_static_restrictions():
    """BF16 static fields use the default element tile only and are refused
    in a batch."""
    case = synth.homogeneous(nt=20, nx=4, ny=4, nz=3, n_rec=2)
    a, _ = api.build_case(case, precision="fp32", bf16_static=True, autotune=True)
    b, _ = api.build_case(case, precision="fp32")
    with pytest.raises(api.S3dError, match="BF16"):
        api.Batch([a, b])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_shared_model_sources_match_independent_runs(precision):
    """Sources on one shared copy of the model (s3d_create_shared), batched
    source-innermost, give bit for bit the traces of independent simulations
    that each own the model; a sharer stepped on its own does too; the
    model's simulation cannot be destroyed while sharers use it."""
    from dataclasses import replace
    case = synth.layered(seed=7, nt=40, nx=6, ny=5, nz=4, lpd=4, n_rec=4, n_src=3)
    solo = [api.build_case(replace(case, sources=[s]), precision=precision)
            for s in case.sources]
    for s, _ in solo:
        s.step(35)
    sims, recs = api.build_shared_sources(case, precision=precision)
    assert len(sims) == 3 and len(recs) == len(solo[0][1])
    b = api.Batch(sims)
    b.step(3)
    b.step(16)
    b.step(16)
    b.synchronize()
    for (s1, _), s2 in zip(solo, sims):
        assert s2.it == 35
        for r in range(len(recs)):
            for ch in range(3):
                assert np.array_equal(s2.trace(r, ch), s1.trace(r, ch)), (r, ch)
        assert np.array_equal(s2.get(api.VZ), s1.get(api.VZ))
    b.close()
    lone = api.Simulation.shared(sims[0])
    api._add_channels(lone, case, case.sources[1:2])
    lone.step(35)
    for r in range(len(recs)):
        assert np.array_equal(lone.trace(r, 2), solo[1][0].trace(r, 2))
    with pytest.raises(api.S3dError, match="still use"):
        sims[0].close()
    lone.close()
    for s in sims[1:]:
        s.close()
    sims[0].close()


@pytest.mark.parametrize("tile", [1, 2, 4])
@pytest.mark.parametrize("shape", [(12, 10, 6), (5, 5, 3)])
def test_set_tile_identical(tile, shape):
    """A tile chosen with s3d_set_tile computes the same per-node arithmetic:
    traces and fields equal the default tile's bit for bit (5x5x3 = 75
    elements leaves a partly filled last tile)."""
    nx, ny, nz = shape
    case = synth.layered(seed=6, nt=24, nx=nx, ny=ny, nz=nz, lpd=4, n_rec=4)
    ref, recs = api.build_case(case, precision="fp32", kahan_sigma="bf16")
    ref.step(24)
    s, _ = api.build_case(case, precision="fp32", kahan_sigma="bf16")
    s.set_tile(tile)
    s.step(24)
    for r in range(len(recs)):
        for ch in range(3):
            assert np.array_equal(s.trace(r, ch), ref.trace(r, ch)), (tile, r, ch)
    for f in (api.VX, api.SXX, api.MXX + 17):
        assert np.array_equal(s.get(f), ref.get(f)), (tile, f)


@pytest.mark.parametrize("tile", [1, 4])
@pytest.mark.parametrize("lpd", [2, 4])
def test_fp64_tiles_identical(tile, lpd):
    """FP64 element tiles (1 and 4 elements per workgroup besides the
    default) give the default tile's traces and fields bit for bit, solo
    and in a batch, whose FP64 groups launch the default tile."""
    case = synth.layered(seed=9, nt=24, nx=5, ny=5, nz=3, lpd=lpd, n_rec=4)
    ref, recs = api.build_case(case, precision="fp64")
    ref.step(24)
    s, _ = api.build_case(case, precision="fp64")
    s.set_tile(tile)
    assert s.tuning()[0] == tile
    s.step(24)
    for r in range(len(recs)):
        for ch in range(3):
            assert np.array_equal(s.trace(r, ch), ref.trace(r, ch)), (tile, r, ch)
    for f in (api.VX, api.SXX, api.MXX + 17):
        assert np.array_equal(s.get(f), ref.get(f)), (tile, f)
    m, _ = api.build_case(case, precision="fp64")
    m.set_tile(tile)
    b = api.Batch([m])
    b.step(24)
    b.synchronize()
    b.close()
    for r in range(len(recs)):
        assert np.array_equal(m.trace(r, 2), ref.trace(r, 2)), (tile, r)


@pytest.mark.parametrize("tile", [1, 2, 4])
def test_lds_layouts_identical(tile):
    """The node-interleaved LDS layout stages the same values and contracts
    them in the same order as the planar one: traces and fields equal bit for
    bit, with every tile, in a batch too; the autotuner reports its choice."""
    case = synth.layered(seed=8, nt=24, nx=9, ny=7, nz=4, lpd=4, n_rec=4)
    ref, recs = api.build_case(case, precision="fp32", kahan_sigma="bf16")
    ref.set_tile(tile)
    ref.step(24)
    s, _ = api.build_case(case, precision="fp32", kahan_sigma="bf16")
    s.set_tile(tile)
    s.set_lds_layout(1)
    assert s.tuning()[2] == 1
    s.step(24)
    for r in range(len(recs)):
        for ch in range(3):
            assert np.array_equal(s.trace(r, ch), ref.trace(r, ch)), (tile, r, ch)
    for f in (api.VX, api.SYZ, api.MXX + 7):
        assert np.array_equal(s.get(f), ref.get(f)), (tile, f)
    b_sim, _ = api.build_case(case, precision="fp32", kahan_sigma="bf16")
    b_sim.set_tile(tile)
    b_sim.set_lds_layout(1)
    b = api.Batch([b_sim])
    b.step(24)
    b.synchronize()
    for r in range(len(recs)):
        assert np.array_equal(b_sim.trace(r, 2), ref.trace(r, 2)), (tile, r)
    b.close()
    t, _, lay = api.build_case(case, precision="fp32", autotune=True)[0].tuning()
    assert t in (1, 2, 4) and lay in (0, 1)


@pytest.mark.parametrize("lpd,mode", [(4, {"kahan_sigma": "bf16"}), (6, {}), (3, {"kahan_sigma": True})])
def test_early_state_stores_identical(lpd, mode):
    """Storing the new state before the weak divergence changes where the
    stores sit, not what they store: traces and fields equal the default
    order bit for bit, with two state blocks, with one, and in a batch."""
    case = synth.layered(seed=12, nt=20, nx=5, ny=4, nz=3, lpd=lpd, n_rec=3)
    for nbuf in (2, 1):
        ref, recs = api.build_case(case, precision="fp32", state_buffers=nbuf, **mode)
        ref.step(20)
        s, _ = api.build_case(case, precision="fp32", state_buffers=nbuf, **mode)
        s.set_store_order(1)
        assert s.store_order() == 1
        s.step(20)
        for r in range(len(recs)):
            for ch in range(3):
                assert np.array_equal(s.trace(r, ch), ref.trace(r, ch)), (nbuf, r, ch)
        for f in (api.VX, api.SXZ, api.MXX + 11):
            assert np.array_equal(s.get(f), ref.get(f)), (nbuf, f)
    m, _ = api.build_case(case, precision="fp32", **mode)
    m.set_store_order(1)
    b = api.Batch([m])
    b.step(20)
    b.synchronize()
    b.close()
    for r in range(len(recs)):
        assert np.array_equal(m.trace(r, 2), ref.trace(r, 2)), r


def test_store_order_checks():
    case = synth.homogeneous(nt=4, nx=4, ny=4, nz=3, n_rec=1)
    s, _ = api.build_case(case, precision="fp32")
    with pytest.raises(api.S3dError, match="store order must be"):
        s.set_store_order(2)
    s.set_lds_layout(1)
    with pytest.raises(api.S3dError, match="node-interleaved"):
        s.set_store_order(1)
    s.set_lds_layout(0)
    s.set_store_order(1)
    with pytest.raises(api.S3dError, match="node-interleaved"):
        s.set_lds_layout(1)
    d, _ = api.build_case(case, precision="fp64")
    with pytest.raises(api.S3dError, match="FP64"):
        d.set_store_order(1)
    t, _ = api.build_case(case, precision="fp32", autotune=True)
    assert t.store_order() in (0, 1)


def test_set_tile_checks():
    elastic = synth.homogeneous(nt=10, nx=5, ny=5, nz=4, n_rec=1)
    s, _ = api.build_case(elastic, precision="fp32")
    with pytest.raises(api.S3dError, match="tile must be"):
        s.set_tile(3)
    with pytest.raises(api.S3dError, match="tile must be"):
        s.set_tile(104)
    s.set_tile(2)
    d, _ = api.build_case(elastic, precision="fp64")
    d.set_tile(4)
    assert d.tuning()[0] == 4
    d.set_tile(0)
    h, _ = api.build_case(elastic, precision="fp32", bf16_static=True)
    with pytest.raises(api.S3dError, match="default tile"):
        h.set_tile(4)
    with pytest.raises(api.S3dError, match="planar LDS layout"):
        d.set_lds_layout(1)
    with pytest.raises(api.S3dError, match="LDS layout must be"):
        s.set_lds_layout(2)
    assert d.tuning() == (2, 1, 0)


def test_batch_members_with_different_tiles_and_layouts_match_solo_runs():
    """A batch launches members of one kind together with the tile and LDS
    layout of its largest member; every tile and layout gives the same
    results, so each member still equals its own solo run bit for bit."""
    cases = [synth.layered(seed=6 + q, nt=24, nx=nx, ny=6, nz=4, lpd=4, n_rec=3)
             for q, nx in enumerate((8, 5, 9, 7))]
    cfgs = [(1, 0), (2, 1), (4, 0), (4, 1)]
    solo, members, recs = [], [], []
    for case, (tile, lay) in zip(cases, cfgs):
        for lst in (solo, members):
            sim, rr = api.build_case(case, precision="fp32", kahan_sigma="bf16")
            sim.set_tile(tile)
            sim.set_lds_layout(lay)
            lst.append(sim)
        recs.append(rr)
    for sim in solo:
        sim.step(24)
    b = api.Batch(members)
    b.step(24)
    b.synchronize()
    b.close()
    for q, (a, m) in enumerate(zip(solo, members)):
        for r in range(len(recs[q])):
            for ch in range(3):
                assert np.array_equal(a.trace(r, ch), m.trace(r, ch)), (q, r, ch)
        assert np.array_equal(a.get(api.VX), m.get(api.VX)), q
        assert np.array_equal(a.get(api.SYZ), m.get(api.SYZ)), q


def test_batch_member_outlives_batch():
    case = synth.homogeneous(seed=3, nt=10, nx=4, ny=4, nz=4, n_rec=1)
    s, _ = api.build_case(case, precision="fp32")
    b = api.Batch([s])
    b.step(2)
    with pytest.raises(api.S3dError, match="member of 1 batch"):
        s.close()
    b.close()
    s.close()


def test_batch_with_bf16_compensation_matches_individual_runs():
    """Batched launches of the BF16-compensated mode (the timed one) on
    shared-model sources and on an independent simulation of another order
    give the traces of the simulations stepped alone, bit for bit."""
    from dataclasses import replace
    case = synth.layered(seed=9, nt=30, nx=6, ny=6, nz=4, lpd=4, n_rec=3, n_src=2)
    other = synth.homogeneous(seed=2, nt=30, nx=4, ny=5, nz=4, lpd=6, n_rec=2)
    kw = dict(precision="fp32", kahan_sigma="bf16")
    solo = [api.build_case(replace(case, sources=[s]), **kw) for s in case.sources]
    solo.append(api.build_case(other, **kw))
    for s, _ in solo:
        s.step(30)
    sims, recs = api.build_shared_sources(case, **kw)
    o2, recs2 = api.build_case(other, **kw)
    b = api.Batch(sims + [o2])
    b.step(30)
    b.synchronize()
    for (s1, r1), s2 in zip(solo, sims + [o2]):
        for r in range(len(r1)):
            for ch in range(3):
                assert np.array_equal(s2.trace(r, ch), s1.trace(r, ch))
    b.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-4)])
def test_volume_curl_fields_match_oracle(precision, tol):
    """Snapshot attributes rot_vx, rot_vy, rot_vz (src/volume_snapshot.f90:
    238-245): the strain rates the step formed from its start velocity (the
    oracle's strain-rate fields after the step, pinned by the closed forms
    of test_oracle_physics.py) combined with the velocity at its end."""
    grid = small_grid(nx=4, ny=3, nz=3, taper_width=1)
    rng = np.random.default_rng(31)
    o, s = _pair(grid, random_model(grid, rng, visco=True), precision=precision)
    _random_state(o, s, grid, rng, visco=True)
    o.step(3)
    s.step(3)
    s.capture_curl_strain()
    o.step(1)
    s.step(1)
    th, r = o.get(oracle.THETA), o.get(oracle.R)
    vx, vy = o.get(oracle.VX), o.get(oracle.VY)
    rs, ct = r * np.sin(th), 1.0 / np.tan(th)
    ref = {api.ROT_VX: -vy / r + o.get(oracle.EYZ) / rs + o.get(oracle.EZY),
           api.ROT_VY: -o.get(oracle.EZX) + (vx - o.get(oracle.EXZ)) / r,
           api.ROT_VZ: (o.get(oracle.EXY) + vy * ct - o.get(oracle.EYX) / np.sin(th)) / r}
    for f, want in ref.items():
        assert rel_l2(s.get(f), want) < tol, f
    s.step(1)   # the captured parts belong to the previous step now
    with pytest.raises(api.S3dError, match="capture_curl_strain"):
        s.get(api.ROT_VX)


@pytest.mark.parametrize("precision,tol,lpd", [("fp64", 1e-10, 4), ("fp32", 1e-4, 4),
                                               ("fp64", 1e-10, 2), ("fp64", 1e-10, 7)])
def test_curl_channels_match_oracle(precision, tol, lpd):
    """rot_vx, rot_vy, rot_vz (src/config/channel_mod.f90:620-774): the strain
    rates formed from the step's start velocity combined with the velocity at
    its end, against the FP64 oracle's record of the same channels; the run
    is longer than one captured chunk, so the graph path records them too."""
    from dataclasses import replace
    case = synth.homogeneous(nt=120, nx=6, ny=6, nz=5, lpd=lpd, n_rec=4, visco=True)
    case = replace(case, receivers=[replace(r, channels=("vx", "rot_vx", "rot_vy", "rot_vz"))
                                    for r in case.receivers])
    o, recs = oracle.build_case(case)
    o.step(case.grid.nt)
    s, recs2 = api.build_case(case, precision=precision)
    assert len(recs2) == len(recs)
    s.step(case.grid.nt)
    for ri in range(len(recs)):
        for ci in range(4):
            ref = o.trace(ri, ci)
            assert np.linalg.norm(ref) > 0
            assert rel_l2(s.trace(ri, ci), ref) < tol, (ri, ci)
    s.close()


def test_curl_channels_in_a_batch_match_solo_runs():
    """A batch launches the curl members' strain parts before its element
    kernels: curl traces of batched members equal their solo runs bit for
    bit, beside a member without curl receivers, through a graph chunk."""
    from dataclasses import replace
    case = synth.homogeneous(nt=40, nx=6, ny=6, nz=5, n_rec=4, visco=True)
    curl = replace(case, receivers=[replace(r, channels=("rot_vx", "vz", "rot_vy", "rot_vz"))
                                    for r in case.receivers])
    kinds = (curl, case, curl)
    solo, members, recs = [], [], []
    for c in kinds:
        for lst in (solo, members):
            sim, rr = api.build_case(c, precision="fp32", kahan_sigma="bf16")
            lst.append(sim)
        recs.append(rr)
    for sim in solo:
        sim.step(37)
    b = api.Batch(members)
    b.step(37)
    b.synchronize()
    b.close()
    for q, (a, m) in enumerate(zip(solo, members)):
        for r, rec in enumerate(recs[q]):
            for ch in range(len(rec.channels)):
                assert np.array_equal(a.trace(r, ch), m.trace(r, ch)), (q, r, ch)
    assert np.linalg.norm(members[0].trace(0, 0)) > 0
