#!/usr/bin/env python3
"""Benchmark of the MI355X forward time step: GLL-node updates per second.

Workload (default): the layered visco-elastic precision-study mesh,
192 x 192 x 96 elements, GLL order 4 (442 368 000 element-local nodes), three
SLS relaxation mechanisms, one moment-tensor source, 256 surface receivers,
synthetic seeded model (synth.layered). One "step" is one iteration of the
reference's forward loop (src/ses3d-nt.F90:240-270): evolution, source add,
receiver record.

Prints ONE JSON line. `value` is element-local GLL-node updates per second of
the FP32 solver (compensated velocity integration) with every input resident
in HBM. The same mesh is also stepped in FP64 on the GPU for the FP32 / FP64
speedup, and the FP64 oracle (oracle/liboracle.so, C + OpenMP) is timed on a
bounded sample of the same workload on the host cores as `cpu_baseline`.
Kernel times come from HIP events recorded on the library's own stream.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ses3d-nt_amd"))

import numpy as np  # noqa: E402
from dataclasses import replace  # noqa: E402

from ses3d_amd import api, synth  # noqa: E402

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
FP64_PEAK = 78.6e12        # MI355X vector FP64, flop/s
FP32_PEAK = 157.3e12       # MI355X vector FP32, flop/s
# Arithmetic per element-local node of the visco-elastic element kernel at
# order 4 with compensated stress (counted from the kernel source; an FMA
# counts 2): FP64 90 for the strain-rate contraction (45 FMA), 24 for the
# spherical completion and symmetrisation, 90 for the divergence contraction
# and 37 for its coefficients and the force; FP32 44 for the stress rate, 24
# for its compensated integration, 6 for the effective stress, 73 for the
# memory variables and 24 for the tapered stores. With five memory words per
# mechanism (the deviatoric part) the stress rate takes 39 and the memory
# variables 66 (60 for the 15 words, 6 for tau/3 and the deviatoric strain
# rate).
ELEM_FP64_FLOP = 241
ELEM_FP32_FLOP = {6: 171, 5: 159}
CPU_SAMPLE = dict(nx=24, ny=24, nz=12)   # oracle sample of the same workload
MISFIT_STEPS = 1200        # FP32 vs FP64 misfit leg after the timing


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Precision modes: Simulation options and the element sizes of the words
# each one moves (state, velocity compensation, stress compensation, static).
MODES = {
    "fp64": dict(precision="fp64", kahan_v=False, kahan_sigma=0, bf16_static=False),
    "fp32": dict(precision="fp32", kahan_v=False, kahan_sigma=0, bf16_static=False),
    "fp32+kahan_v": dict(precision="fp32", kahan_v=True, kahan_sigma=0, bf16_static=False),
    "fp32+kahan_v+kahan_sigma": dict(precision="fp32", kahan_v=True, kahan_sigma=1,
                                     bf16_static=False),
    "fp32+kahan_v+kahan_sigma_bf16": dict(precision="fp32", kahan_v=True, kahan_sigma=2,
                                          bf16_static=False),
    "fp32+kahan_v_bf16+kahan_sigma_bf16": dict(precision="fp32", kahan_v=2, kahan_sigma=2,
                                               bf16_static=False),
    "fp32+kahan_v+bf16_static": dict(precision="fp32", kahan_v=True, kahan_sigma=0,
                                     bf16_static=True),
}


MODE_TEXT = {
    "fp64": "FP64",
    "fp32": "FP32",
    "fp32+kahan_v": "FP32 with compensated velocity integration",
    "fp32+kahan_v+kahan_sigma": "FP32 with compensated velocity and stress integration",
    "fp32+kahan_v+kahan_sigma_bf16": "FP32 with compensated velocity integration and "
                                     "BF16-compensated stress integration",
    "fp32+kahan_v_bf16+kahan_sigma_bf16": "FP32 with BF16-compensated velocity and stress "
                                          "integration",
    "fp32+kahan_v+bf16_static": "FP32 with compensated velocity integration and BF16 "
                                "moduli / 1/M",
}
DEFAULT_MODE = "fp32+kahan_v+kahan_sigma_bf16"


def mode_sizes(mode):
    """(state, velocity compensation, stress compensation, static) bytes."""
    m = MODES[mode]
    s = 8 if m["precision"] == "fp64" else 4
    return (s, {0: 0, 1: s, 2: 2}[int(m["kahan_v"])], {0: 0, 1: 4, 2: 2}[m["kahan_sigma"]],
            2 if m["bf16_static"] else s)


def step_bytes(mode, u, visco, L=3, S=1, nm=6):
    """Algorithmic HBM bytes per element-local node per step:
    s_state*[12 + 2 nm L visco + 6 u] + s_lo_v*6 u + s_lo_sigma*12
    + s_static*[(2 + visco) + u]/S, with nm memory words per mechanism (6,
    or 5 for the deviatoric form)"""
    s, lv, ls, st = mode_sizes(mode)
    return s * (12 + 2 * nm * L * visco + 6 * u) + lv * 6 * u + ls * 12 + st * ((2 + visco) + u) / S


def elem_kernel_bytes(mode, u, visco, nm=6):
    """The element kernel's share of the step's algorithmic bytes per
    element-local node: stress r+w (12 words, +12 compensation words),
    memory variables r+w (2 x 3 nm words: 36, or 30 in the five-word
    deviatoric form), lambda, mu (+tau) reads and the
    velocity read once per unique node (3u words); the rest of step_bytes
    (velocity write, its compensation r+w, dt/M) is the node kernel's. The
    element force that the two kernels exchange (3 words written here, read
    by the node kernel) is not among the algorithmic bytes."""
    s, _, ls, st = mode_sizes(mode)
    return s * (12 + 6 * nm * visco + 3 * u) + ls * 12 + st * (2 + visco)


def load_sims(case, modes, slab, remap=True, tune=None, state_buffers=0, memory_vars=0):
    """Build one Simulation per precision mode from a single pass over the
    model slabs; returns a list of (mode, sim). tune = (tile, LDS layout)
    sets the FP32 simulation's element tile and layout instead of the
    autotuner. The first mode gets `state_buffers` and `memory_vars`;
    further modes update their state in place (one state buffer), so that
    an FP32 and an FP64 simulation fit on one GPU together at 192x192x96,
    and keep six memory words per mechanism."""
    g = case.grid
    sims = [(name, api.Simulation(g, xcd_remap=remap,
                                  state_buffers=state_buffers if q == 0 else 1,
                                  memory_vars=memory_vars if q == 0 else 0, **MODES[name]))
            for q, name in enumerate(modes)]
    t0 = time.time()
    for ez0 in range(0, g.nz, slab):
        nez = min(slab, g.nz - ez0)
        m = case.model_slab(ez0, nez)
        for _, sim in sims:
            sim.set_model_slab(api.RHOINV, ez0, m["rhoinv"])
            sim.set_model_slab(api.LAMBDA, ez0, m["lambda"])
            sim.set_model_slab(api.MU, ez0, m["mu"])
            if "q" in m:
                sim.set_model_slab(api.Q, ez0, m["q"])
        if ez0 % (8 * slab) == 0:
            log("[bench] model layers %d..%d uploaded (%.0f s)" % (ez0, ez0 + nez - 1, time.time() - t0))
    for name, sim in sims:
        sim.init()
        if tune is None or name == "fp64":
            tile, tms = sim.autotune()
            log("[bench] %s autotune: %d elements per workgroup, element kernel %.3f ms" % (name, tile, tms))
        else:
            sim.set_tile(tune[0])
            sim.set_lds_layout(tune[1])
        sim.tile, sim.remap, sim.layout = sim.tuning()
        sim.early = sim.store_order()
        log("[bench] %s tuning: %d elements per workgroup, XCD order %d, LDS layout %d, "
            "early stores %d" % (name, sim.tile, sim.remap, sim.layout, sim.early))
        nrec = 0
        for s_ in case.sources:
            loc = sim.locate(s_.lat, s_.lon, s_.depth)
            if loc is not None:
                f = api.wavelet(s_.wavelet, s_.onset, s_.width, g.dt, g.nt)
                sim.add_source_mt(loc[0], loc[1], np.outer(np.asarray(s_.moment_tensor), f))
        for r in case.receivers:
            loc = sim.locate(r.lat, r.lon, r.depth)
            if loc is not None:
                sim.add_receiver(loc[0], loc[1], r.channels)
                nrec += 1
        sim.synchronize()
        sim.nrec = nrec
        log("[bench] %s ready: device %.1f GB, courant %.3f, %d receivers (%.0f s)"
            % (name, sim.device_bytes() / 1e9, sim.cfl(), nrec, time.time() - t0))
    return sims


def time_gpu(sim, steps, warmup):
    sim.step(warmup)
    sim.synchronize()
    _torch_sync()
    sim.timing(True)
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    _torch_sync()
    wall = time.perf_counter() - t0
    ms, cnt = sim.kernel_times()
    sim.timing(False)
    vr = sim.vrange()
    return wall, ms, cnt, vr


def ring_receivers(src, radii=(0.5, 1.0, 1.5, 2.0), azimuths=8):
    """Surface receivers on rings around a source (degrees): they see the
    whole source wavelet within the misfit leg."""
    out = []
    for d in radii:
        for q in range(azimuths):
            a = 2 * np.pi * q / azimuths
            out.append(synth.Receiver(src.lat + d * np.cos(a), src.lon + d * np.sin(a), 0.0,
                                      "R%02d_%d" % (int(10 * d), q)))
    return out


def traces_at(sim, steps):
    """Continue `sim` to step `steps`; the vx, vy, vz traces of every
    receiver."""
    left = steps - sim.it
    while left > 0:
        n = min(400, left)
        sim.step(n)
        sim.synchronize()
        left -= n
    return np.array([[sim.trace(r, c) for c in range(3)] for r in range(sim.nrec)])


def misfit(t32, t64, steps):
    """Relative L2 per trace of t32 against t64 over the traces with signal
    (norm above 1e-3 of the largest)."""
    norms = np.linalg.norm(t64, axis=2)
    keep = norms > 1e-3 * norms.max()
    err = np.linalg.norm(t32 - t64, axis=2)[keep] / norms[keep]
    return {"steps": steps, "traces": int(keep.sum()), "trace_relL2_median": float(np.median(err)),
            "trace_relL2_max": float(err.max()), "reference": "FP64 GPU path, same mesh"}


def _torch_sync():
    t = sys.modules.get("torch")
    if t is not None and t.cuda.is_available():
        t.cuda.synchronize()


def copy_probe(gb):
    """Run ses3d-nt_amd/bin/copy_probe: a plain device copy of `gb` GB
    (three launches) that calibrates the FETCH_SIZE / WRITE_SIZE counters of
    a profiling run and measures the attainable copy bandwidth."""
    import subprocess
    exe = os.path.join(ROOT, "ses3d-nt_amd", "bin", "copy_probe")
    r = subprocess.run([exe, "%g" % gb, "3"], capture_output=True, text=True)
    log(r.stdout.strip())
    if r.returncode != 0:
        raise RuntimeError("copy_probe failed: " + r.stderr)


def contraction_probe():
    """ses3d-nt_amd/bin/contraction_probe: the GLL derivative contraction on
    its own, FP64-coefficient VALU (the element kernel's form), FP32 VALU and
    FP32 matrix cores with split coefficients, for one element's 25 lines x 3
    components x 16 sources per workgroup. Prints one JSON line."""
    import subprocess
    exe = os.path.join(ROOT, "ses3d-nt_amd", "bin", "contraction_probe")
    r = subprocess.run([exe, "layout"], capture_output=True, text=True)
    log(r.stdout.strip())
    if r.returncode != 0:
        raise RuntimeError("contraction_probe layout check failed: " + r.stdout + r.stderr)
    res = {}
    for mode in ("valu64", "valu32", "mfma32"):
        r = subprocess.run([exe, mode, "1200", "200", "2048"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("contraction_probe %s failed: %s" % (mode, r.stderr))
        res[mode] = json.loads(r.stdout)
        log("[contraction] %s" % r.stdout.strip())
    print(json.dumps({"probe": "derivative contraction, GLL order 4", "results": res}), flush=True)


def cpu_calibration():
    """The oracle against the reference on the reference's CPU-runnable case
    (12x12x8 homogeneous, FP64), both timed on the same CPU type."""
    p = os.path.join(ROOT, "profiles", "cpu_calibration_homogeneous_12x12x8_p4.json")
    return json.load(open(p)) if os.path.exists(p) else None


def cpu_baseline(case_kw, target_s=15.0):
    """FP64 oracle on a bounded sub-mesh of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    case = synth.layered(seed=case_kw["seed"], nt=100000, nx=CPU_SAMPLE["nx"],
                         ny=CPU_SAMPLE["ny"], nz=CPU_SAMPLE["nz"], n_rec=16)
    threads = oracle.lib().ora_set_threads(0)
    o, _ = oracle.build_case(case)
    o.step(2)
    t0 = time.perf_counter()
    o.step(3)
    per = (time.perf_counter() - t0) / 3
    n = max(3, int(target_s / max(per, 1e-6)))
    t0 = time.perf_counter()
    o.step(n)
    wall = time.perf_counter() - t0
    o.close()
    g = case.grid
    out = {"value": g.nodes * n / wall, "unit": "GLL-node updates/s",
           "cores": int(threads), "kind": "port",
           "sample": "FP64 oracle (C, OpenMP) on a %dx%dx%d-element, p=4 sub-mesh of the "
                     "same layered visco-elastic workload, %d steps, %.1f s"
                     % (g.nx, g.ny, g.nz, n, wall)}
    cal = cpu_calibration()
    if cal:
        r = cal["oracle_over_reference"]["8_cores"]
        out["calibration"] = {
            "oracle_over_reference_on_12x12x8_8_cores": r,
            "reference_node_updates_per_s_12x12x8_8_ranks": cal["reference"]["node_updates_per_s_8_ranks"],
            "oracle_node_updates_per_s_12x12x8_8_threads": cal["oracle"]["node_updates_per_s_8_threads"],
            "reference_equivalent_value": out["value"] / r,
            "note": "the reference's Fortran/MPI path cannot run here; on its own CPU-runnable "
                    "case the oracle is %.2fx faster than it on the same cores" % r}
    return out


def precision_study(args):
    """FP32 variants against the FP64 GPU path on the same seeded mesh:
    relative L2 misfit of every receiver trace (vx, vy, vz) and of the final
    velocity snapshot. Prints one JSON line."""
    case = synth.layered(seed=1, nt=args.study_steps, nx=args.nx, ny=args.ny, nz=args.nz,
                         n_rec=args.study_receivers)
    g = case.grid
    variants = ["fp64"] + [m for m in MODES if m != "fp64"]
    if args.study_variants:
        keep = args.study_variants.split(",")
        for m in keep:
            if m not in MODES:
                raise SystemExit("unknown precision mode %r (known: %s)" % (m, ", ".join(MODES)))
        variants = ["fp64"] + [m for m in keep if m != "fp64"]
    res = {}
    ref_tr = ref_snap = None
    for name in variants:
        t0 = time.time()
        sim, recs = api.build_case(case, slab=args.slab,
                                   memory_vars=6 if name == "fp64" else args.memory_vars, **MODES[name])
        chunk = 500
        done = 0
        while done < g.nt:
            n = min(chunk, g.nt - done)
            sim.step(n)
            done += n
            sim.synchronize()
            log("[study] %s step %d/%d (%.0f s)" % (name, done, g.nt, time.time() - t0))
        tr = np.array([[sim.trace(r, c) for c in range(3)] for r in range(len(recs))])
        snap = [sim.get(f) for f in (api.VX, api.VY, api.VZ)]
        sim.close()
        if ref_tr is None:
            ref_tr, ref_snap = tr, snap
            continue
        e = []
        for r in range(tr.shape[0]):
            for c in range(3):
                d = np.linalg.norm(ref_tr[r, c])
                if d > 0:
                    e.append(np.linalg.norm(tr[r, c] - ref_tr[r, c]) / d)
        e = np.array(e)
        sn = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(snap, ref_snap)]
        res[name] = {"trace_relL2_median": float(np.median(e)), "trace_relL2_max": float(e.max()),
                     "trace_relL2_p95": float(np.percentile(e, 95)), "traces": int(e.size),
                     "snapshot_relL2": sn}
        log("[study] %s: median %.3e max %.3e snapshot %s" % (name, np.median(e), e.max(), sn))
    print(json.dumps({"study": "precision", "workload": "layered visco-elastic %dx%dx%d p=4"
                      % (g.nx, g.ny, g.nz), "steps": g.nt, "dt": g.dt,
                      "reference": "FP64 GPU path (six memory words)",
                      "memory_vars": args.memory_vars, "results": res}), flush=True)


def batch_bench(args):
    """Uneven batch: independent regional simulations (synth.uneven_batch)
    advanced together by s3d_batch_step (batched launches, graph-replayed
    chunks), against the same simulations stepped one after another."""
    cases = synth.uneven_batch(seed=4, n_sims=args.batch_sims, nt=args.warmup + 3 * args.steps + 1)
    N = sum(c.grid.nodes for c in cases)
    t0 = time.time()
    sims = [api.build_case(c, precision="fp32", autotune=True, state_buffers=args.state_buffers,
                           memory_vars=args.memory_vars)[0]
            for c in cases]
    log("[batch] %d simulations, %d GLL nodes, orders %s, setup %.0f s"
        % (len(sims), N, sorted(set(c.grid.lpd for c in cases)), time.time() - t0))
    b = api.Batch(sims)
    b.step(args.warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(args.steps)
    b.synchronize()
    wall_b = time.perf_counter() - t0
    b.close()
    # one simulation at a time, all its steps (graph-replayed chunks): a small
    # simulation's state stays in the 256 MB Infinity Cache between its steps
    t0 = time.perf_counter()
    for s_ in sims:
        s_.step(args.steps)
    for s_ in sims:
        s_.synchronize()
    wall_s = time.perf_counter() - t0
    # round robin: every simulation one step per pass, as the batch advances
    # them (all states stream from HBM every step), launched one by one
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for s_ in sims:
            s_.step(1)
    for s_ in sims:
        s_.synchronize()
    wall_r = time.perf_counter() - t0
    out = {"metric": "GLL-node updates/s", "workload": "uneven batch of %d regional simulations "
           "(8x8x8 .. 48x48x24 elements, GLL order 4 or 6, layered visco-elastic, FP32 + Kahan v)"
           % len(sims), "gll_nodes": N, "steps": args.steps,
           "batched": {"value": N * args.steps / wall_b, "ms_per_step": 1e3 * wall_b / args.steps},
           "round_robin": {"value": N * args.steps / wall_r, "ms_per_step": 1e3 * wall_r / args.steps},
           "sequential": {"value": N * args.steps / wall_s, "ms_per_step": 1e3 * wall_s / args.steps},
           "speedup_batched_vs_round_robin": wall_r / wall_b,
           "state_buffers": 1 if args.state_buffers == 1 else 2,
           "memory_vars": args.memory_vars}
    log("[batch] batched %.3f ms/step, round robin %.3f ms/step, sequential %.3f ms/step"
        % (1e3 * wall_b / args.steps, 1e3 * wall_r / args.steps, 1e3 * wall_s / args.steps))
    print(json.dumps(out), flush=True)


def multi_source_bench(args):
    """Many sources sharing one Earth model: one simulation per source on a
    single copy of the model (Simulation.shared), advanced by s3d_batch_step
    with the members' workgroups source-innermost, against the same
    simulations stepped round robin (whole mesh per source, no sharing of
    the static-field reads). More sources than --per-batch run as successive
    batches of --per-batch (the state of 32 sources at 128x128x48 exceeds the
    GPU's memory); the rates count every source over the summed times."""
    mode = args.mode
    kw = dict(MODES[mode])
    if kw.pop("bf16_static"):
        raise SystemExit("BF16 static fields are not supported in a batch")
    # one state block per source unless asked otherwise: with source-innermost
    # batches the second block does not pay (80.05 against 80.64 ms/step for
    # 16 sources in batches of 8), and one block fits 16 sources per batch
    nbuf = args.state_buffers or 1
    kw["state_buffers"] = nbuf
    kw["memory_vars"] = args.memory_vars
    case = synth.multi_source(seed=5, n_src=args.sources, nt=args.warmup + 2 * args.steps + 1)
    g = case.grid
    N = g.nodes
    S = len(case.sources)
    u = g.unique_nodes / N
    per_default = 16 if nbuf == 1 else 8
    per = max(1, min(args.per_batch or per_default, S))
    t0 = time.time()
    owner, recs = api.build_case(replace(case, sources=[]), slab=args.slab, autotune=True, **kw)
    log("[multi-source] %dx%dx%d p=%d, %d sources in batches of %d, %d receivers, %d GLL nodes "
        "per source, model set up in %.0f s" % (g.nx, g.ny, g.nz, g.lpd, S, per, len(recs), N,
                                                time.time() - t0))
    wall_b = wall_r = 0.0
    dev = 0.0
    for first in range(0, S, per):
        srcs = case.sources[first:first + per]
        sims = []
        for s_ in srcs:
            sim = api.Simulation.shared(owner)
            api._add_channels(sim, case, [s_])
            sims.append(sim)
        dev = max(dev, owner.device_bytes() + sum(x.device_bytes() for x in sims))
        b = api.Batch(sims)
        b.step(args.warmup)
        b.synchronize()
        t0 = time.perf_counter()
        b.step(args.steps)
        b.synchronize()
        wb = time.perf_counter() - t0
        b.close()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            for s_ in sims:
                s_.step(1)
        for s_ in sims:
            s_.synchronize()
        wr = time.perf_counter() - t0
        log("[multi-source] sources %d..%d: batched %.3f ms/step, round robin %.3f ms/step, "
            "device %.0f GB" % (first, first + len(srcs) - 1, 1e3 * wb / args.steps,
                                1e3 * wr / args.steps, dev / 1e9))
        wall_b += wb
        wall_r += wr
        for s_ in sims:
            s_.close()
    owner.close()
    B = step_bytes(mode, u, True, S=per, nm=args.memory_vars)
    value = N * S * args.steps / wall_b
    out = {"metric": "GLL-node updates/s (node x source)",
           "workload": "%d moment-tensor sources on one layered visco-elastic model, %dx%dx%d "
                       "elements, GLL order %d, %s, batches of %d" % (S, g.nx, g.ny, g.nz, g.lpd, mode,
                                                                     per),
           "gll_nodes_per_source": N, "sources": S, "per_batch": per, "steps": args.steps,
           "state_buffers": nbuf, "memory_vars": args.memory_vars,
           "batched_shared": {"value": value, "ms_per_step": 1e3 * wall_b / args.steps},
           "round_robin": {"value": N * S * args.steps / wall_r,
                           "ms_per_step": 1e3 * wall_r / args.steps},
           "speedup_batched_vs_round_robin": wall_r / wall_b,
           "step_roofline": {"bytes_per_node_source": B, "achieved_GBps": B * value / 1e9,
                             "frac": B * value / HBM_PEAK},
           "device_bytes_peak": dev}
    log("[multi-source] %d sources: batched %.3f ms/step (%.3e node-source updates/s, %.1f %% of "
        "8 TB/s at %.1f B), round robin %.3f ms/step" % (S, 1e3 * wall_b / args.steps, value,
                                                        100 * B * value / HBM_PEAK, B,
                                                        1e3 * wall_r / args.steps))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=192)
    ap.add_argument("--ny", type=int, default=192)
    ap.add_argument("--nz", type=int, default=96)
    ap.add_argument("--lpd", type=int, default=4, choices=range(2, 8),
                    help="GLL order of the timed mesh (the workload is quoted at 4)")
    ap.add_argument("--slab", type=int, default=4)
    ap.add_argument("--mode", default=DEFAULT_MODE, choices=list(MODES),
                    help="precision mode of the timed simulation")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-remap", action="store_true", help="plain workgroup order (no XCD remap)")
    ap.add_argument("--tile", type=int, default=None,
                    help="element tile of the timed mode instead of the autotuner's (0, 1, 2, 4)")
    ap.add_argument("--lds-layout", type=int, default=0,
                    help="LDS layout with --tile: 0 planar, 1 node-interleaved")
    ap.add_argument("--batch", action="store_true",
                    help="uneven batch of independent simulations instead of the single mesh")
    ap.add_argument("--batch-sims", type=int, default=64)
    ap.add_argument("--multi-source", action="store_true",
                    help="many sources sharing one model (128x128x48, p=4) instead of the single mesh")
    ap.add_argument("--sources", type=int, default=32,
                    help="sources sharing the model (default 32)")
    ap.add_argument("--per-batch", type=int, default=None,
                    help="sources per shared-model batch (default: what fits in HBM at "
                         "128x128x48, 16 with one state block, 8 with two)")
    ap.add_argument("--state-buffers", type=int, default=0, choices=(0, 1, 2),
                    help="state blocks per simulation: 0 default (2; 1 for --multi-source), "
                         "1 updated in place, 2 read one and write the other")
    ap.add_argument("--memory-vars", type=int, default=5, choices=(5, 6),
                    help="memory words per relaxation mechanism of the FP32 timed mode: 5, "
                         "the deviatoric part (default), or 6, the reference's form; the "
                         "FP64 leg keeps 6")
    ap.add_argument("--precision-study", action="store_true",
                    help="FP32 variants vs FP64 seismogram / snapshot misfit instead of timing")
    ap.add_argument("--study-steps", type=int, default=4000)
    ap.add_argument("--study-receivers", type=int, default=256)
    ap.add_argument("--study-variants", default="")
    ap.add_argument("--misfit-steps", type=int, default=MISFIT_STEPS,
                    help="after the timing, continue the timed and the FP64 simulation to this "
                         "step and report the trace misfit (0: skip)")
    ap.add_argument("--contraction-probe", action="store_true",
                    help="time the derivative contraction on the VALU and on the matrix cores")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="first run a plain device copy of this many GB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs the whole mesh on a single "
              "MI355X and has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if api.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.precision_study:
        precision_study(args)
        return
    if args.contraction_probe:
        contraction_probe()
        return
    if args.batch:
        batch_bench(args)
        return
    if args.multi_source:
        multi_source_bench(args)
        return

    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
    nt = args.warmup + args.steps
    misfit_steps = 0 if (args.no_fp64 or args.mode == "fp64") else args.misfit_steps
    case = synth.layered(seed=1, nt=max(nt, misfit_steps), nx=args.nx, ny=args.ny, nz=args.nz,
                         lpd=args.lpd, n_rec=256)
    case = replace(case, receivers=ring_receivers(case.sources[0]) + case.receivers)
    g = case.grid
    N = g.nodes
    u = g.unique_nodes / N
    log("[bench] mesh %dx%dx%d p=%d: %d element-local nodes, u=%.4f, dt=%.4f s"
        % (g.nx, g.ny, g.nz, g.lpd, N, u, g.dt))

    mode = args.mode
    modes = [mode] + ([] if args.no_fp64 or mode == "fp64" else ["fp64"])
    tune = None if args.tile is None else (args.tile, args.lds_layout)
    # the FP64 path runs after the timed mode, each alone on the GPU with its
    # default state blocks (both at once would not fit at 192x192x96)
    nm = args.memory_vars if mode != "fp64" else 6
    sims = dict(load_sims(case, modes[:1], args.slab, remap=not args.no_remap, tune=tune,
                          state_buffers=args.state_buffers, memory_vars=nm))
    tile = sims[mode].tile
    wall, ms, cnt, vr = time_gpu(sims[mode], args.steps, args.warmup)
    value = N * args.steps / wall
    log("[bench] %s %.3f ms/step, elem %.3f ms, node %.3f ms, v range %s"
        % (mode, 1e3 * wall / args.steps, ms[0] / cnt[0], ms[1] / cnt[1], vr))

    elem_ms = ms[0] / max(cnt[0], 1)
    eb = elem_kernel_bytes(mode, u, 1, nm) * N
    achieved = eb / (elem_ms * 1e-3)
    Bstep = step_bytes(mode, u, 1, nm=nm)
    rtype = "double" if mode == "fp64" else "float"
    s_state = mode_sizes(mode)[0]
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": None,
            "kernel": "elem_kernel<%d,%s,visco>, %d elements per workgroup, %s LDS layout, %s "
                      "workgroup order, state stored %s the divergence"
                      % (g.lpd, rtype, tile, "node-interleaved" if sims[mode].layout else "planar",
                         "XCD-aware" if sims[mode].remap else "plain",
                         "before" if sims[mode].early else "after"),
            "algorithmic_bytes_per_launch": eb,
            "algorithmic_bytes_per_node": eb / N,
            # the element force the element kernel writes for the node kernel
            # is outside the algorithmic bytes; with it the kernel moves:
            "force_write_bytes_per_node": 3 * s_state,
            "achieved_with_force_write": (eb + 3 * s_state * N) / (elem_ms * 1e-3) / 1e9}
    roof["achieved"] = achieved / 1e9
    roof["peak"] = HBM_PEAK / 1e9
    # HBM traffic per launch measured by rocprofv3 (FETCH_SIZE, WRITE_SIZE,
    # calibrated on a plain copy kernel) for this workload, when recorded
    tf = os.path.join(ROOT, "profiles", "traffic_layered_192x192x96_p4_%s.json" % mode)
    recorded = (g.nx, g.ny, g.nz, g.lpd) == (192, 192, 96, 4)
    tj = json.load(open(tf)) if recorded and os.path.exists(tf) else None
    if tj is not None and tj.get("memory_vars", 6) == nm:
        roof["traffic"] = tj["traffic_bytes_per_launch"]["elem_kernel"]
        roof["traffic_kind"] = ("FETCH_SIZE + WRITE_SIZE of the element kernel (memory-side "
                                "bytes below L2, including what the Infinity Cache serves), "
                                "calibrated on a plain copy kernel; includes the force write")
    step_frac = Bstep * value / HBM_PEAK

    out = {
        "metric": "GLL-node updates/s, FP32 vs FP64 GPU at <=1e-5 rel L2 seismogram misfit",
        "value": value, "unit": "GLL-node updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if mode == "fp64" else "f32",
        "data": "synthetic (seeded layered visco-elastic model, GRF vs perturbation)",
        "config": {"workload": "layered visco-elastic section 40x40 deg x 1400 km, "
                               "%dx%dx%d elements, GLL order %d, 3 SLS mechanisms%s, %s, "
                               "one source, 256 surface receivers and 32 on rings around "
                               "the source" % (g.nx, g.ny, g.nz, g.lpd,
                                               " (memory variables held as their deviatoric "
                                               "part, 5 words per mechanism)" if nm == 5 else "",
                                               MODE_TEXT[mode]),
                   "elements": g.n_elem, "gll_nodes": N, "unique_nodes": g.unique_nodes,
                   "precision": mode, "parallelism": "single GPU",
                   "state_buffers": 1 if args.state_buffers == 1 else 2,
                   "memory_vars": nm},
        "roofline": roof,
        "step_roofline": {"bytes_per_node": Bstep, "achieved_GBps": Bstep * value / 1e9,
                          "frac": step_frac,
                          "elem_ms": elem_ms, "node_ms": ms[1] / max(cnt[1], 1),
                          "record_ms": ms[2] / max(cnt[2], 1)},
    }
    if mode != "fp64" and g.lpd == 4:
        # the element kernel's arithmetic against the vector peaks (counted
        # for order 4): far from both, the kernel is bound by HBM
        f64 = ELEM_FP64_FLOP * N / (elem_ms * 1e-3)
        f32 = ELEM_FP32_FLOP[nm] * N / (elem_ms * 1e-3)
        out["flop_rate"] = {"kernel": "element kernel", "unit": "TFLOP/s",
                            "fp64_flop_per_node": ELEM_FP64_FLOP, "fp32_flop_per_node": ELEM_FP32_FLOP[nm],
                            "fp64_achieved": f64 / 1e12, "fp64_peak": FP64_PEAK / 1e12,
                            "fp64_frac": f64 / FP64_PEAK,
                            "fp32_achieved": f32 / 1e12, "fp32_peak": FP32_PEAK / 1e12,
                            "fp32_frac": f32 / FP32_PEAK}
    if "fp64" in modes and mode != "fp64":
        t32 = traces_at(sims[mode], misfit_steps) if misfit_steps > 0 else None
        for sim in sims.values():
            sim.close()
        sims = dict(load_sims(case, ["fp64"], args.slab, remap=not args.no_remap,
                              state_buffers=args.state_buffers))
        wall64, ms64, cnt64, _ = time_gpu(sims["fp64"], args.steps, args.warmup)
        v64 = N * args.steps / wall64
        out["fp64_gpu"] = {"value": v64, "ms_per_step": 1e3 * wall64 / args.steps,
                           "elem_ms": ms64[0] / max(cnt64[0], 1),
                           "node_ms": ms64[1] / max(cnt64[1], 1),
                           "speedup_fp32_over_fp64": value / v64,
                           "state_buffers": 1 if args.state_buffers == 1 else 2}
        log("[bench] fp64 %.3f ms/step" % (1e3 * wall64 / args.steps))
        if t32 is not None:
            t0 = time.time()
            out["fp32_vs_fp64_misfit"] = misfit(t32, traces_at(sims["fp64"], misfit_steps),
                                                misfit_steps)
            log("[bench] misfit leg to step %d in %.0f s: %s" % (misfit_steps, time.time() - t0,
                                                               out["fp32_vs_fp64_misfit"]))
    for sim in sims.values():
        sim.close()
    if not args.no_cpu:
        cb = cpu_baseline({"seed": 1})
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
        log("[bench] cpu baseline %.3e node-updates/s on %d cores" % (cb["value"], cb["cores"]))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
