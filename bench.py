#!/usr/bin/env python3
"""Benchmark: 2D isotropic acoustic FWI gradient on one MI355X.

Workload (one "step" = one FWI gradient over all shots): 401 x 401 grid,
dx = dz = 10 m, R = 4 (8th-order space), dt = 1 ms, 2000 steps, 40-cell
CFS-PML; true Vp = 1500->4000 m/s linear gradient with a seeded +-5 %
Gaussian-smoothed perturbation (8-cell correlation), Gardner rho; initial
model = true model smoothed with sigma = 10 cells; 32 Ricker (10 Hz)
explosion shots at z = 20 m, x = 140 + 120 k m; 391 receivers at z = 20 m
every 10 m; L2 misfit.  Observed data are modelled on the GPU from the true
model before the timed region; all inputs are resident in HBM when it starts.

Headline: the batched + checkpointed gradient -- all 32 shots in every
launch, exact discrete adjoint, forward states recomputed from 40 binomial
checkpoints (FP32) -- in grid-point updates per second (useful updates: one
forward and one adjoint sweep per step and shot; the recomputation is extra
work inside the timed region) and seconds per shot.  Next to it: FP16
checkpoints, the plain per-shot GPU path (one shot per launch, full FP32
storage, no checkpointing), OWL's own gradient algorithm (batched and per
shot), the same gradient through torch.autograd, the 3D acoustic batched
forward, the dominant kernels' HBM rooflines from HIP events over the timed
regions, and OWL's algorithm restated in C timed on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# HBM bytes per launch of each workload's dominant kernel, measured on one
# MI355X by two rocprofv3 counter passes over the current kernels:
# 2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE (KiB per dispatch), the factors that a
# 1 GiB copy, elementwise add and reduction give.  Counters cannot run inside
# a timed process, so the values are carried here; profiles/ holds the
# commands and the raw counts.
TRAFFIC = {
    # the fused 2D kernels at the z rows per tile the autotuner picks on this
    # workload (acoustic2d_fwi_gradient_32shots.md)
    ("acoustic2d_gradient", "step_fused", 16): (2.0 * 144.3e3 + 116.1e3) * 1024,
    ("acoustic2d_gradient", "adjoint_fused", 8): (2.0 * 405.5e3 + 304.9e3) * 1024,
    ("acoustic2d_gradient", "step_fused", 12): (2.0 * 151.9e3 + 115.1e3) * 1024,
    ("acoustic2d_gradient", "adjoint_fused", 6): (2.0 * 433.9e3 + 306.3e3) * 1024,
    ("acoustic3d_forward", "update_v"): (2.0 * 4.159e6 + 4.199e6) * 1024,    # acoustic3d_forward_256cube_16shots.md
    ("acoustic3d_forward", "update_p"): (2.0 * 4.318e6 + 1.765e6) * 1024,
    # (z planes, y rows per tile): elastic3d_forward_200x200x150_8shots.md
    ("elastic3d_forward", "stress", (64, 4)): (2.0 * 6.017e6 + 2.686e6) * 1024,
    ("elastic3d_forward", "velocity", (64, 4)): (2.0 * 5.066e6 + 1.602e6) * 1024,
    ("elastic3d_forward", "stress", (8, 4)): (2.0 * 3.730e6 + 2.581e6) * 1024,
    ("elastic3d_forward", "velocity", (8, 4)): (2.0 * 2.977e6 + 1.501e6) * 1024,
    # (rows per tile column, tiles per block) of the FSG sweeps:
    # fsg2d_forward_free_surface_801x401_64shots.md
    ("fsg2d_forward", "stress", (32, 4)): (2.0 * 7.750e5 + 3.767e5) * 1024,
    ("fsg2d_forward", "velocity", (32, 4)): (2.0 * 4.983e5 + 2.834e5) * 1024,
    ("fsg2d_forward", "stress", (8, 2)): (2.0 * 9.090e5 + 3.793e5) * 1024,
    ("fsg2d_forward", "velocity", (8, 2)): (2.0 * 6.567e5 + 2.862e5) * 1024,
    ("fsg2d_forward", "stress", (4, 2)): (2.0 * 9.996e5 + 3.790e5) * 1024,
    ("fsg2d_forward", "velocity", (4, 2)): (2.0 * 8.005e5 + 2.863e5) * 1024,
}


def with_traffic(workload, roof, rows=None):
    """The roofline with the measured HBM bytes per launch of its kernel
    (None where no counter pass covers the kernel at this launch shape)."""
    key = (workload, roof["kernel"]) if rows is None else (workload, roof["kernel"], rows)
    return dict(roof, traffic=TRAFFIC.get(key))


CFG = dict(nx=401, nz=401, dx=10.0, dz=10.0, dt=0.001, tmax=1.999, pml=40, fdhalf=4)
SNAPSHOTS = 40


def workload():
    from owl_amd.synthetic import acoustic2d_initial_model, acoustic2d_shots, acoustic2d_true_model
    vp, rho = acoustic2d_true_model(401, 401, 10.0, seed=0)
    vp0, rho0 = acoustic2d_initial_model(vp, rho, 10.0)
    shots = acoustic2d_shots(32)
    return vp, rho, vp0, rho0, shots


def make_batch(vp, rho, vp0, rho0, shots, tmax=None):
    from owl_amd import Acoustic2DBatch
    b = Acoustic2DBatch(CFG["nx"], CFG["nz"], CFG["dx"], CFG["dz"], vp, rho,
                        [(g, None) for g in shots], CFG["dt"], tmax or CFG["tmax"], pml=CFG["pml"],
                        fdhalf=CFG["fdhalf"])
    b.observe()                # observed data from the true model
    b.set_model(vp0, rho0)     # gradient at the initial model
    return b


def roofline_of(stats):
    """Dominant kernel (by device time among those with modelled bytes):
    algorithmic bytes per launch over its average launch time."""
    name, st = max(((k, v) for k, v in stats.items() if v["bytes_total"] > 0 and v["launches"]),
                   key=lambda kv: kv[1]["total_ms"])
    avg_ms = st["total_ms"] / st["launches"]
    bytes_launch = st["bytes_total"] / st["launches"]
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    return name, dict(bound="hbm", kernel=name, achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=achieved / HBM_PEAK_GBS, avg_launch_ms=avg_ms,
                      algorithmic_bytes_per_launch=bytes_launch)


def kernel_table(stats, steps):
    return {k: {"ms": round(v["total_ms"] / steps, 3), "launches": v["launches"] // steps,
                "GB/s": round(v["bytes_total"] / (v["total_ms"] * 1e-3) / 1e9, 1) if v["bytes_total"] else None}
            for k, v in stats.items() if v["launches"]}


def time_gradient(b, steps, warmup, sync, **kw):
    """Warm up (an exact gradient runs eagerly once, then its HIP graph is
    captured on the next call), time `steps` gradients with profiling off
    (graph replays), then run one more gradient eagerly with per-kernel HIP
    events for the roofline and the kernel table."""
    from owl_amd import OWL_GRAD_EXACT
    if kw.get("mode") == OWL_GRAD_EXACT:
        warmup = max(warmup, 2)
    for _ in range(warmup):
        b.gradient(want_traces=False, **kw)
    sync()
    t0 = time.perf_counter()
    updates = 0.0
    for _ in range(steps):
        _, gv, _, _ = b.gradient(want_traces=False, **kw)
        updates += b.last_stats()[1]
    sync()
    wall = time.perf_counter() - t0
    b.set_profiling(True)
    b.gradient(want_traces=False, **kw)
    stats = b.kernel_stats()
    b.set_profiling(False)
    return wall, updates, stats, gv


def host_facts():
    """CPU model, sockets, physical cores and the CPUs this process may use."""
    model, phys = None, set()
    cur = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if cur:
                        phys.add((cur.get("physical id"), cur.get("core id")))
                    cur = {}
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                cur[k] = v
                if k == "model name" and model is None:
                    model = v
        if cur:
            phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    sockets = len({p for p, _ in phys if p is not None}) or None
    return dict(cpu_model=model, sockets=sockets, physical_cores=len(phys) or None,
                logical_cpus=os.cpu_count(), usable_cpus=len(os.sched_getaffinity(0)))


def _cpu_gradient_shots(cpus, models, shot_list, nt, barrier, results):
    """Worker: OWL's gradient algorithm restated in C (the oracle, float32,
    OpenMP) on the given shots, one after another: forward for the misfit,
    forward with boundary saving, OWL's reverse pass.  Runs in a fresh
    process so that its OpenMP settings take effect; the observed data are
    modelled before the barrier, outside the timing."""
    os.sched_setaffinity(0, cpus)   # disjoint CPU sets per process: OpenMP places bind within them
    os.environ["OMP_NUM_THREADS"] = str(len(cpus))
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    vp, rho, vp0, rho0 = models
    o = Oracle(32)
    g = dict(nx=CFG["nx"], nz=CFG["nz"], dx=CFG["dx"], dz=CFG["dz"], dt=CFG["dt"], nt=nt,
             pml=CFG["pml"], R=CFG["fdhalf"])
    obs = [o.acoustic2d_forward(g, vp, rho, src, rec) for src, rec in shot_list]
    barrier.wait()
    t0 = time.perf_counter()
    for (src, rec), ob in zip(shot_list, obs):
        syn = o.acoustic2d_forward(g, vp0, rho0, src, rec)
        _, adj = o.adjoint_source_l2([x["weight"] for x in rec], syn, ob)
        o.acoustic2d_gradient_owl(g, vp0, rho0, src, rec, adj)
    results.put(time.perf_counter() - t0)


def _cpu_layout(procs, threads, models, shot_list, nt):
    """Wall time of `procs` worker processes with `threads` threads each on
    disjoint CPU sets, the shots dealt round-robin, all starting at one
    barrier."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    barrier, results = ctx.Barrier(procs), ctx.Queue()
    cpus = sorted(os.sched_getaffinity(0))
    ps = [ctx.Process(target=_cpu_gradient_shots,
                      args=(cpus[k * threads:(k + 1) * threads], models, shot_list[k::procs], nt, barrier,
                            results)) for k in range(procs)]
    for p in ps:
        p.start()
    times = [results.get() for _ in ps]
    for p in ps:
        p.join()
    return max(times)


def cpu_baseline(vp, rho, vp0, rho0, shots, nt, nshots=4, procs=4):
    """OWL's gradient algorithm restated in C (oracle, float32, OpenMP) on a
    bounded sample of the gradient workload, in the two layouts of OWL's
    runs: one process with every thread and the shots in sequence, and
    `procs` processes with threads/procs threads each and the shots split
    across them (OWL's shot groups).  Reports the faster layout."""
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
    threads = min(threads, len(os.sched_getaffinity(0)))
    procs = max(1, min(procs, threads, nshots))
    sample = [shots[(k * len(shots)) // nshots] for k in range(nshots)]
    shot_list = [([dict(x=s.x, z=s.z, f0=s.f0) for s in g.srcr],
                  [dict(x=r.x, z=r.z, weight=r.weight) for r in g.recr]) for g in sample]
    models = (vp, rho, vp0, rho0)
    layouts = {}
    for p in sorted({1, procs}):
        wall = _cpu_layout(p, threads // p, models, shot_list, nt)
        layouts[f"{p}x{threads // p}"] = dict(processes=p, threads_each=threads // p, s_per_shot=wall / nshots)
    best = min(layouts, key=lambda k: layouts[k]["s_per_shot"])
    s_shot = layouts[best]["s_per_shot"]
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    _, hnt = Oracle(32).stf(shots[0].srcr[0].f0, CFG["dt"])
    pts = (CFG["nx"] + 2 * CFG["pml"]) * (CFG["nz"] + 2 * CFG["pml"])
    updates_shot = pts * (2 * nt + 2 * (nt - hnt + 1))
    return dict(value=updates_shot / s_shot, unit="grid-pt updates/s", cores=threads, kind="port",
                sample=f"{nshots} of the 32 shots, full {nt} steps each: forward for the misfit, forward with "
                       f"boundary saving, OWL reverse pass; the faster of the two layouts",
                layouts=layouts, best_layout=best, s_per_shot=s_shot,
                gradient_32_shots_s=(32 * s_shot, "extrapolated from the sample" if nshots < 32 else "measured"),
                omp=dict(OMP_NUM_THREADS=threads, OMP_PROC_BIND="close", OMP_PLACES="cores"),
                host=host_facts(),
                io="boundary strips kept in memory: the per-step scratch-file I/O of OWL's reconstruction is "
                   "neither reproduced nor timed")


def bench_3d(sync, steps=1500, shape=None):
    """3D acoustic batched forward: 256^3, 10 m, R = 4, 1500 steps, 20-cell
    PML, 5-layer model, 16 shots with uneven 128..256 windows, 100 x 100
    receivers each.  dt = 0.9 ms: 1 ms exceeds the scheme's stable step
    (0.966 ms at vmax = 4500 m/s), which the library refuses."""
    from owl_amd import Acoustic3DBatch
    from owl_amd.synthetic import acoustic3d_layered_model, acoustic3d_shots
    vp, rho = acoustic3d_layered_model()
    shots = acoustic3d_shots()
    dt = 0.0009
    b = Acoustic3DBatch(256, 256, 256, 10.0, 10.0, 10.0, vp, rho, shots, dt, (steps - 1) * dt, pml=20, fdhalf=4)
    if shape:
        kc, var = shape
        b.set_launch(kc, var)
    else:
        kc, var = b.autotune(3)
    # per-kernel figures from one eager forward with HIP events
    b.set_profiling(True)
    sync()
    t0 = time.perf_counter()
    b.forward(want_traces=False)
    sync()
    wall_eager = time.perf_counter() - t0
    stats = b.kernel_stats()
    b.set_profiling(False)
    # the timed forward: the whole time loop as one captured HIP graph
    # (captured by the untimed call, replayed by the timed one)
    b.set_graphs(True)
    b.forward(want_traces=False)
    sync()
    t0 = time.perf_counter()
    b.forward(want_traces=False)
    sync()
    wall = time.perf_counter() - t0
    ms, updates = b.last_stats()
    b.close()
    _, roof = roofline_of(stats)
    return {"workload": f"3D acoustic batched forward, 256^3, 16 shots (windows 128-256), {steps} steps, "
                        "dt 0.9 ms, R 4, pml 20, one HIP-graph launch of the time loop",
            "grid_pt_updates_per_s": updates / wall, "seconds": wall, "device_ms": ms,
            "eager_seconds_with_events": wall_eager,
            "launch_shape": {"planes": kc, "variant": var}, "roofline": with_traffic("acoustic3d_forward", roof),
            "kernels": kernel_table(stats, 1)}


def bench_elastic3d(sync, steps, gradient_shots=4, shape=None):
    """3D isotropic elastic batched forward: 200 x 200 x 150, 10 m, R = 4,
    dt 0.8 ms, 20-cell CFS-MPML on five faces and a flat free surface,
    vp a smooth random field in 2000..5000 m/s, vs = vp/1.73, Gardner rho,
    8 vertical-force shots, 50 x 50 three-component receivers, at the
    autotuned launch shape: one eager run with per-kernel events (kernel table
    and roofline), then the loop as one captured HIP graph (captured by an
    untimed run, the replay timed).  Then OWL's
    gradient of vp, vs, rho over the first `gradient_shots` shots (all 8 by
    default): boundary records of every step in HBM as BF16 with one scale
    per shot and step (13 GB a shot), reverse reconstruction and adjoint,
    FP64 imaging."""
    from owl_amd import OWL_STORE_BF16, Elastic3DBatch
    from owl_amd.synthetic import elastic3d_model, elastic3d_shots
    vp, vs, rho = elastic3d_model()
    shots = elastic3d_shots()
    dt = 0.0008
    b = Elastic3DBatch(200, 200, 150, 10.0, 10.0, 10.0, vp, vs, rho, shots, dt, (steps - 1) * dt, pml=20,
                       fdhalf=4, free_surface=True)
    if shape:
        b.set_launch(*shape)
    else:
        shape = b.autotune(3)
    b.set_profiling(True)
    sync()
    t0 = time.perf_counter()
    b.forward(want_traces=False)
    sync()
    ev_wall = time.perf_counter() - t0
    stats = b.kernel_stats()
    ev_ms, _ = b.last_stats()
    b.set_profiling(False)
    b.forward(want_traces=False)  # captures the time loop
    sync()
    t0 = time.perf_counter()
    b.forward(want_traces=False)
    sync()
    wall = time.perf_counter() - t0
    ms, updates = b.last_stats()
    b.close()
    grad = None
    if gradient_shots:
        sub = shots[:gradient_shots]
        bg = Elastic3DBatch(200, 200, 150, 10.0, 10.0, 10.0, vp, vs, rho, sub, dt, (steps - 1) * dt, pml=20,
                            fdhalf=4, free_surface=True)
        bg.set_launch(*shape)
        bg.set_record_storage(OWL_STORE_BF16)
        obs = [tuple(np.zeros((nr, bg.nt), np.float32) for _ in range(3)) for nr in bg.nr]
        sync()
        t0 = time.perf_counter()
        _, gv, gs, gr, _ = bg.gradient(obs)
        sync()
        gw = time.perf_counter() - t0
        gms, gup = bg.last_stats()
        bg.close()
        grad = {"shots": len(sub), "seconds": gw, "s_per_shot": gw / len(sub), "device_ms": gms,
                "grid_pt_updates_per_s": gup / gw,
                "algorithm": "OWL's gradient of vp, vs, rho: forward with boundary records in HBM (BF16, one "
                             "scale per shot and step), L2 adjoint source per component (observed traces zero), "
                             "reverse reconstruction and adjoint, FP64 imaging, merged on the model; first call "
                             "(allocations included)",
                "records": "bf16",
                "finite": bool(np.isfinite(gv).all() and np.isfinite(gs).all() and np.isfinite(gr).all())}
    _, roof = roofline_of(stats)
    return {"workload": f"3D isotropic elastic batched forward, 200x200x150, 10 m, dt 0.8 ms, R 4, pml 20 on five "
                        f"faces, flat free surface, 8 shots, 50x50 3-C receivers, {steps} steps",
            "timing": "one captured HIP graph of the time loop, replayed (captured by an untimed run)",
            "grid_pt_updates_per_s": updates / wall, "seconds": wall, "device_ms": ms,
            "evented_eager": {"seconds": ev_wall, "device_ms": ev_ms},
            "launch_shape": list(shape), "roofline": with_traffic("elastic3d_forward", roof, tuple(shape)),
            "kernels": kernel_table(stats, 1), "gradient": grad}


def bench_fsg(sync, steps, gradient=True, shape=None):
    """2D tilted-TI elastic on the Lebedev fully staggered grid with OWL's
    flat free surface on the depth-refined mesh (free_surface_dz_refine 4,
    dz_max 1.5 dz), batched: 801 x 401, 5 m, R = 4, dt 0.4 ms, 20-cell
    CFS-MPML with the Christoffel-scan damping ratios, 64 shots with uneven
    200..801-column windows, vertical forces, two-component receivers every
    10 m.  The sweeps' launch shape is the autotuner's (or `shape`).  The
    forward runs once eagerly with per-kernel events (the kernel table and
    roofline), then as one captured HIP graph of the time loop: captured by
    an untimed run, the replay is timed.  Then OWL's gradient
    (boundary records in HBM as FP16 with one scale per shot and step,
    reconstruction, c11 .. c55 and rho, FP64 imaging, mapped back to the
    regular windows and summed) over all shots, without events."""
    from owl_amd import OWL_STORE_F16, FSG2DBatch
    from owl_amd.synthetic import fsg2d_shots, fsg2d_tti_model
    cij, rho = fsg2d_tti_model()
    shots = fsg2d_shots()
    dt = 0.0004
    t0 = time.perf_counter()
    b = FSG2DBatch(801, 401, 5.0, 5.0, cij, rho, shots, dt, (steps - 1) * dt, pml=20, fdhalf=4,
                   surface=dict(refine=4.0))
    plan_s = time.perf_counter() - t0
    rows = [b.rows(s) for s in range(len(shots))]
    if shape:
        b.set_launch(*shape)
    else:
        shape = b.autotune(10)
    b.set_profiling(True)
    sync()
    t0 = time.perf_counter()
    b.forward(want_traces=False)
    sync()
    ev_wall = time.perf_counter() - t0
    stats = b.kernel_stats()
    ev_ms, _ = b.last_stats()
    b.set_profiling(False)
    b.forward(want_traces=False)  # captures the time loop
    sync()
    t0 = time.perf_counter()
    b.forward(want_traces=False)
    sync()
    wall = time.perf_counter() - t0
    ms, updates = b.last_stats()
    grad = None
    if gradient:
        b.set_record_storage(OWL_STORE_F16)
        obs = [tuple(np.zeros((nr, b.nt), np.float32) for _ in range(2)) for nr in b.nr]
        b.gradient(obs)  # untimed warm-up: buffers, strips
        sync()
        t0 = time.perf_counter()
        mis, g, _ = b.gradient(obs)
        sync()
        gw = time.perf_counter() - t0
        gms, gup = b.last_stats()
        grad = {"seconds": gw, "s_per_shot": gw / len(shots), "device_ms": gms,
                "grid_pt_updates_per_s": gup / gw,
                "algorithm": "OWL's gradient: forward with boundary records in HBM (FP16, one scale per shot "
                             "and step), L2 adjoint source per component (observed traces zero), reverse "
                             "reconstruction and adjoint, FP64 imaging of c11..c55 and rho on the refined "
                             "meshes, mapped to the regular windows and summed on the host",
                "records": "fp16",
                "finite": bool(all(np.isfinite(v).all() for v in g.values()))}
    b.close()
    _, roof = roofline_of(stats)
    return {"workload": f"2D tilted-TI FSG batched forward with OWL's flat free surface on the depth-refined "
                        f"mesh (refinement 4, dz_max 7.5 m; {min(rows)}-{max(rows)} refined rows per window), "
                        f"801x401, 5 m, dt 0.4 ms, R 4, pml 20, 64 shots (windows 200-801), {steps} steps",
            "launch_shape": list(shape),
            "timing": "one captured HIP graph of the time loop, replayed (captured by an untimed run)",
            "grid_pt_updates_per_s": updates / wall, "seconds": wall, "device_ms": ms,
            "evented_eager": {"seconds": ev_wall, "device_ms": ev_ms},
            "plan_seconds": round(plan_s, 3), "roofline": with_traffic("fsg2d_forward", roof, tuple(shape)),
            "kernels": kernel_table(stats, 1), "gradient": grad}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed FWI gradients")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-shots", type=int, default=4,
                    help="shots timed through the plain per-shot path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-shots", type=int, default=4, help="shots of the CPU baseline sample (32: all)")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D forward")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the batched exact gradient (for profiler runs)")
    ap.add_argument("--eager", action="store_true", help="no HIP graphs (for counter runs)")
    ap.add_argument("--rows", default="", metavar="FWD,ADJ",
                    help="fixed z rows of the fused forward and adjoint steps instead of the autotuner's "
                         "(counter runs without the tuner's trial launches)")
    ap.add_argument("--time-steps", type=int, default=0,
                    help="with --headline-only: that many time steps instead of 2000 (short counter runs "
                         "with the same launches)")
    ap.add_argument("--only-3d", action="store_true", help="only the 3D batched forward (for profiler runs)")
    ap.add_argument("--steps-3d", type=int, default=1500,
                    help="with --only-3d: time steps of the 3D forward (short counter runs)")
    ap.add_argument("--shape-3d", default="", metavar="PLANES,VARIANT",
                    help="with --only-3d: this launch shape instead of the autotuner's")
    ap.add_argument("--no-fsg", action="store_true", help="skip the 2D FSG forward")
    ap.add_argument("--no-fsg-gradient", action="store_true", help="skip the 2D FSG gradient")
    ap.add_argument("--no-elastic3d", action="store_true", help="skip the 3D elastic forward")
    ap.add_argument("--only-elastic3d", action="store_true", help="only the 3D elastic forward (for profiler runs)")
    ap.add_argument("--elastic3d-steps", type=int, default=2500, help="time steps of the 3D elastic forward")
    ap.add_argument("--elastic3d-shape", default="", metavar="PLANES,ROWS",
                    help="this launch shape of the 3D elastic sweeps instead of the autotuner's (counter runs)")
    ap.add_argument("--elastic3d-gradient-shots", type=int, default=None,
                    help="shots of the timed 3D elastic gradient (default 8; 0 skips it; with "
                         "--only-elastic3d default 0)")
    ap.add_argument("--only-fsg", action="store_true", help="only the 2D FSG forward (for profiler runs)")
    ap.add_argument("--fsg-steps", type=int, default=5000, help="time steps of the 2D FSG forward")
    ap.add_argument("--fsg-shape", default="", metavar="PLANES,ROWS",
                    help="this launch shape of the FSG sweeps instead of the autotuner's (counter runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device five times (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(shots would shard across GPUs with one gradient all-reduce, not built here)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        c = torch.empty_like(a)
        for _ in range(5):
            c.copy_(a)
        # the same bytes through torch's elementwise kernel (read n, write n)
        # and a reduction (read n)
        for _ in range(5):
            torch.add(a, 1.0, out=c)
        for _ in range(5):
            a.sum()
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5,
                          "elementwise_add": {"bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 5},
                          "sum": {"bytes_read": 4 * n, "launches": 5}}))
        return
    from owl_amd import (OWL_CKPT_BINOMIAL, OWL_CKPT_FULL, OWL_GRAD_EXACT, OWL_GRAD_OWL, OWL_STORE_F16,
                         OWL_STORE_F32)
    sync = torch.cuda.synchronize
    log = lambda m: print("bench: " + m, file=sys.stderr, flush=True)  # noqa: E731
    fsg_shape = tuple(int(v) for v in args.fsg_shape.split(",")) if args.fsg_shape else None
    el3_shape = tuple(int(v) for v in args.elastic3d_shape.split(",")) if args.elastic3d_shape else None
    if args.only_3d:
        shape = tuple(int(v) for v in args.shape_3d.split(",")) if args.shape_3d else None
        print(json.dumps({"acoustic3d_forward": bench_3d(sync, args.steps_3d, shape)}))
        return
    if args.only_elastic3d:
        print(json.dumps({"elastic3d_forward": bench_elastic3d(sync, args.elastic3d_steps,
                                                               args.elastic3d_gradient_shots or 0, el3_shape)}))
        return
    if args.only_fsg:
        print(json.dumps({"fsg2d_forward": bench_fsg(sync, args.fsg_steps, not args.no_fsg_gradient,
                                                     fsg_shape)}))
        return
    log("building workload")
    vp, rho, vp0, rho0, shots = workload()
    nshot = len(shots)
    exact_b40 = dict(mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_BINOMIAL, ckpt_param=SNAPSHOTS, store=OWL_STORE_F32)

    # ---- batched + checkpointed (headline)
    b = make_batch(vp, rho, vp0, rho0, shots,
                   tmax=(args.time_steps - 1) * CFG["dt"] if args.headline_only and args.time_steps > 1 else None)
    b.set_graphs(not args.eager)
    if args.rows:
        zf, za = (int(v) for v in args.rows.split(","))
        b.set_step_kernel("auto", zf)
        b.set_adjoint_rows(za)
    nt = b.nt
    log("batched exact gradient, binomial 40 snapshots, FP32")
    wall, updates, stats, g32 = time_gradient(b, args.steps, args.warmup, sync, **exact_b40)
    rows = b.fused_rows()
    value = updates / wall
    s_shot = wall / args.steps / nshot
    name, roof = roofline_of(stats)
    if args.headline_only:
        print(json.dumps({"value": value, "s_per_shot": s_shot, "roofline": roof, "kernels": kernel_table(stats, 1)}))
        b.close()
        return
    log("batched exact gradient, binomial 40 snapshots, FP16")
    w16, _, st16, g16 = time_gradient(b, 1, 1, sync, **dict(exact_b40, store=OWL_STORE_F16))
    log("batched OWL-algorithm gradient")
    wowl, _, stowl, _ = time_gradient(b, 1, 1, sync, mode=OWL_GRAD_OWL)

    log("torch.autograd gradient (binomial 40, FP32), device tensors and host tensors")
    from owl_amd.autograd import propagate
    b.set_model(vp, rho)
    obs_h = torch.from_numpy(np.concatenate(b.forward(), axis=0))

    def autograd_once(device):
        obs = obs_h.to(device)
        tv = torch.tensor(vp0, requires_grad=True, device=device)
        tr_ = torch.tensor(rho0, requires_grad=True, device=device)
        sync()
        t0 = time.perf_counter()
        syn = propagate(tv, tr_, b, OWL_CKPT_BINOMIAL, SNAPSHOTS, OWL_STORE_F32)
        loss = 0.5 * ((syn - obs) ** 2).sum()
        loss.backward()
        sync()
        return time.perf_counter() - t0

    autograd_once("cuda")  # untimed warm-up
    w_autograd = autograd_once("cuda")
    w_autograd_host = autograd_once("cpu")
    b.close()

    # ---- plain per-shot: one shot per launch, full FP32 storage
    log("plain per-shot exact gradient (full storage) and OWL gradient")
    plain, plain_owl, plain_graph = [], [], []
    for k in range(min(args.plain_shots, nshot)):
        bs = make_batch(vp, rho, vp0, rho0, [shots[k]])
        bs.set_graphs(False)   # the plain path launches eagerly
        if k == 0:
            bs.gradient(want_traces=False, mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_FULL, store=OWL_STORE_F32)
            bs.gradient(want_traces=False, mode=OWL_GRAD_OWL)
        sync()
        t0 = time.perf_counter()
        bs.gradient(want_traces=False, mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_FULL, store=OWL_STORE_F32)
        sync()
        plain.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        bs.gradient(want_traces=False, mode=OWL_GRAD_OWL)
        sync()
        plain_owl.append(time.perf_counter() - t0)
        # the same per-shot path with its launches replayed from a HIP graph:
        # the eager call above records the signature, the next one captures
        bs.set_graphs(True)
        for _ in range(2):
            bs.gradient(want_traces=False, mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_FULL, store=OWL_STORE_F32)
        sync()
        t0 = time.perf_counter()
        bs.gradient(want_traces=False, mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_FULL, store=OWL_STORE_F32)
        sync()
        plain_graph.append(time.perf_counter() - t0)
        bs.close()
    plain_s = float(np.mean(plain))
    plain_owl_s = float(np.mean(plain_owl))
    plain_graph_s = float(np.mean(plain_graph))

    three_d = None
    if not args.no_3d:
        log("3D batched forward")
        three_d = bench_3d(sync)

    el3 = None
    if not args.no_elastic3d:
        log("3D elastic batched forward")
        el3 = bench_elastic3d(sync, args.elastic3d_steps,
                              8 if args.elastic3d_gradient_shots is None else args.elastic3d_gradient_shots, el3_shape)

    fsg = None
    if not args.no_fsg:
        log("2D FSG batched forward and gradient")
        fsg = bench_fsg(sync, args.fsg_steps, not args.no_fsg_gradient, fsg_shape)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline")
        cpu = cpu_baseline(vp, rho, vp0, rho0, shots, nt, nshots=args.cpu_shots)

    rel16 = float(np.linalg.norm(g16 - g32) / np.linalg.norm(g32))
    out = {
        "metric": "grid-pt updates/s & FWI gradient s/shot, batched+ckpt vs plain, 1 MI355X",
        "value": value,
        "unit": "grid-pt updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "2D acoustic FWI gradient, 32 shots batched, exact adjoint, binomial "
                               "checkpoints (40 snapshots, FP32)",
                   "grid": "401x401", "dx_m": 10.0, "fd_order": 8, "dt_s": 0.001, "nt": nt,
                   "pml": 40, "shots": nshot, "receivers": 391, "misfit": "L2",
                   "accumulation": "fp64", "fused_rows": {"step": rows[0], "adjoint": rows[1]}},
        "fwi_gradient_s_per_shot": s_shot,
        "batched_fp16_checkpoints": {"s_per_shot": w16 / nshot,
                                     "gradient_rel_l2_vs_fp32": rel16},
        "plain_per_shot": {"s_per_shot": plain_s, "shots_timed": len(plain),
                           "path": "one shot per launch, exact adjoint, full FP32 storage, eager launches"},
        "speedup_batched_ckpt_vs_plain": plain_s / s_shot,
        "plain_per_shot_graph": {"s_per_shot": plain_graph_s, "shots_timed": len(plain_graph),
                                 "path": "one shot per launch, exact adjoint, full FP32 storage, HIP-graph "
                                         "replay of the whole gradient"},
        "speedup_batched_ckpt_vs_plain_graph": plain_graph_s / s_shot,
        "owl_algorithm": {"batched_s_per_shot": wowl / nshot, "plain_s_per_shot": plain_owl_s,
                          "speedup": plain_owl_s / (wowl / nshot)},
        "autograd_s_per_shot": w_autograd / nshot,
        "autograd_host_tensors_s_per_shot": w_autograd_host / nshot,
        "roofline": with_traffic("acoustic2d_gradient", roof,
                                 rows[0] if roof["kernel"] == "step_fused" else rows[1]),
        "kernels": kernel_table(stats, 1),
        "timing": "value and ms_per_step: graph-replayed gradients; roofline and kernels: HIP events "
                  "of one eager gradient run right after the timed region",
        "acoustic3d_forward": three_d,
        "elastic3d_forward": el3,
        "fsg2d_forward": fsg,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
