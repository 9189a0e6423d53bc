#!/usr/bin/env python3
"""Benchmark of the GraSPH2 derivative + leapfrog step on one MI355X.

A step is the reference's main-loop body (src/main.cu:256-257):
computeDerivatives(pb) then integrate(pb, true), here one
grasph_compute_derivatives + one grasph_integrate on device-resident
particles.  Default workload: the two-sphere solid collision, N = 131072
(SPH solid + von Mises + Balsara + artificial viscosity + self-gravity,
FP64 position drift).  Prints one JSON line.

value = pair interactions per second = N (N - 1) / (seconds per step): every
step evaluates softened gravity on all ordered pairs i != j.  The timed
pass launches the steps one by one, or, for gravity alone (whose step holds
no Morton re-sort), replays them from a HIP graph of config.graph_steps
steps captured on a stream of its own (the same kernels and arguments).
roofline: the dominant kernel, FP32 vector ALU bound, at 20 algorithmic
FLOPs per ordered gravity pair (FMA = 2, rsqrt = 1), divided by its mean
duration from HIP events recorded around every launch on the stream it
runs on, over a second pass of as many steps (the timed pass records no
events: they lengthen a step by ~2%).  From 16384 particles (the symmetric
sweep)
that kernel is the far field (far_kernel): the ordered pairs of particles in
different groups of 1024 slots, N (N - 1) minus the in-group pairs, 99% of
the pairs at N = 131072; the near field, the list merge and the SPH
passes run beside it on a second stream.  Below, it
is the one-sided sweep over all N (N - 1) pairs (its neighbour search rides
on the same r^2 and adds no counted FLOPs).
cpu_baseline: the CPU oracle in FP32 (the same op structure as the
reference's host path) on a bounded sample of target particles i against
all j, all three passes; one OpenMP thread per CPU of the process's
affinity set (capped by OMP_NUM_THREADS when set), with nproc and the CPU
model recorded beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (AMD published)
PEAK_HBM_GBPS = 8000.0     # MI355X HBM3E (AMD published)
# The library's default for the single-instance sweep (option "sweep_symmetric" = 2, include/grasph2.h):
# from this many particles on, gravity between groups of 1024 is evaluated once per pair for both sides.
# The roofline stays algorithmic either way: 20 FLOP per ordered pair.
SYM_AUTO_MIN_LISTS = 16384    # with neighbour lists (SPH on)
SYM_AUTO_MIN_GRAVITY = 16384  # gravity alone
SYM_GROUP = 1024  # slots per group of the symmetric sweep; pairs inside a group are the near field's
FLOP_GRAVITY_PAIR = 20
# Per neighbour pair, the FLOPs of the passes' formulas (FMA 2; add, sub, mul 1; sqrt, rsqrt, div 1;
# compensation 0), data-dependent branches at their typical share: W's inner branch (q < 1/2, 1/8 of
# the neighbours by volume) costs 4 more than the outer; the artificial viscosity of approaching
# pairs (about half) 11 more than that of receding ones.
FLOP_DENSITY_NBR = 17.5        # r 3, r^2 5, sqrt, q, 1 - q, W 3 (7 inside), prefactor, m_j W, sum
FLOP_STRAIN_NBR_SOLID = 75     # r 3, dW/r 12, g 3, v 3, m_j / rho_j, -vol / 2, 9 rate entries x 5, div v 7
FLOP_STRAIN_NBR_LIQUID = 44    # the curl (15) instead of the rate tensors (46); 29 with div v alone
FLOP_ACCEL_NBR_SOLID = 72.5    # r 3, r^2 5, sqrt, gradient 15, stress 6 + 18, v 3, AV 6 + 11 / 2 + 7, sums 3
FLOP_ACCEL_NBR_LIQUID = 55.5   # the stress term 7 instead of 24
INTEGRATOR_BYTES_SOLID = 168   # x64 24 + v 12 + a 12 + S 24 + dS 24 read; x64 24 + xf 12 + v 12 + S 24 written
INTEGRATOR_BYTES_NBODY = 96    # x64 24 + v 12 + a 12 read; x64 24 + xf 12 + v 12 written


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed steps (default 10; 100 for the Plummer workload, whose steps take ~0.1 ms)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="collision", choices=["collision", "contact", "plummer", "liquid", "large"])
    ap.add_argument("--advance", type=int, default=0,
                    help="untimed steps before the warm-up (e.g. to carry the collision into contact)")
    ap.add_argument("--n", type=int, default=0, help="override the particle count")
    ap.add_argument("--graph", type=int, default=-1, metavar="G",
                    help="steps per captured HIP graph in the timed pass (0: plain launches); default 10 for "
                         "gravity alone, whose step holds no Morton re-sort, else 0 (the re-sort every "
                         "reorder_interval steps stays out of graphs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU time of the baseline sample")
    ap.add_argument("--ensemble", type=int, default=512,
                    help="instances of the batched-ensemble measurement (0 skips it)")
    ap.add_argument("--ensemble-steps", type=int, default=5)
    ap.add_argument("--loop-steps", type=int, default=2)
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="context option (grasph_set_option) for every context the run creates; repeatable")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run device-to-device copies of MIB MiB (the reference kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    return ap.parse_args()


def ensemble_measure(ctx, count, steps, loop_steps):
    """Batched ensemble of solid collisions, N_k ~ U[1024, 16384]: steps/s of
    the captured-graph batched step against the plain per-instance launch
    loop on the same GPU (same kernels, same results)."""
    import numpy as np

    import grasph2_amd as g
    from grasph2_amd import ics

    inst, _ = ics.ensemble(count)
    host, offsets = g.pack_batch([h for h, _ in inst])
    params = [p for _, p in inst]
    del inst
    b = ctx.batch(offsets, params)
    sizes = np.diff(offsets).astype(np.float64)
    pairs = float(np.sum(sizes * (sizes - 1)))
    d = g.DeviceParticles.from_host(host)
    t = time.perf_counter()
    ctx.autotune(d, b)
    tune_s = time.perf_counter() - t
    ctx.batch_step(d, b, 1)  # warm-up: builds the order, lists and the graph
    g.hip.synchronize()
    t0 = time.perf_counter()
    ctx.batch_step(d, b, steps)
    g.hip.synchronize()
    t1 = time.perf_counter()
    graph_sps = steps / (t1 - t0)
    d.free()
    dl = g.DeviceParticles.from_host(host)
    ctx.batch_step_loop(dl, b, 1)  # warm-up of the per-instance contexts
    g.hip.synchronize()
    t0 = time.perf_counter()
    ctx.batch_step_loop(dl, b, loop_steps)
    g.hip.synchronize()
    t1 = time.perf_counter()
    loop_sps = loop_steps / (t1 - t0)
    dl.free()
    return {"instances": count, "particles": int(offsets[-1]), "pairs_per_step": pairs,
            "steps": steps, "steps_per_s": graph_sps, "pair_interactions_per_s": pairs * graph_sps,
            "ms_per_step": 1e3 / graph_sps, "loop_steps": loop_steps, "loop_steps_per_s": loop_sps,
            "loop_ms_per_step": 1e3 / loop_sps, "speedup_vs_loop": graph_sps / loop_sps,
            "tuning": ctx.batch_tuning(), "autotune_s": tune_s}


def main():
    args = parse()
    if args.steps <= 0:
        args.steps = 100 if args.workload == "plummer" else 10
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (the sweep would shard by "
              "target particle, with an all-gather of the positions per step); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import grasph2_amd as g
    from grasph2_amd import ics

    if args.calibrate_copy > 0:
        nbytes = args.calibrate_copy << 20
        src, dst = g.hip.DeviceArray(nbytes), g.hip.DeviceArray(nbytes)
        src.zero()
        for _ in range(3):
            dst.copy_from(src)
        g.hip.synchronize()
        src.free()
        dst.free()

    kw = {"n": args.n} if args.n else {}
    h, prm = ics.workload(args.workload, **kw)
    n = h["posm"].shape[0]
    solid = bool(prm.flags & g.SOLIDS)
    sph = bool(prm.flags & g.SPH)

    ctx = g.Context(0)
    for kv in args.option:
        key, val = kv.split("=", 1)
        ctx.set_option(key, int(val))
    d = g.DeviceParticles.from_host(h)
    g.hip.synchronize()
    graph_steps = args.graph if args.graph >= 0 else (0 if sph else 10)
    if graph_steps > 0 and args.steps % graph_steps:
        graph_steps = 1
    st = g.hip.Stream() if graph_steps > 0 else None  # a graph is captured from a stream of its own

    def step(stream=st):
        ctx.compute_derivatives(d, prm, stream)
        ctx.integrate(d, prm, True, stream)

    def sync():
        if st is not None:
            st.synchronize()
        g.hip.synchronize()

    for k in range(args.advance):
        step()
        if (k + 1) % 200 == 0:
            sync()
            print("advanced %d of %d steps" % (k + 1, args.advance), file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    sync()
    if graph_steps > 0:
        # the same steps, replayed: graph_steps steps per launch
        graph = g.hip.Graph(st, lambda s: [step(s) for _ in range(graph_steps)])
        graph.launch(st)  # a replay's worth of warm-up
        sync()
        t0 = time.perf_counter()
        for _ in range(args.steps // graph_steps):
            graph.launch(st)
        sync()
        t1 = time.perf_counter()
        del graph
    else:
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        sync()
        t1 = time.perf_counter()
    # per-kernel times from HIP events in a second timed pass of as many
    # steps (recording events around every launch lengthens a step, so
    # ms_per_step comes from the pass above, which records none)
    ctx.set_option("profile", 1)
    ctx.kernel_times(reset=True)
    for _ in range(args.steps):
        step()
    sync()
    kt = ctx.kernel_times(reset=True)
    ctx.set_option("profile", 0)
    nb_pairs, nb_max, nb_over = ctx.neighbour_stats() if sph else (0, 0, 0)
    plastic = plastic_fraction(ctx, d, prm) if solid else None

    ms_step = (t1 - t0) * 1e3 / args.steps
    pairs = float(n) * (n - 1)
    value = pairs / (ms_step * 1e-3)

    sym = kt["far"][1] > 0  # the symmetric sweep ran (its far field was timed)
    # all kernels of the all-pairs sweep, summed per evaluation (with the symmetric sweep, intervals that overlap)
    sweep_ms = (kt["sweep"][0] + kt["far"][0]) / args.steps
    sweep_flops = FLOP_GRAVITY_PAIR * pairs
    if sym:
        sizes = np.full(n // SYM_GROUP, SYM_GROUP, dtype=np.float64)
        if n % SYM_GROUP:
            sizes = np.append(sizes, n % SYM_GROUP)
        dom_pairs = pairs - float(np.sum(sizes * (sizes - 1)))
        dom_ms = kt["far"][0] / max(kt["far"][1], 1)
        dom_name = "far_kernel (symmetric sweep, far field)"
    else:
        dom_pairs = pairs
        dom_ms = sweep_ms
        dom_name = "sweep (one-sided)"
    dom_flops = FLOP_GRAVITY_PAIR * dom_pairs
    achieved = dom_flops / (dom_ms * 1e-3) / 1e12
    integ_ms = kt["integrate"][0] / max(kt["integrate"][1], 1)
    integ_bytes = n * (INTEGRATOR_BYTES_SOLID if solid else INTEGRATOR_BYTES_NBODY)
    step_flops = sweep_flops
    if sph:
        fl = prm.flags
        dens = 0 if fl & g.INTEGRATE_DENSITY else FLOP_DENSITY_NBR
        strain = 0
        if solid:
            strain = FLOP_STRAIN_NBR_SOLID
        elif fl & g.BALSARA_SWITCH:
            strain = FLOP_STRAIN_NBR_LIQUID
        elif fl & g.INTEGRATE_DENSITY:
            strain = FLOP_STRAIN_NBR_LIQUID - 15
        step_flops += nb_pairs * (dens + strain + (FLOP_ACCEL_NBR_SOLID if solid else FLOP_ACCEL_NBR_LIQUID))

    traffic = None
    # HBM bytes per launch of the dominant kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)
    tpath = os.path.join(ROOT, "profiles", "%s_%s_%d_traffic.json" % ("far_kernel" if sym else "sweep", args.workload, n))
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": "pair-interactions/s & %FP32 VALU peak at N=131k; batched-ensemble steps/s",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded initial conditions of the workload)",
        "config": {"workload": args.workload, "n": n, "flags": int(prm.flags), "H": prm.H,
                   "dt": prm.fixed_timestep, "fp64_drift": True, "advance_steps": args.advance,
                   "graph_steps": graph_steps},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel": dom_name, "kernel_ms": dom_ms, "flops_per_launch": dom_flops,
                     "pairs_per_launch": dom_pairs},
        "sweep": {"kind": "symmetric" if sym else "one-sided", "kernel_ms": sweep_ms,
                  "tflops_all_pairs": sweep_flops / (sweep_ms * 1e-3) / 1e12,
                  "frac_all_pairs": sweep_flops / (sweep_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in kt.items() if v[1]},
        "step_tflops": step_flops / (ms_step * 1e-3) / 1e12,
        "neighbour_pairs": nb_pairs,
        "max_neighbours_seen": nb_max,
        "list_overflows": nb_over,
        "plastic_fraction": plastic,
        "integrator": {"ms": integ_ms, "bytes_per_particle": INTEGRATOR_BYTES_SOLID if solid else INTEGRATOR_BYTES_NBODY,
                       "gbps": integ_bytes / (integ_ms * 1e-3) / 1e9 if integ_ms > 0 else None,
                       "hbm_frac": integ_bytes / (integ_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS if integ_ms > 0 else None},
    }

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(d, prm, n, args.cpu_seconds)
    d.free()
    if args.ensemble > 0:
        out["ensemble"] = ensemble_measure(ctx, args.ensemble, args.ensemble_steps, args.loop_steps)
    ctx.close()
    print(json.dumps(out), flush=True)


def plastic_fraction(ctx, d, prm):
    """Share of particles whose stress the next drift carries past yield
    (von Mises: 3 J2(S + dS/dt dt) > Y^2, so the plasticity step scales S),
    after one more (untimed) derivative evaluation."""
    import numpy as np

    import grasph2_amd as g

    if not (prm.flags & g.PLASTICITY_MIESE):
        return None
    ctx.compute_derivatives(d, prm)
    g.hip.synchronize()
    st = d.download(["dstress", "dstress_dt"])
    S = st["dstress"].astype(np.float64) + st["dstress_dt"].astype(np.float64) * float(prm.fixed_timestep)
    J2 = 0.5 * (S[0] ** 2 + S[1] ** 2 + S[2] ** 2) + S[3] ** 2 + S[4] ** 2 + S[5] ** 2
    return float(np.mean(3.0 * J2 > float(prm.Y) ** 2))


def cpu_baseline(d, prm, n, target_s):
    """Time the FP32 oracle's passes of computeDerivatives for a sample of i
    against all j: the ones the flags select (computeDerivatives.h:315-328):
    density and strain with SPH, acceleration always."""
    import numpy as np

    import grasph2_amd as g
    from oracle import pyoracle as O

    host = host_cpu()
    threads = host["threads_used"]
    O.set_threads(threads)
    state = d.download()
    rng = np.random.default_rng(0)
    f = int(prm.flags)
    sph = bool(f & g.SPH)
    dens = sph and not (f & g.INTEGRATE_DENSITY)
    bal = sph and bool(f & (g.SOLIDS | g.INTEGRATE_DENSITY | g.BALSARA_SWITCH))
    passes_run = (["calcDensity"] if dens else []) + (["calcBalsaraDensityDTDStressDT"] if bal else []) + \
        ["calcAcceleration"]

    def run(count):
        sample = np.sort(rng.choice(n, size=count, replace=False))
        t = time.perf_counter()
        if dens:
            O.calc_density(prm, state, "f32", sample)
        if bal:
            O.calc_balsara(prm, state, "f32", sample)
        O.calc_acceleration(prm, state, "f32", sample)
        return time.perf_counter() - t

    # two probes (the first one pays thread start-up and cold caches), then
    # a sample sized to target_s; when that exceeds all n targets, whole
    # passes over every target are repeated until target_s is reached
    probe = min(n, max(64, threads * 8))
    run(probe)
    probe = min(n, max(probe, threads * 256))
    tp = run(probe)
    count = int(max(probe, probe * target_s / max(tp, 1e-6)))
    if count <= n:
        t = run(count)
        passes, what = 1, "%d of %d target particles" % (count, n)
    else:
        t, passes = 0.0, 0
        while passes == 0 or t < target_s:
            t += run(n)
            passes += 1
        count = n
        what = "%d pass(es) over all %d target particles" % (passes, n)
    out = {"value": passes * count * (n - 1) / t, "unit": "pair-interactions/s", "cores": threads, "kind": "port",
           "sample": "FP32 oracle, %s for %s against all %d partners, %.1f s, %d OpenMP threads"
                     % (" + ".join(passes_run), what, n, t, threads)}
    out.update(host)
    return out


def host_cpu():
    """The host the CPU baseline runs on: logical CPUs (nproc), the CPUs this
    process may run on, the model name, and the OpenMP threads used: every
    CPU of the affinity set, capped by OMP_NUM_THREADS when the environment
    sets it (a shared GPU box gives each GPU's jobs that many)."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = affinity
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return {"nproc": nproc, "affinity_cpus": affinity, "cpu_model": model, "omp_num_threads_env": omp or None,
            "threads_used": threads}


if __name__ == "__main__":
    main()
