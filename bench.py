"""Throughput benchmark: PCISPH particle-updates/s on one MI355X.

Workload: the 1M-particle PCISPH throughput scene (0.005 m lattice, h =
0.01 m, Morton re-sort every step, 2x1x1 m tank sampled as one boundary
layer; fluids_amd.scenes.c3_pcisph_throughput). A step is one full
simulate() including every PCISPH correction; value = N * K / wall time of
K timed steps (200 by default, as the throughput run is quoted) after W
warm-up steps, with all state resident in HBM. The cost per step grows over
the run as the fluid boxes collapse and spread over the floor.

Roofline: algorithmic bytes per particle-update follow the unique-bytes model
of the reference dataflow (each array a reference kernel reads or writes
counted once per launch, neighbour gathers as cache hits, list indices int32
at the mean neighbour counts K and Kb over the timed steps). The
dominant kernel's
achieved bandwidth is its algorithmic bytes per launch over its mean launch
time measured with HIP events on the solver's stream inside the timed region.

cpu_baseline: the CPU restatement of the reference kernels (oracle/, fp32,
same kernel sequence, OpenMP), built with -march=native on the host that runs
the benchmark, on the same scene, timed over a few steps with every host CPU
this process may use up to the 16 a job is granted (nproc, the allowed CPUs
and the CPU model are recorded).

fast_math: the same scene and step count with SPH_FAST_MATH (approximate
reciprocal and square root in the PCISPH kernels, as the reference compiles
them with -cl-fast-relaxed-math). A side figure: the value is the IEEE build,
whose fields are bit-identical to the oracle's.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

from fluids_amd.profiling import read_traffic  # noqa: E402


def kernel_bytes(name, K, Kb, n_it):
    """Algorithmic bytes per particle of one launch of each kernel: the
    unique bytes of the reference kernels that the launch replaces, with
    list indices as int32 at the measured mean neighbour counts K and Kb and
    n_it PCISPH corrections per step."""
    table = {
        "hash": 24.0,
        "hash_moved": 24.0,
        "resort_place": 16.0,  # the reference's sort: 16 B/p over all its passes
        "radix_scatter": 16.0 / 4,  # the reference's sort counts 16 B/p over all passes
        "radix_hist": 0.0,
        "radix_rowscan": 0.0,
        "gather_ranges": 72.0 + 21.0,
        "lists_density": 68.0 + 8 * K + 8 * Kb,
        "normals": 40.0 + 4 * K,
        "forces_predict": 72.0 + 80.0 + 16.0 + 4 * K,
        "update_pressure": 40.0 + 4 * K + 4 * Kb + 4.0 / max(n_it, 1),
        "pressure_force": 48.0 + 4 * K + 4 * Kb + (80.0 * (n_it - 1) + 96.0) / max(n_it, 1),
    }
    return table.get(name)


def step_bytes(K, Kb, n_it, N, B, B_moving):
    """B_pci per particle-update, reference dataflow."""
    b = 429 + 168 * n_it + (16 + 8 * n_it) * K + (8 + 8 * n_it) * Kb
    return b + (108.0 * B + 132.0 * B_moving) / N


def cpu_description():
    """nproc, the CPUs this process may run on, and the host's CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "allowed_cpus": allowed, "model": model}


def native_oracle():
    """Builds the oracle with -march=native for this host (a temporary copy
    beside the shipped x86-64-v3 build) and points the oracle loader at it.
    Returns the build flags used, or None if the native build failed."""
    import subprocess
    import tempfile
    from oracle import build as ob
    out = os.path.join(tempfile.mkdtemp(prefix="oracle_native_"), "liboracle_native.so")
    flags = [f if not f.startswith("-march=") else "-march=native" for f in ob.FLAGS]
    r = subprocess.run(["g++", *flags, os.path.join(ROOT, "oracle", "sph_oracle.cpp"), "-o", out],
                       capture_output=True, text=True)
    if r.returncode != 0:
        return None
    os.environ["SPH_ORACLE_LIB"] = out
    return " ".join(flags)


def run_cpu_baseline(sc, steps, threads):
    import oracle
    oracle.set_threads(threads)
    o = oracle.Oracle(sc.settings)
    o.set_rect_limits(*sc.container)
    for v in sc.volumes:
        o.add_volume(v)
    for b in sc.bodies:
        o.add_boundary(b.samples, b.can_move, b.mass, b.R, b.com)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return sc.n_fluid * steps / dt, dt


def run_fast_math(n, steps, warmup, max_iter=None):
    """The same scene with SPH_FAST_MATH (the reference's -cl-fast-relaxed-math
    arithmetic: approximate reciprocal and square root in the PCISPH kernels),
    timed like the headline run; a reported side figure, not the value."""
    from fluids_amd import FAST_MATH, scenes
    sc = scenes.c3_pcisph_throughput(n)
    sc.settings["flags"] = FAST_MATH
    if max_iter is not None:
        sc.settings["pcisph_max_iterations"] = max_iter
    sim = scenes.build(sc)
    sim.step(warmup)
    sim.reset_kernel_times()
    sim.set_profiling(True)
    sim.step(3)
    sim.synchronize()
    sim.set_profiling(False)
    kt = sim.kernel_times()
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    st = sim.step_stats()
    return {"value": sc.n_fluid * steps / wall, "unit": "particle-updates/s", "steps": steps,
            "ms_per_step": 1e3 * wall / steps, "corrections": st["iterations"],
            "kernels_avg_us": {k: round(1e3 * ms / max(c, 1), 1) for k, (c, ms) in kt.items()},
            "note": "flags = SPH_FAST_MATH; lists bit-exact, fields within 5e-5 max-relative of the IEEE "
                    "build (pressure, pressure force and velocity 1.3e-5 at 1M, the rest below 1e-6; "
                    "tests/test_gpu_fast_math.py); the headline value is the IEEE build"}


def run_batched(steps, warmup, plain_steps, count=512, tune_cache=None):
    """Batched WCSPH: `count` seeded scenes (fluids_amd.scenes.c4_batched) in
    one packed batch, HIP-graph steps, autotuned launch shapes; against the
    same scenes as separate single-scene solvers stepped in turn (each on its
    own stream, so their launches may overlap). With `tune_cache`, launch
    shapes come from (or, when the file holds none for this batch, are tuned
    and appended to) that file, so separate runs use the same shapes."""
    from fluids_amd import scenes
    scs = scenes.c4_batched(count=count)
    total = sum(sc.n_fluid for sc in scs)
    t0 = time.perf_counter()
    b = scenes.build_batch(scs, autotune=True, cache=tune_cache, trials=6)
    t_tune = time.perf_counter() - t0
    b.step(warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    wall = time.perf_counter() - t0
    shapes = {name: b.launch(name) for name in ("batch_hash", "batch_gather_ranges", "batch_lists_density",
                                                "batch_normals", "batch_accel_integrate")}
    # mean fluid neighbour count over every particle of the batch after the
    # timed steps (no boundary particles: Kb = 0)
    K = sum(float(b.neighbour_counts(k).sum()) for k in range(len(scs))) / total
    b.close()
    sims = [scenes.build(sc) for sc in scs]
    for sm in sims:
        sm.step(1)
    for sm in sims:
        sm.synchronize()
    t1 = time.perf_counter()
    for _ in range(plain_steps):
        for sm in sims:
            sm.step(1)
    for sm in sims:
        sm.synchronize()
    wall_plain = time.perf_counter() - t1
    for sm in sims:
        sm.close()
    v = total * steps / wall
    vp = total * plain_steps / wall_plain
    b_wc = 473.0 + 16.0 * K  # B_wc per particle-update, Kb = 0
    roof = {"bound": "hbm", "alg_bytes_per_pu": b_wc, "mean_fluid_neighbours": K,
            "achieved": b_wc * v / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": b_wc * v / HBM_PEAK}
    return {"workload": f"batched_wcsph_{count}", "scenes": count, "particles": total, "steps": steps,
            "value": v, "unit": "particle-updates/s", "ms_per_step": 1e3 * wall / steps, "roofline": roof,
            "plain_value": vp, "plain_steps": plain_steps, "plain_ms_per_step": 1e3 * wall_plain / plain_steps,
            "speedup_vs_plain": v / vp, "launch_shapes": shapes, "build_and_tune_s": t_tune}


def run_adjoint(repeats=3):
    """WCSPH adjoint on the 16 384-particle inverse problem
    (fluids_amd.scenes.c5_adjoint: 200 steps, checkpoint every 20): wall time
    of the plain solver's 200 steps, of the adjoint's forward pass (same
    steps, plus tapes and checkpoints) and of its backward pass (recomputing
    each segment from its checkpoint, then the adjoint kernels), best of
    `repeats`."""
    import numpy as np
    from fluids_amd import scenes
    from fluids_amd.adjoint import WCSPHAdjoint
    case = scenes.c5_adjoint()
    sc = case.scene
    n = sc.n_fluid
    v4 = np.concatenate([case.v0_hidden, np.zeros((n, 1), np.float32)], 1)
    plain = []
    for _ in range(repeats):
        sim = scenes.build(sc)
        sim.set_velocities(v4)
        sim.set_half_velocities(v4)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.step(case.steps)
        sim.synchronize()
        plain.append(time.perf_counter() - t0)
        sim.close()
    adj = WCSPHAdjoint(sc.settings, sc.container, sc.volumes[0], v4, v4, checkpoint_every=case.checkpoint_every)
    fwd, bwd = [], []
    g = np.ones((n, 3), np.float32) / n
    for _ in range(repeats):
        adj.forward(case.steps)
        adj.backward(g)
        f, b = adj.timings()
        fwd.append(f / 1e3)
        bwd.append(b / 1e3)
    adj.close()
    tp, tf, tb = min(plain), min(fwd), min(bwd)
    return {"workload": "wcsph_adjoint_16k", "particles": n, "steps": case.steps,
            "checkpoint_every": case.checkpoint_every, "plain_forward_ms": 1e3 * tp,
            "adjoint_forward_ms": 1e3 * tf, "backward_ms": 1e3 * tb,
            "backward_over_plain_forward": tb / tp, "forward_plus_backward_over_plain_forward": (tf + tb) / tp}


def run_next_rows(steps=50, frame_steps=20):
    """Side figures of the code around the path (not the value): the rigid
    coupling loop on the configured double dam break (0.01 m lattice, the
    12 kg cube; sph_coupled_step reads the body wrench every step, so the
    plain step is also synchronised every step for the comparison); the 1M
    throughput scene with its positions written to a frame file after every
    step (after the writer's first frames, which allocate its pinned
    slots), against the same steps without; the reference's many_objects.json
    scene file loaded and built (samplers and the PCISPH relaxation step
    included), and the teapot mesh sampled at 0.01 m."""
    import os
    import tempfile
    from fluids_amd import load_scene, load_settings, model_sampling, scenes
    from fluids_amd.output import FrameWriter
    from fluids_amd.rigid import RigidWorld, coupled_step
    golden = os.path.join(ROOT, "tests", "golden")
    out = {}

    sc = scenes.c2_pcisph_double_dambreak(spacing=0.01)
    sim = scenes.build(sc)
    world = RigidWorld(-9.8)
    world.set_container(*sc.container)
    cube = sc.bodies[0]
    pairs = [(0, world.add_cube(0.3, cube.mass, cube.com, (0, 1, 0, 1)))]
    for _ in range(5):
        coupled_step(sim, world, pairs)
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        coupled_step(sim, world, pairs)
    sim.synchronize()
    t_coupled = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step(1)
        sim.synchronize()
    t_plain = (time.perf_counter() - t0) / steps
    out["coupled_double_dambreak"] = {"fluid_particles": sc.n_fluid, "boundary_particles": sc.n_boundary,
                                      "steps": steps, "coupled_ms_per_step": 1e3 * t_coupled,
                                      "plain_synchronised_ms_per_step": 1e3 * t_plain,
                                      "coupling_overhead": t_coupled / t_plain - 1.0}
    sim.close()
    world.close()

    sc = scenes.c3_pcisph_throughput(1_000_000)
    sim = scenes.build(sc)
    sim.step(5)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(frame_steps)
    sim.synchronize()
    t_steps = (time.perf_counter() - t0) / frame_steps
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "frames.bin")
        fw = FrameWriter(sim, path, positions=True)
        for _ in range(3):  # the writer's three pinned slots are allocated by its first frames
            sim.step(1)
            fw.write()
        sim.synchronize()
        time.sleep(0.2)
        t0 = time.perf_counter()
        for _ in range(frame_steps):
            sim.step(1)
            fw.write()
        frames = fw.close() - 3
        t_frames = (time.perf_counter() - t0) / frame_steps
        size = os.path.getsize(path)
    out["frames_1000k"] = {"steps": frame_steps, "frames": frames, "bytes_per_frame": (size - 16) / (frames + 3),
                           "step_ms": 1e3 * t_steps, "step_and_frame_ms": 1e3 * t_frames}
    sim.close()

    t0 = time.perf_counter()
    phys, simset = load_settings(os.path.join(golden, "config.pci"))
    ls = load_scene(os.path.join(golden, "many_objects.json"), phys, simset)
    ls.simulation.synchronize()
    t_load = time.perf_counter() - t0
    t0 = time.perf_counter()
    teapot = model_sampling(os.path.join(golden, "teapot.obj"), 0.01)
    t_model = time.perf_counter() - t0
    out["scene_front_end"] = {"many_objects_load_and_build_s": t_load,
                              "fluid_particles": ls.simulation.particle_count(),
                              "teapot_samples": int(teapot.shape[0]), "teapot_sampling_s": t_model}
    ls.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)  # the throughput run is quoted over 200 steps
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--max-iter", type=int, default=None, help="override pcisph_max_iterations")
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-events", action="store_true", help="time without per-kernel events")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--full-sort", action="store_true",
                    help="sort every key each step with the radix sort (SPH_FULL_SORT) instead of placing "
                         "only the keys that changed cell")
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-scenes measurement")
    ap.add_argument("--batched-only", action="store_true", help="only the batched-scenes measurement")
    ap.add_argument("--batched-steps", type=int, default=50)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--tune-cache", default=None,
                    help="batched launch shapes: read from this file, or tuned and appended to it")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timing")
    ap.add_argument("--no-fast-math", action="store_true", help="skip the SPH_FAST_MATH side run")
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint timing")
    ap.add_argument("--next-rows", action="store_true",
                    help="only the side figures of the rigid coupling, frame output and scene loading")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "pmc_pcisph_throughput_1000k.md"),
                    help="rocprofv3 FETCH_SIZE/WRITE_SIZE summary (python -m fluids_amd.profiling pmc) "
                         "that gives the dominant kernel's HBM bytes per launch")
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the solver has no multi-GPU path. "
              "A scene would shard into Morton-contiguous slabs with halo exchanges; the benchmark "
              "measures one GPU (--n 125000 runs one GPU's share of the 1M scene on 8).", file=sys.stderr)
        sys.exit(2)

    from fluids_amd import scenes
    from fluids_amd._lib import lib
    lib()
    if a.adjoint_only:
        print(json.dumps({"adjoint": run_adjoint()}), flush=True)
        return
    if a.next_rows:
        print(json.dumps({"next_rows": run_next_rows()}), flush=True)
        return
    if a.batched_only:
        print(json.dumps({"batched": run_batched(a.batched_steps, 5, a.plain_steps, tune_cache=a.tune_cache),
                          "adjoint": None if a.no_adjoint else run_adjoint()}), flush=True)
        return
    t_init = time.perf_counter()
    sc = scenes.c3_pcisph_throughput(a.n)
    if a.max_iter is not None:
        sc.settings["pcisph_max_iterations"] = a.max_iter
    if a.full_sort:
        sc.settings["flags"] = sc.settings.get("flags", 0) | 4  # SPH_FULL_SORT
    sim = scenes.build(sc)
    sim.synchronize()
    t_init = time.perf_counter() - t_init
    print(f"# scene {sc.name}: N={sc.n_fluid} B={sc.n_boundary} init {t_init:.1f}s", file=sys.stderr, flush=True)

    shapes = None
    if not a.no_autotune:
        # launch shapes of the list kernels, timed on this scene (restores the state)
        sim.autotune(None, trials=3)
        shapes = {k: sim.launch(k) for k in sim.TUNED_KERNELS}
        print(f"# tuned launch shapes: {shapes}", file=sys.stderr, flush=True)

    sim.step(a.warmup)
    sim.synchronize()
    st = sim.step_stats()
    print(f"# warmup done: {st}", file=sys.stderr, flush=True)

    # per-kernel breakdown over a few untimed steps, to find the dominant kernel
    sim.reset_kernel_times()
    sim.set_profiling(True)
    sim.step(max(3, a.warmup // 2))
    sim.synchronize()
    sim.set_profiling(False)
    breakdown = sim.kernel_times()
    dom = max(breakdown, key=lambda k: breakdown[k][1]) if breakdown else None

    # timed region: HIP events bracket only the dominant kernel's launches
    sim.reset_kernel_times()
    sim.set_profiling(not a.no_events and dom is not None, kernel=dom)
    # the state the timed steps start from, to replay them for their counts
    pos0, vel0 = sim.positions(), sim.velocities()
    sim.set_positions(pos0)
    sim.set_velocities(vel0)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(a.steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    sim.set_profiling(False)
    kt = sim.kernel_times()
    st = sim.step_stats()
    pos_end = sim.positions()

    # K, Kb and the corrections as the mean over the timed steps: the same
    # steps replayed (untimed) from the saved state give the same bits, with
    # the step statistics read after each
    N, B = sc.n_fluid, sc.n_boundary
    sim.set_positions(pos0)
    sim.set_velocities(vel0)
    fl = bl = its = 0
    for _ in range(a.steps):
        sim.step(1)
        s1 = sim.step_stats()
        fl += s1["fluid_neighbours"]
        bl += s1["boundary_neighbours"]
        its += s1["iterations"]
    replay_identical = bool((sim.positions() == pos_end).all())
    n_it = its / a.steps
    K = fl / (N * a.steps)
    Kb = bl / (N * a.steps)
    pu_s = N * a.steps / wall
    bpp = step_bytes(K, Kb, n_it, N, B, 0)

    # breakdown: share of each kernel (untimed profiling pass); dominant
    # kernel's launch time measured inside the timed region
    per = {}
    tot = sum(ms for _, ms in breakdown.values()) or 1.0
    for name, (launches, ms) in breakdown.items():
        kb = kernel_bytes(name, K, Kb, n_it)
        per[name] = dict(launches=launches, avg_us=1e3 * ms / max(launches, 1), share=ms / tot,
                         alg_bytes_per_launch=None if kb is None else kb * N)
    roof = None
    if dom is not None and dom in kt and kernel_bytes(dom, K, Kb, n_it):
        launches, ms = kt[dom]
        avg_s = 1e-3 * ms / max(launches, 1)
        alg = kernel_bytes(dom, K, Kb, n_it) * N
        achieved = alg / avg_s
        roof = {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": read_traffic(a.traffic_file, dom),
                "traffic_source": os.path.relpath(a.traffic_file, ROOT), "alg_bytes_per_launch": alg,
                "avg_launch_us": 1e6 * avg_s, "timed_launches": launches}

    cpu = None
    if not a.no_cpu:
        host = cpu_description()
        # a job on the GPU host is granted at most 16 host threads, whatever
        # nproc shows; every CPU this process may use, up to that share
        threads = max(1, min(host["allowed_cpus"], 16))
        flags = native_oracle()
        v, dt = run_cpu_baseline(sc, a.cpu_steps, threads)
        cpu = {"value": v, "unit": "particle-updates/s", "cores": threads, "kind": "port",
               "host": host, "build": flags or "shipped x86-64-v3 build (native build failed)",
               "sample": f"{a.cpu_steps} PCISPH steps of the same {N}-particle scene after its relaxation step "
                         f"({dt:.1f} s), fp32 oracle, OpenMP with {threads} threads"}

    del sim
    fast = None if a.no_fast_math else run_fast_math(a.n, a.steps, a.warmup, a.max_iter)

    batched = None
    if not a.no_batched:
        batched = run_batched(a.batched_steps, 5, a.plain_steps, tune_cache=a.tune_cache)
    adjoint = None if a.no_adjoint else run_adjoint()

    rec = {
        "metric": "particle-updates/s (PCISPH 1M fp32), %HBM roofline; batched & adjoint vs plain",
        "value": pu_s, "unit": "particle-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"pcisph_throughput_{N // 1000}k", "fluid_particles": N, "boundary_particles": B,
                   "spacing_m": 0.005, "support_radius_m": 0.01, "corrections": round(n_it, 3),
                   "mean_fluid_neighbours": round(K, 3), "mean_boundary_neighbours": round(Kb, 3),
                   "sort": "radix sort of every key" if a.full_sort else "moved keys placed into the previous order",
                   "counts": "means over the timed steps (replayed untimed from the same state; "
                             f"replay bit-identical: {replay_identical})"},
        "roofline": roof,
        "step_roofline": {"alg_bytes_per_pu": bpp, "achieved_GBps": bpp * pu_s / 1e9, "frac": bpp * pu_s / HBM_PEAK},
        "cpu_baseline": cpu,
        "fast_math": fast,
        "batched": batched,
        "adjoint": adjoint,
        "kernels": per,
        "launch_shapes": shapes,
        "events_in_timed_region": not a.no_events,
    }
    line = json.dumps(rec)
    print(line, flush=True)
    if a.json_out:
        with open(a.json_out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
