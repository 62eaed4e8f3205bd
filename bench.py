"""Benchmark of the batched step path on one MI355X.

Prints ONE JSON line.  A step is one pass of the hot path over every world of
the workload (collide -> assemble -> [warm-start inference] -> SOR-LCP sweeps
until the world's LCP residual is <= 1e-4 -> integrate); value = world-steps
per second of the whole job with inputs resident in HBM.

Workloads (seeded synthetic scenes, ode_amd/scenes.py):
  towers  4,096 worlds of a 32-box tower plus a projectile sphere; the learned
          warm start (trained first on the GPU on 2M samples of the same
          configuration with other seeds; gated: contacts matched to a
          previous-step contact keep its impulses, the network predicts the
          new ones) vs cold PGS at the same residual.
          value = learned warm-start world-steps/s.
  uneven  4,096 worlds of 8-64 spheres/boxes dropped on a plane, 20 sweeps
          per step (--tol 0 --max-iters 20 unless given), 500 steps, cold PGS.
  chains  1,024 worlds of a 16-capsule chain on ball joints (anchored at one
          end) with 8 free boxes, 40 sweeps per step (--tol 0 --max-iters 40
          unless given), cold PGS.
  large   one world of 200,000 spheres poured into an open box (hash space,
          30 sweeps per step, each contact's rows together in one pass over
          the colours: --row-order 2); value = world-steps/s.
  boxes   one world of 1,000 boxes dropped on a plane, stepped through the
          drop-in ODE API by a C++ frame loop, 300 steps after the warmup;
          value = world-steps/s.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload towers]
"""
import argparse
import json
import os
import sys
import time
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "world-steps/sec at LCP residual ≤1e-4: warm-start vs cold PGS, 1 MI355X"
HBM_PEAK_GBS = 8000.0
# LDS bandwidth of the chip: 256 CUs x 128 B per clock x 2.4 GHz
LDS_PEAK_GBS = 256 * 128 * 2.4
BF16_PEAK_TFLOPS = 2500.0
ROW_SWEEP_BYTES = 228  # per row per PGS sweep: row 124 B, lambda r/w 8 B, fc gather+scatter 96 B
MLP_FLOP_PER_CONTACT = 2 * (16 * 128 + 128 * 128 + 128 * 3)
# HBM bytes k_solve_small must move per launch by its design (rows are read
# once and swept in LDS): per row the 128 B row record, its schedule entries
# (item, perm, inv: 12 B) and lambda in and out (8 B); per body its seven
# float4 streams and 128 B of body scratch read, pos / quat / lvel / avel
# written; per contact the 48 B previous-contact record written
SOLVE_ROW_BYTES = 128 + 12 + 8
SOLVE_BODY_BYTES = 7 * 16 + 128 + 4 * 16
SOLVE_CONTACT_BYTES = 48


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "available": avail, "cpu_model": model}


def build_worlds(workload, nworlds, seed):
    from ode_amd import scenes as S
    if workload == "uneven":
        return S.uneven_batch(nworlds, seed=seed)
    if workload == "towers":
        return S.tower_batch(nworlds, seed=seed)
    if workload == "chains":
        return S.chain_batch(nworlds, seed=seed)
    raise SystemExit("unknown workload %r" % workload)


def _cpu_worker(task):
    """One host process of the all-core CPU baseline: the oracle's float32 step
    over its share of the worlds until the time budget runs out; returns
    (world-steps done, seconds)."""
    geoms, balls, bodies, prm, steps, budget = task
    import oracle as O
    t0 = time.perf_counter()
    done = 0
    for g, bl, b in zip(geoms, balls, bodies):
        b = b.astype(np.float32)
        for _ in range(steps):
            b, _, _ = O.world_step(b, g, prm, balls=bl, dtype=np.float32)
        done += steps
        if time.perf_counter() - t0 > budget:
            break
    return done, time.perf_counter() - t0


def cpu_baseline(worlds, bodies_now, body_start, prm, budget_s=12.0, steps_per_world=5):
    """The oracle's sequential float32 step (cold start, same stopping rule) on
    a sample of the same worlds from the GPU's state after warmup: one core,
    then one world at a time per host process on the host's cores (at most 16,
    the GPU's share of the host; processes, so that no interpreter lock
    serialises the worlds' host-side work)."""
    import concurrent.futures as cf
    import multiprocessing as mp
    import oracle as O
    info = cpu_info()

    def run(wi):
        w = worlds[wi]
        b = bodies_now[body_start[wi]:body_start[wi + 1]].astype(np.float32)
        for _ in range(steps_per_world):
            b, _, _ = O.world_step(b, w.geoms, prm, balls=w.balls, dtype=np.float32)
        return steps_per_world

    # one core
    done, nw = 0, 0
    t0 = time.perf_counter()
    for wi in range(len(worlds)):
        done += run(wi)
        nw += 1
        if time.perf_counter() - t0 > budget_s:
            break
    one = done / (time.perf_counter() - t0)
    # all cores: worlds dealt round-robin to the processes, each stepping its
    # share for the same budget; the rate is the sum of the processes' rates
    cores = max(1, min(16, info["available"]))
    per = max(1, min(len(worlds), 4 * nw) // cores)
    tasks = []
    for c in range(cores):
        idx = list(range(c, len(worlds), cores))[:per]
        tasks.append(([worlds[i].geoms for i in idx], [worlds[i].balls for i in idx],
                      [bodies_now[body_start[i]:body_start[i + 1]] for i in idx], prm, steps_per_world, budget_s))
    # fresh interpreters (spawn): the children never touch the GPU
    with cf.ProcessPoolExecutor(max_workers=cores, mp_context=mp.get_context("spawn")) as ex:
        res = list(ex.map(_cpu_worker, tasks))
    rate = sum(d / max(t, 1e-9) for d, t in res)
    nw2 = sum(d for d, _ in res) // steps_per_world
    dt = max(t for _, t in res)
    return {"value": rate, "unit": "world-steps/s", "cores": cores, "kind": "port",
            "single_core": one, "nproc": info["nproc"], "cpu_model": info["cpu_model"],
            "sample": "the oracle's float32 step (cold start, reference row order, same residual stopping rule), "
                      "%d steps per world from the GPU state after warmup: %d worlds on one core, then %d worlds "
                      "on %d processes, one world at a time each (%.1f s)" % (steps_per_world, nw, nw2, cores, dt)}


def fp64_residual(B, worlds, prm64, sample=64):
    """The LCP residual of the GPU's impulses re-evaluated in float64 by the
    oracle (rows rebuilt in double from the step-start state) for a sample of
    worlds over one more step, beside the kernel's own float32 residual."""
    import oracle as O
    idx = np.unique(np.linspace(0, len(worlds) - 1, sample).astype(int))
    before = B.bodies()
    B.step(1)
    st = B.stats()
    res = []
    for wi in idx:
        bw = before[B.body_start[wi]:B.body_start[wi + 1]].astype(np.float64)
        j = B.joints(int(wi)).astype(np.float64)
        lam = B.lam(int(wi)).astype(np.float64)
        if len(j) == 0:
            continue
        res.append(O.lcp_residual(bw, j, prm64, lam))
    res = np.array(res)
    return {"worlds": int(len(res)), "max": float(res.max()), "median": float(np.median(res)),
            "gpu_f32_max": float(np.max(st["residual"][idx])),
            "definition": "max_i |clamp(l_i - w_i/A_ii) - l_i| / max(max|l|, 1e-3), A = J M^-1 J^T + cfm/h, "
                          "w = A l - b, rebuilt in float64 from the step-start state"}


def measure(B, args, warm, nworlds):
    """Warmup, timed graph replays, then per-kernel hipEvent durations.  The
    LDS row budget is sized during warmup to the largest world seen (rounded
    up to 16 rows, within the 64 KB limit) so that every world sweeps from LDS."""
    B.set_solver(order=args.order, max_iters=args.max_iters, tol=args.tol, check_every=args.check_every, warm=warm)
    B.set_solve_options(restart_after=args.restart_after, lds_rows=args.lds_rows if args.lds_rows >= 0 else 400)
    half = args.warmup // 2
    B.step(half)
    max_rows = int(B.stats()["rows"].max())
    for _ in range(half, args.warmup):
        B.step(1)
        max_rows = max(max_rows, int(B.stats()["rows"].max()))
    lds_rows = args.lds_rows if args.lds_rows >= 0 else 400
    if args.lds_rows < 0:
        # the largest budget that still fits five worlds in a CU's 160 KB of
        # LDS, or 1.25x the largest world seen if that is more: rows keep
        # growing after the warmup (towers settle into more contacts), so the
        # budget is not sized to the warmup's largest world alone
        nb_max = int(np.max(np.diff(B.body_start)))
        def lds_bytes(r):  # solve_lds_floats * 4
            return 4 * ((nb_max + 1) * 16 + r * 17 + r * 3 // 4 + 16 + 65 + 64)
        five = max(r for r in range(16, 2049) if 5 * lds_bytes(r) <= 160 * 1024)
        want = max(five, -(-int(max_rows * 1.25) // 16) * 16)
        while want > max_rows:
            try:
                B.set_solve_options(restart_after=args.restart_after, lds_rows=want)
                lds_rows = want
                break
            except Exception:
                want -= 16
    B.sync()
    schedule = {"splits": args.splits, "graph_steps": args.graph_steps, "fused": 0, "autotuned": False}
    if not args.no_autotune:
        # the launch schedule is chosen by timing the candidates on these worlds
        sch, _ = B.autotune(steps=args.tune_steps, reps=2)
        schedule = dict(sch, autotuned=True)
    bodies_after_warmup = B.bodies()
    total_ms = B.time(args.steps)
    st = B.stats()
    # kernel durations per launch on the schedule the timed run uses (each
    # world range on its own stream), with the per-launch algorithmic bytes
    nb_w = np.diff(B.body_start)
    launches = []
    sweeps, contacts, inferred = [], [], []
    for _ in range(args.kernel_steps):
        ks = B.schedule_kernel_times(1)
        s = B.stats()
        rows = s["rows"].astype(np.float64)
        hbm = SOLVE_ROW_BYTES * rows + SOLVE_BODY_BYTES * nb_w + SOLVE_CONTACT_BYTES * s["contacts"]
        lds = ROW_SWEEP_BYTES * rows * s["iters"]
        for w0, w1, ms in ks:
            launches.append((ms[0], ms[1], ms[2], float(np.sum(hbm[w0:w1])), float(np.sum(lds[w0:w1])), w1 - w0))
        sweeps.append(float(np.mean(s["iters"])))
        contacts.append(float(np.sum(s["contacts"])))
        if warm == 3:
            inferred.append(len(B.samples()[0]))
    L = np.array(launches)
    solve_ms = float(np.mean(L[:, 2]))
    ms = total_ms / args.steps
    out = {
        "world_steps_per_s": nworlds / (ms * 1e-3),
        "ms_per_step": ms,
        "mean_sweeps": float(np.mean(st["iters"])),
        "max_sweeps": int(np.max(st["iters"])),
        "max_residual": float(np.max(st["residual"])),
        "worlds_at_target": float(np.mean(st["residual"] <= args.tol)),
        "contacts_per_step": float(np.sum(st["contacts"])),
        "kernel_ms_per_launch": {"prepare": float(np.mean(L[:, 0])), "infer": float(np.mean(L[:, 1])),
                                 "solve": solve_ms},
        "launches_per_step": len(launches) // args.kernel_steps,
        "solve_hbm_bytes_per_launch": float(np.mean(L[:, 3])),
        "solve_hbm_gbs": float(np.mean(L[:, 3])) / (solve_ms * 1e-3) / 1e9,
        "solve_lds_gbs": float(np.mean(L[:, 4])) / (solve_ms * 1e-3) / 1e9,
        "kernel_window_contacts": float(np.mean(contacts)),
        "kernel_window_sweeps": float(np.mean(sweeps)),
        "overflow": [int(x) for x in B.overflow()],
        "lds_rows": lds_rows,
        "schedule": schedule,
        "worlds_over_lds_budget": int(np.sum(st["rows"] > lds_rows)),
        "_bodies": bodies_after_warmup,
    }
    if warm == 3 and inferred:
        ki = float(np.mean(L[:, 1]))
        per_launch = float(np.mean(inferred)) / max(1, out["launches_per_step"])
        out["inferred_contacts_per_step"] = float(np.mean(inferred))
        if ki > 0:
            out["mlp_infer_tflops"] = per_launch * MLP_FLOP_PER_CONTACT / (ki * 1e-3) / 1e12
    return out


def bench_large(args):
    """One world of 200,000 spheres (radius 0.05-0.15 m) poured from a lattice
    into a 10x10 m open box of 5 static planes; hash-space broadphase, mu=0.3,
    30 sweeps per step.  The pour runs `pour` + `warmup` steps (the lattice's
    top layer lands after about 340 steps; the pile has settled by about 800)
    before the timed `steps`."""
    from ode_amd import odx
    from ode_amd import scenes as S
    import oracle as O
    bodies = args.bodies
    world = S.sphere_pour(bodies, seed=args.seed)
    prm = odx.params_vector(iters=args.large_iters, w=args.sor_w, mu=0.3, space=1, row_order=args.row_order)
    L = odx.Large(world, prm)
    L.set_solve_blocks(args.solve_blocks)
    t0 = time.perf_counter()
    done = 0
    total = args.pour + args.warmup
    while done < total:
        n = min(50, total - done)
        L.step(n)
        L.sync()
        done += n
        print("bench: pour %d/%d steps, %d contacts" % (done, total, L.counts()["contacts"]),
              file=sys.stderr, flush=True)
    pour_s = time.perf_counter() - t0
    total_ms = L.time(args.steps)
    cnt = L.counts()
    ks = [L.kernel_times(1) for _ in range(args.kernel_steps)]
    c2 = L.counts()
    km = np.mean(np.array(ks), axis=0)
    rows = c2["rows"]
    colour_sizes = np.bincount(L.colours(), minlength=c2["colours"]).tolist()
    state_crc = zlib.crc32(np.ascontiguousarray(L.bodies()).tobytes())   # the state after every step run
    solve_gbs = rows * args.large_iters * ROW_SWEEP_BYTES / (km[2] * 1e-3) / 1e9
    ms = total_ms / args.steps
    out = {
        "metric": "world-steps/sec, one large world, %d sweeps per step, 1 MI355X" % args.large_iters,
        "value": 1000.0 / ms,
        "unit": "world-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "large", "bodies": bodies, "geoms": bodies + 5, "space": "hash",
                   "sweeps": args.large_iters, "mu": 0.3, "erp": 0.2, "cfm": 1e-5, "h": 0.01, "sor_w": args.sor_w,
                   "seed": args.seed, "pour_steps": args.pour + args.warmup, "row_order": args.row_order},
        "contacts_per_step": cnt["contacts"],
        "contacts_per_s": cnt["contacts"] / (ms * 1e-3),
        "counts": cnt,
        "kernel_ms": {"prepare": float(km[0]), "colour": float(km[1]), "solve": float(km[2])},
        "colour_sizes": colour_sizes,
        "state_crc32": state_crc,
        "pour_s": pour_s,
        "roofline": {
            "bound": "hbm", "achieved": solve_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": solve_gbs / HBM_PEAK_GBS, "traffic": None,
            "kernel": "k_lg_solve (%d rows x %d sweeps x 228 B per launch)" % (rows, args.large_iters),
        },
    }
    tpath = os.path.join(ROOT, "profiles", "k_lg_solve_large%d.json" % bodies)
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        if t.get("row_order", 0) == args.row_order:
            out["roofline"]["traffic"] = t["traffic_bytes_per_launch"]
            out["roofline"]["traffic_source"] = os.path.relpath(tpath, ROOT)
        else:
            out["roofline"]["traffic_note"] = ("%s holds row order %d's counters" %
                                               (os.path.relpath(tpath, ROOT), t.get("row_order", 0)))
    if not args.no_cpu:
        # the oracle's QuickStep (rows, 30 sweeps, integration) over the GPU's
        # contacts of the current state; broadphase and narrowphase are left out
        # (the oracle's hash pairs are all-pairs), so the CPU rate is optimistic
        b = L.bodies().astype(np.float64)
        j = L.joints().astype(np.float64)
        p64 = O.params(space=1, mu=0.3, iters=args.large_iters, w=args.sor_w, row_order=args.row_order)
        t0 = time.perf_counter()
        O.quickstep(b, j, p64, dtype=np.float32)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "world-steps/s", "cores": 1, "kind": "port",
                               "sample": "1 step of the whole world: the oracle's float32 QuickStep over the "
                                         "GPU's %d contacts (rows, %d sweeps, integration; no broadphase or "
                                         "narrowphase), %.1f s" % (len(j), args.large_iters, dt)}
    L.close()
    print(json.dumps(out), flush=True)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration << 20)


def bench_boxes(args):
    """One world of 1,000 boxes (side 0.5 m, density 1000) dropped from a
    20x20x20 m volume onto a plane, stepped through the drop-in ODE API by a
    C++ frame loop (ode_amd/examples/frame_loop: sweep-and-prune space, near
    callback with dCollide and dJointCreateContact, dWorldQuickStep with 20
    iterations, dJointGroupEmpty).  The timed region includes the host side
    of the API (island order, callbacks) and its PCIe transfers."""
    import subprocess
    import tempfile
    from ode_amd import scenes as S
    import oracle as O
    world = S.boxes_on_plane(args.bodies, seed=args.seed)
    exe = os.path.join(ROOT, "ode_amd", "examples", "frame_loop")
    if not os.path.exists(exe):
        raise SystemExit("ode_amd/examples/frame_loop is not built (python -m ode_amd.build)")
    with tempfile.NamedTemporaryFile(suffix=".scene", delete=False) as f:
        np.array([len(world.bodies), len(world.geoms)], np.int32).tofile(f)
        np.ascontiguousarray(world.bodies, np.float32).tofile(f)
        np.ascontiguousarray(world.geoms, np.float32).tofile(f)
        path = f.name
    try:
        res = subprocess.run([exe, path, str(args.warmup), str(args.steps), str(args.large_iters), "0.5",
                              str(args.row_order), "0", str(args.sweep_mode)],
                             capture_output=True, text=True, timeout=900)
    finally:
        os.unlink(path)
    if res.returncode != 0:
        raise SystemExit("frame_loop failed (%d): %s" % (res.returncode, res.stderr[-2000:]))
    r = json.loads(res.stdout.strip().splitlines()[-1])
    ms = r["ms_per_step"]
    kms = r["kernel_ms_per_step"]
    gbs = r["row_sweeps_per_step"] * ROW_SWEEP_BYTES / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    out = {
        "metric": "world-steps/sec, one %d-box world through the ODE API, %d iterations, 1 MI355X"
                  % (args.bodies, args.large_iters),
        "value": 1000.0 / ms,
        "unit": "world-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "boxes", "bodies": args.bodies, "space": "sweep-and-prune (XYZ)",
                   "iterations": args.large_iters, "mu": 0.5, "erp": 0.2, "cfm": 1e-5, "h": 0.01, "sor_w": 1.3,
                   "row_order": "reference (dependency levels)" if args.row_order == 0 else "colour phases",
                   "dynamic_iterations": "off (fixed count)",
                   "seed": args.seed},
        "frame": r,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_single_step (row-sweeps x 228 B per launch, hipEvents inside dWorldQuickStep)"},
    }
    if not args.no_cpu:
        # the oracle's float32 canonical step (SAP pairs, dCollide, QuickStep in
        # island order with 20 iterations) over the same warmup + timed steps
        prm = O.params(iters=args.large_iters, space=0)
        b = np.ascontiguousarray(world.bodies, np.float32)
        for _ in range(args.warmup):
            b, _, _ = O.world_step(b, world.geoms, prm, dtype=np.float32, ccap=16 * args.bodies)
        t0 = time.perf_counter()
        n = 0
        while n < args.steps and time.perf_counter() - t0 < 30.0:
            b, _, _ = O.world_step(b, world.geoms, prm, dtype=np.float32, ccap=16 * args.bodies)
            n += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / dt, "unit": "world-steps/s", "cores": 1, "kind": "port",
                               "sample": "steps %d..%d of the same scene: the oracle's float32 step (SAP pairs, "
                                         "dCollide, QuickStep, %d iterations), %.1f s" % (args.warmup, args.warmup + n,
                                                                                         args.large_iters, dt)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="towers")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--max-iters", type=int, default=5000)
    ap.add_argument("--check-every", type=int, default=5)
    ap.add_argument("--order", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-steps", type=int, default=10)
    ap.add_argument("--train-samples", type=int, default=2_000_000)
    ap.add_argument("--train-steps", type=int, default=6000)
    ap.add_argument("--modes", default="cold,plain,learned")
    ap.add_argument("--gate", type=int, default=1,
                    help="learned mode: 1 keeps matched contacts at their previous impulses and runs the "
                         "network on new contacts only (odxMlpSetGate); 0 adds its predicted change everywhere")
    ap.add_argument("--restart-after", type=int, default=600,
                    help="sweeps after which an unconverged warm start restarts from lambda = 0")
    ap.add_argument("--lds-rows", type=int, default=-1, help="LDS row budget per world (-1: sized during warmup)")
    ap.add_argument("--sor-w", type=float, default=None,
                    help="SOR relaxation (default: ODE's 1.3 for the fixed-sweep workloads; 1.0 for towers, where "
                         "1.3 sends some warm-started solves into limit cycles and cold convergence is the same)")
    ap.add_argument("--splits", type=int, default=4,
                    help="world ranges stepped on parallel streams (with --no-autotune)")
    ap.add_argument("--tune-steps", type=int, default=40, help="steps per autotuner trial")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep --splits and staged launches instead of timing the launch schedules after warmup")
    ap.add_argument("--graph-steps", type=int, default=10, help="steps per graph launch")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="finally copy MIB MiB device to device (a reference kernel of known traffic for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--row-order", type=int, default=None,
                    help="boxes workload: 0 the reference's row order (default), 1 colour phases; large "
                         "workload: 0 ReorderPrep's order over the colours (two passes per sweep), 2 each "
                         "contact's rows together (one pass per sweep, the default)")
    ap.add_argument("--sweep-mode", type=int, default=0,
                    help="boxes workload: 0 automatic, 1 workgroup-wide levels, 2 one wave (odxWorldSetSweepMode)")
    ap.add_argument("--bodies", type=int, default=None, help="bodies of the large (200,000) or boxes (1,000) workload")
    ap.add_argument("--solve-blocks", type=int, default=1,
                    help="large workload: workgroups per CU of the cooperative solve (odxLargeSetSolveBlocks)")
    ap.add_argument("--pour", type=int, default=800,
                    help="large workload: steps poured (untimed) before the warmup and the timed steps")
    ap.add_argument("--large-iters", type=int, default=None,
                    help="sweeps per step of the large (30) or boxes (20) workload")
    args = ap.parse_args()

    if args.row_order is None:
        args.row_order = 2 if args.workload == "large" else 0
    if args.sor_w is None:
        args.sor_w = 1.0 if args.workload == "towers" else 1.3
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (worlds shard by index "
              "across GPUs with no exchange)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload in ("large", "boxes"):
        large = args.workload == "large"
        if args.bodies is None:
            args.bodies = 200_000 if large else 1000
        if not large and "--steps" not in sys.argv:
            args.steps = 300  # the 1,000-box world's 300 steps (after the warmup)
        if args.large_iters is None:
            args.large_iters = 30 if large else 20
    if args.workload == "large":
        return bench_large(args)
    if args.workload == "boxes":
        return bench_boxes(args)
    if args.workload in ("chains", "uneven"):
        # fixed sweeps per step: 40 (chains), 20 (uneven, the solver of the
        # 1,000-box world); uneven runs its 500 steps
        chains = args.workload == "chains"
        if chains and "--worlds" not in sys.argv:
            args.worlds = 1024
        if "--tol" not in sys.argv:
            args.tol = 0.0
        if "--max-iters" not in sys.argv:
            args.max_iters = 40 if chains else 20
        if not chains and "--steps" not in sys.argv:
            args.steps = 500

    from ode_amd import odx
    import oracle as O

    prm = odx.params_vector(iters=args.max_iters, w=args.sor_w)
    modes = args.modes.split(",") if args.workload == "towers" else ["cold"]
    train_info = None
    mlp = None
    if "learned" in modes:
        from ode_amd import warmstart as WS
        t0 = time.perf_counter()
        train_worlds = build_worlds(args.workload, args.worlds, args.seed + 1000)
        Bt = odx.Batch(train_worlds, prm)
        per_step = max(1, args.train_samples // 25)
        xs, ts = [], []
        Bt.set_mlp(None)
        Bt.set_solver(order=args.order, max_iters=args.max_iters, tol=1e-5, check_every=args.check_every, warm=3)
        rng = np.random.default_rng(0)
        for s in range(100):
            Bt.step(1)
            if s % 4 == 3:
                X, T = Bt.samples()
                keep = rng.permutation(len(X))[:per_step]
                xs.append(X[keep])
                ts.append(T[keep])
        Bt.close()
        X, T = np.concatenate(xs), np.concatenate(ts)
        t1 = time.perf_counter()
        mlp, losses = WS.train(X, T, steps=args.train_steps, batch=4096)
        mlp.set_gate(args.gate)
        t2 = time.perf_counter()
        train_info = {"samples": int(len(X)), "collect_s": t1 - t0, "train_s": t2 - t1,
                      "train_steps": args.train_steps, "gate": args.gate, "loss_first": float(losses[:20].mean()),
                      "loss_last": float(losses[-20:].mean())}

    worlds = build_worlds(args.workload, args.worlds, args.seed)
    results = {}
    bodies0 = None
    for mode in modes:
        B = odx.Batch(worlds, prm)
        B.set_splits(args.splits, args.graph_steps)
        if mode == "learned":
            B.set_mlp(mlp)
        r = measure(B, args, {"cold": 0, "plain": 2, "learned": 3}[mode], args.worlds)
        if mode == "cold":
            bodies0 = r["_bodies"]
            body_start = B.body_start
        r.pop("_bodies")
        r["fp64_residual"] = fp64_residual(B, worlds, O.params(iters=args.max_iters, w=args.sor_w))
        results[mode] = r
        B.close()

    head = "learned" if "learned" in results else modes[0]
    R = results[head]
    metric = METRIC if args.tol > 0 else "world-steps/sec, %d sweeps per step, 1 MI355X" % args.max_iters
    solver = ("%s start, coloured contact blocks, SOR w=%g, sweeps until LCP residual <= %g (checked every %d sweeps, "
              "cap %d, warm starts restart cold after %d)" % (head, args.sor_w, args.tol, args.check_every, args.max_iters,
                                                             args.restart_after)
              if args.tol > 0 else
              "%s start, coloured contact blocks, SOR w=%g, %d sweeps per step" % (head, args.sor_w, args.max_iters))
    out = {
        "metric": metric,
        "value": R["world_steps_per_s"],
        "unit": "world-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": R["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "worlds": args.worlds,
            "solver": solver,
            "launch": ("schedule autotuned per mode after warmup (modes[*].schedule)" if not args.no_autotune else
                       "%d world ranges on parallel streams, %d steps per hipGraph launch" % (args.splits, args.graph_steps)),
            "h": 0.01, "mu": 0.5, "erp": 0.2, "cfm": 1e-5, "seed": args.seed,
        },
        "modes": results,
    }
    # roofline of the dominant kernel of the plain GPU path (cold PGS: the
    # solve), per launch on the schedule the timed run used
    RR = results["cold"] if "cold" in results else R
    out["roofline"] = {
        "bound": "hbm",
        "achieved": RR["solve_hbm_gbs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": RR["solve_hbm_gbs"] / HBM_PEAK_GBS,
        "traffic": None,
        "kernel": "k_solve_small, cold mode, %d launches per step on the run's schedule, %.3f ms and %.1f MB "
                  "algorithmic per launch" % (RR["launches_per_step"], RR["kernel_ms_per_launch"]["solve"],
                                              RR["solve_hbm_bytes_per_launch"] / 1e6),
        "algorithmic": "per launch: 148 B per row (row record, schedule entries, lambda in and out), 304 B per body, "
                       "48 B per contact; the rows are swept in LDS, not re-read from HBM",
        "lds": {"achieved": RR["solve_lds_gbs"], "peak": LDS_PEAK_GBS, "unit": "GB/s",
                "frac": RR["solve_lds_gbs"] / LDS_PEAK_GBS,
                "model": "228 B per row per sweep (row 124 B, lambda 8 B, two bodies' fc 96 B) through LDS"},
        "occupancy": "one wave per world; LDS budget %d rows per world" % RR["lds_rows"],
    }
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE over
    # the k_solve_small dispatches of the same schedule, scaled by a copy of
    # known size)
    tpath = os.path.join(ROOT, "profiles", "k_solve_small_%s4096_cold.json" % args.workload)
    if args.worlds == 4096 and os.path.exists(tpath):
        t = json.load(open(tpath))
        tl = t.get("launches_per_step")
        if tl == RR["launches_per_step"]:
            out["roofline"]["traffic"] = t["traffic_bytes_per_launch"]
            out["roofline"]["traffic_source"] = os.path.relpath(tpath, ROOT)
        elif tl:
            # the profile's launches each cover 4096 / tl worlds: the bytes
            # scale with the worlds a launch covers
            out["roofline"]["traffic"] = t["traffic_bytes_per_launch"] * tl / RR["launches_per_step"]
            out["roofline"]["traffic_source"] = ("%s (measured at %d launches per step, scaled by worlds per launch)"
                                                 % (os.path.relpath(tpath, ROOT), tl))
    out["fp64_residual"] = R.get("fp64_residual")
    if "cold" in results and head != "cold":
        out["speedup_vs_cold"] = R["world_steps_per_s"] / results["cold"]["world_steps_per_s"]
    if "plain" in results and head != "plain":
        out["speedup_vs_plain_warm"] = R["world_steps_per_s"] / results["plain"]["world_steps_per_s"]
    if train_info:
        out["training"] = train_info
        if "mlp_infer_tflops" in R:
            out["mlp_infer_tflops"] = R["mlp_infer_tflops"]
            out["mlp_infer_frac_bf16_peak"] = R["mlp_infer_tflops"] / BF16_PEAK_TFLOPS
            out["mlp_inferred_contacts_per_step"] = R["inferred_contacts_per_step"]
    if not args.no_cpu and bodies0 is not None:
        out["cpu_baseline"] = cpu_baseline(worlds, bodies0, body_start,
                                           O.params(iters=args.max_iters, tol=args.tol, check_every=args.check_every, w=args.sor_w))
    print(json.dumps(out), flush=True)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration << 20)


def copy_calibration(nbytes):
    """One device-to-device hipMemcpy of nbytes (the runtime's copy kernel
    reads and writes nbytes), straight through the HIP runtime."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)):
        raise SystemExit("copy calibration: hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


if __name__ == "__main__":
    main()
