#!/usr/bin/env python3
"""Benchmark: DEM direct shear, 1e5 particles, FP64, on one MI355X.

Prints one JSON line.  A step is one iteration of the fluid-free loop body
of DEM::startTime (hash, sort, cells, contact search + interaction,
integration, wall reduction and wall integration) over all particles; the
metric is particle-steps/s over the timed steps, with the state resident in
HBM before timing starts.

Workload: 100,000 spheres, lognormal radii (median 0.01 m, sigma 0.2),
density 2600 kg/m^3, in a 1.6 x 1.6 m cell periodic in x and y (about
0.35 m deep), prepacked by seeded gravity settling
(sphere_amd.scenes.settled_shear_cell), consolidated
under a 10 kPa stress-controlled top wall, then sheared with the top layer at
1e-3 m/s; linear visco-elastic contacts (contactmodel 2) with Coulomb
mu = 0.5, k_n = k_t = 1.16e7 N/m, gamma = 0.5 critical, dt = 0.075
sqrt(m_min/k_n).

The same line also carries the metric's other two parts, measured in the
same run: "batched", the batched sweep (256 consolidation beds of
log-uniform size in [500, 20000], radii uniform in [0.005, 0.01] m, each
with its own normal stress, stiffness, friction and dt) as one batched
context against the plain loop over per-simulation contexts; and
"adjoint", the column-collapse calibration gradient (2,000 spheres, 500
steps sampled every 50) as forward + backward against the forward.
"""

import argparse
import hashlib
import json
import os
import platform
import struct
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_model():
    """The CPU's full model name (/proc/cpuinfo, else lscpu)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_threads():
    """The threads this process may use: OMP_NUM_THREADS when set (the
    machine's share for this job), else the CPUs in its affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.strip().isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def kernel_source_hash():
    h = hashlib.sha256()
    for f in ("dem_kernels.hip", "dem_math.hpp", "dem_types.hpp", "dem_arrays.hpp"):
        with open(os.path.join(ROOT, "sphere_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def measured_traffic(z, nw):
    """(HBM bytes per launch of the contact kernel, the command that measured
    them) from the committed PMC summary, when it was measured on the
    current kernel sources; else None.  Not measured in this run."""
    fn = os.path.join(ROOT, "profiles", "traffic_contact.json")
    if not os.path.exists(fn):
        return None
    try:
        with open(fn) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    if t.get("kernel_source_sha256") != kernel_source_hash():
        return None
    return t.get("bytes_per_launch"), t.get("command", "")


def build_workload(args):
    from sphere_amd import engine, scenes
    if args.load_image:
        with open(args.load_image, "rb") as fh:
            return fh.read()
    t0 = time.time()

    def opener(img):
        return engine.Context(img)

    s = scenes.settled_shear_cell(args.np, args.seed, open_stepper=opener,
                                  settle_steps=args.settle_steps,
                                  consolidate_steps=args.consolidate_steps, log=log)
    img = s.tobytes()
    log("workload prepacked in %.1f s" % (time.time() - t0))
    if args.save_image:
        with open(args.save_image, "wb") as fh:
            fh.write(img)
    return img


def apply_settings(ctx, args):
    for item in args.set:
        name, _, val = item.partition("=")
        ctx.set_scalar(name.strip(), float(val))
        log("set %s = %g" % (name.strip(), float(val)))


def batch_images(args, nsims):
    """The batched sweep: `nsims` consolidation beds of log-uniform size in
    [500, 20000] particles, each a settled 500-particle cell (radii uniform
    in [0.005, 0.01] m, one of four seeds, periodic in x and y, under a
    stress-controlled top wall) tiled in x and y to its size, with its own
    normal stress (uniform in [5, 160] kPa), stiffness (log-uniform in
    [1e6, 1e8] N/m, with its own critical damping and dt) and friction
    (uniform in [0.3, 0.7]).  kind "gas": granular gases collapsing in
    closed boxes from a lattice start."""
    from sphere_amd import engine, scenes
    sizes = scenes.batch_sizes(nsims, seed=args.seed)
    if args.batch_kind == "gas":
        images = [scenes.lattice_box(int(n), seed=args.seed + k).tobytes()
                  for k, n in enumerate(sizes)]
        return images, "batched granular gases, lattice start, contactmodel 2"
    t0 = time.time()

    def opener(img):
        return engine.Context(img)

    cells = scenes.consolidation_batch(nsims, args.seed, opener,
                                       cell_settle_steps=args.cell_settle_steps,
                                       cell_consolidate_steps=args.cell_consolidate_steps)
    log("batch of %d consolidation beds prepacked in %.1f s" % (nsims, time.time() - t0))
    return ([c.tobytes() for c in cells],
            "batched consolidation beds (tiled settled 500-particle cells, radii U[0.005, "
            "0.01] m), contactmodel 2, per-simulation normal stress U[5, 160] kPa, k_n "
            "log-uniform [1e6, 1e8] N/m with its own critical damping and dt, mu U[0.3, 0.7]")


def measure_batch(args, images, steps, warmup, graph_steps):
    """One batched context against the plain per-simulation loop (one
    context per simulation, stepped one after another) and against one
    context and stream per simulation all enqueued before one
    synchronisation; the same engine settings for all."""
    from sphere_amd import engine
    # np is the u32 after the f64 version and the i32 nd (src/file_io.cpp:430-433).
    sizes = [struct.unpack_from("<I", im, 12)[0] for im in images]
    total = int(sum(sizes))

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0

    def setup(c):
        if args.contact_lanes:
            c.set_scalar("contact_lanes", args.contact_lanes)
        c.set_scalar("graph_steps", graph_steps)
        apply_settings(c, args)

    ctx = engine.Context(images=images)
    setup(ctx)
    ctx.step(warmup, sync=True)
    t_batch = timed(lambda: ctx.step(steps, sync=True))
    z = sum(int((ctx.get("contacts", sim=k) != sizes[k]).sum()) for k in range(len(images)))
    ctx.profile(True)
    ctx.step(min(steps, 50), sync=True)
    rep = ctx.profile_report()
    ctx.profile(False)
    ctx.close()
    ctxs = [engine.Context(im) for im in images]
    for c in ctxs:
        setup(c)
        c.step(warmup, sync=False)
    for c in ctxs:
        c.sync()

    def overlapped():
        for c in ctxs:
            c.step(steps, sync=False)
        for c in ctxs:
            c.sync()

    def sequential():
        for c in ctxs:
            c.step(steps, sync=True)

    t_ovl = timed(overlapped)
    t_seq = timed(sequential)
    for c in ctxs:
        c.close()
    value = total * steps / t_batch
    log("batch of %d (%d particles): batched %.4f s, per-simulation loop %.4f s, enqueued "
        "together %.4f s for %d steps" % (len(images), total, t_batch, t_seq, t_ovl, steps))
    return {
        "value": round(value, 1), "unit": "particle-steps/s", "steps": steps,
        "nsims": len(images), "np_total": total, "np_min": int(min(sizes)),
        "np_max": int(max(sizes)), "batched_s": round(t_batch, 5),
        "per_sim_loop_s": round(t_seq, 5), "overlapped_loop_s": round(t_ovl, 5),
        "speedup_vs_per_sim_loop": round(t_seq / t_batch, 3),
        "speedup_vs_overlapped_loop": round(t_ovl / t_batch, 3),
        "contacts_per_particle_end": round(z / total, 3), "graph_steps": graph_steps,
        "kernel_us": {k: round(v[0], 3) for k, v in rep.items()},
    }


def batch_bench(args):
    """The batched sweep alone, as its own JSON line."""
    images, workload = batch_images(args, args.batch)
    r = measure_batch(args, images, args.steps, args.warmup, args.graph_steps)
    line = {
        "metric": "particle-steps/s, batch of %d simulations FP64" % len(images),
        "value": r["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * r["batched_s"] / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "nsims": r["nsims"], "np_total": r["np_total"],
                   "np_min": r["np_min"], "np_max": r["np_max"], "seed": args.seed},
        "batched": r,
    }
    print(json.dumps(line), flush=True)
    return 0


def adjoint_bench(args, img):
    """Wall time of sph_adjoint over K steps (record forward + reverse, with
    the seed upload and the gradient download) against K plain forward
    steps, on a context that already holds its tape."""
    import numpy
    from sphere_amd import engine
    K = args.adjoint
    ctx = engine.Context(img)
    n = int(ctx.scalar("np"))
    ctx.step(args.warmup, sync=True)
    state = ctx.tobytes()
    ctx.close()
    fwd = engine.Context(state)
    fwd.step(1, sync=True)
    t0 = time.perf_counter()
    fwd.step(K, sync=True)
    t_fwd = time.perf_counter() - t0
    fwd.close()
    rng = numpy.random.default_rng(args.seed)
    seed_v = rng.standard_normal((n, 3))

    def run_adjoint(k):
        # The first call allocates the context's tape; the second, timed,
        # continues from the state the first one left.
        adj = engine.Context(state)
        apply_settings(adj, args)
        adj.step(1, sync=True)
        adj.adjoint(k, None, seed_v, None)
        adj.sync()
        t0 = time.perf_counter()
        g = adj.adjoint(k, None, seed_v, None)
        adj.sync()
        t = time.perf_counter() - t0
        adj.close()
        return t, g

    t_adj, g = run_adjoint(K)
    k_short = max(1, K // 10)
    t_short, _ = run_adjoint(k_short)
    per_step = (t_adj - t_short) / (K - k_short) if K > k_short else float("nan")
    fixed = t_adj - K * per_step
    log("forward %d steps %.4f s, adjoint %.4f s (ratio %.2f; %.1f us per step + %.2f ms "
        "per call); dL/dk_n %.4e"
        % (K, t_fwd, t_adj, t_adj / t_fwd, per_step * 1e6, fixed * 1e3, g["k_n"]))
    print(json.dumps({"metric": "adjoint cost / forward cost", "value": round(t_adj / t_fwd, 3),
                      "unit": "ratio", "n_gpus": 1, "steps": K, "higher_is_better": False,
                      "np": n, "forward_s": round(t_fwd, 5), "adjoint_s": round(t_adj, 5),
                      "adjoint_us_per_step": round(per_step * 1e6, 2),
                      "adjoint_ms_per_call": round(fixed * 1e3, 3),
                      "forward_us_per_step": round(t_fwd / K * 1e6, 2),
                      "data": "synthetic"}), flush=True)
    return 0


def collapse_column(args, n, seed):
    from sphere_amd import engine, scenes
    return scenes.collapse_column(n, seed=seed, k_n=1.0e7, damping_frac=0.2,
                                  settle_steps=args.column_settle_steps,
                                  open_stepper=lambda img: engine.Context(img))


def measure_collapse(args, reps=5):
    """The differentiable column collapse: 2,000 spheres (radii uniform in
    [0.008, 0.012] m) settled in a 0.3 m column and released onto a
    frictional floor; 500 steps sampled every 50; the MSE of positions
    against a target made with k_n* = 2e7, gamma_n* = 0.3 critical, from
    k_n = 1e7, gamma_n = 0.2 critical.  The torch forward (the trajectory)
    against forward + backward (the loss gradient with respect to k_n,
    gamma_n, mu and the initial velocities), best of `reps`."""
    import torch
    from sphere_amd import autograd, scenes
    s = collapse_column(args, 2000, args.seed)
    traj = autograd.Trajectory(s, 50, 10, params=("k_n", "gamma_n", "mu"))
    r_min, rho = float(s.radius.min()), float(s.rho[0])

    def crit(k, f):
        return f * scenes.critical_damping(r_min, rho, k)

    with torch.no_grad():
        target = traj(torch.tensor([2.0e7, crit(2.0e7, 0.3), 0.5], dtype=torch.float64))
    theta0 = [1.0e7, crit(1.0e7, 0.2), 0.5]
    v0 = torch.as_tensor(s.vel[:, :3].copy())
    t_fwd = t_all = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        with torch.no_grad():
            loss = ((traj(torch.tensor(theta0, dtype=torch.float64), v0=v0) - target) ** 2).mean()
        t_fwd = min(t_fwd, time.perf_counter() - t0)
        theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        v = v0.clone().requires_grad_(True)
        t0 = time.perf_counter()
        loss = ((traj(theta, v0=v) - target) ** 2).mean()
        loss.backward()
        t_all = min(t_all, time.perf_counter() - t0)
    g = theta.grad.numpy()
    log("column collapse: forward %.4f s, forward + backward %.4f s (ratio %.2f); loss %.4e, "
        "dL/dk_n %.3e, dL/dgamma_n %.3e, dL/dmu %.3e, |dL/dv0| %.3e"
        % (t_fwd, t_all, t_all / t_fwd, float(loss.detach()), g[0], g[1], g[2],
           float(v.grad.norm())))
    return {"value": round(t_all / t_fwd, 3), "unit": "(forward + backward) / forward",
            "higher_is_better": False, "forward_s": round(t_fwd, 5),
            "forward_backward_s": round(t_all, 5), "np": 2000, "steps": 500,
            "workload": "granular column collapse, 2000 spheres, 500 steps sampled every 50, "
                        "MSE vs k_n*=2e7, gamma_n*=0.3 crit; gradient w.r.t. k_n, gamma_n, "
                        "mu, v0"}


def collapse_bench(args):
    r = measure_collapse(args)
    print(json.dumps({"metric": "trajectory-loss gradient cost / forward cost",
                      "value": r["value"], "unit": "ratio", "n_gpus": 1, "steps": 500,
                      "higher_is_better": False, "np": 2000, "forward_s": r["forward_s"],
                      "forward_backward_s": r["forward_backward_s"], "data": "synthetic",
                      "config": {"workload": r["workload"],
                                 "params": ["k_n", "gamma_n", "mu", "v0"]}}), flush=True)
    return 0


def measure_batch_calibration(args, nsims=64, n=4000, steps=300, every=50):
    """Many calibrations at once: `nsims` collapse columns of `n` spheres
    (four settled columns, each simulation with its own seeded initial
    velocities N(0, 0.01 m/s) and its own target parameters k_n* log-uniform
    in [1e7, 4e7] N/m, gamma_n* in [0.2, 0.4] critical, mu* in [0.3, 0.7]),
    `steps` steps sampled every `every`, the MSE of positions of each
    against its own target; one gradient step of all the (k_n, gamma_n,
    mu) fits as a batched trajectory + batched adjoint, against the plain
    loop of one trajectory + adjoint per simulation."""
    import numpy
    import torch
    from sphere_amd import autograd, scenes
    from sphere_amd.sim import sim
    bases = [collapse_column(args, n, args.seed + 10 * k) for k in range(4)]
    rng = numpy.random.default_rng(args.seed)
    sims, stars = [], []
    for k in range(nsims):
        s = sim().frombytes(bases[k % 4].tobytes())
        s.vel[:, :3] = rng.normal(0.0, 0.01, (n, 3))
        r_min, rho = float(s.radius.min()), float(s.rho[0])
        ks = float(numpy.exp(rng.uniform(numpy.log(1.0e7), numpy.log(4.0e7))))
        stars.append([ks, rng.uniform(0.2, 0.4) * scenes.critical_damping(r_min, rho, ks),
                      rng.uniform(0.3, 0.7)])
        sims.append(s)
    nsamp = steps // every
    batch = autograd.BatchTrajectory(sims, every, nsamp, params=("k_n", "gamma_n", "mu"))
    with torch.no_grad():
        target = batch(torch.tensor(stars, dtype=torch.float64))
    theta0 = numpy.array([[float(s.k_n[0]), float(s.gamma_n[0]), float(s.mu_s[0])]
                          for s in sims])

    # Each simulation's loss is its own MSE; their sum has, for each
    # simulation's parameters, that simulation's gradient.
    scale = 1.0 / (nsamp * n * 3)

    def batched():
        th = torch.tensor(theta0, requires_grad=True)
        (((batch(th) - target) ** 2).sum() * scale).backward()
        return th.grad.numpy().copy()

    def forward_only():
        with torch.no_grad():
            batch(torch.tensor(theta0))

    singles = [autograd.Trajectory(s, every, nsamp, params=("k_n", "gamma_n", "mu"))
               for s in sims]

    def loop():
        out = []
        for k, t in enumerate(singles):
            th = torch.tensor(theta0[k], requires_grad=True)
            a, b = batch.offsets[k], batch.offsets[k + 1]
            (((t(th) - target[:, a:b]) ** 2).sum() * scale).backward()
            out.append(th.grad.numpy().copy())
        return numpy.array(out)

    def timed(fn):
        t0 = time.perf_counter()
        r = fn()
        return time.perf_counter() - t0, r

    batched()  # warm-up (allocations)
    t_b, g_b = timed(batched)
    t_f, _ = timed(forward_only)
    t_l, g_l = timed(loop)
    same = bool(numpy.array_equal(g_b, g_l))
    log("batched calibration of %d x %d: batched %.3f s (forward %.3f s), loop %.3f s, "
        "gradients identical: %s" % (nsims, n, t_b, t_f, t_l, same))
    return {"nsims": nsims, "np_each": n, "steps": steps, "every": every,
            "batched_s": round(t_b, 4), "batched_forward_s": round(t_f, 4),
            "per_sim_loop_s": round(t_l, 4), "speedup_vs_per_sim_loop": round(t_l / t_b, 3),
            "adjoint_ratio": round(t_b / t_f, 3), "gradients_bit_identical": same}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--np", type=int, default=100000)
    ap.add_argument("--seed", type=int, default=2001)
    ap.add_argument("--settle-steps", type=int, default=40000)
    ap.add_argument("--consolidate-steps", type=int, default=10000)
    ap.add_argument("--cell-settle-steps", type=int, default=20000,
                    help="settling steps of the batched sweep's 500-particle base cells")
    ap.add_argument("--cell-consolidate-steps", type=int, default=5000)
    ap.add_argument("--column-settle-steps", type=int, default=100000)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=600,
                    help="oracle steps of the all-threads CPU baseline sample")
    ap.add_argument("--cpu-steps-1t", type=int, default=40,
                    help="oracle steps of the single-thread CPU baseline sample")
    ap.add_argument("--check-steps", type=int, default=20,
                    help="steps of the full-size bit-exact check of the timed state")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="leave the batched and adjoint parts out of the default line")
    ap.add_argument("--batch-steps", type=int, default=100,
                    help="steps of the batched-sweep part of the default line")
    ap.add_argument("--contact-lanes", type=int, default=0,
                    help="work-items per particle in the contact kernel (1, 4, 8, 16)")
    ap.add_argument("--sort-mode", type=int, default=-1,
                    help="0: sort launches every step, 1: end-of-step one-workgroup "
                         "re-sort, 2: automatic (default)")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="set an engine scalar before timing (e.g. xcd_order=1, "
                         "graph_steps=0); repeatable")
    ap.add_argument("--collapse", action="store_true",
                    help="instead: time the column-collapse calibration gradient (torch "
                         "autograd over the trajectory adjoint) against its forward pass")
    ap.add_argument("--batch-calibration", type=int, default=0, metavar="S",
                    help="instead: S batched calibrations of 4000-sphere collapse columns "
                         "over 300 steps against the per-simulation loop")
    ap.add_argument("--autotune", type=int, default=200, metavar="K",
                    help="before the warm-up, run sph_autotune with K steps per trial (0: skip)")
    ap.add_argument("--graph-steps", type=int, default=100,
                    help="steady-state chunk replayed from a captured HIP graph "
                         "(identical results; 0 = plain launches)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (5 times): the "
                         "reference kernel for scaling HBM byte counters")
    ap.add_argument("--batch-kind", choices=("consolidation", "gas"), default="consolidation")
    ap.add_argument("--batch", type=int, default=0, metavar="S",
                    help="instead: S independent simulations (log-uniform 500..20000 "
                         "particles) stepped as one batched context vs one context each")
    ap.add_argument("--adjoint", type=int, default=0, metavar="K",
                    help="instead: time the adjoint of K steps against K forward steps")
    ap.add_argument("--save-image")
    ap.add_argument("--load-image")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.mul(src, 1.0, out=dst)  # one elementwise kernel: reads and writes nbytes
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                          "launches": 5}), flush=True)
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this engine runs the DEM step on one GPU; --gpus %d is not supported: "
              "a batch of simulations or one large bed is measured on a single GPU."
              % max(args.gpus, world), file=sys.stderr)
        return 2

    if args.batch:
        return batch_bench(args)
    if args.collapse:
        return collapse_bench(args)
    if args.batch_calibration:
        r = measure_batch_calibration(args, nsims=args.batch_calibration)
        print(json.dumps({"metric": "batched calibration speedup vs per-simulation loop",
                          "value": r["speedup_vs_per_sim_loop"], "unit": "ratio", "n_gpus": 1,
                          "higher_is_better": True, "data": "synthetic", "config": r}),
              flush=True)
        return 0

    from sphere_amd import engine, scenes

    img = build_workload(args)
    if args.adjoint:
        return adjoint_bench(args, img)
    ctx = engine.Context(img)
    if args.contact_lanes:
        ctx.set_scalar("contact_lanes", args.contact_lanes)
    if args.sort_mode >= 0:
        ctx.set_scalar("sort_mode", args.sort_mode)
    ctx.set_scalar("graph_steps", args.graph_steps)
    apply_settings(ctx, args)
    graph_steps = int(ctx.scalar("graph_steps"))
    n = int(ctx.scalar("np"))
    nw = int(ctx.scalar("nw"))
    dt = ctx.scalar("dt")
    log("np %d, nw %d, ncells %d, dt %.4g s" % (n, nw, int(ctx.scalar("ncells")), dt))
    tuning = None
    if args.autotune:
        # Tuned on a second context from the same image (the trials advance
        # the simulation); the timed context takes the chosen settings and
        # starts from the image itself.
        tctx = engine.Context(img)
        tctx.set_scalar("graph_steps", args.graph_steps)
        apply_settings(tctx, args)
        us = tctx.autotune(args.autotune)
        knobs = ("contact_lanes", "contact_wg", "vl_cap", "radix_bits", "xcd_order")
        tuning = {k: int(tctx.scalar(k)) for k in knobs}
        tctx.close()
        for k in knobs:
            ctx.set_scalar(k, tuning[k])
        tuning["trial_steps"] = args.autotune
        tuning["contact_us"] = round(us, 3)
        log("autotune: contact kernel %.2f us with %s" % (us, tuning))
    ctx.step(args.warmup, sync=True)
    z = scenes.contact_number(ctx, n)

    sorts0, builds0 = ctx.scalar("sorts"), ctx.scalar("vl_builds")
    # Timed region: K steps, synchronised on both sides.
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(args.steps, sync=False)
    ctx.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = n * args.steps / elapsed
    z_end = scenes.contact_number(ctx, n)
    sorts = int(ctx.scalar("sorts") - sorts0)
    builds = int(ctx.scalar("vl_builds") - builds0)
    log("timed %d steps in %.4f s: %.4e particle-steps/s (contacts per particle %.2f -> %.2f, "
        "%d re-sorts, %d list rebuilds, contact lanes %d, sort mode %d, graph steps %d)"
        % (args.steps, elapsed, value, z, z_end, sorts, builds,
           int(ctx.scalar("contact_lanes")), int(ctx.scalar("sort_mode")), graph_steps))

    # Kernel timing pass with HIP events on the context stream.
    ctx.profile(True)
    ctx.step(args.profile_steps, sync=True)
    rep = ctx.profile_report()
    ctx.profile(False)
    log("kernel means (us): " + ", ".join("%s %.2f" % (k, v[0]) for k, v in rep.items()))
    zz = 0.5 * (z + z_end)
    # Algorithmic bytes of the contact kernel per particle: the search
    # (36 + 128 + 36 Z) plus the interaction (84 + 128 + 72 Z + 104 + 8 nw);
    # when it also integrates, plus the integration's 256 + 184.
    fused = "contact_integrate" in rep
    b_contact = 480.0 + 108.0 * zz + 8.0 * nw + (440.0 if fused else 0.0)
    t_contact = rep.get("contact_integrate" if fused else "contact", (0.0, 0))[0] * 1e-6
    timed = t_contact > 0.0  # no profiling pass (--profile-steps 0): no launch time
    achieved = n * b_contact / t_contact / 1e9 if timed else None
    traffic = measured_traffic(zz, nw)
    roofline = {"bound": "hbm", "kernel": "k_contact_g<2, 4, true>" if fused else "k_contact_g",
                "achieved": round(achieved, 2) if timed else None,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4) if timed else None,
                "traffic": traffic[0] if traffic else None,
                "traffic_source": ("not measured in this run: rocprofv3 --pmc FETCH_SIZE and "
                                   "WRITE_SIZE of the same kernel sources, "
                                   "profiles/traffic_contact.json" if traffic else None),
                "algorithmic_bytes_per_launch": round(n * b_contact),
                "mean_launch_us": round(t_contact * 1e6, 3) if timed else None,
                "mean_launch_us_source": "HIP events around each launch on the context stream",
                "contacts_per_particle": round(zz, 3)}
    # Whole-step algorithmic floor, B_min = 528 + 72 Z bytes per particle-step.
    step_bytes = 528.0 + 72.0 * zz
    step_gbs = step_bytes * value / 1e9
    state = ctx.tobytes()
    ctx.close()

    cpu = None
    check = None
    if not args.no_cpu:
        import numpy
        import oracle
        threads = cpu_threads()
        model = cpu_model()

        def cpu_rate(nthreads, steps):
            oracle.set_threads(nthreads)
            o = oracle.Oracle(state)
            o.step(1)
            c0 = time.perf_counter()
            o.step(steps)
            c1 = time.perf_counter()
            return n * steps / (c1 - c0), c1 - c0, o

        rate_all, s_all, o = cpu_rate(threads, args.cpu_steps)
        rate_1, s_1, _ = cpu_rate(1, args.cpu_steps_1t)
        oracle.set_threads(threads)
        cpu = {"value": round(rate_all, 1), "unit": "particle-steps/s", "cores": threads,
               "kind": "port",
               "single_thread": {"value": round(rate_1, 1), "cores": 1,
                                 "steps": args.cpu_steps_1t, "s": round(s_1, 2)},
               "cpu": model,
               "sample": "CPU oracle (oracle/sphere_oracle.cpp, g++ -O3 -fopenmp "
                         "-ffp-contract=off) from the %d-particle shear state after the GPU "
                         "timed region: %d steps on %d threads (%.1f s) and %d steps on one "
                         "thread (%.1f s)" % (n, args.cpu_steps, threads, s_all,
                                              args.cpu_steps_1t, s_1)}
        log("cpu baseline: %.4e particle-steps/s on %d threads, %.4e on one (%s)"
            % (rate_all, threads, rate_1, model))
        # The timed bed at full size: a fresh GPU context and the oracle from
        # the same image, the same steps, every field compared bit for bit.
        ref = oracle.Oracle(state)
        ref.step(args.check_steps)
        fields = ("perm", "cellid", "contacts", "x", "vel", "angvel", "force", "torque",
                  "delta_t", "walls_nx", "walls_mvfd")
        with engine.Context(state) as chk:
            chk.step(args.check_steps)
            diff = [f for f in fields if not numpy.array_equal(chk.get(f), ref.get(f))]
        check = {"steps": args.check_steps, "np": n, "fields": list(fields),
                 "bit_exact": not diff, "differing": diff}
        log("full-size check of the timed bed over %d steps: %s"
            % (args.check_steps, "bit-exact" if not diff else "DIFFERS in %s" % diff))

    batched = adjoint = None
    if not args.no_extras:
        images, _ = batch_images(args, 256)
        batched = measure_batch(args, images, max(args.batch_steps, 50), 100, 50)
        batched["workload"] = ("256 consolidation beds of log-uniform size in [500, 20000] "
                               "(radii U[0.005, 0.01] m, normal stress U[5, 160] kPa, k_n "
                               "log-uniform [1e6, 1e8] N/m, mu U[0.3, 0.7], own dt each); "
                               "one batched context vs one context per simulation stepped "
                               "one after another")
        adjoint = measure_collapse(args, reps=3)

    line = {
        "metric": "particle-steps/s at 1e5 particles FP64",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "direct shear, contactmodel 2, settled lognormal bed",
                   "np": n, "nw": nw, "periodic": "xy", "normal_stress_Pa": 1.0e4,
                   "shear_velocity_m_s": 1.0e-3, "k_n": 1.16e7, "mu": 0.5, "dt_s": dt,
                   "contacts_per_particle": round(zz, 3), "seed": args.seed,
                   "graph_steps": graph_steps},
        "roofline": roofline,
        "step_roofline": {"bytes_per_particle_step": round(step_bytes, 1),
                          "achieved_GBs": round(step_gbs, 2),
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
        "state_check": check,
        "batched": batched,
        "adjoint": adjoint,
        "kernel_us": {k: v[0] for k, v in rep.items()},
        "timed_events": {"re_sorts": sorts, "list_rebuilds": builds},
        "tuning": tuning,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
