#!/usr/bin/env python3
"""Throughput of the adaptive sparse-grid DG time step on one MI355X.

Default workload: the batching configuration, 256 independent 3D
variable-coefficient advection solves (seeds 1..256: speeds, bump, eps per
instance), NMAX 7, k=2, Lagrange P3, t=0.25.  A "step" is one full solve of
the workload from its initial state (restored from a device-resident
snapshot): every time step is predict (forward Euler RHS), refine, three
SSP-RK3 stages, coarsen.  The metric counts, for every RK stage, the active
degrees of freedom it updates (DoF-stage updates); the predict stage is done
but not counted.  `value` = DoF-stage updates / wall seconds of exactly
--steps timed solves, after --warmup untimed solves, with all data resident on
the GPU.

Prints one JSON line.  The roofline object describes the dominant kernel of
the chosen form, timed with HIP events on the library's streams during the
timed region; the cpu_baseline object times the CPU oracle (a restatement of
the reference's algorithm with its data structures) on the same instances,
one instance per host thread, for a bounded wall time.
"""
import argparse
import ctypes
import glob
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "adaptive-multiresolution-dg_amd"))

import pyadg  # noqa: E402
import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "DoF-stage updates/s on 1 GPU: plain FP64 vs chosen form at equal L2 error"


def _oracle_for(w):
    import pyoracle
    o = pyoracle.Oracle(w.dim, w.k, w.m, w.nmax, w.sparse, w.level_init)
    o._keep.append(w.ic)
    o.init_adaptive_intp_c(pyadg.ic_eval_ptr(), ctypes.cast(ctypes.pointer(w.ic), ctypes.c_void_p), w.init_eps)
    return o


def cpu_baseline(ws, max_seconds, threads):
    """Times the CPU oracle on the workload's own instances: min(threads,
    instances) host threads, each solving instances from t=0 one time step per
    call with threads / (host threads) OpenMP threads per oracle (all of them
    for a single instance), until max_seconds of wall time have passed.  For
    the linear workload: time steps of the one instance on `threads` OpenMP
    threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    import runs

    w = ws[0]
    if w.linear:
        pyoracle.set_threads(threads)
        o = _oracle_for(w)
        sec, dof_stages, steps = 0.0, 0, 0
        while sec < max_seconds and steps < w.nsteps:
            n = min(20, w.nsteps - steps)
            t0 = time.perf_counter()
            dof_stages += runs.run_linear(o, w.coef, w.dt, n)
            sec += time.perf_counter() - t0
            steps += n
        return {"value": dof_stages / sec, "unit": "DoF-stage updates/s", "cores": threads, "kind": "port",
                "sample": "CPU oracle, first %d of %d time steps of the solve on %d OpenMP threads: %d DoF-stage "
                          "updates in %.2f s" % (steps, w.nsteps, threads, dof_stages, sec)}
    lock = threading.Lock()
    state = {"next": 0, "dof": 0, "steps": 0, "solves": 0, "partial": 0}
    deadline = [None]

    nworkers = max(1, min(threads, len(ws)))
    omp = max(1, threads // nworkers)

    def worker():
        pyoracle.set_threads(omp)
        while True:
            with lock:
                i = state["next"]
                state["next"] += 1
            if i >= len(ws) or time.perf_counter() >= deadline[0]:
                return
            wi = ws[i]
            o = _oracle_for(wi)
            prob = wi.oracle_problem()
            prob.t_start, prob.max_steps = 0.0, 1
            done = True
            while prob.t_start < prob.t_final:
                if time.perf_counter() >= deadline[0]:
                    done = False
                    break
                n, ds, t = runs.run_hyperbolic(o, prob)
                prob.t_start = t
                with lock:
                    state["dof"] += ds
                    state["steps"] += n
            with lock:
                state["solves" if done else "partial"] += 1
            o.close()

    th = [threading.Thread(target=worker) for _ in range(nworkers)]
    t0 = time.perf_counter()
    deadline[0] = t0 + max_seconds
    for t in th:
        t.start()
    for t in th:
        t.join()
    sec = time.perf_counter() - t0
    return {"value": state["dof"] / sec, "unit": "DoF-stage updates/s", "cores": threads, "kind": "port",
            "sample": "CPU oracle on the workload's instances (%d host threads, one instance at a time each on %d "
                      "OpenMP threads, reference data structures, g++ -O3 -fopenmp): %d whole solves and %d stopped "
                      "at the %.0f s wall limit, %d time steps, %d DoF-stage updates in %.2f s wall (oracle setup "
                      "included in the wall time)" % (nworkers, omp, state["solves"], state["partial"], max_seconds,
                                                      state["steps"], state["dof"], sec)}


def imex_sample(args):
    """IMEX43 with the assembled artificial viscosity (viscosity option 1 of
    example/02_hyperbolic_05_burgers_adapt.cpp:343-420) on a fixed grid: the
    rotating-wave initial value of c5 (seed 1) on the NMAX 6 grid its adaptive
    initialisation builds, shock detector kappa 0, viscosity -nu0 h with
    nu0 = 2, h = 2^-6, LF flux terms as the explicit part.  A time step is the
    detector, the assembly and the five IMEX43 stages; the GPU and the CPU
    oracle run the same steps from the same state.  Prints one JSON line."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    import runs

    w = workloads.c5_rotating(seed=1, nmax=6, level_init=4)
    prob = pyadg.Problem.from_buffer_copy(w.problem)
    dt, coef = 0.05 / (2 * 0.25 * 2 ** 6), -2.0 / 2 ** 6

    def step(s):
        s.viscosity_elements(0.0)
        s.viscosity_operator(coef, 1)
        s.imex43_init(dt)
        it = 0
        for st in range(5):
            if st >= 2:
                runs.hyperbolic_rhs(s, prob)
            it += max(s.imex43_stage(st), 0)
        return it

    g = workloads.make_solution(w)
    step(g)  # warm-up
    g.sync()
    n_gpu, t0 = 20, time.perf_counter()
    iters = sum(step(g) for _ in range(n_gpu))
    g.sync()
    gpu_sec = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pyoracle.set_threads(threads)
    o = pyoracle.Oracle(w.dim, w.k, w.m, w.nmax, w.sparse, w.level_init)
    o._keep.append(w.ic)
    o.init_adaptive_intp_c(pyadg.ic_eval_ptr(), ctypes.cast(ctypes.pointer(w.ic), ctypes.c_void_p), w.init_eps)
    step(o)  # the same first step as the GPU's warm-up
    n_cpu, cpu_sec = 0, 0.0
    while n_cpu < n_gpu and cpu_sec < args.cpu_seconds:
        t0 = time.perf_counter()
        step(o)
        cpu_sec += time.perf_counter() - t0
        n_cpu += 1
    dofs = g.num_elem * g.K
    print(json.dumps({
        "metric": "IMEX43 time steps/s with the assembled artificial viscosity (option 1), fixed grid",
        "value": n_gpu / gpu_sec, "unit": "time steps/s", "higher_is_better": True, "dtype": "f64",
        "data": "synthetic", "config": {"workload": "c5 rotating-wave initial value (seed 1), NMAX 6, level_init 4",
                                        "elements": g.num_elem, "dofs": dofs, "dt": dt, "viscosity": coef,
                                        "gpu_steps": n_gpu, "cg_iterations_per_step": iters / n_gpu},
        "ms_per_step": 1e3 * gpu_sec / n_gpu,
        "cpu_baseline": {"value": n_cpu / cpu_sec, "unit": "time steps/s", "cores": threads, "kind": "port",
                         "sample": "CPU oracle, %d IMEX43 steps of the same problem after the same first step "
                                   "(assembly as std::map of entries, conjugate gradient in index order)" % n_cpu},
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c4", choices=sorted(workloads.WORKLOADS))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="wall time of the CPU oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-events", action="store_true", help="time without per-kernel HIP events")
    ap.add_argument("--time-steps", type=int, default=0,
                    help="cap the time steps of every solve (0: run to the workload's final time)")
    ap.add_argument("--streams", type=int, default=3,
                    help="contexts (HIP streams, host threads) the batch is split over in the chosen form")
    ap.add_argument("--batch", type=int, default=None,
                    help="independent seeded instances solved together in the chosen form (c4: 256, the "
                         "configuration's own batch; c5: 256 and c2: 1024 seeds of the configuration)")
    ap.add_argument("--refine", default="predict", choices=("predict", "learned"),
                    help="refinement of the chosen form's adaptive steps: forward-Euler prediction (the reference) "
                         "or the learned proposer with the threshold judge")
    ap.add_argument("--threshold", type=float, default=0.01, help="proposer probability threshold")
    ap.add_argument("--judge", default="redo", choices=("redo", "defer"),
                    help="learned refinement: the judge redoes the instances it refines (redo) or gives the "
                         "missed leaves their children after the step (defer)")
    ap.add_argument("--train-steps", type=int, default=40, help="time steps of each logged training run")
    ap.add_argument("--train-samples", type=int, default=2000000, help="cap on the logged decisions trained on")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "fp32"),
                    help="storage precision of the chosen form's fused stages (fp32: split FP32 tables, FP64 "
                         "accumulation, RK and refine/coarsen norms)")
    ap.add_argument("--variant", type=int, default=-1, choices=range(-1, 10),
                    help="chosen form: thread mapping of every fused kernel (-1: the autotuner picks per class)")
    ap.add_argument("--mappings", default=None,
                    help="chosen form: thread mapping per fused kernel class (eval-up phases, eval-up final, "
                         "right-hand-side phases, right-hand-side final), e.g. 3,3,7,7, or 'auto' for the "
                         "autotuner; default 3,3,7,7 on c4 (the best per class in a sweep of all ten mappings "
                         "at the per-context size), auto elsewhere")
    ap.add_argument("--graph", type=int, default=0, choices=(0, 1),
                    help="chosen form: record whole RK steps into HIP graphs and replay them (adg_set_graph_step)")
    ap.add_argument("--adapt", default="device", choices=("host", "device"),
                    help="chosen form: refine / coarsen and the element-table rebuild on the GPU "
                         "(adg_set_device_adapt, default) or on the host (decisions from device norms)")
    ap.add_argument("--imex", action="store_true",
                    help="instead of the main measurement: IMEX43 with the assembled artificial viscosity, GPU "
                         "against the CPU oracle on one fixed grid")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.imex:
        imex_sample(args)
        return
    make = workloads.WORKLOADS[args.workload]
    seeded = args.workload in ("c2", "c4", "c5")
    if args.batch is None:
        # C5's grids hold ~45 k elements per instance at NMAX 9: 256 instances
        args.batch = 256 if args.workload in ("c4", "c5") else (1024 if seeded else 1)
    batched = seeded and args.batch > 1

    def workload(seed):
        return make(seed=seed) if seeded else make()

    w = workload(args.seed)
    print(f"[bench] workload {w.name}: init", file=sys.stderr, flush=True)
    sol = workloads.make_solution(w)
    if batched:
        # beside a batched chosen form, the single-instance fused form (a
        # comparison point) runs the per-element mapping 3 throughout, so the
        # chosen form's kernel instances are its own in kernel traces; the
        # comparison forms adapt on the host, as the reference does
        sol.set_kernel_variant(3)
        sol.set_device_adapt(0)
    print(f"[bench] init done: {sol.num_elem} elements; warmup {args.warmup} steps", file=sys.stderr, flush=True)
    sol.save_state(0)
    bsols = []
    if batched:
        ws = [workload(args.seed + i) for i in range(args.batch)]
        print(f"[bench] batch of {args.batch} instances (seeds {args.seed}..{args.seed + args.batch - 1}) in "
              f"{args.streams} context(s): init", file=sys.stderr, flush=True)
        per = (args.batch + args.streams - 1) // args.streams
        for c in range(args.streams):
            part = ws[c * per:(c + 1) * per]
            if part:
                b = workloads.make_batch(part)
                b.save_state(0)
                bsols.append(b)
    prec = pyadg.PREC_FP32_SPLIT if args.precision == "fp32" else pyadg.PREC_FP64
    if args.mappings is None:
        args.mappings = "3,3,7,7" if args.workload == "c4" else "auto"
    mappings = [] if args.mappings == "auto" else [int(v) for v in args.mappings.split(",")]
    if mappings and prec == pyadg.PREC_FP32_SPLIT:  # classes 4..7 are the FP32 ones
        mappings = [-1] * 4 + mappings
    chosen_sols = bsols if batched else ([] if w.linear else [workloads.make_solution(w)])
    for c in chosen_sols:
        c.set_precision(prec)
        if c is not sol:
            c.save_state(0)
    # launch settings of the chosen form; without a batch the single-instance
    # fused form measured on `sol` is the chosen form in FP64
    for c in chosen_sols + ([] if batched else [sol]):
        c.set_graph_step(args.graph)
        c.set_device_adapt(1 if args.adapt == "device" else 0)
        c.set_kernel_variant(args.variant)
        if mappings:
            c.set_class_variants(mappings)
    if batched:
        print(f"[bench] batch init done: {sum(b.num_elem for b in bsols)} elements", file=sys.stderr, flush=True)
    learned = None
    if args.refine == "learned" and w.problem is not None and w.problem.adaptive:
        # decision logs of plain runs of 64 other seeds train the proposer (not timed)
        t0 = time.perf_counter()
        train = workloads.make_batch([workload(10001 + i) for i in range(64)])
        train.decision_log(True)
        tp = pyadg.Problem.from_buffer_copy(w.problem)
        tp.form, tp.max_steps = 1, args.train_steps
        train.run_hyperbolic(tp)
        x, y = train.decision_log_get()
        if len(y) > args.train_samples:  # a seeded subsample bounds the training time
            keep = np.sort(np.random.default_rng(args.seed).choice(len(y), args.train_samples, replace=False))
            x, y = x[keep], y[keep]
        print(f"[bench] proposer training on {len(y)} logged decisions", file=sys.stderr, flush=True)
        train.proposer_init(args.seed)
        loss = train.proposer_train(x, y, epochs=20, batch=1024, lr=2e-3, seed=args.seed)
        params = train.proposer_params()
        del train
        for c in chosen_sols:
            c.set_proposer_params(params)
            c.set_refine_mode(pyadg.REFINE_LEARNED, args.threshold)
            c.set_judge_defer(1 if args.judge == "defer" else 0)
        learned = {"training_samples": int(len(y)), "refined_fraction": float(y.mean()), "final_loss": loss,
                   "training_seconds": time.perf_counter() - t0, "threshold": args.threshold,
                   "training_steps_per_instance": args.train_steps, "judge": args.judge}
        print(f"[bench] proposer trained on {len(y)} logged decisions in {learned['training_seconds']:.1f} s",
              file=sys.stderr, flush=True)

    def measure(sols, form, nsolves=None):
        """Time nsolves (default --steps) full solves (after --warmup) of one form; returns figures and
        HIP-event profile.
        Several contexts are solved concurrently, one host thread each (each context has its own
        stream; the host work of one overlaps the kernels of the others)."""
        if not isinstance(sols, list):
            sols = [sols]
        probs = [None if w.linear else pyadg.Problem.from_buffer_copy(w.problem) for _ in sols]

        def solve(i):
            s, prob = sols[i], probs[i]
            s.restore_state(0)
            if w.linear:
                return s.run_linear(w.coef, w.dt, min(w.nsteps, args.time_steps) if args.time_steps else w.nsteps)
            prob.t_start, prob.max_steps, prob.form = 0.0, args.time_steps, form
            return s.run_hyperbolic(prob)

        results = [[] for _ in sols]

        def worker(i, n):
            for j in range(n):
                results[i].append(solve(i))
                print(f"[bench] form {form}, context {i}: solve {j + 1}/{n} done", file=sys.stderr, flush=True)

        def run_all(n):
            th = [threading.Thread(target=worker, args=(i, n)) for i in range(len(sols))]
            for t in th:
                t.start()
            for t in th:
                t.join()

        # the first context warms up alone, so that the autotuner (whose
        # decisions the contexts share) times its kernels without the others
        if args.warmup > 0 and len(sols) > 1:
            worker(0, args.warmup)
            th = [threading.Thread(target=worker, args=(i, args.warmup)) for i in range(1, len(sols))]
            for t in th:
                t.start()
            for t in th:
                t.join()
        else:
            run_all(args.warmup)
        for s in sols:
            if not args.no_events:
                s.profile(True)
            s.sync()
        for r in results:
            r.clear()
        counts0 = [(s.stage_launch_count(), s.graph_step_count()) for s in sols]
        t0 = time.perf_counter()
        nsolves = args.steps if nsolves is None else nsolves
        run_all(nsolves)
        for s in sols:
            s.sync()
        sec = time.perf_counter() - t0
        counts1 = [(s.stage_launch_count(), s.graph_step_count()) for s in sols]
        launches = sum(b[0][0] - a[0][0] for a, b in zip(counts0, counts1))
        stages = sum(b[0][1] - a[0][1] for a, b in zip(counts0, counts1))
        graphs = (sum(b[1][0] - a[1][0] for a, b in zip(counts0, counts1)),
                  sum(b[1][1] - a[1][1] for a, b in zip(counts0, counts1)))
        dof_stages = sum(st.dof_stages for r in results for st in r)
        redone = sum(st.redone for r in results for st in r)
        time_steps = max(sum(st.steps for st in r) for r in results)
        max_elem = sum(max(st.max_elem for st in r) for r in results)
        prof = {k: {"ms": 0.0, "count": 0, "bytes": 0.0} for k in sols[0].profile_get()}
        host = {}
        for s in sols:
            for k, v in s.profile_get().items():
                for f in ("ms", "count", "bytes"):
                    prof[k][f] += v[f]
            for k, v in s.profile_host().items():
                host[k] = host.get(k, 0.0) + v
            s.profile(False)
        s0 = sols[0]
        u = s0.instance_get(0) if s0.num_instances > 1 else s0.get(pyadg.FIELD_UCOE)
        keys = s0.instance_grid(0)[2] if s0.num_instances > 1 else s0.grid()[2]
        ninst = sum(s.num_instances for s in sols)
        print(f"[bench] form {form}, {ninst} instance(s) in {len(sols)} context(s): {nsolves} solves "
              f"({time_steps} time steps) in {sec:.3f} s, {dof_stages / sec:.4g} DoF-stage/s", file=sys.stderr,
              flush=True)
        return {"value": dof_stages / sec, "sec": sec, "dof_stages": dof_stages, "time_steps": time_steps,
                "nsolves": nsolves,
                "max_elem": max_elem, "prof": prof, "host": host, "u": u, "keys": keys, "redone": redone,
                "sim_time": results[0][-1].final_time, "launches_per_stage": launches / stages if stages else None,
                "graph_steps_recorded": graphs[0], "graph_steps_replayed": graphs[1]}

    # the comparison forms (plain, single-instance fused) are timed on a few
    # solves; the chosen form, whose rate is `value`, on --steps solves
    ncmp = min(args.steps, 3) if batched else args.steps
    plain = measure(sol, 0, ncmp)
    fused = plain if w.linear else measure(sol, 1, ncmp if batched else None)
    if batched:
        chosen = measure(bsols, 1)
    elif w.linear or (prec == pyadg.PREC_FP64 and learned is None):
        chosen = fused
    else:
        chosen = measure(chosen_sols, 1)
    value = chosen["value"]
    dstats = [c.device_adapt_stats(reasons=True) for c in (bsols if batched else chosen_sols or [sol])]
    if learned is not None:
        learned["judge_deferred_leaves"] = int(sum(c.proposer_stats()[2] for c in chosen_sols))
    # instance 0 of the batch is the plain problem itself: equal L2 error means equal coefficients
    same_set = bool(np.array_equal(plain["keys"], chosen["keys"]))
    identical = same_set and bool(np.array_equal(plain["u"], chosen["u"]))
    if same_set:
        l2_diff = float(np.sqrt(np.sum((plain["u"] - chosen["u"]) ** 2)))
        l2_rel = l2_diff / float(np.sqrt(np.sum(plain["u"] ** 2)))
    else:
        l2_diff = l2_rel = None

    roofline = None
    prof = chosen["prof"]
    # the dominant kernel class of the chosen form, named by the mapping the
    # autotuner chose for it (one kernel instance per launch in the phase
    # classes; the final classes launch their tiles, penalties and combine)
    cats = ("sweep",) if chosen is plain else ("sweep", "rhs_phase", "flux", "rk", "surplus")
    cat = max(cats, key=lambda c: prof[c]["ms"])
    kernel = {"sweep": "k_sweep", "surplus": "k_surplus<double>", "flux": "eval-up final", "rk": "rhs final"}.get(cat)
    tuning = {}
    if chosen is not plain and cat in ("sweep", "rhs_phase", "flux", "rk"):
        cls = {"sweep": 0, "flux": 1, "rhs_phase": 2, "rk": 3}[cat] + (4 if prec == pyadg.PREC_FP32_SPLIT else 0)
        fixed = args.variant if args.variant >= 0 else (mappings[cls] if cls < len(mappings) else -1)
        if fixed >= 0:
            tuning = {"all": fixed}
        tune_src = bsols if batched else ([sol] if chosen is fused else chosen_sols)
        for c in tune_src if fixed < 0 else []:
            for (kc, bucket), v in c.kernel_tuning().items():
                if kc == cls:
                    tuning[bucket] = v
        if cat in ("sweep", "rhs_phase") and len(set(tuning.values())) == 1:
            v = next(iter(tuning.values()))
            pf, pt = (w.k + 1, w.m + 1) if cat == "sweep" else (w.m + 1, w.k + 1)
            ty = "float" if prec == pyadg.PREC_FP32_SPLIT else "double"
            kernel = ("k_tile<%d, %d, %s, %d, false>" % (pf, pt, ty, 4 if v >= 7 else 1) if v >= 4
                      else "k_phase<%d, %d, %s, %d>" % (pf, pt, ty, v))
        elif kernel is None:
            kernel = "phase kernels (mappings %s)" % sorted(set(tuning.values()))
    pc = prof[cat]
    if pc["count"] > 0 and pc["ms"] > 0:
        avg_ms = pc["ms"] / pc["count"]
        avg_bytes = pc["bytes"] / pc["count"]
        achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
        # HBM bytes per launch: the traffic / algorithmic ratio measured by the FETCH_SIZE and
        # WRITE_SIZE counter passes recorded under profiles/ for this kernel and workload,
        # applied to this run's launches
        traffic, traffic_src = None, None
        here = os.path.dirname(os.path.abspath(__file__))
        for tf in sorted(glob.glob(os.path.join(here, "profiles", "traffic_*.json"))):
            with open(tf) as f:
                t = json.load(f)
            if t.get("workload") == args.workload and t.get("kernel") == kernel:
                traffic = t["traffic_over_algorithmic"] * avg_bytes
                traffic_src = "profiles/%s (ratio %.2f)" % (os.path.basename(tf), t["traffic_over_algorithmic"])
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                    "kernel": kernel, "kernel_class": cat, "tuned_mapping_by_grid_size_class": tuning or None,
                    "avg_launch_us": avg_ms * 1e3, "avg_algorithmic_bytes": avg_bytes, "launches": pc["count"],
                    "kernel_ms_by_category": {k: round(v["ms"], 3) for k, v in prof.items()},
                    "launches_by_category": {k: v["count"] for k, v in prof.items()},
                    "avg_algorithmic_bytes_by_category": {k: (v["bytes"] / v["count"] if v["count"] else 0.)
                                                          for k, v in prof.items()},
                    "host_ms_inclusive": {k: round(v, 3) for k, v in chosen["host"].items()}}

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        print(f"[bench] cpu baseline on {threads} threads", file=sys.stderr, flush=True)
        cws = ws if batched else [w]
        cpu = cpu_baseline(cws, args.cpu_seconds, threads)
        one = cpu_baseline(cws, args.cpu_seconds / 4, 1)
        cpu["one_thread_value"] = one["value"]

    where = "GPU" if args.adapt == "device" else "host"
    pname = ("FP32 (split FP32 tables; FP64 right-hand-side accumulation, RK and norms)"
             if prec == pyadg.PREC_FP32_SPLIT and not w.linear else "FP64")
    line = {
        "metric": METRIC, "value": value, "unit": "DoF-stage updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": chosen["sec"] / max(args.steps, 1) * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == pyadg.PREC_FP64 or w.linear else "f32",
        "data": "synthetic",
        "config": dict(w.config(), step="one full solve from the saved initial state",
                       form=("chosen: fused %s stages, %d seeded instances (seeds %d..%d) in %d context(s) solved "
                             "concurrently, refine/coarsen on the %s"
                             % (pname, args.batch, args.seed, args.seed + args.batch - 1, len(bsols), where))
                       if batched else "chosen: fused %s stages, refine/coarsen on the %s" % (pname, where),
                       launches_per_stage=chosen["launches_per_stage"], graph_step=args.graph,
                       kernel_variant=args.variant, mappings=args.mappings,
                       graph_steps_recorded=chosen["graph_steps_recorded"],
                       graph_steps_replayed=chosen["graph_steps_replayed"],
                       precision=args.precision if not w.linear else "fp64",
                       contexts=len(bsols) if batched else 1,
                       plain_form="plain FP64, one kernel per 1D sweep, one instance, refine/coarsen on the %s"
                       % ("host" if batched else where),
                       fused_single_instance_form=("fused FP64 stages, one instance, per-element mapping 3, "
                                                   "refine/coarsen on the host"
                                                   if batched else "fused FP64 stages, one instance (chosen)"),
                       plain_value=plain["value"], fused_single_instance_value=fused["value"],
                       chosen_value=chosen["value"], batch=args.batch if batched else 1,
                       speedup_vs_plain=chosen["value"] / plain["value"],
                       forms_bit_identical=identical, element_sets_identical=same_set,
                       l2_difference_between_forms=l2_diff, relative_l2_difference_between_forms=l2_rel,
                       time_steps_per_solve=plain["time_steps"] // max(plain["nsolves"], 1),
                       comparison_form_solves=plain["nsolves"],
                       time_step_cap=args.time_steps or None,
                       max_elements=int(chosen["max_elem"]), dof_stage_updates=int(chosen["dof_stages"]),
                       refine=args.refine if learned else "predict", learned_proposer=learned,
                       chosen_time_steps_per_solve=chosen["time_steps"] // max(args.steps, 1),
                       judge_redone_instance_steps=int(chosen["redone"]),
                       adapt=args.adapt, device_adapt_calls=int(sum(d[0] for d in dstats)),
                       device_adapt_host_fallbacks=int(sum(d[1] for d in dstats)),
                       device_adapt_fallback_reasons=int(np.bitwise_or.reduce([d[2] for d in dstats] or [0])),
                       host_grid_share_per_context=(chosen["host"].get("grid", 0.) * 1e-3
                                                    / (chosen["sec"] * (len(bsols) if batched else 1))),
                       sim_time=chosen["sim_time"]),
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
