"""Benchmark: 3D Yee FDTD step throughput on one MI355X.

Workload (default): 256^3 FP32 grid, 10-cell CPML (m=3, kappa_max=5, alpha_max=0.05), 40 seeded
dielectric spheres stored as a uint8 material index, oblique TF/SF plane wave at seeded angles,
64 probes recording Ez and Hy every step, Ez snapshot every 250 steps.  A step is one full time
step of the grid (H update, TF/SF, CPML, E update, TF/SF, CPML, probes): the TF/SF face-sheet
fill, the row kernel's interior part (rows P+1 .. J-P-1 of the planes P .. K-P-1, the step's
dominant launch), its slab part (the other rows and planes, y/z CPML inline) and the probe gather.

Prints one JSON line: Gcell-updates/s over exactly --steps timed steps (inputs resident in HBM),
the interior part's HBM roofline fraction from HIP-event times around its launches, the CPU
oracle timed on a bounded sample of the same workload on this host's cores, and the GPU's probe
record and fields after the oracle's timed steps compared with the oracle's (float32, and a
shorter float64 run).

The CPU oracle (oracle/) is used only by the CPU-baseline leg, as the baseline and the checker:
cpu_baseline and the parity checks (oracle_check, currents_parity, state_diff), all after the
timed region and all skipped with --no-cpu; nothing timed runs on it.

--copy-calibration runs a plain device copy of known size (the scale of the HBM byte counters).
--workload batch runs the 96-grid workload through the batched engine and, for comparison, each
grid alone; --workload serve runs a Poisson stream of solve requests through the request server
and, for comparison, one after another through the plain path (secondary measurements; the
default line is the 256^3 workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HIP events bracket every 16th step of the timed run (and, inside it, the row kernel's interior
# part): each event pair costs the stream about 4 us, so timing every step would slow the run
# that `value` measures by ~7%.
TIMING_EVERY = 16
# parity bar of a float32 GPU result against the float32 CPU oracle (relative L2)
REL_L2_TOL = 1e-5
FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
# states (probe records, field components) whose float64 RMS is below this (the incident plane
# wave has unit amplitude) hold rounding residue only -- e.g. a low-frequency pulse that has not
# started within the run's steps, or a component the polarisation never excites -- and a relative
# L2 against them measures nothing; they are listed, not counted in the relative maxima
F64_SIGNAL_FLOOR = 1e-9


def bytes_per_cell(sc, material_bytes=1):
    """Compulsory HBM bytes per cell-update: 6 components read + written, material, 12 CPML psi
    arrays read + written over the face slabs."""
    P = sc.P
    return 48.0 + material_bytes + 64.0 * P * (1.0 / sc.I + 1.0 / sc.J + 1.0 / sc.K)


def row_kernel_bytes(sc, material_bytes=1):
    """Compulsory HBM bytes of one launch of the row kernel's interior part: the rows j in
    [P+1, J-P) of the planes k in [P, K-P), every column: 6 components read + written and the
    material byte per cell, plus the x-slab psi (Hyx, Hzx, Eyx, Ezx read + written) in the 2P
    x-slab columns of each row."""
    P = sc.P
    rows = (sc.J - 2 * P - 1) * (sc.K - 2 * P)
    return rows * (sc.I * (48.0 + material_bytes) + 2 * P * 32.0)


def run_roofline(nbytes, seconds):
    """A whole run's compulsory HBM bytes (bytes_per_cell of every cell update) over its wall time:
    the bandwidth the batched and served lines reach, against the 8 TB/s peak."""
    achieved = nbytes / seconds / 1e9
    return {"kernel": "whole run (compulsory bytes of every step of every grid / wall time)", "bound": "hbm",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes": int(nbytes)}


def copy_calibration(nbytes=1 << 30, reps=10):
    """A plain device-to-device copy of nbytes, reps times (reads and writes nbytes per launch)."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": nbytes, "launches": reps}), flush=True)


def run_batch(args):
    """Many-grid workload (96 independent grids, 32..268 cells per edge, 300..1500 steps each):
    every grid run alone one after another with fdtd_run; all grids queued at once on their own
    streams; all grids advanced together by the batched engine (merged launches, the device work
    queue, the engine's automatic choice).  --batch-legs picks the legs; reported are each leg's
    rate, the batched / sequential ratio, and whether every grid's probe record of each batched run
    equals that of its run alone (when the sequential leg ran).  Unless --no-cpu, the probe records
    of a sample of the grids (the least-work grid of each class of largest edge, then the least-work
    others up to --oracle-cells cell-updates) from the leg that gives `value` are compared with the
    CPU oracle run in float32 and float64 (`parity`)."""
    import numpy as np

    from maxwell_solver_amd import Batch, Simulation
    from maxwell_solver_amd import scenarios as S
    work = S.config4(args.grids, args.min_edge, args.max_edge)
    sims = []
    for sc, steps in work:
        g = Simulation()
        S.setup(g, sc)
        sims.append((g, sc, steps))
    total = sum(sc.cells * steps for sc, steps in work)
    # warm both paths (kernel modules load on first use), then rewind
    g0, sc0, _ = sims[0]
    g0.run(0, 2, sc0.mode)
    g0.synchronize()
    b = Batch()
    b.add(g0, 0, 2, sc0.mode)
    b.run()
    b.clear()
    g0.reset()
    g0.synchronize()
    legs = [x for x in args.batch_legs.split(",") if x]
    bad = set(legs) - {"sequential", "streams", "merged", "queue", "auto"}
    if bad or not legs:
        raise SystemExit(f"--batch-legs: unknown leg(s) {sorted(bad)} (sequential, streams, merged, queue, auto)")
    times, rec, same = {}, None, True
    # the leg that gives `value`, and the grids whose probe records from it go to the CPU oracle
    order = [x for x in ("sequential", "streams", "merged", "queue", "auto") if x in legs]
    val_leg = "auto" if "auto" in legs else order[0]
    # one grid per class of largest edge (32, 64, 96, 128, 192 and up), then the least work
    sample = sample_by_work([sc.cells * steps for sc, steps in work], args.oracle_cells,
                            [size_class(sc, (0, 64, 96, 128, 192)) for sc, _ in work])
    kept = {}

    def rewind(leg):
        if leg == val_leg:
            for q in sample:
                kept[q] = sims[q][0].probes().copy()
        for g, _, _ in sims:
            g.reset()
            g.synchronize()

    if "sequential" in legs:
        # one grid after another (each run completes before the next starts)
        t0 = time.perf_counter()
        for g, sc, steps in sims:
            g.run(0, steps, sc.mode)
            g.synchronize()
        times["sequential"] = time.perf_counter() - t0
        rec = [g.probes().copy() for g, _, _ in sims]
        rewind("sequential")
    if "streams" in legs:
        # every grid queued at once on its own stream (the GPU overlaps them as its queues allow)
        t0 = time.perf_counter()
        for g, sc, steps in sims:
            g.run(0, steps, sc.mode)
        for g, _, _ in sims:
            g.synchronize()
        times["streams"] = time.perf_counter() - t0
        rewind("streams")
    # merged launches (each kernel of the step once for all grids, four interleaved sub-batches),
    # the device work queue (persistent launches pulling tiles from per-XCD counters), and the
    # engine's automatic choice (the measured value)
    for leg, strategy in (("merged", 1), ("queue", 3), ("auto", 0)):
        if leg not in legs:
            continue
        b.clear()
        for g, sc, steps in sims:
            b.add(g, 0, steps, sc.mode)
        b.set_strategy(strategy)
        t0 = time.perf_counter()
        b.run()
        times[leg] = time.perf_counter() - t0
        if rec is not None:
            same = same and all(np.array_equal(g.probes(), r) for (g, _, _), r in zip(sims, rec))
        rewind(leg)
    t_bat = times.get("auto", next(iter(times.values())))
    grid_steps = sum(steps for _, steps in work)
    out = {
        "metric": "Gcell-updates/s",
        "value": round(total / t_bat / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": grid_steps,
        "warmup": 2,
        "ms_per_step": round(t_bat * 1e3 / max(s for _, s in work), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded grids, spheres and plane-wave angles)",
        "config": {
            "workload": f"{len(work)} independent grids (edges 32..{max(max(sc.I, sc.J, sc.K) for sc, _ in work)}, "
                        "300..1500 steps, 8-cell CPML, oblique TF/SF, 0..20 spheres), batched engine",
            "grids": len(work),
            "cell_updates": total,
            "legs": legs,
            "batched_s": round(t_bat, 3),
        },
    }
    for leg, t in times.items():
        out["config"][f"{leg}_value"] = round(total / t / 1e9, 3)
        out["config"][f"{leg}_s"] = round(t, 3)
    if "sequential" in times:
        out["config"]["speedup_vs_sequential"] = round(times["sequential"] / t_bat, 3)
        out["config"]["batched_equals_sequential"] = bool(same)
    out["roofline"] = run_roofline(sum(sc.cells * steps * bytes_per_cell(sc) for sc, steps in work), t_bat)
    b.close()
    for g, _, _ in sims:
        g.close()
    if not args.no_cpu:
        out["parity"] = oracle_check([(f"grid {q} ({val_leg} leg)", work[q][0], work[q][1], kept[q]) for q in sample])
    print(json.dumps(out), flush=True)


def run_serve(args):
    """Served workload: a Poisson stream of solve requests (edges 48..320, 500..3000 steps, spheres
    or layered ground, own CPML and plane wave each) through the request server, against the same
    requests run one after another through the plain path.  The arrival rate is sized from the
    plain path's rate to keep the GPU near saturation (0.9 of it).  Reports aggregate
    Gcell-updates/s, p50 / p99 time to completion, and whether every served probe record equals the
    plain run's.  Unless --no-cpu, the probe records of a sample of the requests (the least-work
    request of each class of largest edge and material kind, then the least-work others up to
    --oracle-cells cell-updates) from the leg that gives `value` are compared with the CPU oracle
    run in float32 and float64 (`parity`: per-request relative L2).

    --serve-legs picks the legs (sequential, open, closed); a run with more requests than one
    process should take at once runs the sequential leg alone with --serve-digests FILE (the plain
    rate, latencies and a digest of every probe record are written there), then the open and
    closed legs with the same FILE (the arrival rate comes from its plain rate, and each served
    record is compared with its digest)."""
    import hashlib

    import numpy as np

    from maxwell_solver_amd import Server, Simulation
    from maxwell_solver_amd import scenarios as S
    legs = set(args.serve_legs.split(","))
    reqs = S.config5(args.requests)
    total = sum(sc.cells * steps for sc, steps, _ in reqs)
    # the leg that gives `value`, and the requests whose probe records from it go to the CPU oracle
    val_leg = "open" if "open" in legs else "closed" if "closed" in legs else "sequential"
    # one request per class of largest edge (48, 80, 128, 200 and up) and material kind (spheres
    # with an oblique plane wave, mode 5; layered ground with a vertical pulse, mode 2), then the
    # least work
    sample = set(sample_by_work([sc.cells * steps for sc, steps, _ in reqs], args.oracle_cells,
                                [(size_class(sc, (0, 80, 128, 200)), sc.mode) for sc, _, _ in reqs]))
    kept = {}

    def digest(a):
        return hashlib.sha1(np.ascontiguousarray(a).tobytes()).hexdigest()

    # warm the kernels
    sc0, _, _ = reqs[0]
    g = Simulation()
    S.setup(g, sc0)
    g.run(0, 2, sc0.mode)
    g.synchronize()
    g.close()
    seq = None
    if "sequential" in legs:
        # plain path, one request after another (set-up on the GPU, steps, completion)
        rec, lat_seq, t_seq = [], [], 0.0
        for q, (sc, steps, _) in enumerate(reqs):
            t0 = time.perf_counter()
            g = Simulation()
            S.setup(g, sc)
            g.run(0, steps, sc.mode)
            g.synchronize()
            dt = time.perf_counter() - t0
            t_seq += dt
            lat_seq.append(t_seq)
            rec.append(digest(g.probes()))
            if val_leg == "sequential" and q in sample:
                kept[q] = g.probes().copy()
            g.close()
            if q % 100 == 99:
                print(f"sequential: {q + 1} requests, {t_seq:.1f} s", file=sys.stderr, flush=True)
        seq = {"requests": len(reqs), "t_seq": t_seq, "lat_seq": lat_seq, "digests": rec}
        if args.serve_digests:
            with open(args.serve_digests, "w") as f:
                json.dump(seq, f)
    elif args.serve_digests:
        with open(args.serve_digests) as f:
            seq = json.load(f)
        if seq["requests"] != len(reqs):
            raise SystemExit("--serve-digests holds another request count")
    else:
        raise SystemExit("the open / closed legs need the sequential leg or --serve-digests")
    t_seq = seq["t_seq"]
    rate_plain = total / t_seq
    mean_work = total / len(reqs)
    scale = mean_work / (0.9 * rate_plain)
    arrivals = np.cumsum([gap for _, _, gap in reqs]) * scale
    arrivals -= arrivals[0]

    def serve(arr, label):
        srv = Server(hbm_budget=int(args.hbm_budget_gb * (1 << 30)), max_active=args.max_active,
                     steps_per_round=args.steps_per_round)
        sims = []
        for sc, steps, _ in reqs:  # host-side set-up, as the clients that submit them
            sim = Simulation()
            S.setup_host(sim, sc)
            sims.append(sim)
        ids = {}
        lat, same = [], True
        nxt = 0
        t_start = srv.now()
        t_end = t_start
        while len(lat) < len(reqs):
            now = srv.now() - t_start
            while nxt < len(reqs) and arr[nxt] <= now:
                sc, steps, _ = reqs[nxt]
                ids[srv.submit(sims[nxt], steps, sc.mode, source=sc.source, t_arrive=t_start + arr[nxt])] = nxt
                nxt += 1
            active, queued = srv.round()
            for rid, ta, tadm, tdone in srv.collect():
                q = ids[rid]
                pr = srv.simulation(rid).probes()
                same = same and digest(pr) == seq["digests"][q]
                if label == val_leg and q in sample:
                    kept[q] = pr.copy()
                lat.append(tdone - ta)
                t_end = max(t_end, tdone)
                srv.release(rid)
                if len(lat) % 100 == 0:
                    print(f"{label}: {len(lat)} requests done, {t_end - t_start:.1f} s", file=sys.stderr, flush=True)
            if not active and not queued and nxt < len(reqs):
                time.sleep(max(0.0, min(0.05, arr[nxt] - (srv.now() - t_start))))
        peak = srv.stats()[1]
        srv.close()
        return t_end - t_start, np.asarray(lat), same, peak

    cfg = {
        "workload": f"{len(reqs)} solve requests (edges 48..320, 500..3000 steps, spheres or layered ground), "
                    "Poisson arrivals at 0.9 of the plain path's rate, request server with continuous batching",
        "requests": len(reqs),
        "cell_updates": total,
        "sequential_value": round(rate_plain / 1e9, 3),
        "sequential_s": round(t_seq, 3),
        "sequential_p50_s": round(float(np.percentile(seq["lat_seq"], 50)), 3),
        "sequential_p99_s": round(float(np.percentile(seq["lat_seq"], 99)), 3),
        "max_active": args.max_active,
        "steps_per_round": args.steps_per_round,
        "legs": sorted(legs),
    }
    value = None
    same = True
    if "open" in legs:
        span_open, lat_open, same_open, peak_open = serve(arrivals, "open")
        value = total / span_open / 1e9
        cfg.update({"p50_s": round(float(np.percentile(lat_open, 50)), 3),
                    "p99_s": round(float(np.percentile(lat_open, 99)), 3),
                    "span_s": round(span_open, 3), "peak_hbm_gb": round(peak_open / (1 << 30), 2)})
        same = same and same_open
    if "closed" in legs:
        span_closed, lat_closed, same_closed, _ = serve(np.zeros(len(reqs)), "closed")
        cfg.update({"closed_value": round(total / span_closed / 1e9, 3),
                    "closed_p50_s": round(float(np.percentile(lat_closed, 50)), 3),
                    "closed_p99_s": round(float(np.percentile(lat_closed, 99)), 3),
                    "closed_s": round(span_closed, 3),
                    "closed_speedup_vs_sequential": round(t_seq / span_closed, 3)})
        same = same and same_closed
        if value is None:
            value = total / span_closed / 1e9
    if value is None:
        value = rate_plain / 1e9
    cfg["served_equals_sequential"] = bool(same) if ("open" in legs or "closed" in legs) else None
    parity = None
    if not args.no_cpu and kept:  # the sample's records from the leg that gives `value`
        parity = oracle_check([(f"request {q} ({val_leg} leg)", reqs[q][0], reqs[q][1], kept[q]) for q in sorted(kept)])
    out = {
        "metric": "Gcell-updates/s",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": int(sum(steps for _, steps, _ in reqs)),
        "warmup": 2,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded request stream: sizes, materials, plane waves, Poisson arrivals)",
        "config": cfg,
    }
    t_val = span_open if "open" in legs else span_closed if "closed" in legs else t_seq
    out["roofline"] = run_roofline(sum(sc.cells * steps * bytes_per_cell(sc) for sc, steps, _ in reqs), t_val)
    if parity is not None:
        out["parity"] = parity
    print(json.dumps(out), flush=True)


def host_cpu():
    """CPU model, physical core count and the logical CPUs this process may use (/proc/cpuinfo)."""
    model, cores, phys, core = None, set(), None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in list(f) + [""]:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not line.strip():
                    if core is not None:
                        cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model or "unknown", len(cores) or None, usable


def oracle_threads():
    """The threads the CPU oracle may use here: the usable logical CPUs, at most OMP_NUM_THREADS."""
    usable = host_cpu()[2]
    omp = os.environ.get("OMP_NUM_THREADS")
    return min(usable, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else usable


def _diff(a, b):
    """(relative L2, max-abs) of a against b, in float64."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    nb = float(np.linalg.norm(b))
    d = float(np.linalg.norm(a - b))
    return (d / nb if nb > 0 else d), (float(np.max(np.abs(a - b))) if a.size else 0.0)


def gpu_state(sc, steps, init_line=True):
    """Probe record and the six field components of `sc` after `steps` steps from step 0 on the
    GPU (a fresh simulation, no snapshots)."""
    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S
    g = Simulation()
    S.setup(g, sc, init_line=init_line)
    g.run(0, steps, sc.mode)
    g.synchronize()
    out = g.probes().copy(), {c: g.field(c) for c in FIELDS}
    g.close()
    return out


def state_diff(gpu, o):
    """Probe and field differences of a GPU state (gpu_state) against an oracle simulation."""
    import numpy as np
    probes, fields = gpu
    op = o.probes() if np.asarray(probes).size else np.zeros(0)  # no antennas: no record
    pr, pa = _diff(probes, op)
    of = {c: o.field(c) for c in FIELDS}
    fr = {c: _diff(fields[c], of[c]) for c in FIELDS}
    # RMS of the oracle's state: the magnitude the differences are relative to.  A component whose
    # RMS is below F64_SIGNAL_FLOOR is zero in exact arithmetic here (e.g. a component the
    # polarisation never excites) and holds rounding residue only: it is listed, and left out of
    # the per-component maximum; the whole-state figure covers all six components.
    ss = {c: float(np.sum(np.square(of[c], dtype=np.float64))) for c in FIELDS}
    n = sum(of[c].size for c in FIELDS)
    dd = sum(float(np.sum(np.square(fields[c].astype(np.float64) - of[c]))) for c in FIELDS)
    tot = sum(ss.values())
    quiet = [c for c in FIELDS if np.sqrt(ss[c] / max(of[c].size, 1)) < F64_SIGNAL_FLOOR]
    rms_p = float(np.sqrt(np.mean(np.square(op, dtype=np.float64)))) if op.size else 0.0
    return {"probe_rel_l2": pr, "probe_max_abs": pa, "probe_rms": rms_p,
            "field_rel_l2": float(np.sqrt(dd / tot)) if tot > 0 else float(np.sqrt(dd)),
            "field_rel_l2_max": max((fr[c][0] for c in FIELDS if c not in quiet), default=0.0),
            "field_max_abs": max(a for _, a in fr.values()),
            "field_rms": float(np.sqrt(tot / n)), "field_components_below_floor": quiet}


def currents_parity(sc, steps):
    """A workload with a line current after every step and no plane wave (GIC, graded mesh): the
    GPU's fused run against the oracle's reference op sequence (UpdateH_All, UpdateH_PML,
    UpdateE_All, UpdateE_PML, AddCurrents) over the same steps from step 0, in float32 and in
    float64."""
    import numpy as np

    from oracle import oracle as orc
    from maxwell_solver_amd import scenarios as S
    threads = oracle_threads()
    orc.set_threads(threads)
    gpu = gpu_state(sc, steps, init_line=False)
    pol, box = sc.currents
    out = {"steps": steps, "oracle": "CPU oracle, reference op sequence with AddCurrents after every step",
           "oracle_threads": threads}
    t0 = time.perf_counter()
    for dt, key in ((np.float32, ""), (np.float64, "_f64")):
        o = orc.OracleSim(dt)
        S.setup(o, sc, init_line=False)
        for n in range(steps):
            o.update_h_all()
            o.update_h_pml()
            o.update_e_all()
            o.update_e_pml()
            o.add_currents(n, pol, list(box))
        out.update({k + key: v for k, v in state_diff(gpu, o).items()})
        o.close()
    out["seconds"] = round(time.perf_counter() - t0, 2)
    out["tol"] = REL_L2_TOL
    out["within_tol"] = bool(out["probe_rel_l2"] <= REL_L2_TOL and out["field_rel_l2"] <= REL_L2_TOL
                             and out["field_rel_l2_max"] <= REL_L2_TOL)
    return out


def sample_by_work(works, budget_cells, groups=None):
    """Indices of a sample of the cases: with `groups` (a class label per case) first the case with
    the least work (cell-updates) of every class, whatever its work; then the remaining cases with
    the least work first, as many as fit in budget_cells in all; at least one case."""
    pick, tot = [], 0
    if groups is not None:
        for g in sorted(set(groups), key=str):
            q = min((q for q in range(len(works)) if groups[q] == g), key=lambda q: works[q])
            pick.append(q)
            tot += works[q]
    for q in sorted(range(len(works)), key=lambda q: works[q]):
        if q in pick:
            continue
        if pick and tot + works[q] > budget_cells:
            break
        pick.append(q)
        tot += works[q]
    return sorted(pick, key=lambda q: works[q])


def size_class(sc, bounds):
    """Index of the class [bounds[c], bounds[c+1]) that holds the grid's largest edge."""
    e = max(sc.I, sc.J, sc.K)
    return max(c for c in range(len(bounds)) if bounds[c] <= e)


def oracle_check(cases):
    """Probe records of GPU runs against the CPU oracle run on this host from step 0, in float32
    (the parity bar: relative L2 at most 1e-5) and in float64 (reported beside it; the maximum over
    the cases whose float64 record is above F64_SIGNAL_FLOOR in RMS).  `cases` holds (label,
    scenario, steps, gpu probe record) tuples."""
    import numpy as np

    from oracle import oracle as orc
    from maxwell_solver_amd import scenarios as S
    threads = oracle_threads()
    orc.set_threads(threads)
    t0 = time.perf_counter()
    rows = []
    for label, sc, steps, rec in cases:
        row = {"case": label, "grid": [sc.I, sc.J, sc.K], "steps": steps}
        for dt, key in ((np.float32, "f32"), (np.float64, "f64")):
            o = orc.OracleSim(dt)
            S.setup(o, sc)
            o.run(0, steps, sc.mode)
            pr = o.probes()
            r, a = _diff(rec, pr)
            o.close()
            row[f"probe_rel_l2_{key}"] = r
            row[f"probe_max_abs_{key}"] = a
            if key == "f64":
                row["probe_rms_f64"] = float(np.sqrt(np.mean(np.square(pr, dtype=np.float64)))) if pr.size else 0.0
        rows.append(row)
    worst = max(r["probe_rel_l2_f32"] for r in rows)
    sig = [r for r in rows if r["probe_rms_f64"] >= F64_SIGNAL_FLOOR]
    return {
        "oracle": "CPU oracle on this host, float32 and float64, same scenario from step 0",
        "oracle_threads": threads,
        "cases": len(rows),
        "cell_updates": int(sum(sc.cells * steps for _, sc, steps, _ in cases)),
        "probe_rel_l2_f32_max": worst,
        "probe_rel_l2_f64_max": max((r["probe_rel_l2_f64"] for r in sig), default=None),
        "probe_max_abs_f64_max": max(r["probe_max_abs_f64"] for r in rows),
        "f64_signal_floor": F64_SIGNAL_FLOOR,
        "cases_below_f64_floor": [r["case"] for r in rows if r["probe_rms_f64"] < F64_SIGNAL_FLOOR],
        "tol": REL_L2_TOL,
        "within_tol": bool(worst <= REL_L2_TOL),
        "seconds": round(time.perf_counter() - t0, 2),
        "per_case": rows,
    }


def cpu_baseline(sc, budget_s=15.0, parity=True):
    """The CPU oracle (float32, OpenMP over k planes) on the same workload, on this host's cores:
    built for this CPU (-march=native) when g++ can, timed with 1 thread and with all the threads
    the process may use (at most OMP_NUM_THREADS when set), each the median of 3 timed samples of a
    few steps; about budget_s seconds in all.

    The all-thread samples run contiguously from step 0, so with `parity` the oracle's state after
    them is compared with the GPU's after the same steps (probe record and fields), and a float64
    oracle run over the same steps (about budget_s more seconds) likewise; returned as the second
    value (None without `parity`)."""
    import statistics
    import tempfile

    import numpy as np

    from oracle import oracle as orc
    from maxwell_solver_amd import scenarios as S
    model, phys, usable = host_cpu()
    build = "-O3 -march=native -fopenmp -ffp-contract=off (built on this host)"
    tmp = tempfile.mkdtemp(prefix="oracle_native_")
    try:
        orc.use_library(orc.build_native(tmp))
    except Exception as e:  # no compiler or no -march=native: the in-tree x86-64-v3 build
        build = f"-O3 -march=x86-64-v3 -fopenmp -ffp-contract=off (in-tree build; native build failed: {type(e).__name__})"
    threads = oracle_threads()
    o = orc.OracleSim(np.float32)
    S.setup(o, sc)
    n = 0

    def rate(nthreads, steps):
        nonlocal n
        orc.set_threads(nthreads)
        if n + steps > sc.T:  # the source has T samples: reuse the time indices (timing only)
            n = 0
        t0 = time.perf_counter()
        o.run(n, steps, sc.mode)
        dt = time.perf_counter() - t0
        n += steps
        return sc.cells * steps / dt / 1e9, dt

    per = budget_s / 6.0  # three samples each of two thread counts
    out, par = {}, None
    for label, nt in (("all", threads), ("one", 1)):
        r0, dt0 = rate(nt, 1)  # calibration step (also warms the caches)
        # the all-thread samples stay within the source's T samples: steps 0 .. n-1 in order
        steps = max(1, min((sc.T - 1) // 3, int(per / dt0)))
        runs = [rate(nt, steps)[0] for _ in range(3)]
        out[label] = (statistics.median(runs), runs, steps)
        if label == "all" and parity:
            par = {"steps": n, "oracle": "float32 CPU oracle (the all-thread samples, steps 0 .. n-1 in order)"}
            par.update(state_diff(gpu_state(sc, n), o))
            par["tol"] = REL_L2_TOL
            par["within_tol"] = bool(par["probe_rel_l2"] <= REL_L2_TOL and par["field_rel_l2"] <= REL_L2_TOL
                                     and par["field_rel_l2_max"] <= REL_L2_TOL)
    o.close()
    if par is not None:
        # float64 oracle over the same steps (about half the float32 rate: ~budget_s seconds)
        n64 = par["steps"]
        orc.set_threads(threads)
        o64 = orc.OracleSim(np.float64)
        S.setup(o64, sc)
        t0 = time.perf_counter()
        o64.run(0, n64, sc.mode)
        f64 = state_diff(gpu_state(sc, n64), o64)
        o64.close()
        par.update({"f64_steps": n64, "f64_seconds": round(time.perf_counter() - t0, 2)})
        par.update({f"{k}_f64": v for k, v in f64.items()})
        # relative figures against a float64 state that holds rounding residue only measure nothing
        par["f64_below_floor"] = {"probes": f64["probe_rms"] < F64_SIGNAL_FLOOR,
                                  "fields": f64["field_rms"] < F64_SIGNAL_FLOOR}
    base = {
        "value": round(out["all"][0], 4),
        "unit": "Gcell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sc.I}x{sc.J}x{sc.K} workload, median of 3 samples of {out['all'][2]} steps with {threads} "
                  f"OpenMP threads and of {out['one'][2]} steps with 1 thread (float32)",
        "one_core_value": round(out["one"][0], 4),
        "all_core_runs": [round(x, 4) for x in out["all"][1]],
        "one_core_runs": [round(x, 4) for x in out["one"][1]],
        "cpu_model": model,
        "physical_cores": phys,
        "logical_cpus_usable": usable,
        "build": build,
    }
    return base, par


def run_gic(args):
    """GICsimulation.m workload (250 x 280 x 50 cells, dx = 10 km, layered and graded ground
    conductivity, an x-directed 100 kA line current added after every step, no plane wave): all
    --gic-steps steps of the run with graph replay (--gic-graph steps per recorded graph), and a
    sample of --gic-eager steps launched one by one for comparison.  Unless --no-cpu, the first
    --parity-steps steps are compared with the CPU oracle (`parity`, float32 and float64)."""
    import numpy as np

    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S
    steps, eager = args.gic_steps, args.gic_eager
    sc = S.config_gic(T=args.warmup + steps + eager)
    g = Simulation()
    S.setup(g, sc, init_line=False)
    g.set_graph_steps(args.gic_graph)
    g.run(0, args.warmup, 0)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(args.warmup, steps, 0)
    g.synchronize()
    dt = time.perf_counter() - t0
    g.set_graph_steps(0)
    t0 = time.perf_counter()
    g.run(args.warmup + steps, eager, 0)
    g.synchronize()
    dt_eager = time.perf_counter() - t0
    ez = g.field("Ex")
    finite = bool(np.isfinite(ez).all())
    g.close()
    out = {
        "metric": "Gcell-updates/s",
        "value": round(sc.cells * steps / dt / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (the GICsimulation.m geometry, conductivities and source)",
        "config": {
            "workload": f"GIC {sc.I}x{sc.J}x{sc.K}, dx 10 km, 10-cell CPML, {len(sc.mat_table)} conductivity "
                        "rows (uint8 index), line current (AddCurrents Ex) after every step, no plane wave",
            "grid": [sc.I, sc.J, sc.K],
            "cells": sc.cells,
            "graph_steps": args.gic_graph,
            "eager_steps": eager,
            "eager_value": round(sc.cells * eager / dt_eager / 1e9, 3),
            "eager_ms_per_step": round(dt_eager * 1e3 / eager, 5),
            "graph_vs_eager": round((dt_eager / eager) / (dt / steps), 3),
            "fields_finite": finite,
        },
    }
    if not args.no_cpu:
        out["parity"] = currents_parity(sc, min(args.parity_steps, sc.T))
    print(json.dumps(out), flush=True)


def run_graded(args):
    """The 256^3 sphere grid on a graded mesh (spacings from dx to 1.3 dx, dx in the CPML) with an
    x-directed line current added after every step (a graded mesh takes no plane wave): --steps
    timed steps after --warmup, the same run on the uniform grid beside it.  The row kernel's
    interior part runs the full slab march on a graded mesh (every row and plane applies the
    per-axis spacing factors).  Unless --no-cpu, the first --parity-steps steps of the graded run
    are compared with the CPU oracle (`parity`, float32 and float64)."""
    import numpy as np

    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S
    T = args.warmup + args.steps
    legs = {}
    for graded in (True, False):
        sc = S.config_graded(T=T, N=args.size, graded=graded)
        g = Simulation()
        S.setup(g, sc, init_line=False)
        g.run(0, args.warmup, 0)
        g.synchronize()
        g.set_timing(True, every=TIMING_EVERY)
        t0 = time.perf_counter()
        g.run(args.warmup, args.steps, 0)
        g.synchronize()
        dt = time.perf_counter() - t0
        step_ms, _ = g.kernel_time()
        rows_ms = g.row_kernel_time()
        finite = bool(np.isfinite(g.field("Ez")).all())
        g.close()
        rows_bytes = row_kernel_bytes(sc)
        legs[graded] = {"value": round(sc.cells * args.steps / dt / 1e9, 3), "ms_per_step": round(dt * 1e3 / args.steps, 4),
                        "interior_kernel_ms": round(rows_ms, 5), "step_kernels_ms": round(step_ms, 5),
                        "interior_achieved_gbs": round(rows_bytes / (rows_ms * 1e-3) / 1e9, 1) if rows_ms > 0 else None,
                        "fields_finite": finite}
        cells = sc.cells
    gr, un = legs[True], legs[False]
    out = {
        "metric": "Gcell-updates/s",
        "value": gr["value"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gr["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded spheres of the 256^3 workload, a smooth graded mesh, a line-current pulse)",
        "config": {
            "workload": f"{args.size}^3 FP32 Yee FDTD on a graded mesh (spacings from dx to 1.3 dx, dx in the 10-cell "
                        "CPML), 40 dielectric spheres (uint8 index), x-directed line current (AddCurrents Ex) after "
                        "every step, 64 probes Ez/Hy",
            "cells": cells,
            "graded": gr,
            "uniform": un,
            "graded_vs_uniform": round(gr["value"] / un["value"], 3),
        },
    }
    if not args.no_cpu:
        out["parity"] = currents_parity(S.config_graded(T=T, N=args.size, graded=True), min(args.parity_steps, T))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--rows-per-group", type=int, default=0)
    ap.add_argument("--cols-per-lane", type=int, default=0)
    ap.add_argument("--z-chunk", type=int, default=0)
    ap.add_argument("--slab-chunk", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--parity-steps", type=int, default=200,
                    help="gic / graded: steps of the GPU run compared with the CPU oracle (float32 and float64)")
    ap.add_argument("--oracle-cells", type=float, default=4e9,
                    help="batch / serve: the probe records of the least-work grid or request of each size "
                         "class (and material kind), then of the least-work others up to this many cell-updates "
                         "in all, are checked against the CPU oracle (float32 and float64)")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--workload", choices=["grid", "layered", "batch", "serve", "gic", "graded"], default="grid",
                    help="grid: the 256^3 workload (default); layered: the 384x384x256 lossy layered-ground "
                         "workload (4000 steps with --steps 4000); batch: 96 independent grids, batched engine; "
                         "serve: a stream of solve requests through the request server; gic: the "
                         "GICsimulation.m run (line current every step, graph replay); graded: the 256^3 "
                         "sphere grid on a graded mesh with a line current, beside the same run uniform")
    ap.add_argument("--gic-steps", type=int, default=300000)
    ap.add_argument("--gic-eager", type=int, default=20000)
    ap.add_argument("--gic-graph", type=int, default=64)
    ap.add_argument("--requests", type=int, default=32)
    ap.add_argument("--serve-legs", default="sequential,open,closed")
    ap.add_argument("--serve-digests", default=None)
    # served requests advanced together: more slots raise the closed rate and the p99 time to
    # completion together (2000 requests: 8 slots closed 52.8 Gcell/s, open p99 8.15 s; 32 slots
    # 54.4, 10.97 s); the default keeps the latency
    ap.add_argument("--max-active", type=int, default=8)
    ap.add_argument("--steps-per-round", type=int, default=64)
    ap.add_argument("--hbm-budget-gb", type=float, default=288.0)
    ap.add_argument("--grids", type=int, default=96)
    ap.add_argument("--min-edge", type=float, default=32)
    ap.add_argument("--max-edge", type=float, default=192)
    ap.add_argument("--batch-legs", default="sequential,streams,merged,auto",
                    help="legs of the batched workload, in order: sequential, streams, merged, queue, auto "
                         "(value = auto, else the first leg run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU z-slab "
              "decomposition is a design, not a build, in this version.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload == "gic":
        run_gic(args)
        return
    if args.workload == "graded":
        run_graded(args)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "serve":
        run_serve(args)
        return

    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S

    T = args.warmup + args.steps
    if args.workload == "layered":
        sc = S.config3(T=max(T, 1))
    else:
        sc = S.config2(T=max(T, 1))
        if args.size != 256:
            sc = S.small_variant(sc, args.size, T)
            sc.snap_every = 250
    g = Simulation()
    S.setup(g, sc)
    g.set_tuning(args.rows_per_group, args.z_chunk, args.cols_per_lane)
    if args.slab_chunk:
        g.set_slab_chunk(args.slab_chunk)

    # warmup (untimed), then exactly --steps timed steps
    if args.warmup:
        g.run(0, args.warmup, sc.mode)
    g.synchronize()
    g.set_timing(True, every=TIMING_EVERY)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(args.warmup, args.steps, sc.mode, snap_every=sc.snap_every, snap_comp=sc.snap_comp)
    g.synchronize()
    dt = time.perf_counter() - t0
    step_ms, launches = g.kernel_time()
    rows_ms = g.row_kernel_time()
    # the stream ceiling of this grid: 49 bytes per cell (six components read and written, the
    # material byte) in one pass with no stencil, the step's layout and store policy
    ceil_ms = g.stream_ceiling(20)
    ceil_gbs = 49.0 * sc.cells / (ceil_ms * 1e-3) / 1e9

    cell_updates = sc.cells * args.steps
    value = cell_updates / dt / 1e9
    bpc = bytes_per_cell(sc)
    rows_bytes = row_kernel_bytes(sc)
    achieved = rows_bytes / (rows_ms * 1e-3) / 1e9 if rows_ms > 0 else 0.0
    step_achieved = bpc * sc.cells / (step_ms * 1e-3) / 1e9 if step_ms > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_k_rows_c1_{sc.I}.json")  # rocprofv3 FETCH/WRITE, copy-calibrated
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
        except Exception:
            traffic = None
    rec = {
        "metric": "Gcell-updates/s",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic (seeded layer thicknesses, conductivities, permittivities and pulse frequency)"
                 if args.workload == "layered" else "synthetic (seeded spheres, seeded plane-wave angles)"),
        "config": {
            "workload": (f"{sc.I}x{sc.J}x{sc.K} FP32 Yee FDTD, lossy layered ground (5 layers, uint8 index), "
                         f"12-cell CPML, vertically incident x-polarised pulse (layered 1D line), "
                         f"{len(sc.probes)} surface probes Ex/Hy" if args.workload == "layered" else
                         f"{sc.I}^3 FP32 Yee FDTD, 10-cell CPML, 40 dielectric spheres (uint8 index), oblique TF/SF "
                         f"plane wave, 64 probes Ez/Hy, Ez snapshot every {sc.snap_every} steps"),
            "grid": [sc.I, sc.J, sc.K],
            "cells": sc.cells,
            "time_steps": args.steps,
            "kernels": "two launches per step: k_rows_c1 (row kernel, interior part), k_rows<..., 1> (row "
                       "kernel, slab part: y/z CPML rows and planes, plus the next step's TF/SF face sheets "
                       "and the probe gather as tail workgroups)",
        },
        "roofline": {
            "kernel": "k_rows_c1 (row kernel, interior part)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": int(rows_bytes),
            "kernel_ms_avg": round(rows_ms, 5),
            "kernel_launches": launches,
            "timed_every": TIMING_EVERY,
            "step_ms_avg": round(step_ms, 5),
            "step_bytes_per_cell": round(bpc, 2),
            "step_achieved": round(step_achieved, 1),
            "step_frac": round(step_achieved / HBM_PEAK_GBS, 4),
            "stream_ceiling": {"kernel": "k_stream (49 B per cell: 6 components + material byte read, "
                                         "6 components written, one pass, no stencil)",
                               "ms": round(ceil_ms, 5), "achieved": round(ceil_gbs, 1),
                               "frac": round(ceil_gbs / HBM_PEAK_GBS, 4),
                               "kernel_frac_of_ceiling": round(achieved / ceil_gbs, 4) if ceil_gbs > 0 else None},
        },
    }
    g.close()
    if not args.no_cpu:
        base = (S.config3(T=600) if args.workload == "layered" else
                S.config2(T=2000) if args.size == 256 else sc)
        rec["cpu_baseline"], rec["parity"] = cpu_baseline(base, budget_s=args.cpu_budget)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
