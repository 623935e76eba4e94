#!/usr/bin/env python3
"""Throughput of pas-cg's per-step loop on one MI355X.

A step is one time step of every system of the batch: NL verify (and
rebuild when the displacement check fails), per-term list updates, all force
terms, the QA contact bookkeeping and the Langevin Gear-5 predictor-corrector
(cg/src/cg/simul/thread_main.cpp:373-536). The timed workload is the ragged
batch: 256 distinct synthetic proteins with N ~ U[30, 600] on Go-annealed
natives, 10 % of them pulled, full force field, forces in FP32 and
integration in FP64, advanced by the persistent workgroup-per-system kernel.
Exactly --steps steps are timed after exactly --warmup steps. Beside it, on
the same workload:

* the whole configured run of the ragged batch (5e4 steps) when --steps is
  shorter;
* the plain single-protein path (one protein at a time, hipGraph-captured
  stage kernels) on a systematic sample of the proteins;
* the max relative force error per term (QA terms included) of the mixed-
  precision kernels against the FP64 GPU path on the same positions;
* a served Poisson stream (continuous batching, HBM-budget admission) at
  three arrival rates, with p50/p99 job latency at each;
* the FP64 CPU restatement of the loop on bounded samples (1 core, and one
  process per host core).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import pcg_import  # noqa: E402

pcg_import.load()
from pas_cg_amd import model as M  # noqa: E402
from pas_cg_amd import workloads as W  # noqa: E402

METRIC = "residue-steps/s batched vs plain @force err≤1e-5 vs FP64; p50/p99 job latency"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(HERE, "profiles", "k_persistent_ragged256_traffic.txt")
COUNTERS_FILE = os.path.join(HERE, "profiles", "k_persistent_ragged256_counters.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(spl: int):
    """HBM bytes per launch of `spl` steps from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, copy-kernel calibrated), with the
    configuration it was measured at; (None, None) without it."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, None
    vals = {}
    for ln in open(TRAFFIC_FILE):
        if ":" in ln and not ln.startswith("#"):
            k, v = ln.split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        cfg = {"pad_A": float(vals["pad_A"]), "block": int(vals["block"]),
               "steps_per_launch": int(vals["steps_per_launch"]),
               "over_algorithmic": float(vals["traffic_over_algorithmic"])}
        return float(vals["hbm_bytes_per_step"]) * spl, cfg
    except (KeyError, ValueError):
        return None, None


def read_valu_active():
    """VALU-active fraction of wave cycles of the persistent kernel, from the
    committed rocprofv3 counter summary (None without it)."""
    if not os.path.exists(COUNTERS_FILE):
        return None
    for ln in open(COUNTERS_FILE):
        if ln.startswith("* VALU active per wave cycle"):
            try:
                return float(ln.rsplit(":", 1)[1])
            except ValueError:
                return None
    return None


def batched(p, models, steps, warmup, spl, block, label="batched"):
    """Exactly `warmup` untimed steps, then exactly `steps` timed steps (whole
    launches of `spl` steps and a shorter last one), wall clock around the
    timed run (synchronised on both sides) and HIP events on the batch's
    stream around its launches."""
    from pas_cg_amd import pcg as G
    n_res = sum(int(m["n"]) for m in models)
    b = G.Batch(p, models, precision=G.MIXED, philox_seed=448, block=block,
                steps_per_launch=max(1, min(spl, steps)))
    info = b.info()
    spl = info["steps_per_launch"]
    log(f"{label}: {len(models)} systems, {n_res} residues, block {info['block']}, "
        f"{spl} steps/launch, {info['bytes'] / 2**20:.1f} MiB; warmup {warmup}, timed {steps}")
    if warmup > 0:
        b.run(warmup)
    b.sync()
    reb0 = [b.sizes(s)["rebuilds"] for s in range(len(models))]
    t0 = time.perf_counter()
    b.run(steps)
    b.sync()
    wall = time.perf_counter() - t0
    ms, launches = b.last_ms()
    sizes = [b.sizes(s) for s in range(len(models))]
    rebuilds = [sz["rebuilds"] - r0 for sz, r0 in zip(sizes, reb0)]
    step_bytes = sum(W.step_bytes(p, sz, rb / steps) for sz, rb in zip(sizes, rebuilds))
    Vs = [b.energy(s)[0] for s in range(len(models))]
    if not all(math.isfinite(v) for v in Vs):
        raise RuntimeError(f"non-finite energy in the {label} run")
    log(f"{label}: {steps} steps in {wall:.3f} s wall, {ms:.1f} ms in {launches} launches, "
        f"{sum(rebuilds)} NL rebuilds")
    return {"b": b, "steps": steps, "warmup": warmup, "wall": wall, "ms": ms,
            "launches": launches, "spl": spl, "block": info["block"], "n_res": n_res,
            "step_bytes": step_bytes, "sizes": sizes, "rebuilds": rebuilds,
            "nl_pairs": sum(sz["nl"] for sz in sizes)}


def force_error(p, models, bat, picks, qa_rounds=3):
    """Mixed-precision per-term force and energy error against the FP64 GPU
    path, on the batch's current positions. The QA terms are compared on
    identical QA states: after the list rebuild, `qa_rounds` rounds of
    prepare_nh, the candidates (repulsion and the pairs that qualify), the QA
    finish (every decision exact in both precisions) and the contacts with
    their saturations."""
    from pas_cg_amd import pcg as G
    worst_f, worst_v = {}, {}

    def note(t, F32, F64, V32, V64):
        scale = np.abs(F64).max()
        if scale == 0:
            return
        worst_f[t] = max(worst_f.get(t, 0.0), float(np.abs(F32 - F64).max() / scale))
        if V64 != 0:
            worst_v[t] = max(worst_v.get(t, 0.0), abs(V32 - V64) / abs(V64))

    for s in picks:
        r = bat["b"].get(s, 0)
        ctx = {}
        for prec in (G.FP64, G.MIXED):
            c = G.Context(p, models[s], precision=prec, noise=G.PHILOX, seed=s)
            c.set(0, r)
            c.fix_nl()
            ctx[prec] = c
        for t in G.TERMS:
            if t.startswith("qa"):
                continue
            F64, V64 = ctx[G.FP64].eval_term(t)
            F32, V32 = ctx[G.MIXED].eval_term(t)
            note(t, F32, F64, V32, V64)
        if p["qa_on"]:
            for _ in range(qa_rounds):
                for c in ctx.values():
                    c.prepare_nh()
                F64, V64 = ctx[G.FP64].eval_term("qa_cand")
                F32, V32 = ctx[G.MIXED].eval_term("qa_cand")
                note("qa_cand", F32, F64, V32, V64)
                for c in ctx.values():
                    c.qa_finish()
                F64, V64 = ctx[G.FP64].eval_term("qa_cont")
                F32, V32 = ctx[G.MIXED].eval_term("qa_cont")
                note("qa_cont", F32, F64, V32, V64)
        for c in ctx.values():
            c.close()
    return worst_f, worst_v


def batched_force_error(p, models, bat, picks):
    """Total force of the batched kernel in mixed precision against the same
    kernel in FP64, on the timed batch's current positions (one step with
    the noise off evaluates the forces there; energies off, so that step
    takes the FP32 force paths of every other step)."""
    from pas_cg_amd import pcg as G
    sub = [models[s] for s in picks]
    pos = [bat["b"].get(s, 0) for s in picks]
    F = {}
    for prec in (G.FP64, G.MIXED):
        b = G.Batch(p, sub, precision=prec, philox_seed=448, block=bat["block"],
                    steps_per_launch=1)
        b.set_noise_scale(0.0)
        b.energies(False)
        for k, r in enumerate(pos):
            b.set(k, 0, r)
        b.run(1)
        F[prec] = [b.get(k, 6) for k in range(len(sub))]
        b.close()
    return max(float(np.abs(a - c).max() / np.abs(c).max()) for a, c in zip(F[G.MIXED], F[G.FP64]))


def plain(p, models, picks, steps=200, graph=True, form=None):
    """The drop-in path: one protein per plain context through pcg_step, in
    the context's own choice of form (the persistent single-system kernel for
    these proteins) or in the stage kernels (form G.FORM_STAGES)."""
    from pas_cg_amd import pcg as G
    rs, tt, per = 0, 0.0, []
    for s in picks:
        c = G.Context(p, models[s], precision=G.MIXED, noise=G.PHILOX, seed=s)
        c.use_graph(graph)
        c.step_form(G.FORM_AUTO if form is None else form)
        c.step(20)
        t0 = time.perf_counter()
        c.step(steps)
        dt = time.perf_counter() - t0
        n = int(models[s]["n"])
        rs += n * steps
        tt += dt
        per.append((n, dt / steps))
        c.close()
    log(f"plain: {len(picks)} proteins, {rs / tt:.3e} residue-steps/s")
    return rs / tt, per


def other_configs(c2, c1, graph=True):
    """The other configurations, briefly: 1024 replicas of the 56-residue
    chain in one batch, and the 6400-residue periodic box through the plain
    (stage-kernel) path."""
    from pas_cg_amd import pcg as G
    out = {}
    p, models = c2
    b = G.Batch(p, models, precision=G.MIXED, philox_seed=448, seeds=list(range(len(models))),
                steps_per_launch=500)
    block, _ = b.autotune(500)
    b.run(500)
    b.sync()
    t0 = time.perf_counter()
    b.run(2000)
    b.sync()
    dt = time.perf_counter() - t0
    n = sum(int(m["n"]) for m in models)
    out["replicas_1024x56"] = {"residue_steps_per_s": round(n * 2000 / dt, 1), "block": block,
                               "steps": 2000}
    b.close()
    p, m = c1
    c = G.Context(p, m, precision=G.MIXED, noise=G.PHILOX, seed=1)
    c.use_graph(graph)
    c.step(50)
    t0 = time.perf_counter()
    c.step(300)
    dt = time.perf_counter() - t0
    out["box_64x100_1mM_plain"] = {"residue_steps_per_s": round(m["n"] * 300 / dt, 1),
                                   "steps": 300, "nl_rebuilds": c.nl_rebuilds()}
    c.close()
    log(f"other: {out}")
    return out


def fifo_latency(jobs, per):
    """One-job-at-a-time plain GPU path as a FIFO queue, each job's service
    time modelled from the plain per-step times measured on the ragged sample
    (a line in N): a 1e6-step job alone takes minutes on that path, so the
    stream is not run on it."""
    ns = np.array([n for n, _ in per], float)
    ts = np.array([t for _, t in per], float)
    a, b0 = np.polyfit(ns, ts, 1)
    free = 0.0
    lat = []
    for j in jobs:
        start = max(free, j["arrival"])
        free = start + j["steps"] * max(a * j["n"] + b0, ts.min())
        lat.append(free - j["arrival"])
    return np.percentile(lat, 50), np.percentile(lat, 99)


# the served stream as configured: step counts log-uniform in [1e4, 1e6]
SERVE_STEPS = (10_000, 1_000_000)
SERVE_POOL = 600
SERVE_RATE_FRACS = (0.2, 0.4, 0.7)


def serve_pool(workers: int):
    """Jobs of the served stream at unit arrival rate, and their proteins, made
    before the GPU is touched (the generator forks worker processes); the
    natives are annealed on the GPU later (W.c4_finish)."""
    jobs = W.c4_jobs(SERVE_POOL, 1.0, *SERVE_STEPS)
    if workers > 1:
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(workers) as ex:
            models = list(ex.map(W.c4_model, jobs, chunksize=8))
    else:
        models = [W.c4_model(j) for j in jobs]
    return jobs, models


def mean_job_work(jobs):
    return float(np.mean([j["n"] * j["steps"] for j in jobs]))


def served_one(p, pool, rate, window_s, per_plain, chunk):
    """One Poisson stream at `rate` jobs/s: the jobs that arrive within
    `window_s` seconds (at most the pool), each run to its full step count;
    continuous batching with HBM-budget admission. Latency is submit to end
    from the server's own clock."""
    from pas_cg_amd import pcg as G
    jobs_all, models_all = pool
    count = max(8, min(len(jobs_all), int(rate * window_s)))
    jobs = [dict(j, arrival=j["arrival"] / rate) for j in jobs_all[:count]]
    models = models_all[:count]
    srv = G.Server(p, precision=G.MIXED, hbm_budget_bytes=64 << 30, max_resident=1024,
                   chunk_steps=chunk, philox_seed=448)
    t0 = time.perf_counter()
    ids = {}
    nxt = 0
    rounds = 0
    peak_resident = 0
    next_note = 10.0
    while True:
        now = time.perf_counter() - t0
        while nxt < count and jobs[nxt]["arrival"] <= now:
            j = jobs[nxt]
            ids[srv.submit(models[nxt], j["steps"], temperature=j["T"], seed=j["id"])] = j
            nxt += 1
        active = srv.round()
        rounds += 1
        peak_resident = max(peak_resident, srv.stats()["resident"])
        if now > next_note:
            log(f"served: {now:.0f} s, {nxt}/{count} submitted, {active} active")
            next_note += 10.0
        if nxt == count and active == 0:
            break
        if active == 0:
            time.sleep(max(0.0, jobs[nxt]["arrival"] - (time.perf_counter() - t0)))
    span = time.perf_counter() - t0
    lat, work, failed = [], 0, 0
    for jid, j in ids.items():
        st, done = srv.status(jid)
        if st != "done":
            failed += 1
            continue
        sub, _start, end = srv.times(jid)
        lat.append(end - sub)
        work += j["n"] * j["steps"]
    srv.close()
    p50, p99 = np.percentile(lat, 50), np.percentile(lat, 99)
    f50, f99 = fifo_latency(jobs, per_plain) if per_plain else (float("nan"), float("nan"))
    log(f"served: {count} jobs at {rate:.2f}/s, {work / span:.3e} residue-steps/s, "
        f"p50 {p50:.2f} s p99 {p99:.2f} s, {rounds} rounds, peak {peak_resident} resident")
    return {"jobs": count, "arrival_rate_per_s": round(rate, 3),
            "failed": failed, "span_s": round(span, 2), "rounds": rounds,
            "peak_resident": peak_resident,
            "residue_steps_per_s": work / span, "p50_latency_s": round(float(p50), 3),
            "p99_latency_s": round(float(p99), 3),
            "plain_fifo_model_p50_latency_s": None if math.isnan(f50) else round(float(f50), 1),
            "plain_fifo_model_p99_latency_s": None if math.isnan(f99) else round(float(f99), 1)}


def served_backlog(p, pool, jobs_n=256, steps=4000, chunk=200):
    """Capacity of the served form against the batched form on the same job
    set: `jobs_n` jobs of the stream (their lengths, temperatures, pulling),
    `steps` steps each, all submitted at once; served residue-steps/s over
    the time to drain them, against one batch of the same proteins run for
    the same steps."""
    from pas_cg_amd import pcg as G
    jobs, models = pool[0][:jobs_n], pool[1][:jobs_n]
    work = sum(j["n"] for j in jobs) * steps
    srv = G.Server(p, precision=G.MIXED, hbm_budget_bytes=64 << 30, max_resident=1024,
                   chunk_steps=chunk, philox_seed=448)
    t0 = time.perf_counter()
    ids = [srv.submit(m, steps, temperature=j["T"], seed=j["id"]) for j, m in zip(jobs, models)]
    while srv.round():
        pass
    span = time.perf_counter() - t0
    ok = sum(srv.status(i)[0] == "done" for i in ids)
    lat = [srv.times(i)[2] - srv.times(i)[0] for i in ids]
    srv.close()
    b = G.Batch(p, models, temps=[j["T"] for j in jobs], seeds=[j["id"] for j in jobs],
                precision=G.MIXED, philox_seed=448, steps_per_launch=500)
    b.run(1)
    b.sync()
    t0 = time.perf_counter()
    b.run(steps)
    b.sync()
    bspan = time.perf_counter() - t0
    b.close()
    out = {"jobs": jobs_n, "steps_each": steps, "done": ok,
           "served_residue_steps_per_s": work / span,
           "batched_residue_steps_per_s": work / bspan,
           "served_over_batched": round(bspan / span, 3),
           "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
           "p99_latency_s": round(float(np.percentile(lat, 99)), 3)}
    log(f"served backlog: {out}")
    return out


def served(p, capacity_rs, pool, per_plain, window_s, chunk):
    """The served stream as configured (lengths lognormal with median 150 and
    at most 2000 residues, steps log-uniform in [1e4, 1e6], 20 % pulled, T
    uniform in [0.3, 0.45]) at arrival rates of SERVE_RATE_FRACS of the
    batched capacity in jobs per second."""
    work = mean_job_work(pool[0])
    rates = [f * capacity_rs / work for f in SERVE_RATE_FRACS]
    out = {"steps_range": list(SERVE_STEPS), "median_n": 150, "max_n": 2000,
           "pulled_frac": 0.2, "arrival_window_s": window_s, "chunk_steps": chunk,
           "mean_job_residue_steps": round(work), "rates": []}
    for f, r in zip(SERVE_RATE_FRACS, rates):
        row = served_one(p, pool, r, window_s, per_plain, chunk)
        row["rate_frac_of_batched_capacity"] = f
        out["rates"].append(row)
    return out


def host_cpu():
    """The host's CPU model and the cores this process may run on."""
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "nproc": usable, "cpu_count": os.cpu_count()}


def cpu_pool(workers: int):
    """Worker processes for the all-cores CPU baseline, forked now, before the
    GPU is touched (each runs one oracle system later)."""
    from concurrent.futures import ProcessPoolExecutor
    ex = ProcessPoolExecutor(workers)
    list(ex.map(time.sleep, [0.3] * workers))  # start every worker now
    return ex


def _cpu_one(a):
    """One oracle system stepped for budget_s seconds (a worker of cpu_pool)."""
    p, m, seed, budget_s = a
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from oracle_ffi import Oracle
    o = Oracle(p, m, seed=seed)
    o.step(2)
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.step(20)
        steps += 20
    return int(m["n"]) * steps, time.perf_counter() - t0


def cpu_baseline(p, models, picks, budget_s=15.0, ex=None, picks_all=None, budget_all=10.0):
    """The FP64 CPU restatement of the loop (oracle/) on bounded samples of the
    same proteins: one core, and one process per core (every usable host core
    up to the pool's size) at once."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from oracle_ffi import Oracle
    ors = [Oracle(p, models[s], seed=s) for s in picks]
    for o in ors:
        o.step(2)
    steps, t0 = 0, time.perf_counter()
    while True:
        for o in ors:
            o.step(20)
        steps += 20
        dt = time.perf_counter() - t0
        if dt >= budget_s:
            break
    rs = sum(int(models[s]["n"]) for s in picks) * steps
    ns = [int(models[s]["n"]) for s in picks]
    log(f"cpu: {rs / dt:.3e} residue-steps/s over {dt:.1f} s")
    host = host_cpu()
    one = {"value": rs / dt, "unit": "residue-steps/s", "cores": 1, "kind": "port",
           "host": host,
           "sample": f"{len(picks)} proteins of the ragged workload (N={ns}), "
                     f"{steps} steps each after 2 warm-up steps, FP64, one thread"}
    if ex is None or not picks_all:
        return one
    k = ex._max_workers
    picks_all = picks_all[:k]  # one system per worker, all at once
    t0 = time.perf_counter()
    res = list(ex.map(_cpu_one, [(p, models[s], s, budget_all) for s in picks_all]))
    rs_all = sum(r for r, _ in res)
    dt_all = time.perf_counter() - t0
    log(f"cpu: {rs_all / dt_all:.3e} residue-steps/s on {k} cores over {dt_all:.1f} s")
    return {"value": rs_all / dt_all, "unit": "residue-steps/s", "cores": k, "kind": "port",
            "host": host,
            "sample": f"{len(picks_all)} proteins of the ragged workload (spread over the "
                      f"lengths, N={[int(models[s]['n']) for s in picks_all]}), one process per "
                      f"protein on {k} cores at once, {budget_all:.0f} s each after 2 warm-up "
                      f"steps, FP64; value = their residue-steps / wall time",
            "one_core": one}


C3_STEPS = 50_000  # the ragged workload's configured run


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=C3_STEPS,
                    help="steps timed (exactly); default the configured 5e4")
    ap.add_argument("--warmup", type=int, default=1000, help="untimed steps before (exactly)")
    ap.add_argument("--steps-per-launch", type=int, default=500)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--pad", type=float, default=2.0, help="NL pad (skin), A, with --no-tune")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (pad and block from the flags)")
    ap.add_argument("--serve-window", type=float, default=6.0,
                    help="seconds of arrivals of each served stream")
    ap.add_argument("--serve-chunk", type=int, default=200, help="steps per scheduling round")
    ap.add_argument("--no-full", action="store_true",
                    help="skip the configured 5e4-step run when --steps is shorter")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-stages", action="store_true",
                    help="plain contexts launch each step's stage kernels without a hipGraph "
                         "(for kernel tracers)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-backlog-only", action="store_true",
                    help="of the served legs only the saturated backlog (for counter profiles)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the 1024-replica and periodic-box configurations")
    ap.add_argument("--gen-workers", type=int, default=8,
                    help="processes generating the synthetic proteins (1: in-process)")
    ap.add_argument("--copy-only", type=int, default=0, metavar="BYTES",
                    help="run only the streaming-copy calibration kernel on BYTES")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} (world size {world}) is not "
              "supported: the workload's systems are independent and would shard by job, "
              "with no exchange during stepping", file=sys.stderr)
        sys.exit(2)
    if args.steps <= 0 or args.warmup < 0:
        print("--steps must be > 0 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)

    if args.copy_only:
        from pas_cg_amd import pcg as G
        gbs, ms = G.stream_copy(args.copy_only, 20)
        print(json.dumps({"kernel": "k_stream_copy", "bytes_read": args.copy_only,
                          "bytes_written": args.copy_only, "launches": 21,
                          "GBs": round(gbs, 1), "ms_per_copy": round(ms, 4)}), flush=True)
        return
    # every process is forked before the GPU is touched
    p, raw = W.c3(workers=args.gen_workers, pulled=False)
    extra_raw = None if args.no_other else (W.c0()[1], W.c1_chain())
    pool = None if args.no_serve else serve_pool(min(args.gen_workers, os.cpu_count() or 1))
    cpu_ex = None if args.no_cpu else cpu_pool(min(16, host_cpu()["nproc"] or 1))
    # Go-model annealing of every native on the GPU (input preparation)
    t_an = time.perf_counter()
    models = [M.add_pulling(m) if j % 10 == 0 else m
              for j, m in enumerate(W.anneal_natives(raw))]
    extra = None
    if extra_raw is not None:
        c0m, c1m = W.anneal_natives(list(extra_raw))
        extra = (W.c2(chain=c0m), W.c1(chain=c1m))
    if pool is not None:
        pool = (pool[0], W.c4_finish(*pool))
    log(f"annealed the natives in {time.perf_counter() - t_an:.1f} s")
    tuning = None
    if args.no_tune:
        pad, block = args.pad, args.block
    else:
        from pas_cg_amd import tune
        tuning = tune.tune(p, models, trial_steps=args.steps_per_launch, log=log)
        pad, block = tuning["pad_A"], tuning["block"]
    args.pad = pad
    p.update(nl_pad=pad * W.A)
    bat = batched(p, models, args.steps, args.warmup, args.steps_per_launch, block)
    value = bat["n_res"] * bat["steps"] / bat["wall"]
    avg_launch_s = bat["ms"] * 1e-3 / bat["launches"]
    bytes_launch = bat["step_bytes"] * bat["steps"] / bat["launches"]
    achieved = bytes_launch / avg_launch_s / 1e9
    from pas_cg_amd import pcg as G
    copy_gbs, _ = G.stream_copy(1 << 30, 20)
    order = np.argsort([int(m["n"]) for m in models])
    traffic, traffic_cfg = read_traffic(bat["spl"])
    nsys = len(models)
    out = {"metric": METRIC, "value": round(value, 1), "unit": "residue-steps/s",
           "n_gpus": 1, "steps": bat["steps"], "warmup": bat["warmup"],
           "ms_per_step": round(bat["wall"] * 1e3 / bat["steps"], 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 forces, f64 integration", "data": "synthetic",
           "config": {"workload": "ragged-256: 256 proteins, N~U[30,600] on Go-annealed natives, "
                                  "10% pulled, full force field, 1 persistent workgroup per "
                                  "protein",
                      "residues": bat["n_res"], "block": bat["block"],
                      "steps_per_launch": bat["spl"], "launches": bat["launches"],
                      "nl_pad_A": args.pad, "nl_pairs": bat["nl_pairs"],
                      "nl_rebuilds_in_timed_steps": int(sum(bat["rebuilds"])),
                      "nl_rebuilds_per_system_per_1000_steps":
                          round(1000.0 * sum(bat["rebuilds"]) / nsys / bat["steps"], 3),
                      "energies": "summed on the last step of each launch (every system's V "
                                  "once a launch); the other steps evaluate forces only",
                      "autotune": tuning and tuning["table"]},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": traffic,
                        "traffic_frac": (round(traffic / avg_launch_s / 1e9 / HBM_PEAK_GBS, 4)
                                         if traffic else None),
                        "traffic_measured_at": traffic_cfg,
                        "valu_active": read_valu_active(),
                        "limiter": "latency: one workgroup per protein, the largest protein's "
                                   "CU sets the launch time; neither HBM nor VALU is saturated",
                        "kernel": "k_persistent",
                        "stream_copy_GBs": round(copy_gbs, 1),
                        "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                        "algorithmic_bytes_per_launch": round(bytes_launch)}}
    bat["b"].close()
    if not args.no_full and args.steps != C3_STEPS:
        full = batched(p, models, C3_STEPS, 0, args.steps_per_launch, bat["block"],
                       label="full 5e4")
        out["c3_full_run"] = {
            "steps": C3_STEPS, "warmup": 0, "wall_s": round(full["wall"], 4),
            "residue_steps_per_s": round(full["n_res"] * C3_STEPS / full["wall"], 1),
            "nl_rebuilds": int(sum(full["rebuilds"])),
            "avg_launch_ms": round(full["ms"] / full["launches"], 4),
            "launches": full["launches"],
            "roofline_frac": round(full["step_bytes"] * C3_STEPS / (full["ms"] * 1e-3) / 1e9
                                   / HBM_PEAK_GBS, 4)}
        full["b"].close()
    if not args.no_check:
        picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
        chk = batched(p, models, 200, 0, 200, bat["block"], label="check")
        ef, ev = force_error(p, models, chk, picks)
        out["force_err_vs_fp64"] = {"max": max(ef.values()),
                                    "batched_total": batched_force_error(p, models, chk, picks),
                                    "per_term": {k: float(f"{v:.3g}") for k, v in ef.items()},
                                    "energy_per_term": {k: float(f"{v:.3g}")
                                                        for k, v in ev.items()}}
        chk["b"].close()
    per_plain = None
    if not args.no_plain:
        from pas_cg_amd import pcg as G
        picks = [int(i) for i in order[8::16]]
        prate, per_plain = plain(p, models, picks, graph=not args.plain_stages)
        srate, _ = plain(p, models, picks, graph=not args.plain_stages, form=G.FORM_STAGES)
        out["plain"] = {"value": round(prate, 1), "unit": "residue-steps/s",
                        "sample": "16 proteins (every 16th by length), one plain context at a "
                                  "time, 200 steps each through pcg_step in its own form: the "
                                  "persistent kernel for one system (one workgroup)",
                        "stage_kernels": round(srate, 1),
                        "stage_kernels_sample": "the same, forced to the stage kernels (one "
                                                "launch per stage, " +
                                                ("plain launches" if args.plain_stages else
                                                 "hipGraph-captured step") + ")"}
        out["speedup_vs_plain"] = round(value / prate, 2)
        out["speedup_vs_plain_stages"] = round(value / srate, 2)
    if not args.no_serve:
        sp = W.full_params(args.pad)
        out["served"] = ({} if args.serve_backlog_only else
                         served(sp, value, pool, per_plain, args.serve_window, args.serve_chunk))
        out["served"]["backlog"] = served_backlog(sp, pool, chunk=args.serve_chunk)
    if extra is not None:
        out["other_configs"] = other_configs(*extra, graph=not args.plain_stages)
    if not args.no_cpu:
        k = cpu_ex._max_workers
        spread = [int(order[(2 * j + 1) * len(order) // (2 * k)]) for j in range(k)]
        out["cpu_baseline"] = cpu_baseline(p, models, [int(i) for i in order[16::32]], ex=cpu_ex,
                                           picks_all=spread)
        cpu_ex.shutdown()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
