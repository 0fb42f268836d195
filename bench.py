"""Benchmark: plain batched KMC on one MI355X.

Workload (the plain batched configuration): 1024 independent replicas of FCC Al with
20 x 20 x 20 cells (32000 sites), one vacancy each, Mg and Zn fractions drawn uniformly in
[1, 5] at.%, T uniform in [400, 700] K, synthetic cluster-expansion coefficients (E0 barrier
model), FP64 energies, rates, prefix sums and clock. A bench "step" is one KMC step of every
replica; the persistent kernel advances all replicas `--chunk` steps per launch.

Prints one JSON line: KMC steps/s summed over all replicas, the roofline of the dominant kernel
(algorithmic bytes: 137 B per replica-step, measured with HIP events on the launch stream) and
the CPU baseline (the oracle's KineticMcFirst restatement on this host's cores, bounded sample
of the single-replica 10^3-cell workload).

Other workloads (--workload): uneven (uneven-size KMC beside canonical MC), large (64 replicas of
50^3 cells with the quaternary surrogate), chain (second-order KMC), anneal (simulated
annealing), exittime (ExitTime's per-site jump scan and energy changes), cmc (CanonicalMcSerial and
CanonicalMcOmp trials). Each prints one JSON line.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic HBM bytes of one replica-step: the species bytes of the 135 distinct sites within
# 3NN of the vacancy's 12 jump pairs, read once, plus the 2 bytes the jump writes. Site ids are
# closed forms of the coordinates here, so no neighbour-index bytes are read.
ALGO_BYTES_PER_STEP = 137
ALGO_BYTES_NOTE = ("137 B per replica-step = 135 distinct occupancy bytes read + 2 written; site ids "
                   "are closed forms, so the 4 B neighbour index per site of a table-driven design "
                   "(677 B per step) is not read. The kernel's working set is cache/LDS resident "
                   "(traffic is measured HBM bytes per launch), so HBM is not its bound: the step is "
                   "VALU/LDS-issue and latency bound (issue)")
FORMS = ["dense", "solvent-baseline", "dense-wide"]  # kernel shapes of the autotuner
METRIC = "KMC steps/s over all replicas, surrogate vs plain GPU at \u22645 meV barrier MAE"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(R, n, chunk, model="e0"):
    """HBM bytes per launch of the KMC kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (copy-calibrated), scaled to `chunk` steps per launch; None when no
    summary exists for this replica count and lattice size."""
    path = os.path.join(ROOT, "profiles", f"kmc_{model}_{R}x{4 * n ** 3}_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_replica_step:"):
                return round(float(line.split(":")[1]) * R * chunk)
    return None


def measured_issue(R, n, model="e0"):
    """VALU issue utilisation and instructions per replica-step of the KMC kernel from the
    committed rocprofv3 SQ-counter summary of this workload; None when none exists."""
    path = os.path.join(ROOT, "profiles", f"kmc_{model}_{R}x{4 * n ** 3}_sq_counters.txt")
    if not os.path.exists(path):
        return None
    out = {"source": os.path.relpath(path, ROOT)}
    with open(path) as f:
        for line in f:
            if line.startswith("per replica-step:"):
                v = line.split(":")[1].replace(",", "").split()
                out["valu_insts_per_replica_step"] = int(v[0])
                out["lds_insts_per_replica_step"] = int(v[2])
            elif line.startswith("VALU issue:"):
                out["valu_issue_frac"] = float(line.split("=")[-1].split()[0])
    return out


def launch_plan(steps, chunk):
    """Launch sizes that add up to exactly `steps` KMC steps: steps // chunk launches of `chunk`
    steps (chunk clamped to steps), then one shorter launch for the remainder. Returns a list of
    (steps per launch, launches)."""
    if steps <= 0 or chunk <= 0:
        raise ValueError("steps and chunk must be positive")
    chunk = min(chunk, steps)
    full, rest = divmod(steps, chunk)
    plan = [(chunk, full)]
    if rest:
        plan.append((rest, 1))
    return plan


def uses_graph(plan, eager=False):
    """Full launches come from a captured HIP graph when there are at least two of them (a
    one-launch graph saves nothing) and --eager is not set."""
    return not eager and plan[0][1] >= 2


def run_plan(drv, plan, stream, eager=False):
    """Issues the launches of `plan` on `stream`: the full launches (replayed from the graph
    prepare_plan captured, or eagerly), then the remainder launch eagerly."""
    (chunk, full), rest = plan[0], plan[1:]
    if uses_graph(plan, eager):
        drv.run_graph(chunk, full, stream)
    else:
        for _ in range(full):
            drv.run(chunk, stream)
    for c, _ in rest:
        drv.run(c, stream)


def prepare_plan(drv, plan, stream, eager=False):
    """Captures the HIP graph of the plan's full launches, outside any timed region."""
    assert sum(c * k for c, k in plan) > 0
    if uses_graph(plan, eager):
        drv.build_graph(plan[0][0], plan[0][1])


def host_cores():
    """Threads the CPU baselines use, and the host facts behind the figure. The count is the
    process's affinity mask, capped by OMP_NUM_THREADS when that is set (the thread allotment of
    a host that is shared between GPUs)."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    cores = min(aff, int(omp)) if omp.isdigit() and int(omp) > 0 else aff
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, {"nproc": os.cpu_count(), "affinity_cpus": aff,
                   "omp_num_threads": omp or None, "cpu_model": model,
                   "threads_used_rule": "affinity mask, capped by OMP_NUM_THREADS when set"}


def make_workload(R, n, seed):
    from latticemontecarlo_amd import synth
    rng = np.random.default_rng(seed)
    species = ["Al", "Mg", "X", "Zn"]
    N = 4 * n ** 3
    occs, temps = [], []
    for r in range(R):
        xm, xz = rng.uniform(0.01, 0.05, 2)
        occs.append(synth.random_occupancy(n, species, {"Mg": round(xm * N), "Zn": round(xz * N)},
                                           seed=seed * 100003 + r))
        temps.append(rng.uniform(400.0, 700.0))
    return occs, np.array(temps)


def cpu_baseline(coeff_path, budget_s=12.0, els=("Al", "Mg", "Zn"), kind="E0"):
    """Oracle KineticMcFirst on the host cores: single replica (n=10, 4000 sites, 80 atoms of
    each solute, T=500 K) with min(12, cores) OpenMP threads on the 12-event loop, and `cores`
    independent single-thread replicas; returns the larger aggregate steps/s. `kind` is the
    barrier model, E0 or Quartic (the reference's KMC driver holds the Quartic LRU predictor)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from latticemontecarlo_amd import synth
    cores, host = host_cores()
    n = 10
    els = list(els)
    lat = O.Lattice(n)
    bm = O.Model(lat, coeff_path, els, O.QUARTIC if kind == "Quartic" else O.E0)
    em = O.Model(lat, coeff_path, els, O.ENERGY)
    occ = synth.random_occupancy(n, sorted(els + ["X"]), {e: 80 for e in els if e != "Al"}, seed=1)
    # calibration (also builds every pair table once)
    secs, _ = O.kmc_baseline(bm, em, lat, occ, 500.0, 7, 1, 1, 100000, 200)
    per_step = max(secs / 200, 1e-6)
    steps = int(max(200, budget_s / 2 / per_step))
    secs1, _ = O.kmc_baseline(bm, em, lat, occ, 500.0, 7, 1, min(12, cores), 100000, steps)
    single = steps / secs1
    steps_t = int(max(100, steps / 2))
    secs2, _ = O.kmc_baseline(bm, em, lat, occ, 500.0, 7, cores, 1, 100000, steps_t)
    thru = cores * steps_t / secs2
    log(f"cpu baseline: single replica {single:.1f} steps/s ({min(12, cores)} threads), "
        f"throughput {thru:.1f} steps/s ({cores} replicas)")
    best = max(single, thru)
    return {"value": round(best, 2), "unit": "KMC steps/s", "cores": cores, "kind": "port",
            "host": host, "single_replica_steps_per_s": round(single, 2),
            "throughput_steps_per_s": round(thru, 2),
            "sample": (f"oracle KineticMcFirst ({kind} {'/'.join(els)}, LRU 100000), 10^3 cells x 4000 sites, "
                       f"T=500 K: 1 replica x {steps} steps with {min(12, cores)} threads "
                       f"({single:.1f} steps/s) and {cores} replicas x {steps_t} steps "
                       f"single-threaded ({thru:.1f} steps/s); value = larger")}


SURROGATE_EPOCHS = 160


def train_surrogate(L, bm, s_occs, s_temps, n, seed, epochs=SURROGATE_EPOCHS):
    """Four BF16 MLP members trained on the GPU on exact samples of KMC trajectories of the
    replicas `s_occs` (n^3 cells; 160 samples of the 12 jumps of every replica, 12 steps apart,
    x4 by the jump symmetries), Adam with a cosine schedule. Returns the surrogate and a
    summary."""
    import torch
    t0 = time.time()
    R = len(s_occs)
    drv = L.KineticMcFirst(bm, None, [n] * R, s_occs, s_temps, seed=seed + 1, maximum_steps=10 ** 9)
    drv.run(20)
    E, K, D = [], [], []
    for _ in range(160):
        drv.run(12)
        e, k, d = drv.samples()
        E.append(e)
        K.append(k)
        D.append(d)
    env, ek, de = L.augment(torch.cat(E), torch.cat(K), torch.cat(D))
    del drv, E, K, D
    n_samples = env.shape[0] // 4
    sur = L.Surrogate(bm, n_members=4, seed=seed)
    torch.cuda.synchronize()
    t1 = time.time()
    mse = sur.train(env, ek, de, epochs=epochs, batch=4096, lr=3e-3, seed=seed)
    torch.cuda.synchronize()
    train_s = time.time() - t1
    del env, ek, de
    log(f"surrogate: {n_samples} samples ({time.time() - t0 - train_s:.1f} s), trained in "
        f"{train_s:.1f} s ({epochs} epochs), mse {np.round(mse, 5)}")
    return sur, {"train_samples": n_samples, "train_rows_augmented": 4 * n_samples,
                 "train_epochs": epochs, "train_s": round(train_s, 2)}


def measured_mfma(R, n):
    """MFMA rate of the surrogate kernel from the committed rocprofv3 counter summary of this
    workload (512 x SQ_INSTS_VALU_MFMA_MOPS_BF16 per launch over its duration); None when none
    exists."""
    path = os.path.join(ROOT, "profiles", f"surrogate_kmc_{R}x{4 * n ** 3}_kernel_trace.md")
    if not os.path.exists(path):
        return None
    rates = []
    with open(path) as f:
        for line in f:
            if line.startswith("| dispatch"):
                rates.append(float(line.split("|")[3]))
    if not rates:
        return None
    return {"source": os.path.relpath(path, ROOT), "counter": "SQ_INSTS_VALU_MFMA_MOPS_BF16",
            "tflop_per_s": rates, "frac_of_dense_bf16_peak": round(max(rates) / 2500.0, 4),
            "note": "counted FLOP include the 4 unused rows of each 16-row replica tile"}


def mlp_flops_per_replica_step(sur):
    """Multiply-adds x 2 of the ensemble on the 12 event rows of one replica-step (the rows the
    network sees; the kernel's 16-row tiles carry 4 unused rows per replica)."""
    return 12 * sur.members * 2 * (sur.inputs * 128 + 128 * 128 + 128 * 2)


def time_surrogate(L, bm, sur, R, n, occs, temps, seed, steps, warmup, chunk, plain_value,
                   window=0.0):
    """The surrogate KMC path on the given replicas (same seeds as the plain run): throughput,
    fallback rate, and barrier error of the final states against the exact evaluation.
    window = 0 is the fallback rule as specified (any event whose members disagree by more than
    10 meV, or whose environment is out of distribution, is evaluated exactly); window > 0 lets
    disagreeing events whose rate is negligible (lowest member barrier more than `window` kT
    above the replica's lowest barrier) keep the ensemble's value."""
    import torch
    threshold = 0.010
    drv = L.KineticMcFirst(bm, None, [n] * R, occs, temps, seed=seed, maximum_steps=10 ** 9)
    drv.set_surrogate(sur, threshold=threshold, window=window)
    stream = torch.cuda.current_stream()
    drv.run(warmup, stream)
    torch.cuda.synchronize()
    st0 = drv.surrogate_stats()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    left = steps
    while left > 0:
        c = min(chunk, left)
        drv.run(c, stream)
        left -= c
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    value = R * steps / (ms / 1e3)
    st = drv.surrogate_stats()
    occ_now, state = drv.occupancy(), drv.state()
    batch = L.ReplicaBatch([n] * R, occ_now)
    _, ea, _, fl = [x.cpu().numpy() for x in sur.eval_jumps(bm, batch, state["vacancy"], temps,
                                                                threshold, window)]
    _, ea0, _ = [x.cpu().numpy() for x in L.eval_jumps(bm, batch, state["vacancy"])]
    f = fl.astype(bool)
    kT = 8.617333262145e-5 * np.asarray(temps)[:, None]
    rel = ea0 < ea0.min(1, keepdims=True) + 20.0 * kT
    err = np.abs(ea - ea0)
    flops = mlp_flops_per_replica_step(sur)
    events = st["events"] - st0["events"]
    fallbacks = st["fallbacks"] - st0["fallbacks"]
    out = {"value": round(value, 1), "unit": "KMC steps/s", "steps": steps,
           "speedup_vs_plain": round(value / plain_value, 3),
           "fallback_rule": ("members disagree by > 10 meV or environment out of distribution"
                             if window <= 0 else
                             f"as specified, but only for events within {window} kT of the "
                             "replica's lowest barrier"),
           "fallback_rate": round(fallbacks / max(1, events), 4),
           "mae_ea_meV_all_events": round(float(err.mean()) * 1e3, 3),
           "mae_target_meV": 5.0,
           "mae_ea_meV_surrogate_events": round(float(err[~f].mean()) * 1e3, 3) if (~f).any() else None,
           "mae_ea_meV_relevant_surrogate_events":
               round(float(err[~f & rel].mean()) * 1e3, 3) if (~f & rel).any() else None,
           "mae_sample": f"the 12 jumps of each of the {R} final states ({12 * R} events)",
           "threshold_meV": threshold * 1e3, "relevance_window_kT": window, "members": 4,
           "mlp_flop_per_replica_step": flops,
           "mlp_tflop_per_s": round(flops * value / 1e12, 2),
           "mfma_frac_of_dense_bf16_peak": round(flops * value / 2.5e15, 4),
           "mfma_counters": measured_mfma(R, n),
           "dtype": "bf16 weights and activations, fp32 accumulation"}
    log(f"surrogate KMC (window {window}): {value:.4g} steps/s ({value / plain_value:.2f}x "
        f"plain), fallback {out['fallback_rate'] * 100:.1f}%, MAE Ea "
        f"{out['mae_ea_meV_all_events']} meV (all events), MLP {out['mlp_tflop_per_s']} TFLOP/s")
    return out


def surrogate_leg(L, bm, R, n, occs, temps, seed, steps, warmup, chunk, plain_value,
                  epochs=SURROGATE_EPOCHS):
    """Learned surrogate on the bench workload: trained on trajectories of 1024 replicas of
    10^3 cells of the same composition range, then timed on the bench replicas with the
    specified fallback rule, and once more with the relevance window."""
    s_occs, s_temps = make_workload(1024, 10, seed + 1)
    sur, info = train_surrogate(L, bm, s_occs, s_temps, 10, seed, epochs)
    out = time_surrogate(L, bm, sur, R, n, occs, temps, seed, steps, warmup, chunk, plain_value)
    out.update(info)
    win = time_surrogate(L, bm, sur, R, n, occs, temps, seed, steps, warmup, chunk, plain_value,
                         window=20.0)
    out["with_relevance_window"] = {k: win[k] for k in (
        "value", "speedup_vs_plain", "fallback_rule", "fallback_rate", "mae_ea_meV_all_events",
        "mae_ea_meV_surrogate_events", "mae_ea_meV_relevant_surrogate_events",
        "relevance_window_kT")}
    return out


def timed_plain(L, drv, R, steps, warmup, chunk):
    """Warm-up launches, then exactly `steps` KMC steps of every replica (a HIP graph of the
    full `chunk`-step launches, then a shorter launch for the remainder); returns the seconds
    measured with HIP events on the stream."""
    import torch
    stream = torch.cuda.current_stream()
    left = warmup
    while left > 0:
        c = min(chunk, left)
        drv.run(c, stream)
        left -= c
    plan = launch_plan(steps, chunk)
    prepare_plan(drv, plan, stream)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    run_plan(drv, plan, stream)
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1e3


def large_workload(args):
    """Large surrogate run: 64 replicas of 50^3 cells (500000 sites), Al/Cu/Mg/Zn 94/2/2/2 at.%,
    T = 450 K, quaternary E0 model. The plain path and the surrogate path (trained on exact
    samples of 1024 trajectories of 10^3 cells of the same composition) run on the same seeds;
    the surrogate's fallback rate and barrier error are reported. Prints one JSON line."""
    import latticemontecarlo_amd as L
    from latticemontecarlo_amd import synth
    els = ["Al", "Cu", "Mg", "Zn"]
    species = sorted(els + ["X"])
    tmp = tempfile.mkdtemp()
    coeff = os.path.join(tmp, "coeff_e0_q.json")
    synth.write_coefficients(coeff, els, "E0", seed=args.seed)
    bm = L.Model(coeff, els, L.E0)

    def alloy(R, n, seed):
        N = 4 * n ** 3
        c = round(0.02 * N)
        return [synth.random_occupancy(n, species, {"Cu": c, "Mg": c, "Zn": c}, seed=seed * 100003 + r)
                for r in range(R)]

    R, n = 64, 50
    occs, temps = alloy(R, n, args.seed), np.full(R, 450.0)
    log(f"large: {R} replicas x {4 * n ** 3} sites, Al/Cu/Mg/Zn 94/2/2/2, T 450 K")
    drv = L.KineticMcFirst(bm, None, [n] * R, occs, temps, seed=args.seed, maximum_steps=10 ** 9)
    tune = drv.autotune(steps=200)
    elapsed = timed_plain(L, drv, R, args.steps, args.warmup, args.chunk)
    bm.check_errors()
    value = R * args.steps / elapsed
    log(f"plain: {value:.4g} KMC steps/s over {R} replicas ({elapsed:.2f} s)")
    del drv
    sur, info = train_surrogate(L, bm, alloy(1024, 10, args.seed + 1), np.full(1024, 450.0), 10,
                                args.seed, args.surrogate_epochs)
    s = time_surrogate(L, bm, sur, R, n, occs, temps, args.seed, args.steps, args.warmup,
                       args.chunk, value)
    s.update(info)
    line = {"metric": METRIC, "value": round(value, 1), "unit": "KMC steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(elapsed / args.steps * 1e3, 6), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"large surrogate run, {R} replicas x {n}^3 FCC cells "
                                   f"({4 * n ** 3} sites), Al/Cu/Mg/Zn 94/2/2/2 at.%, T 450 K, "
                                   "quaternary E0 barrier model",
                       "value_path": "plain GPU path (exact); the surrogate path is reported "
                                     "under surrogate",
                       "formulation": FORMS[tune["formulation"]],
                       "event_evaluations_per_s": round(12 * value, 1)},
            "surrogate": s}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(coeff, els=els)
    print(json.dumps(line), flush=True)


def chain_workload(args):
    """Second-order (chain) KMC, KineticMcChainOmpi, on the batched workload (1024 replicas of
    20^3 cells, Mg and Zn 1-5 at.% each, T 400-700 K, E0): every step evaluates the 144 jumps
    out of the vacancy's first neighbours. CPU baseline: the oracle's chain steps, independent
    single-thread replicas of 10^3 cells on the host cores. Prints one JSON line."""
    import latticemontecarlo_amd as L
    from latticemontecarlo_amd import synth
    R, n = args.replicas, args.ncell
    tmp = tempfile.mkdtemp()
    coeff = os.path.join(tmp, "coeff_e0.json")
    synth.write_coefficients(coeff, ["Al", "Mg", "Zn"], "E0", seed=args.seed)
    bm = L.Model(coeff, ["Al", "Mg", "Zn"], L.E0)
    occs, temps = make_workload(R, n, args.seed)
    drv = L.KineticMcChain(bm, None, [n] * R, occs, temps, seed=args.seed, maximum_steps=10 ** 9)
    steps, chunk = args.steps, min(args.chunk, args.steps)
    elapsed = timed_plain(L, drv, R, steps, args.warmup, chunk)
    bm.check_errors()
    value = R * steps / elapsed
    log(f"chain: {value:.4g} KMC steps/s over {R} replicas ({elapsed:.2f} s, "
        f"{144 * value:.4g} jump evaluations/s)")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        cores, host = host_cores()
        lat = O.Lattice(10)
        om = O.Model(lat, coeff, ["Al", "Mg", "Zn"], O.E0)
        occ = synth.random_occupancy(10, ["Al", "Mg", "X", "Zn"], {"Mg": 80, "Zn": 80}, seed=1)
        O.kmc_baseline(om, None, lat, occ, 500.0, 7, 1, 1, 100000, 20, chain=True)  # tables
        secs, _ = O.kmc_baseline(om, None, lat, occ, 500.0, 7, cores, 1, 100000, 200, chain=True)
        thru = cores * 200 / secs
        log(f"cpu baseline (chain): {thru:.1f} steps/s ({cores} single-thread replicas)")
        cpu = {"value": round(thru, 2), "unit": "KMC steps/s", "cores": cores, "kind": "port",
               "host": host,
               "sample": f"oracle KineticMcChainOmpi steps (E0, LRU 100000), 10^3 cells x 4000 "
                         f"sites, T=500 K: {cores} single-thread replicas x 200 steps"}
    line = {"metric": "KMC steps/s over all replicas (second-order chain steps)",
            "value": round(value, 1), "unit": "KMC steps/s", "n_gpus": 1, "steps": steps,
            "warmup": args.warmup, "ms_per_step": round(elapsed / steps * 1e3, 6),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"second-order chain KMC, {R} replicas x {n}^3 FCC cells, Al with "
                                   "1-5 at.% Mg and Zn each, T 400-700 K, E0 barrier model",
                       "jump_evaluations_per_s": round(144 * value, 1)},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def cmc_workload(args):
    """Canonical MC on the canonical workload's replicas: 64 replicas of 30^3 cells (108000 sites,
    Al/Mg/Zn 90/5/5), T linear 800 -> 300 K over --cmc-trials trials each, pair-site swap dE, as
    CanonicalMcSerial trials and as CanonicalMcOmp rounds of 64 trials (lmc_cmc_set_parallel_trials).
    CPU baseline: the oracle's CanonicalMc trials, one single-thread replica per host core. Prints
    one JSON line; value = the CanonicalMcOmp rate."""
    import torch
    import latticemontecarlo_amd as L
    from latticemontecarlo_amd import synth
    tmp = tempfile.mkdtemp()
    coeff = os.path.join(tmp, "coeff_e0.json")
    synth.write_coefficients(coeff, ["Al", "Mg", "Zn"], "E0", seed=args.seed)
    ps = L.Model(coeff, ["Al", "Mg", "Zn"], L.PAIRSITE)
    species = ["Al", "Mg", "X", "Zn"]
    R, n, trials, K = 64, 30, args.cmc_trials, 64
    N = 4 * n ** 3
    occs = [synth.random_occupancy(n, species, {"Mg": round(0.05 * N), "Zn": round(0.05 * N)},
                                   seed=args.seed * 11 + r, vacancy=False) for r in range(R)]
    stream = torch.cuda.current_stream()
    rates = {}
    for mode in ("serial", "omp"):
        drv = L.CanonicalMc(ps, None, [n] * R, occs, 800.0, 300.0, args.seed, maximum_steps=trials)
        if mode == "omp":
            drv.set_parallel_trials(K)
        drv.run(min(2000, trials), stream)  # warm-up
        torch.cuda.synchronize()
        done0 = int(drv.state()["steps"].sum())
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        drv.run(trials, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        secs = e0.elapsed_time(e1) / 1e3
        st = drv.state()
        ps.check_errors()
        done = int(st["steps"].sum()) - done0
        rates[mode] = (done / secs, secs, float(st["accepted"].sum()) / max(1, int(st["steps"].sum())))
        log(f"canonical MC {mode}: {done / secs:.4g} trials/s over {R} replicas ({secs:.3f} s)")
        del drv
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        lat = O.Lattice(n)
        om = O.Model(lat, coeff, ["Al", "Mg", "Zn"], O.PAIRSITE)
        cores, host = host_cores()
        k = 4000
        per = O.cmc_baseline(om, lat, occs[0], 800.0, args.seed, 1, k, trials, False) / k
        k = int(min(trials, max(2000, 12.0 / max(per, 1e-7))))
        secs_c = O.cmc_baseline(om, lat, occs[0], 800.0, args.seed, cores, k, trials, False)
        thru = cores * k / secs_c
        cpu = {"value": round(thru, 1), "unit": "trials/s", "cores": cores, "kind": "port",
               "host": host,
               "sample": f"oracle CanonicalMc trials, {cores} replicas of 30^3 cells run "
                         f"concurrently, one thread each, {k} trials each ({secs_c:.1f} s)"}
        log(f"cpu baseline (canonical MC): {thru:.1f} trials/s ({cores} concurrent replicas)")
    value, secs, accr = rates["omp"]
    line = {"metric": "canonical MC trials/s over all replicas (CanonicalMcOmp rounds)",
            "value": round(value, 1), "unit": "trials/s", "n_gpus": 1, "steps": trials,
            "warmup": min(2000, trials), "ms_per_step": round(secs / trials * 1e3, 6),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"canonical MC, {R} replicas x {n}^3 FCC cells ({N} sites), "
                                   "Al/Mg/Zn 90/5/5, T linear 800 -> 300 K, pair-site swap dE",
                       "trials_per_round": K,
                       "serial_trials_per_s": round(rates["serial"][0], 1),
                       "omp_over_serial": round(value / rates["serial"][0], 3),
                       "acceptance_omp": round(accr, 4),
                       "acceptance_serial": round(rates["serial"][2], 4)},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def anneal_workload(args):
    """Simulated annealing (SimulatedAnnealing.cpp:99-185): 64 replicas of 30^3 cells (108000
    sites, Al/Mg/Zn 90/5/5) from T = 800 K over --cmc-trials trials each, pair-site swap dE.
    CPU baseline: the oracle's annealing trials on the host cores. Prints one JSON line."""
    import torch
    import latticemontecarlo_amd as L
    from latticemontecarlo_amd import synth
    tmp = tempfile.mkdtemp()
    coeff = os.path.join(tmp, "coeff_e0.json")
    synth.write_coefficients(coeff, ["Al", "Mg", "Zn"], "E0", seed=args.seed)
    ps = L.Model(coeff, ["Al", "Mg", "Zn"], L.PAIRSITE)
    species = ["Al", "Mg", "X", "Zn"]
    R, n, trials = 64, 30, args.cmc_trials
    N = 4 * n ** 3
    occs = [synth.random_occupancy(n, species, {"Mg": round(0.05 * N), "Zn": round(0.05 * N)},
                                   seed=args.seed * 13 + r, vacancy=False) for r in range(R)]
    sa = L.SimulatedAnnealing(ps, None, [n] * R, occs, 800.0, args.seed, maximum_steps=trials)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    chunk = max(1, trials // 20)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(max(1, trials // chunk)):
        sa.run(chunk, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    secs = e0.elapsed_time(e1) / 1e3
    st, an = sa.state(), sa.annealing_state()
    ps.check_errors()
    done = int(st["steps"].sum())
    value = done / secs
    log(f"annealing: {value:.4g} trials/s over {R} replicas ({secs:.2f} s), final T "
        f"{an['temperature'].min():.1f}-{an['temperature'].max():.1f} K, reheats {an['reheats'].sum()}")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        lat = O.Lattice(n)
        om = O.Model(lat, coeff, ["Al", "Mg", "Zn"], O.PAIRSITE)
        cores, host = host_cores()
        k = 4000  # calibration, then a sample of about 15 s of concurrent work
        per = O.cmc_baseline(om, lat, occs[0], 800.0, args.seed, 1, k, trials, True) / k
        k = int(min(trials, max(2000, 15.0 / max(per, 1e-7))))
        secs_c = O.cmc_baseline(om, lat, occs[0], 800.0, args.seed, cores, k, trials, True)
        thru = cores * k / secs_c
        cpu = {"value": round(thru, 1), "unit": "trials/s", "cores": cores, "kind": "port",
               "host": host,
               "sample": f"oracle SimulatedAnnealing trials, {cores} replicas of 30^3 cells run "
                         f"concurrently, one thread each, {k} trials each ({secs_c:.1f} s)"}
        log(f"cpu baseline (annealing): {thru:.1f} trials/s ({cores} concurrent replicas)")
    line = {"metric": "simulated-annealing trials/s over all replicas", "value": round(value, 1),
            "unit": "trials/s", "n_gpus": 1, "steps": trials, "warmup": 0,
            "ms_per_step": round(secs / trials * 1e3, 6), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"simulated annealing, {R} replicas x {n}^3 FCC cells "
                                   f"({N} sites), Al/Mg/Zn 90/5/5, T0 800 K, pair-site swap dE",
                       "final_temperature_K": [round(float(an["temperature"].min()), 2),
                                               round(float(an["temperature"].max()), 2)],
                       "reheats": int(an["reheats"].sum()),
                       "acceptance": round(float(st["accepted"].sum()) / max(1, done), 4)},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def exittime_workload(args):
    """ExitTime's per-site jump evaluations (ExitTime::GetBarrierListAndExitTime and
    GetJumpEnergetics, lmc/ansys/src/ExitTime.cpp:216-311): a vacancy placed at every site of one
    --ncell^3-cell frame (Al with 3 at.% Mg and Zn each, E0 model), the frame's own vacancy
    holding Al, 12 jumps per site, in one lmc_vacancy_scan launch over device buffers. CPU
    baseline: the oracle's single-jump evaluation over a sample of sites. Prints one JSON line."""
    import torch
    import latticemontecarlo_amd as L
    from latticemontecarlo_amd import capi, fcc_neighbors, synth
    n = args.ncell
    N = 4 * n ** 3
    els = ["Al", "Mg", "Zn"]
    tmp = tempfile.mkdtemp()
    coeff = os.path.join(tmp, "coeff_e0.json")
    synth.write_coefficients(coeff, els, "E0", seed=args.seed)
    bm = L.Model(coeff, els, L.E0)
    sp = bm.species
    occ = synth.random_occupancy(n, sp, {"Mg": round(0.03 * N), "Zn": round(0.03 * N)},
                                 seed=args.seed)
    vac = int(np.flatnonzero(occ == sp.index("X"))[0])
    dev = torch.device("cuda")
    d_occ = torch.from_numpy(occ).to(dev)
    sites = torch.arange(N, dtype=torch.int64, device=dev)
    nb = torch.empty((N, 12), dtype=torch.int64, device=dev)
    ea = torch.empty((N, 12), dtype=torch.float64, device=dev)
    de = torch.empty((N, 12), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sp_ptr = capi.stream_ptr(stream)

    def scan():
        capi.check(capi.lib().lmc_vacancy_scan(bm.h, capi.dptr(d_occ), n, vac, sp.index("Al"),
                                               capi.dptr(sites), N, capi.dptr(nb), capi.dptr(ea),
                                               capi.dptr(de), sp_ptr))
    for _ in range(max(1, args.warmup // 500)):
        scan()
    reps = max(1, args.steps // 10000)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        scan()
    e1.record(stream)
    torch.cuda.synchronize()
    secs = e0.elapsed_time(e1) / 1e3 / reps
    bm.check_errors()
    value = 12 * N / secs
    log(f"exit-time scan: {value:.4g} jump evaluations/s ({N} sites, {secs * 1e3:.3f} ms per scan)")
    # GetProfileEnergy / GetBindingEnergy's per-site energy changes (ExitTime.cpp:401-512): every
    # site of the frame from a vacancy to Mg, one batched lmc_site_change_de launch
    ps = L.Model(coeff, els, L.PAIRSITE)
    z_off = torch.zeros(N, dtype=torch.int64, device=dev)
    ncs = torch.full((N,), n, dtype=torch.int32, device=dev)
    frm = torch.full((N,), sp.index("X"), dtype=torch.uint8, device=dev)
    to = torch.full((N,), sp.index("Mg"), dtype=torch.uint8, device=dev)
    sde = torch.empty(N, dtype=torch.float64, device=dev)

    def site_changes():
        capi.check(capi.lib().lmc_site_change_de(ps.h, capi.dptr(d_occ), capi.dptr(z_off),
                                                 capi.dptr(ncs), capi.dptr(sites), capi.dptr(frm),
                                                 capi.dptr(to), N, capi.dptr(sde), sp_ptr))
    site_changes()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        site_changes()
    e1.record(stream)
    torch.cuda.synchronize()
    ssecs = e0.elapsed_time(e1) / 1e3 / reps
    ps.check_errors()
    log(f"exit-time site energy changes: {N / ssecs:.4g} sites/s ({ssecs * 1e3:.3f} ms per frame)")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        om = O.Model(O.Lattice(n), coeff, els, O.E0)
        base = occ.copy()
        base[vac] = sp.index("Al")
        sample = range(0, N, max(1, N // 400))
        nbs = {s: [int(y) for y in fcc_neighbors(n, s, 1)] for s in sample}

        def cpu_pass():
            k = 0
            for s in sample:
                keep = base[s]
                base[s] = sp.index("X")
                for y in nbs[s]:
                    om.barrier_de(base, s, y)
                    k += 1
                base[s] = keep
            return k
        cpu_pass()  # builds the oracle's pair tables, as the reference does at construction
        t0 = time.perf_counter()
        k = cpu_pass()
        cv = k / (time.perf_counter() - t0)
        cpu = {"value": round(cv, 1), "unit": "jump evaluations/s", "cores": 1, "kind": "port",
               "sample": f"oracle GetBarrierAndDiffFromLatticeIdPair (E0, no memo) of the 12 jumps "
                         f"of {len(sample)} sites of the same frame, one thread, tables built "
                         f"before timing"}
        log(f"cpu baseline (exit-time scan): {cv:.1f} jump evaluations/s")
    line = {"metric": "ExitTime jump evaluations/s (vacancy at every site of a frame)",
            "value": round(value, 1), "unit": "jump evaluations/s", "n_gpus": 1, "steps": reps,
            "warmup": max(1, args.warmup // 500), "ms_per_step": round(secs * 1e3, 6),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"ExitTime barrier scan, one {n}^3-cell frame ({N} sites, "
                                   f"{12 * N} jumps per scan), Al with 3 at.% Mg and Zn each, "
                                   "E0 barrier model",
                       "site_energy_changes_per_s": round(N / ssecs, 1)},
            # 300 algorithmic bytes per jump: its 60 environment sites x (1 B species + 4 B index)
            "roofline": {"bound": "hbm", "achieved": round(300 * value / 1e9, 3), "peak": 8000.0,
                         "unit": "GB/s", "frac": round(300 * value / 1e9 / 8000.0, 6),
                         "traffic": None},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def uneven_workload(args):
    """Uneven-size batched KMC alongside canonical MC: 256 KMC replicas of n^3 cells with n drawn
    from 8..40 (2048-256000 sites; Mg, Zn 1-5 at.% each, T 400-700 K) on one stream and 64
    CanonicalMc replicas of 30^3 cells (108000 sites, Al/Mg/Zn 90/5/5, T 800 -> 300 K linear
    over the run's trials) on another, concurrently. Prints one JSON line."""
    import torch
    import latticemontecarlo_amd as L
    from latticemontecarlo_amd import synth
    rng = np.random.default_rng(args.seed)
    tmp = tempfile.mkdtemp()
    coeff = os.path.join(tmp, "coeff_e0.json")
    synth.write_coefficients(coeff, ["Al", "Mg", "Zn"], "E0", seed=args.seed)
    bm = L.Model(coeff, ["Al", "Mg", "Zn"], L.E0)
    ps = L.Model(coeff, ["Al", "Mg", "Zn"], L.PAIRSITE)
    species = ["Al", "Mg", "X", "Zn"]
    Rk, Rc, nc = 256, 64, 30
    sizes = rng.integers(8, 41, Rk)
    kocc, ktemp = [], rng.uniform(400.0, 700.0, Rk)
    for r, n in enumerate(sizes):
        N = 4 * int(n) ** 3
        xm, xz = rng.uniform(0.01, 0.05, 2)
        kocc.append(synth.random_occupancy(int(n), species, {"Mg": round(xm * N), "Zn": round(xz * N)},
                                           seed=args.seed * 7 + r))
    Nc = 4 * nc ** 3
    cocc = [synth.random_occupancy(nc, species, {"Mg": round(0.05 * Nc), "Zn": round(0.05 * Nc)},
                                   seed=args.seed * 11 + r, vacancy=False) for r in range(Rc)]
    ksteps, trials = args.steps, args.cmc_trials
    log(f"uneven: {Rk} KMC replicas of {int(sizes.min())}^3..{int(sizes.max())}^3 cells "
        f"({int((4 * sizes.astype(np.int64) ** 3).sum())} sites), {Rc} CMC replicas of {nc}^3")
    kmc = L.KineticMcFirst(bm, None, [int(n) for n in sizes], kocc, ktemp, seed=args.seed,
                           maximum_steps=10 ** 9)
    cmc = L.CanonicalMc(ps, None, [nc] * Rc, cocc, 800.0, 300.0, args.seed, maximum_steps=trials)
    tune = kmc.autotune(steps=200)
    sk, sc = torch.cuda.Stream(), torch.cuda.Stream()  # default priority
    kmc.run(500, sk)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(sk)
    ev[2].record(sc)
    for c, k in launch_plan(ksteps, args.chunk):
        for _ in range(k):
            kmc.run(c, sk)
    cchunk = max(1, trials // 20)
    for _ in range(max(1, trials // cchunk)):
        cmc.run(cchunk, sc)
    ev[1].record(sk)
    ev[3].record(sc)
    torch.cuda.synchronize()
    tk = ev[0].elapsed_time(ev[1]) / 1e3
    tc = ev[2].elapsed_time(ev[3]) / 1e3
    kst = kmc.state()
    cst = cmc.state()
    bm.check_errors()
    kdone = int(kst["steps"].sum()) - 500 * Rk
    cdone = int(cst["steps"].sum())
    kval, cval = kdone / tk, cdone / tc
    log(f"KMC {kval:.4g} steps/s over {Rk} uneven replicas ({tk:.2f} s), CMC {cval:.4g} trials/s "
        f"over {Rc} replicas ({tc:.2f} s), acceptance {cst['accepted'].sum() / max(1, cdone):.3f}")
    line = {"metric": "KMC steps/s over all replicas (uneven sizes, with CanonicalMc alongside)",
            "value": round(kval, 1), "unit": "KMC steps/s", "n_gpus": 1, "steps": ksteps,
            "warmup": 500, "ms_per_step": round(tk / max(1, ksteps) * 1e3, 6),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "uneven-size batched KMC alongside CanonicalMc", "kmc_replicas": Rk,
                       "kmc_cells": [int(sizes.min()), int(sizes.max())],
                       "kmc_formulation": FORMS[tune["formulation"]],
                       "cmc_replicas": Rc, "cmc_sites": Nc, "cmc_trials_per_replica": trials,
                       "cmc_trials_per_s": round(cval, 1),
                       "cmc_acceptance": round(float(cst["accepted"].sum()) / max(1, cdone), 4)}}
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(coeff)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        lat = O.Lattice(nc)
        op = O.Model(lat, coeff, ["Al", "Mg", "Zn"], O.PAIRSITE)
        cores, host = host_cores()
        kc = 4000
        per = O.cmc_baseline(op, lat, cocc[0], 800.0, args.seed, 1, kc, trials, False) / kc
        kc = int(min(trials, max(2000, 10.0 / max(per, 1e-7))))
        secs_c = O.cmc_baseline(op, lat, cocc[0], 800.0, args.seed, cores, kc, trials, False)
        cpu["cmc"] = {"value": round(cores * kc / secs_c, 1), "unit": "trials/s", "cores": cores,
                      "sample": f"oracle CanonicalMc trials, {cores} replicas of {nc}^3 cells run "
                                f"concurrently, one thread each, {kc} trials each at 800 K"}
        log(f"cpu baseline (canonical MC): {cores * kc / secs_c:.1f} trials/s")
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100000, help="timed KMC steps per replica")
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--chunk", type=int, default=5000, help="KMC steps per kernel launch")
    ap.add_argument("--replicas", type=int, default=1024)
    ap.add_argument("--ncell", type=int, default=20)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--eager", action="store_true", help="launch eagerly instead of a HIP graph")
    ap.add_argument("--workload", choices=["batched", "uneven", "large", "chain", "anneal", "exittime",
                                           "cmc"],
                    default="batched")
    ap.add_argument("--cmc-trials", type=int, default=100000)
    ap.add_argument("--surrogate-steps", type=int, default=20000)
    ap.add_argument("--surrogate-epochs", type=int, default=SURROGATE_EPOCHS,
                    help="training epochs of the surrogate (fewer for a profiling run)")
    ap.add_argument("--model", choices=["e0", "quartic"], default="e0",
                    help="barrier model of the batched workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); replicas are "
              "independent, so N GPUs would each run their own share of them.", file=sys.stderr)
        sys.exit(2)

    import torch
    import latticemontecarlo_amd as L
    from latticemontecarlo_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    if args.workload == "uneven":
        uneven_workload(args)
        return
    if args.workload == "large":
        large_workload(args)
        return
    if args.workload == "chain":
        chain_workload(args)
        return
    if args.workload == "anneal":
        anneal_workload(args)
        return
    if args.workload == "cmc":
        cmc_workload(args)
        return
    if args.workload == "exittime":
        exittime_workload(args)
        return
    R, n = args.replicas, args.ncell
    quartic = args.model == "quartic"
    tmp = tempfile.mkdtemp()
    coeff = os.path.join(tmp, f"coeff_{args.model}.json")
    synth.write_coefficients(coeff, ["Al", "Mg", "Zn"], "Quartic" if quartic else "E0",
                             seed=args.seed)
    log(f"setup: {R} replicas x {4 * n ** 3} sites, {args.model} barrier model")
    occs, temps = make_workload(R, n, args.seed)
    bm = L.Model(coeff, ["Al", "Mg", "Zn"], L.QUARTIC if quartic else L.E0)
    em = L.Model(coeff, ["Al", "Mg", "Zn"], L.ENERGY)
    start_steps = 0
    # maximum_steps is far beyond any run, so the timed steps never meet it
    drv = L.KineticMcFirst(bm, em, [n] * R, occs, temps, seed=args.seed, maximum_steps=10 ** 12)
    stream = torch.cuda.current_stream()
    tune = drv.autotune(steps=300)
    log(f"autotune: dense {tune['ms_dense']:.4f} ms, solvent baseline {tune['ms_solvent']:.4f} ms, "
        f"dense wide {tune['ms_wide']:.4f} ms per KMC step of the batch -> "
        f"{FORMS[tune['formulation']]}")
    t0 = time.time()
    for c, k in launch_plan(args.warmup, args.chunk) if args.warmup > 0 else []:
        for _ in range(k):
            drv.run(c, stream)
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} steps: {time.time() - t0:.2f} s")
    steps0 = int(drv.state()["steps"].min())

    # timed region: exactly args.steps KMC steps of every replica -- the full launches of
    # args.chunk steps replayed from a HIP graph captured beforehand, then one shorter launch
    plan = launch_plan(args.steps, args.chunk)
    prepare_plan(drv, plan, stream, args.eager)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    run_plan(drv, plan, stream, args.eager)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    elapsed = start.elapsed_time(end) / 1e3
    n_launch = sum(k for _, k in plan)
    log(f"{n_launch} launches ({plan}, {'HIP graph' if uses_graph(plan, args.eager) else 'eager'}): "
        f"{elapsed * 1e3:.3f} ms, {elapsed / n_launch * 1e3:.3f} ms per launch")
    st = drv.state()
    assert int(st["steps"].min()) == steps0 + args.steps, "every replica must run every timed step"
    assert int(st["steps"].max()) == steps0 + args.steps, "no replica may run past the timed steps"
    bm.check_errors()

    replica_steps = R * args.steps
    value = replica_steps / elapsed
    # algorithmic bytes over the launches' average duration (the launches run back to back on
    # one stream, so the region's time is the sum of the launch durations)
    achieved = ALGO_BYTES_PER_STEP * replica_steps / elapsed / 1e9
    traffic = measured_traffic(R, n, args.steps / n_launch, args.model)
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "KMC steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain batched KMC, {R} replicas x {n}^3 FCC cells "
                               f"({4 * n ** 3} sites), Al with 1-5 at.% Mg and Zn each, "
                               f"T 400-700 K, {'Quartic' if quartic else 'E0'} barrier model, "
                               "one vacancy per replica",
                   "value_path": "plain GPU path (exact); the surrogate path is reported under "
                                 "surrogate",
                   "replicas": R, "sites_per_replica": 4 * n ** 3,
                   "steps_per_replica": args.steps, "launches": [list(x) for x in plan],
                   "launch": "hip graph (+ eager remainder)" if uses_graph(plan, args.eager)
                             else "eager",
                   "formulation": FORMS[tune["formulation"]],
                   "autotune_ms_per_batch_step": {FORMS[f]: (None if tune[k] != tune[k] else
                                                             round(tune[k], 6))
                                                  for f, k in enumerate(["ms_dense", "ms_solvent",
                                                                         "ms_wide"])},
                   "event_evaluations_per_s": round(12 * value, 1),
                   "wall_s": round(wall, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                     "traffic": traffic,
                     "algorithmic_bytes_per_replica_step": ALGO_BYTES_PER_STEP,
                     "note": ALGO_BYTES_NOTE,
                     "issue": measured_issue(R, n, args.model)},
    }
    log(f"GPU: {value:.4g} KMC steps/s ({12 * value:.4g} event evaluations/s), "
        f"{elapsed:.4f} s for {replica_steps} replica-steps")
    if not args.no_surrogate and not quartic:
        line["surrogate"] = surrogate_leg(L, bm, R, n, occs, temps, args.seed,
                                          args.surrogate_steps, min(max(args.warmup, 100), 1000),
                                          args.chunk, value,
                                          args.surrogate_epochs)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(coeff, kind="Quartic" if quartic else "E0")
        if not quartic:
            # the reference's KMC driver holds the Quartic LRU predictor; timed beside E0
            qcoeff = os.path.join(tmp, "coeff_quartic.json")
            synth.write_coefficients(qcoeff, ["Al", "Mg", "Zn"], "Quartic", seed=args.seed)
            line["cpu_baseline"]["quartic"] = cpu_baseline(qcoeff, budget_s=8.0, kind="Quartic")
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
