#!/usr/bin/env python3
"""Benchmark of the KNP network step on one MI355X.

Workload (default): the plain FP32 method at GPU scale - 100,000 default BLIFAT neurons (80% excitatory,
20% inhibitory), fixed in-degree 1,000 (1e8 DeltaSynapses, delays 1..20), Bernoulli(0.005) external drive,
seeded synthetic data. A step is one network step (`_step()`): every population update plus every
projection's spike -> impact expansion, with all inputs generated and resident on the device.

Prints one JSON line: synaptic events/s over the timed steps (events = SynapticImpacts created, counted on
the device), network steps/s, the roofline position of the dominant kernel measured with HIP events on the
library's stream, and the CPU oracle (the restated single-threaded backend) timed on a bounded sample of the
same workload on this host, plus a bit-exact spike comparison of the first steps against that oracle.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10_000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--parity-steps", type=int, default=10_000,
                    help="c1: steps from step 0 whose spike trains are compared with the oracle (the configuration "
                         "runs 10,000); the multi-threaded oracle, equal to the single-threaded one, runs them")
    ap.add_argument("--timing-steps", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample length")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="c1: leave out the batched-instances measurement (the metric's ratio half)")
    ap.add_argument("--workload", choices=["c1", "c0", "c2", "c3", "c4"], default="c1",
                    help="c1 (default): the plain FP32 network at GPU scale, the headline line, with the batched "
                         "STDP instances (instance-steps/s, ratio to one instance on the plain FP32 path, FP64 "
                         "agreement) in its `batched` field. The others print their own measurement line: c0 small "
                         "plain network, c2 single STDP instance, c3 batched STDP instances, c4 uneven instances "
                         "with the FP8 dense projection on MFMA")
    ap.add_argument("--batch", type=int, default=256, help="c3: instances")
    ap.add_argument("--agree-steps", type=int, default=5000,
                    help="c2/c3/c4: steps compared with the FP64 oracle (c3 runs 5,000)")
    ap.add_argument("--agree-instances", type=int, default=4, help="c3: instances compared with the FP64 oracle")
    ap.add_argument("--dense", choices=["full", "skip", "none"], default="skip",
                    help="c3: the plastic input projection expanded on MFMA over all inputs (full: binary spikes x "
                         "BF16 weight copies), over the groups of 8 inputs holding a spike (skip), or by the FP32 "
                         "event-driven scatter (none)")
    ap.add_argument("--no-split", dest="split", action="store_false",
                    help="c3: one BF16 value per weight (default: two parts, hi + lo, 16 significant bits)")
    ap.add_argument("--graph-steps", type=int, default=None,
                    help="steps per captured step graph in long one-launch-per-step calls (0: no graphs); default: "
                         "the library's")
    ap.add_argument("--train-wg-limit", type=int, default=None,
                    help="synaptic-resource training: population size up to which each candidate gets a workgroup "
                         "(beyond it, a wave); default: the library's")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the library's default execution choices (default: the step tuner times the "
                         "candidates from the loaded state, restores it, and keeps the fastest)")
    ap.add_argument("--tune-steps", type=int, default=256, help="steps per step-tuner trial")
    ap.add_argument("--scatter", choices=["auto", "stepwise", "atomic"], default="auto",
                    help="impact summation: tiled populations and persistent launches where eligible (auto), "
                         "tiled with one launch per step (stepwise), or atomics everywhere")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the step of one network runs on one MI355X "
              "(sharding over several GPUs is designed, not implemented). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.workload != "c1":
        other_workload(args)
        return
    line = run_c1(args)
    if not args.no_batched:
        # the metric's second half: batched BF16 instances against one instance on the plain FP32 path
        c3 = dict(vars(args))
        c3.update(steps=1000, warmup=200)
        line["batched"] = measure_batched(argparse.Namespace(**c3))
    print(json.dumps(line), flush=True)


def run_c1(args):
    from knp_amd import abi, workloads
    from knp_amd.backend import MI355XBackend, device_name, KERNEL_IDS

    t_gen = time.perf_counter()
    spec = workloads.plain_gpu_scale()
    pop = spec.meta["pop"]
    n_neurons = spec.pop_sizes[pop]
    log(f"generated {spec.name}: {n_neurons} neurons, {spec.n_synapses} synapses in {time.perf_counter() - t_gen:.1f}s")

    gpu = MI355XBackend(0)
    gpu.set_scatter_mode({"auto": abi.SCATTER_AUTO, "stepwise": abi.SCATTER_STEPWISE,
                          "atomic": abi.SCATTER_ATOMIC}[args.scatter])
    gpu = spec.load_into(gpu)
    t_up = time.perf_counter()
    gpu._init()
    gpu.synchronize()
    log(f"uploaded in {time.perf_counter() - t_up:.1f}s on {device_name(0)}")
    tuning = gpu.autotune(args.tune_steps) if args.autotune else gpu.tuning()
    log(f"execution choices: {tuning}")
    exactness = {"accumulator_frac_bits": gpu.accumulator_bits(pop), "inexact_weights": gpu.inexact_weights(pop)}

    # the compared run: steps 0 .. P-1 recorded (untimed), then the warmup
    P = max(1, args.parity_steps)
    P64 = min(P, 100)  # the prefix also compared with the FP64 oracle (single-threaded)
    gpu.record_spikes(pop, P)
    gpu.step(P64)
    gpu_pot64 = gpu.potentials(pop)
    gpu.step(P - P64)
    gpu.synchronize()
    gpu_pot = gpu.potentials(pop)
    gpu_events_p = gpu.events
    gpu_prefix = gpu.read_spike_record(pop)
    gpu.step(args.warmup)
    gpu.synchronize()
    log(f"{P} recorded steps and {args.warmup} warmup steps done")

    # ---- timed region: K steps, synchronized on both sides
    gpu.synchronize()
    e0, r0 = gpu.counters()
    t0 = time.perf_counter()
    gpu.step(args.steps)
    gpu.synchronize()
    t1 = time.perf_counter()
    e1, r1 = gpu.counters()
    dt = t1 - t0
    events = e1 - e0
    value = events / dt
    log(f"timed {args.steps} steps in {dt:.3f}s: {events} events, {r1 - r0} rows, {value:.3e} events/s")

    # ---- per-kernel durations (HIP events on the library's stream), separate pass
    tsteps = args.timing_steps
    gpu.enable_timing(True)
    e2, r2 = gpu.counters()
    p2, c2 = gpu.tile_counters()
    gpu.step(tsteps)
    gpu.synchronize()
    e3, r3 = gpu.counters()
    p3, c3 = gpu.tile_counters()
    kern = {}
    for name, kid in KERNEL_IDS.items():
        avg_ms, n = gpu.kernel_time(kid)
        if n:
            kern[name] = {"avg_us": avg_ms * 1e3, "launches": n, "us_per_step": avg_ms * 1e3 * n / tsteps}
    gpu.enable_timing(False)
    gpu.close()
    ev_step = (e3 - e2) / tsteps
    rows_step = (r3 - r2) / tsteps
    parts_step = (p3 - p2) / tsteps
    cells_step = (c3 - c2) / tsteps
    # Algorithmic HBM bytes per step (the per-unit figures of the measurement plan, identical for every path):
    #   spike -> impact scatter: 9 B per impact (post index 4, weight 4, delay/type 1) + 8 B per expanded
    #            presynaptic row (row pointer pair); the ring accumulators are on-chip state and not counted
    #   population update (default BLIFAT, FP32 SoA): 24 B per neuron (ring slot read 8 + clear 8, potential
    #            read/write 4 + 4) + 1 bit of spike output
    scatter_bytes = ev_step * 9.0 + rows_step * 8.0
    neuron_bytes = n_neurons * (24.0 + 1.0 / 8.0)
    if "tile_run" in kern:  # one persistent launch runs both
        dominant, kname, step_bytes = "tile_run", "k_tile_run", scatter_bytes + neuron_bytes
    elif "tile_scatter" in kern:
        dominant, kname, step_bytes = "tile_scatter", "k_tile_scatter", scatter_bytes
        if kern.get("neuron", {}).get("launches", 0) < tsteps:  # phase 1 fused into the tile launch
            step_bytes += neuron_bytes
    else:
        dominant, kname, step_bytes = "scatter", "k_scatter", scatter_bytes
    if "neuron" in kern and kern["neuron"]["launches"] >= tsteps:
        kn = kern["neuron"]
        kn.update(bytes_per_step=neuron_bytes, achieved_GBs=neuron_bytes / (kn["us_per_step"] * 1e-6) / 1e9)
    dk = kern[dominant]
    per_launch = tsteps / dk["launches"]  # steps per launch
    dk.update(bytes_per_step=step_bytes, achieved_GBs=step_bytes / (dk["us_per_step"] * 1e-6) / 1e9)
    if dominant != "scatter":
        dk.update(row_parts_per_step=parts_step, ring_cells_per_step=cells_step)
    roofline = {
        "kernel": kname,
        "bound": "hbm",
        "achieved": round(dk["achieved_GBs"], 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(dk["achieved_GBs"] / HBM_PEAK_GBS, 5),
        "traffic": None,
        "avg_launch_us": round(dk["avg_us"], 3),
        "steps_per_launch": per_launch,
        "bytes_per_launch": round(step_bytes * per_launch, 1),
        "us_per_step": round(dk["us_per_step"], 3),
    }
    roofline.update(profiled_traffic(kname, per_launch))

    # ---- parity and CPU baselines
    cpu = cpu_mt = cpu_release = None
    parity = {"steps_compared": 0}
    if not args.no_cpu_baseline:
        from oracle.pyoracle import Oracle

        threads = max(1, min(16, os.cpu_count() or 1))
        # the compared steps on the multi-threaded oracle (its spikes, potentials and events equal the
        # single-threaded oracle's), timed: the multi-threaded CPU baseline over the same steps
        tb = time.perf_counter()
        mt = spec.load_into(Oracle("exact", threads=threads) if threads > 1 else Oracle("exact"))
        log(f"oracle built in {time.perf_counter() - tb:.1f}s")
        mt.record_spikes(pop, P)
        tm = time.perf_counter()
        mt.step(P)
        tmt = time.perf_counter() - tm
        ora_prefix = mt.read_spike_record(pop)
        same = bool(np.array_equal(ora_prefix, gpu_prefix))
        first = np.argwhere(ora_prefix != gpu_prefix)
        parity = {"steps_compared": int(P), "spike_trains_identical": same,
                  "first_differing_step": int(first[0][0]) if len(first) else None,
                  "spikes_compared": int(ora_prefix.sum()), "events_equal": int(mt.events) == int(gpu_events_p),
                  "events_compared": int(mt.events),
                  "potential_rel_err_max": pot_error(gpu_pot, mt.potentials(pop))[0],
                  "oracle": f"the restated single-threaded backend's step in the multi-threaded backend's "
                            f"decomposition ({threads} threads; equal to one thread), KNP-exact float types",
                  "fp64": compare_f64(spec, gpu_prefix[:P64], gpu_pot64, pop, P64)}
        log(f"parity over {P} steps: identical={same}, events equal={parity['events_equal']}, oracle {tmt:.1f}s")
        cpu_mt = {
            "value": mt.events / tmt, "unit": "events/s", "cores": threads, "kind": "port",
            "sample": f"steps 0..{P - 1} of the same seeded network ({mt.events} events, {tmt:.1f}s), {threads} "
                      f"threads: the multi-threaded backend's decomposition (parallel neuron parts and projection "
                      f"chunks; cpu-multi-threaded-backend/impl/backend.cpp:77-323) of the oracle, -O2 strict "
                      f"IEEE, OpenMP",
            "net_steps_per_s": P / tmt,
        }
        mt.close()
        # the single-threaded oracle on a bounded sample (probe, then ~cpu_seconds of steps)
        ora = spec.load_into(Oracle("exact"))
        tp = time.perf_counter()
        ora.step(10)
        probe = time.perf_counter() - tp
        n_cpu = min(5000, max(10, int(args.cpu_seconds / max(probe / 10, 1e-6))))
        ev_b = ora.events
        tc = time.perf_counter()
        ora.step(n_cpu)
        tcpu = time.perf_counter() - tc
        ev_c = ora.events
        cpu = {
            "value": (ev_c - ev_b) / tcpu,
            "unit": "events/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{spec.name}: steps 10..{10 + n_cpu - 1} of the same seeded network ({n_cpu} steps, "
                      f"{ev_c - ev_b} events, {tcpu:.1f}s), single thread, oracle built -O2 strict IEEE",
            "net_steps_per_s": n_cpu / tcpu,
            "host_cpu": cpu_model(),
            "host_threads": os.cpu_count(),
        }
        log(f"cpu baseline {cpu['value']:.3e} events/s over {n_cpu} steps")
        ora.close()
        # the same single-thread step built with Release-like flags (-Ofast, AVX2 / FMA): a speed figure only
        rl = spec.load_into(Oracle("release"))
        rl.step(10)
        eb = rl.events
        tr = time.perf_counter()
        rl.step(n_cpu)
        trl = time.perf_counter() - tr
        cpu_release = {
            "value": (rl.events - eb) / trl, "unit": "events/s", "cores": 1, "kind": "port",
            "sample": f"the same {n_cpu} steps, single thread, oracle built -Ofast -march=x86-64-v3 (the "
                      f"reference's Release flags with a portable target; not bit-exact)",
            "net_steps_per_s": n_cpu / trl,
        }
        rl.close()
        log(f"cpu baseline (Release-like) {cpu_release['value']:.3e} events/s")

    return {
        "metric": "synaptic events/s",
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generators, device Bernoulli inputs)",
        "config": {
            "workload": "plain_gpu_scale: 100k default BLIFAT (80% exc / 20% inh), in-degree 1000 -> 1e8 "
                        "DeltaSynapses, delays 1..20, Bernoulli(0.005) drive w=1 d=1",
            "neurons": n_neurons,
            "synapses": spec.n_synapses,
            "parallelism": "single GPU",
        },
        "net_steps_per_s": args.steps / dt,
        "events_per_step": events / args.steps,
        "summation": {"tile_run": "tile-owner scatter, persistent launch (ring in LDS)",
                      "tile_scatter": "tile-owner scatter, one launch per step",
                      "scatter": "atomic scatter"}[dominant],
        "roofline": roofline,
        "kernels": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kern.items()},
        "cpu_baseline": cpu,
        "cpu_baseline_mt": cpu_mt,
        "cpu_baseline_release": cpu_release,
        "parity": parity,
        "exactness": exactness,
        "tuning": dict(tuning, tuned=bool(args.autotune)),
        "device": device_name(0),
    }


def profiled_traffic(kname, per_launch):
    """HBM traffic per launch from the committed rocprofv3 counter summary of this kernel, used only when that
    summary was taken of the current kernel source (its recorded source digest matches); else null."""
    import hashlib

    out = {"traffic_source": None}
    path = os.path.join(ROOT, "profiles", f"{kname}_plain_gpu_scale_traffic.json")
    try:
        rec = json.load(open(path))
        digest = hashlib.sha256(open(os.path.join(ROOT, "knp_amd", "csrc", "step_kernels.h"), "rb").read()).hexdigest()
        if rec.get("step_kernels_sha256") == digest:
            out["traffic"] = rec["hbm_bytes_per_step"] * per_launch
            out["traffic_source"] = (f"profiles/{os.path.basename(path)}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                                     f"of this kernel source, scaled by a copy kernel of known size")
        else:
            out["traffic_source"] = f"profiles/{os.path.basename(path)} was taken of another kernel source: not used"
    except (OSError, ValueError, KeyError):
        pass
    return out


def agreement(a, b):
    """1 - |A xor B| / max(1, |A or B|) over (step, neuron) spike triples, and the Hamming accuracy."""
    import numpy as np

    union = int(np.logical_or(a, b).sum())
    diff = int(np.logical_xor(a, b).sum())
    return 1.0 - diff / max(1, union), 1.0 - diff / max(1, a.size)


def pot_error(g, o):
    import numpy as np

    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    rel = np.abs(g - o) / np.maximum(np.abs(o), 1e-6)
    return float(rel.max()) if rel.size else 0.0, float(rel.mean()) if rel.size else 0.0


def timed_steps(gpu, steps):
    gpu.synchronize()
    e0, _ = gpu.counters()
    t0 = time.perf_counter()
    gpu.step(steps)
    gpu.synchronize()
    dt = time.perf_counter() - t0
    e1, _ = gpu.counters()
    return dt, e1 - e0


def compare_f64(spec, gpu_rec, gpu_pot, pop, steps, neurons=None):
    """Spike agreement and potential error of a GPU run against the FP64 oracle on the same spec."""
    from oracle.pyoracle import Oracle

    o = spec.load_into(Oracle("f64"))
    o.record_spikes(pop, steps)
    o.step(steps)
    rec = o.read_spike_record(pop)
    pot = o.potentials(pop)
    o.close()
    if neurons is not None:
        rec, pot = rec[:, neurons], pot[neurons]
        gpu_rec, gpu_pot = gpu_rec[:, neurons], gpu_pot[neurons]
    agr, ham = agreement(gpu_rec, rec)
    pmax, pmean = pot_error(gpu_pot, pot)
    return {"steps": steps, "spike_agreement": agr, "hamming_accuracy": ham, "spikes_oracle": int(rec.sum()),
            "spikes_gpu": int(gpu_rec.sum()), "potential_rel_err_max": pmax, "potential_rel_err_mean": pmean}


def _f64_worker(job):
    from oracle.pyoracle import Oracle

    spec, steps = job
    o = spec.load_into(Oracle("f64"))
    o.record_spikes(0, steps)
    o.step(steps)
    out = (o.read_spike_record(0), o.potentials(0))
    o.close()
    return out


def f64_runs(specs, labels, steps):
    """FP64 oracle runs of single-instance specs, in parallel threads (the oracle's step runs without the
    interpreter lock): [(spike record, potentials)]."""
    from multiprocessing.pool import ThreadPool

    t0 = time.perf_counter()
    n = max(1, min(len(specs), 8, os.cpu_count() or 1))
    with ThreadPool(n) as pool:
        out = pool.map(_f64_worker, [(s, steps) for s in specs])
    log(f"fp64 oracle: {len(specs)} instances x {steps} steps in {time.perf_counter() - t0:.1f}s")
    return out


def agreement_report(g_rec, o_rec, g_pot, o_pot, steps):
    agr, ham = agreement(g_rec[:steps], o_rec[:steps])
    first = np.argwhere(g_rec[:steps] != o_rec[:steps])
    pmax, pmean = pot_error(g_pot, o_pot)
    return {"spike_agreement": round(agr, 6), "hamming_accuracy": round(ham, 6),
            "spikes_gpu": int(g_rec[:steps].sum()), "spikes_oracle": int(o_rec[:steps].sum()),
            "first_differing_step": int(first[0][0]) if len(first) else None,
            "potential_rel_err_max_at_end": pmax, "potential_rel_err_mean_at_end": pmean}


def device_loaders(args):
    """(load(spec) -> initialized device network, kernels(g, steps) -> per-kernel HIP-event durations)."""
    from knp_amd import abi
    from knp_amd.backend import KERNEL_IDS, MI355XBackend

    def load(spec):
        g = MI355XBackend(0)
        g.set_scatter_mode({"auto": abi.SCATTER_AUTO, "stepwise": abi.SCATTER_STEPWISE,
                            "atomic": abi.SCATTER_ATOMIC}[args.scatter])
        if args.train_wg_limit is not None:
            g.set_train_workgroup_limit(args.train_wg_limit)
        if args.graph_steps is not None:
            g.set_step_graph(args.graph_steps)
        spec.load_into(g)
        g._init()
        g.synchronize()
        g.tuned = g.autotune(args.tune_steps) if args.autotune else g.tuning()
        log(f"{spec.name}: execution choices {g.tuned}")
        return g

    def kernels(g, steps):
        g.enable_timing(True)
        g.step(steps)
        g.synchronize()
        out = {}
        for name, kid in KERNEL_IDS.items():
            avg_ms, n = g.kernel_time(kid)
            if n:
                out[name] = {"avg_us": round(avg_ms * 1e3, 3), "launches": n,
                             "us_per_step": round(avg_ms * 1e3 * n / steps, 3)}
        g.enable_timing(False)
        return out

    return load, kernels


def plain_stdp_run(spec2, load, kernels, agree, warm, steps):
    """One STDP instance on the plain FP32 path: `agree` recorded steps, warmup to `warm`, `steps` timed."""
    g2 = load(spec2)
    g2.record_spikes(0, agree)
    g2.step(agree)
    rec2, pot2 = g2.read_spike_record(0), g2.potentials(0)
    g2.step(warm - agree)
    dt2, ev2 = timed_steps(g2, steps)
    k2 = kernels(g2, min(steps, 500))
    g2.close()
    log(f"plain fp32 instance: {steps / dt2:.0f} steps/s")
    return ({"net_steps_per_s": steps / dt2, "events_per_s": ev2 / dt2, "steps": steps,
             "ms_per_step": dt2 / steps * 1e3, "kernels": k2, "path": "fp32 event-driven scatter",
             "tuning": g2.tuned}, rec2, pot2)


def measure_batched(args):
    """The batched configuration: args.batch instances of the STDP network stepped together (the plastic input
    projection and the lateral one expanded on MFMA per args.dense / args.split), against one instance on the
    plain FP32 path timed over the same step range; spike agreement of args.agree_instances instances (and of
    the plain path's instance) with the FP64 oracle over args.agree_steps steps, by horizon."""
    from knp_amd import workloads

    load, kernels = device_loaders(args)
    steps = args.steps if args.steps != 10_000 else 1000
    agree = args.agree_steps
    warm = max(args.warmup, agree)  # the plain and the batched runs time the same step range
    n_img = (warm + steps) // 20 + 1
    src_seed = 30 * 7919 + 17
    # the plain FP32 path, one instance (the batch's instance 0): the denominator of the batching ratio
    spec2 = workloads.stdp_learning(seed=30, n_images=n_img, source_seed=src_seed, instance=0)
    c2_line, rec2, pot2 = plain_stdp_run(spec2, load, kernels, agree, warm, steps)
    B = args.batch
    tg = time.perf_counter()
    spec3 = workloads.batched_stdp(batch=B, n_images=n_img, dense=None if args.dense == "none" else args.dense,
                                   split=args.split)
    log(f"generated {B} instances ({spec3.n_synapses} synapses) in {time.perf_counter() - tg:.1f}s")
    tl = time.perf_counter()
    g = load(spec3)
    log(f"loaded in {time.perf_counter() - tl:.1f}s")
    tuning3 = g.tuned
    check = list(range(min(B, args.agree_instances)))
    g.record_spikes(0, agree)
    g.step(agree)
    rec = g.read_spike_record(0)
    pot = g.potentials(0)
    g.step(warm - agree)
    dt3, ev3 = timed_steps(g, steps)
    ksteps = min(steps, 500)
    wb0, full = g.dense_weight_bytes()
    k3 = kernels(g, ksteps)
    wb1, _ = g.dense_weight_bytes()
    g.close()
    inst_steps_per_s = B * steps / dt3
    log(f"batched: {inst_steps_per_s:.3e} instance-steps/s")
    parts = [workloads.stdp_learning(seed=30 + b, n_images=n_img, source_seed=src_seed, instance=b) for b in check]
    f64 = f64_runs(parts, check, agree)
    per = [agreement_report(rec[:, b * 1024:(b + 1) * 1024], f64[i][0], pot[b * 1024:(b + 1) * 1024], f64[i][1], agree)
           for i, b in enumerate(check)]
    allg = np.concatenate([rec[:, b * 1024:(b + 1) * 1024] for b in check], axis=1)
    allo = np.concatenate([x[0] for x in f64], axis=1)
    horizons = [h for h in (100, 500, 1000, 2000, 5000) if h <= agree]
    batch_agr = {str(h): round(agreement(allg[:h], allo[:h])[0], 6) for h in horizons}
    plain_agr = {str(h): round(agreement(rec2[:h], f64[0][0][:h])[0], 6) for h in horizons}
    roof = None
    if "dense_expand" in k3:
        d = k3["dense_expand"]
        bytes_step = full + (wb1 - wb0) / ksteps
        ach = bytes_step / (d["us_per_step"] * 1e-6) / 1e9
        roof = {"kernel": "k_dense_counts + k_dense_expand (both MFMA projections)", "bound": "hbm",
                "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 5), "weight_bytes_per_step": bytes_step,
                "dense_equivalent_bytes_per_step": float(B * 784 * 1024 * (4 if args.split else 2)),
                "useful_mfma_flops_per_step": 2.0 * B * 784 * 1024, "us_per_step": d["us_per_step"]}
    return dict(
        metric="instance-steps/s", value=inst_steps_per_s, unit="instance-steps/s",
        dtype="fp32" if args.dense == "none" else
        ("bf16 hi+lo" if args.split else "bf16") + " weights x spike counts, fp32 accumulate",
        workload=f"batched_stdp: {B} instances of the STDP network (784 -> 1,024 synaptic-resource STDP, lateral "
                 f"inhibition, dopamine), own seeds and image streams",
        dense=args.dense, split=args.split, batch=B, steps=steps, ms_per_step=dt3 / steps * 1e3,
        events_per_s=ev3 / dt3, ratio_to_plain_fp32=inst_steps_per_s / c2_line["net_steps_per_s"],
        plain_fp32=c2_line, kernels=k3, dense_roofline=roof, tuning=tuning3,
        fp64={"instances": check, "steps": agree, "spike_agreement_by_horizon": batch_agr,
              "plain_fp32_instance0_by_horizon": plain_agr, "per_instance": per})


def other_workload(args):
    """Measurement lines of the other configurations (not the headline line)."""
    from knp_amd import workloads
    from knp_amd.backend import device_name

    load, kernels = device_loaders(args)
    line = {"workload": args.workload, "device": device_name(0), "data": "synthetic (seeded generators)"}
    if args.workload == "c0":
        from oracle.pyoracle import Oracle

        spec = workloads.plain_small()
        steps = spec.meta["steps"]
        # the measured run starts at step 0 (its spikes are compared with the oracle's), so the one-time
        # costs of a first launch (code objects loaded on first use) are paid by a throwaway copy
        w = load(spec)
        w.step(max(args.warmup, 2))
        w.synchronize()
        w.close()
        g = load(spec)
        g.record_spikes(0, steps)
        dt, ev = timed_steps(g, steps)
        rec, pot = g.read_spike_record(0), g.potentials(0)
        o = spec.load_into(Oracle("exact"))
        o.record_spikes(0, steps)
        tc = time.perf_counter()
        o.step(steps)
        tcpu = time.perf_counter() - tc
        # the same steps with the multi-threaded decomposition and with Release-like flags
        others = {}
        threads = max(1, min(16, os.cpu_count() or 1))
        for key, kw, what in (("cpu_baseline_mt", {"threads": threads}, f"{threads} threads (the multi-threaded "
                               "backend's decomposition, OpenMP)"),
                              ("cpu_baseline_release", {"policy": "release"}, "single thread, -Ofast "
                               "-march=x86-64-v3 (not bit-exact)")):
            if key == "cpu_baseline_mt" and threads == 1:
                continue
            x = spec.load_into(Oracle(**kw))
            tx = time.perf_counter()
            x.step(steps)
            tx = time.perf_counter() - tx
            others[key] = {"value": x.events / tx, "unit": "events/s",
                           "cores": threads if key == "cpu_baseline_mt" else 1, "kind": "port",
                           "net_steps_per_s": steps / tx, "sample": f"all {steps} steps, {what}"}
            x.close()
        line.update(metric="synaptic events/s", value=ev / dt, unit="events/s", net_steps_per_s=steps / dt,
                    steps=steps, ms_per_step=dt / steps * 1e3, events=ev, dtype="fp32", tuning=g.tuned,
                    parity_knp_exact={"spike_trains_identical": bool(np.array_equal(rec, o.read_spike_record(0))),
                                      "events_equal": ev == o.events,
                                      "potential_rel_err_max": pot_error(pot, o.potentials(0))[0]},
                    fp64=compare_f64(spec, rec, pot, 0, steps),
                    cpu_baseline={"value": o.events / tcpu, "unit": "events/s", "cores": 1, "kind": "port",
                                  "net_steps_per_s": steps / tcpu, "sample": f"all {steps} steps"}, **others)
        o.close()
    elif args.workload == "c2":
        steps = args.steps if args.steps != 10_000 else 1000
        agree = args.agree_steps
        warm = max(args.warmup, agree)
        spec2 = workloads.stdp_learning(seed=30, n_images=(warm + steps) // 20 + 1, source_seed=30 * 7919 + 17,
                                        instance=0)
        c2_line, rec2, pot2 = plain_stdp_run(spec2, load, kernels, agree, warm, steps)
        f64 = f64_runs([spec2], [0], agree)[0]
        line.update(metric="network steps/s", value=c2_line["net_steps_per_s"], unit="net-steps/s", dtype="fp32",
                    fp64=agreement_report(rec2, f64[0], pot2, f64[1], agree), **c2_line)
    elif args.workload == "c3":
        line.update(measure_batched(args))
    elif args.workload == "c4":
        steps = args.steps if args.steps != 10_000 else 2000
        tg = time.perf_counter()
        spec = workloads.uneven_inference(steps=steps)
        log(f"generated c4 ({spec.n_synapses} sparse synapses) in {time.perf_counter() - tg:.1f}s")
        N = spec.pop_sizes[0]
        g = load(spec)
        tuning4 = g.tuned
        g.step(args.warmup)
        dt, ev = timed_steps(g, steps)
        k4 = kernels(g, min(steps, 1024))
        T = g.dense_time_block(0)
        g.close()
        d = k4.get("dense", {})
        # dense product per launch: 2 x T steps x 784 inputs x N neurons (binary spikes x FP8 weights)
        flops = 2.0 * T * 784 * N
        mfma = None
        if d:
            ach = flops / (d["avg_us"] * 1e-6) / 1e12
            mfma = {"kernel": "k_dense_mfma (+ spike bits)", "bound": "mfma", "achieved": round(ach, 2),
                    "peak": 5000.0, "unit": "TFLOP/s", "frac": round(ach / 5000.0, 5), "time_block": T,
                    "flops_per_launch": flops}
        # agreement on a reduced batch (the FP64 oracle expands every dense weight into a synapse): against the
        # FP64 oracle with the same FP8 weights (the kernels' exactness), and with the weights before
        # quantization (what FP8 costs), by horizon
        agree = min(args.agree_steps, 2000)
        small = workloads.uneven_inference(n_instances=16, steps=agree)
        gs = load(small)
        gs.record_spikes(0, agree)
        gs.step(agree)
        rec, pot = gs.read_spike_record(0), gs.potentials(0)
        gs.close()
        runs = f64_runs([small, workloads.unquantized_dense(small)], [0, 1], agree)
        horizons = [h for h in (100, 500, 2000) if h <= agree]
        fp64 = agreement_report(rec, runs[0][0], pot, runs[0][1], agree)
        fp64["instances"] = 16
        unq = agreement_report(rec, runs[1][0], pot, runs[1][1], agree)
        unq["by_horizon"] = {str(h): round(agreement(rec[:h], runs[1][0][:h])[0], 6) for h in horizons}
        unq["instances"] = 16
        line.update(metric="synaptic events/s", value=ev / dt, unit="events/s", dtype="fp32 sparse + fp8 dense",
                    neurons=N, instances=512, steps=steps, ms_per_step=dt / steps * 1e3, net_steps_per_s=steps / dt,
                    instance_steps_per_s=512 * steps / dt, dense_roofline=mfma, kernels=k4, fp64=fp64, tuning=tuning4,
                    fp64_unquantized_weights=unq)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
