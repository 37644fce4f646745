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
    ap.add_argument("--parity-steps", type=int, default=100)
    ap.add_argument("--timing-steps", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample length")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--scatter", choices=["auto", "stepwise", "atomic"], default="auto",
                    help="impact summation: tiled populations and persistent launches where eligible (auto), "
                         "tiled with one launch per step (stepwise), or atomics everywhere")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the step of one network runs on one MI355X "
              "(multi-GPU sharding is a design note, not part of this build). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from knp_amd import workloads
    from knp_amd.backend import MI355XBackend, device_name, KERNEL_IDS

    t_gen = time.perf_counter()
    spec = workloads.plain_gpu_scale()
    pop = spec.meta["pop"]
    n_neurons = spec.pop_sizes[pop]
    log(f"generated {spec.name}: {n_neurons} neurons, {spec.n_synapses} synapses in {time.perf_counter() - t_gen:.1f}s")

    from knp_amd import abi

    gpu = MI355XBackend(0)
    gpu.set_scatter_mode({"auto": abi.SCATTER_AUTO, "stepwise": abi.SCATTER_STEPWISE,
                          "atomic": abi.SCATTER_ATOMIC}[args.scatter])
    gpu = spec.load_into(gpu)
    t_up = time.perf_counter()
    gpu._init()
    gpu.synchronize()
    log(f"uploaded in {time.perf_counter() - t_up:.1f}s on {device_name(0)}")

    parity_steps = min(args.parity_steps, args.warmup)
    gpu.record_spikes(pop, parity_steps)
    gpu.step(args.warmup)
    gpu.synchronize()
    gpu_prefix = gpu.read_spike_record(pop)
    log(f"warmup {args.warmup} steps done")

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
    rows = r1 - r0
    value = events / dt
    log(f"timed {args.steps} steps in {dt:.3f}s: {events} events, {rows} rows, {value:.3e} events/s")

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
    traffic_file = os.path.join(ROOT, "profiles", f"{kname}_plain_gpu_scale_traffic.json")
    if os.path.exists(traffic_file):
        try:
            # HBM bytes per step measured with rocprofv3 counters, scaled to this launch length
            roofline["traffic"] = json.load(open(traffic_file))["hbm_bytes_per_step"] * per_launch
        except (OSError, ValueError, KeyError):
            pass

    # ---- CPU baseline: the oracle (restated single-threaded backend), bounded sample, 1 thread
    cpu = None
    parity = {"steps_compared": 0}
    if not args.no_cpu_baseline:
        import numpy as np

        from oracle.pyoracle import Oracle

        tb = time.perf_counter()
        ora = spec.load_into(Oracle("exact"))
        log(f"oracle built in {time.perf_counter() - tb:.1f}s")
        ora.record_spikes(pop, parity_steps)
        ora.step(parity_steps)
        ora_prefix = ora.read_spike_record(pop)
        same = bool(np.array_equal(ora_prefix, gpu_prefix))
        parity = {"steps_compared": int(parity_steps), "spike_trains_identical": same,
                  "spikes_compared": int(ora_prefix.sum())}
        # probe, then a sample sized to ~cpu_seconds
        ev_a = ora.events
        tp = time.perf_counter()
        ora.step(10)
        probe = time.perf_counter() - tp
        n_cpu = max(10, int(args.cpu_seconds / max(probe / 10, 1e-6)))
        n_cpu = min(n_cpu, 5000)
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
            "sample": f"{spec.name}: steps {parity_steps + 10}..{parity_steps + 10 + n_cpu - 1} of the same seeded "
                      f"network ({n_cpu} steps, {ev_c - ev_b} events, {tcpu:.1f}s), single thread, "
                      f"oracle built -O2 strict IEEE",
            "net_steps_per_s": n_cpu / tcpu,
            "host_cpu": cpu_model(),
            "host_threads": os.cpu_count(),
        }
        _ = ev_a
        log(f"cpu baseline {cpu['value']:.3e} events/s over {n_cpu} steps")
        ora.close()

    line = {
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
        "parity": parity,
        "device": device_name(0),
    }
    gpu.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
