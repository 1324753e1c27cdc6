"""Benchmark of the dynamics hot path on one MI355X.

Workload (the batched-flow configuration): R = 1024 independent Line1d
Cuspy-Laplace replicas of N = 4096 particles (replica r built with seed r),
FP64, m = 1, eta = 2 sqrt(3) / 10, mu = 1, k_interactions = 1, k_frame = 1/N,
dt = 0.1, Weibull(k=2, scale=1) yield increments in chunks of 200 wells,
driven by flowSteps at v_frame = 1e-3. One bench step = one resident launch of
1000 flowSteps over all replicas (4.19e9 particle-steps); the default
--steps 100 is the full 1e5-step run.

Prints one JSON line: whole-job particle-steps/s of the batched path, the
plain one-system path beside it, the roofline of the resident kernel (FP64
vector ALU, 40 algorithmic FLOP per particle-step, timed with HIP events on
the launch stream) and the CPU oracle timed on a bounded sample.
"""

import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

R = 1024
N = 4096
BLOCK = 1000
V_FRAME = 1e-3
FLOP_PER_PARTICLE_STEP = 40.0  # Cuspy + Laplace1d Velocity-Verlet step, see DESIGN.md
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (= FP64 matrix) peak
METRIC = "particle-steps/sec batched vs plain on 1 GPU (FP64) + % HBM/MFMA roofline"

PARAMS = dict(
    m=1.0,
    eta=2.0 * np.sqrt(3.0) / 10.0,
    mu=1.0,
    k_interactions=1.0,
    k_frame=1.0 / N,
    dt=0.1,
    shape=[N],
    seed=0,
    distribution="weibull",
    parameters=[2.0, 1.0, 0.0],
    offset=-100.0,
    nchunk=200,
)


def cpu_baseline(replicas=256, target_s=10.0):
    """The oracle on the host: `replicas` systems of the same workload, one per
    thread, flowSteps sized from a short pilot to take about `target_s`."""
    from oracle import pyoracle

    pyoracle.lib()
    cores = min(16, os.cpu_count() or 1)
    systems = [pyoracle.Line1d.System_Cuspy_Laplace(**dict(PARAMS, seed=r)) for r in range(replicas)]

    def sweep(n):
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(max_workers=cores) as ex:
            list(ex.map(lambda s: s.flowSteps(n, V_FRAME), systems))
        return time.perf_counter() - t0

    pilot = 200
    dt_pilot = sweep(pilot)
    steps = max(pilot, int(pilot * target_s / max(dt_pilot, 1e-3)))
    dt = sweep(steps)
    return {
        "value": replicas * N * steps / dt,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{replicas} replicas x N={N} x {steps} flowSteps of the same workload, "
        f"one replica per thread ({dt:.1f} s)",
    }


def plain_path(steps=20000):
    """One system per launch through the reference's API (System_Cuspy_Laplace)."""
    import torch

    import frictionqpotspringblock_amd as fq

    s = fq.Line1d.System_Cuspy_Laplace(**PARAMS)
    s.flowSteps(1000, V_FRAME)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.flowSteps(steps, V_FRAME)
    torch.cuda.synchronize()
    return N * steps / (time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed resident blocks of 1000 flowSteps")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain one-system path")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: replicas are independent and would shard across GPUs "
            "without communication; run with --gpus 1",
            file=sys.stderr,
        )
        sys.exit(2)

    import torch

    from frictionqpotspringblock_amd.batch import Batch

    torch.cuda.init()
    kw = dict(PARAMS)
    batch = Batch(R, seed_stride=1, potential="Cuspy", interactions="Laplace1d", **kw)
    tuned = {}
    if not args.no_autotune:
        tuned = batch.autotune(V_FRAME, steps=BLOCK)
    for _ in range(args.warmup):
        batch.flowSteps(BLOCK, V_FRAME, block=BLOCK)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev_ms = batch.flowSteps(args.steps * BLOCK, V_FRAME, block=BLOCK, timed=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert np.all(np.isfinite(batch.u_frame))

    particle_steps = R * N * BLOCK * args.steps
    value = particle_steps / wall
    kernel_ms = ev_ms / args.steps  # average duration of one resident launch
    flop = FLOP_PER_PARTICLE_STEP * R * N * BLOCK
    achieved = flop / (kernel_ms * 1e-3) / 1e12

    plain = None if args.no_plain else plain_path()
    cpu = None if args.no_cpu else cpu_baseline()
    ppt, threads = batch.launch
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Weibull yield landscape from pcg32 seeds 0..1023)",
        "config": {
            "workload": f"batched flow: {R} Line1d Cuspy-Laplace replicas x N={N}, "
            f"{BLOCK}-step resident launches, v_frame={V_FRAME}",
            "replicas": R,
            "N": N,
            "flow_steps": BLOCK * args.steps,
            "nchunk": 200,
            "launch": {"particles_per_thread": ppt, "threads": threads},
            "autotune_ms": {f"{p}x{t}": round(ms, 3) for (p, t), ms in tuned.items()},
        },
        "plain_value": plain,
        "batched_over_plain": None if plain is None else value / plain,
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None,
            "kernel": "k_resident<Cuspy, Laplace1d>",
            "kernel_ms": kernel_ms,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
