"""Benchmark of the dynamics hot path on one MI355X.

Default workload (--workload flow, the batched-flow configuration): R = 1024
independent Line1d Cuspy-Laplace replicas of N = 4096 particles (replica r
built with seed r), FP64, m = 1, eta = 2 sqrt(3) / 10, mu = 1,
k_interactions = 1, k_frame = 1/N, dt = 0.1, Weibull(k=2, scale=1) yield
increments in chunks of 200 wells, driven by flowSteps at v_frame = 1e-3. One
bench step = one resident launch of 1000 flowSteps over all replicas (4.19e9
particle-steps); the default --steps 100 is the full 1e5-step run.

--workload adjoint: R = 64 Cuspy-Laplace replicas of N = 4096 (parameters as
the flow workload, random initial velocities); one bench step = one forward
pass of T = 2000 timeSteps with a checkpoint every 100 steps plus its adjoint
(torch.autograd backward) for loss = mean over replicas of |u(T) - u_t|^2,
where u_t comes from replicas with mu * 1.05, k * 0.95, eta * 1.1; gradients
with respect to u0, v0, a0, (mu, k, k_frame, eta) and the yield positions.

--workload ragged: 512 Line1d systems with N = floor(2^U(9, 16)) (seeded draw),
potentials cycling Cuspy / SemiSmooth (kappa 1) / Smooth, 30% Quartic (a1 1,
a2 5) and 70% Laplace (k 1) interactions, plus 8 Line2d Cuspy-Laplace 256 x
256; one bench step = one quasistatic event of the examples' protocol on all
systems (elastic step except for Smooth, kick, minimise), then the same number
of events with the no-passing minimiser on the Cuspy-Laplace chains (up to
8192 particles); value = sum over systems of N x (time steps or sweeps) per
second.

--workload quasistatic: R = 1024 replicas of the reference's CPU configuration
(Line1d Cuspy-Laplace, N = 1024, parameters as above with k_frame = 1/1024,
replica r with seed r), each running the examples' quasistatic protocol
(examples/Line1d_Cuspy_Laplace.py:39-61) at its own pace on the device; one
bench step = one event pair (elastic step to the next yield, then kick and
minimise) on every replica, avalanche sizes recorded; the default --steps 100
is the configuration's 100 events. value = sum over replicas of N x time steps
per second.

--workload longrange: R = 256 replicas of N = 8192 particles with LongRange1d
interactions (alpha = 1.5, k = 1) and random normal forcing (mean 0, stddev
0.1, dinc = dinc_init = 10), otherwise as above; timeSteps with the
interaction forces of all replicas as one FP64 matrix-core product per step.
One bench step = 100 timeSteps (5.4e7 particle-steps); the default --steps
200 is the full 2e4-step run.

Prints one JSON line: whole-job particle-steps/s of the batched path, the
plain one-system path beside it, the roofline of the dominant kernel (timed
with HIP events on the launch stream) against the published MI355X peak, the
CPU oracle timed on a bounded sample, and ("peaks_measured", after the timed
region) the two peaks measured on this GPU: device-to-device copy bandwidth
and FP64 matrix-product throughput.
"""

import argparse
import concurrent.futures as cf
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

R = 1024
N = 4096
BLOCK = 1000
V_FRAME = 1e-3
FLOP_PER_PARTICLE_STEP = 40.0  # Cuspy + Laplace1d Velocity-Verlet step
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (= FP64 matrix) peak
# HBM bytes per 1000-step launch of k_flow<Cuspy, Laplace1d, 4 x 1024> on this
# workload, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch),
# each scaled by the ratio of the true bytes of a 1 GiB device copy to what
# the same counter reads for it (reads x 2.000, writes x 1.000);
# profiles/k_flow_cuspy_laplace_flow_R1024_N4096.md
FLOW_TRAFFIC_BYTES = (520_371 * 2.000 + 869_879 * 1.000) * 1024

LR_R = 256
LR_N = 8192
LR_BLOCK = 100

RG_EPS = 1e-3
RG_TUNE_MIN_PARTICLES = 262144  # groups tuned by the bench: at least this many particles
# a trial choice replaces a group's default only if faster by this fraction:
# the trials time each group alone, and choices that won by 10-15% there
# (shorter blocks) made the concurrent run 2-5% slower
RG_TUNE_MARGIN = 0.25
QS_R = 1024
QS_N = 1024
QS_EPS = 1e-3
FLOP_PER_PARTICLE_STEP_MIN = 44.0  # the Verlet step plus the residual norms of the StopList
QS_SHAPES = [(1, 1024), (2, 512), (4, 256), (8, 128)]  # (particles per thread, threads) for N = 1024

AD_R = 64
AD_N = 4096
AD_T = 2000
AD_K = 100
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "particle-steps/sec batched vs plain on 1 GPU (FP64) + % HBM/MFMA roofline"

# CPUs a one-GPU job may use on the GPU host (the host's share per GPU);
# the CPU baselines run one system per thread on that many threads
CPU_SHARE = 16


def cpu_threads():
    """Threads of the CPU baselines: the CPUs this process may run on, at
    most the one-GPU share of the host."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(CPU_SHARE, avail))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


ORACLE_SRC = os.path.join(ROOT, "oracle", "fqsb_oracle.c")
ORACLE_BUILDS = {
    # name -> (gcc flags, what it is)
    "parity": (None, "oracle/_build (gcc -O3 -ffp-contract=off, generic x86-64): the parity build"),
    "native": (
        ["-O3", "-march=native", "-ffp-contract=fast"],
        "gcc -O3 -march=native -ffp-contract=fast: this host's vector ISA, contractions on (timing only)",
    ),
}


def _oracle_library(build):
    """Path of an oracle build (None: the parity build of oracle/_build); the
    native build is compiled here, for this host's CPU, into a temporary
    directory."""
    flags, _ = ORACLE_BUILDS[build]
    if flags is None:
        return None
    out = os.path.join(tempfile.gettempdir(), f"fqsb_oracle_{build}_{os.getpid()}.so")
    if not os.path.exists(out):
        subprocess.run(["gcc", *flags, "-fPIC", "-std=c11", "-shared", "-o", out, ORACLE_SRC, "-lm", "-pthread"],
                       check=True)
    return out


def cpu_leg(workload, args):
    """The CPU oracle's rate on a bounded sample of `workload`, timed in a
    child process per oracle build (the parity build and the build for this
    host's ISA with contractions on); the faster build is the baseline."""
    builds = {}
    for build in ("parity", "native"):
        env = dict(os.environ)
        env.pop("FQO_LIB", None)
        try:
            lib = _oracle_library(build)
        except (OSError, subprocess.CalledProcessError) as e:
            builds[build] = {"error": str(e)}
            continue
        if lib:
            env["FQO_LIB"] = lib
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", workload, "--cpu-seconds",
               str(args.cpu_seconds)]
        r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            builds[build] = {"error": r.stderr.strip()[-400:]}
            continue
        res = json.loads(r.stdout.strip().splitlines()[-1])
        res["flags"] = ORACLE_BUILDS[build][1]
        builds[build] = res
    ok = {k: v for k, v in builds.items() if "value" in v}
    if not ok:
        return {"error": builds}
    best = max(ok, key=lambda k: ok[k]["value"])
    out = dict(ok[best])
    out["build"] = best
    out["builds"] = {k: ({"value": v["value"], "flags": v["flags"]} if "value" in v else v) for k, v in builds.items()}
    out["cpu_model"] = cpu_model()
    out["host_cpus"] = os.cpu_count()
    out["cpus_allowed"] = cpu_threads()
    return out


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


LR_PARAMS = dict(
    PARAMS,
    k_frame=1.0 / LR_N,
    shape=[LR_N],
    alpha=1.5,
    mean=0.0,
    stddev=0.1,
    seed_forcing=0,
    dinc_init=10,
    dinc=10,
)


def cpu_baseline(replicas=256, target_s=10.0):
    """The oracle on the host: `replicas` systems of the same workload, one per
    thread, flowSteps sized from a short pilot to take about `target_s`."""
    from oracle import pyoracle

    pyoracle.lib()
    cores = cpu_threads()
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


def lr_cpu_baseline(replicas=16, target_s=10.0):
    """The oracle on the host for the long-range workload: `replicas` systems
    (one per thread) of the same physics, timeSteps sized from a pilot to take
    about `target_s`. Each step is the reference's O(N^2) sequential sum."""
    from oracle import pyoracle

    pyoracle.lib()
    cores = cpu_threads()
    replicas = max(replicas, cores)
    kw = dict(LR_PARAMS)
    systems = [
        pyoracle._System(
            rank=1, potential="Cuspy", interactions="LongRange1d", external=1, minimisation=2,
            **dict(kw, seed=r * LR_N, seed_forcing=r * LR_N),
        )
        for r in range(replicas)
    ]

    def sweep(n):
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(max_workers=cores) as ex:
            list(ex.map(lambda s: s.timeSteps(n), systems))
        return time.perf_counter() - t0

    pilot = 2
    dt_pilot = sweep(pilot)
    steps = max(pilot, int(pilot * target_s / max(dt_pilot, 1e-3)))
    dt = sweep(steps)
    return {
        "value": replicas * LR_N * steps / dt,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{replicas} replicas x N={LR_N} x {steps} timeSteps of the same workload, "
        f"one replica per thread ({dt:.1f} s)",
    }


def lr_batch(replicas, method):
    from frictionqpotspringblock_amd.batch import Batch

    b = Batch(
        replicas, seed_stride=LR_N, potential="Cuspy", interactions="LongRange1d",
        external="RandomNormalForcing", **LR_PARAMS,
    )
    b.longrange_method = method
    return b


def lr_plain_path(steps=200):
    """One system per launch: the reference's sequential long-range sum."""
    b = lr_batch(1, "exact")
    b.timeSteps(10)
    ms = b.timeSteps(steps, timed=True)
    return LR_N * steps / (ms * 1e-3)


def run_longrange(args):
    import torch

    torch.cuda.init()
    batch = lr_batch(LR_R, "matrix")
    for _ in range(args.warmup):
        batch.timeSteps(LR_BLOCK, block=LR_BLOCK)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev_ms = batch.timeSteps(args.steps * LR_BLOCK, block=LR_BLOCK, timed=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert np.all(np.isfinite(batch.u[:, :8]))
    product_ms = batch.longrange_product_ms
    steps_total = args.steps * LR_BLOCK
    value = LR_R * LR_N * steps_total / wall
    flop = 2.0 * LR_N * LR_N * LR_R  # per product launch
    achieved = flop / (product_ms * 1e-3) / 1e12
    plain = None if args.no_plain else lr_plain_path()
    cpu = None if args.no_cpu else cpu_leg("longrange", args)
    return {
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
        "data": f"synthetic (Weibull yield landscape and normal forcing from pcg32 seeds r*{LR_N})",
        "config": {
            "workload": f"batched long-range: {LR_R} Line1d Cuspy-LongRange(alpha=1.5) + random-forcing "
            f"replicas x N={LR_N}, timeSteps in blocks of {LR_BLOCK}",
            "replicas": LR_R,
            "N": LR_N,
            "time_steps": steps_total,
            "events_ms": ev_ms,
        },
        "plain_value": plain,
        "batched_over_plain": None if plain is None else value / plain,
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None,
            "kernel": "k_lr_gemm",
            "kernel_ms": product_ms,
        },
        "cpu_baseline": cpu,
    }


LRQ_R = 16
LRQ_EVENTS = 5


def run_lr_quasistatic(args):
    """The reference's long-range protocol (examples/Line1d_Cuspy_Laplace_LongRange.py:
    39-61: eventDrivenStep, then minimise) on the long-range configuration's
    chain (N = 8192, alpha 1.5, without the random forcing, whose class has no
    minimise): a batch of LRQ_R replicas with the forces of every minimise
    step as one FP64 matrix-core product, and one system on the exact
    sequential sum spread over tiles (the reference's API)."""
    import torch

    import frictionqpotspringblock_amd as fq
    from frictionqpotspringblock_amd.batch import Batch

    torch.cuda.init()
    kw = dict(PARAMS, k_frame=1.0 / LR_N, shape=[LR_N], alpha=1.5)
    b = Batch(LRQ_R, seed_stride=LR_N, potential="Cuspy", interactions="LongRange1d", **kw)
    b.quasistatic(QS_EPS, 1)  # event 0
    inc0 = b.inc.copy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    q = b.quasistatic(QS_EPS, 2 * args.steps, first_event=1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    steps = b.inc - inc0
    work = float(LR_N * steps.sum())
    product_ms = b.longrange_product_ms
    plain = None
    if not args.no_plain:
        s = fq.Line1d.System_Cuspy_LongRange(**kw)
        s.u_frame = 0.0
        s.minimise()
        i0 = s.inc
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(max(1, args.steps // 2)):
            s.eventDrivenStep(QS_EPS, False)
            s.eventDrivenStep(QS_EPS, True)
            s.minimise()
        torch.cuda.synchronize()
        plain = LR_N * (s.inc - i0) / (time.perf_counter() - t1)
    return {
        "metric": METRIC,
        "value": work / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 0,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (Weibull yield landscapes from pcg32 seeds r*{LR_N})",
        "config": {
            "workload": f"long-range quasistatic: {LRQ_R} Line1d Cuspy-LongRange(alpha=1.5) replicas x N={LR_N}, "
            f"{args.steps} event pairs of the examples' protocol, minimise in matrix form",
            "time_steps_per_replica_mean": float(steps.mean()),
            "avalanche_size_mean": float(np.abs(q["S"][:, 1::2]).mean()),
        },
        "plain_value": plain,
        "plain": "one System_Cuspy_LongRange through the reference's API (exact sum over tiles of 256 particles)",
        "roofline": None,
        "roofline_note": "minimise steps carry one matrix product each; see the longrange workload for k_lr_gemm",
        "cpu_baseline": None,
    }


def ad_cpu_baseline(replicas=16, target_s=10.0):
    """The oracle's forward timeSteps of the adjoint workload on the host (one
    replica per thread). The oracle has no adjoint; a CPU forward + adjoint
    costs about three forward passes, so this rate is an upper bound."""
    from oracle import pyoracle

    pyoracle.lib()
    cores = cpu_threads()
    replicas = max(replicas, cores)
    rng = np.random.default_rng(0)
    systems = []
    for r in range(replicas):
        s = pyoracle.Line1d.System_Cuspy_Laplace(**dict(PARAMS, seed=r * AD_N))
        s.v = rng.normal(size=AD_N) * 0.5
        systems.append(s)

    def sweep(n):
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(max_workers=cores) as ex:
            list(ex.map(lambda s: s.timeSteps(n), systems))
        return time.perf_counter() - t0

    pilot = 100
    dt_pilot = sweep(pilot)
    steps = max(pilot, int(pilot * target_s / max(dt_pilot, 1e-3)))
    dt = sweep(steps)
    return {
        "value": replicas * AD_N * steps / dt,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{replicas} replicas x N={AD_N} x {steps} forward timeSteps (no adjoint on the CPU), "
        f"one replica per thread ({dt:.1f} s)",
    }


def run_adjoint(args):
    import torch

    from frictionqpotspringblock_amd import adjoint as A
    from frictionqpotspringblock_amd.batch import Batch

    torch.cuda.init()
    kw = dict(PARAMS, k_frame=1.0 / AD_N, shape=[AD_N])
    batch = Batch(AD_R, seed_stride=AD_N, potential="Cuspy", interactions="Laplace1d", **kw)
    target = Batch(AD_R, seed_stride=AD_N, potential="Cuspy", interactions="Laplace1d", **kw)
    rng = np.random.default_rng(0)
    dev = "cuda"
    u0 = A.get_field(batch, "u")
    v0 = torch.tensor(rng.normal(size=(AD_R, AD_N)) * 0.5, device=dev)
    a0 = torch.zeros((AD_R, AD_N), dtype=torch.float64, device=dev)
    par = A.parameters(batch)
    par_t = par.clone()
    par_t[:, 0] *= 1.05
    par_t[:, 1] *= 0.95
    par_t[:, 3] *= 1.1
    y = A.chunk_data(batch)
    with torch.no_grad():
        u_t, _, _ = A.run(target, AD_T, AD_K, u0, v0, a0, par_t, y)
    del target

    def step():
        leaves = [x.clone().requires_grad_(True) for x in (u0, v0, a0, par, y)]
        uT, _, _ = A.run(batch, AD_T, AD_K, *leaves)
        loss = ((uT - u_t) ** 2).sum(dim=1).mean()
        loss.backward()
        return loss.item(), leaves

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, leaves = step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    rec_ms, rev_ms = A.kernel_ms(batch)
    g = leaves[3].grad
    assert torch.all(torch.isfinite(g))
    value = AD_R * AD_N * AD_T * args.steps / wall
    tape_bytes = 28.0 * AD_R * AD_N * AD_K  # u, v, a (FP64) and well index (int32) per particle-step
    achieved = tape_bytes / (rev_ms * 1e-3) / 1e9
    cpu = None if args.no_cpu else cpu_leg("adjoint", args)
    return {
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
        "data": f"synthetic (Weibull yield landscape from pcg32 seeds r*{AD_N}, normal initial velocities)",
        "config": {
            "workload": f"adjoint: {AD_R} Line1d Cuspy-Laplace replicas x N={AD_N}, T={AD_T} timeSteps "
            f"forward + adjoint, checkpoint every {AD_K} steps",
            "replicas": AD_R,
            "N": AD_N,
            "T": AD_T,
            "window": AD_K,
            "loss": loss,
            "grad_params_replica0": [float(x) for x in g[0]],
            "record_ms_per_window": rec_ms,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_adj_reverse",
            "kernel_ms": rev_ms,
        },
        "cpu_baseline": cpu,
    }


def ragged_systems(n1d=512, n2d=8, side=256, nchunk=200):
    """The ragged workload: (class name, constructor arguments) per system."""
    rng = np.random.default_rng(0)
    Ns = np.floor(2.0 ** rng.uniform(9, 16, n1d)).astype(int)
    quartic = rng.uniform(size=n1d) < 0.3
    systems, seed = [], 0
    for i, N in enumerate(Ns):
        pot = ["Cuspy", "SemiSmooth", "Smooth"][i % 3]
        kw = dict(PARAMS, k_frame=1.0 / N, shape=[int(N)], seed=seed, nchunk=nchunk)
        if quartic[i]:
            kw.pop("k_interactions")
            kw.update(a1=1.0, a2=5.0)
        if pot == "SemiSmooth":
            kw["kappa"] = 1.0
        systems.append((f"System_{pot}_{'Quartic' if quartic[i] else 'Laplace'}", kw))
        seed += int(N)
    for _ in range(n2d):
        N = side * side
        systems.append(
            ("Line2d.System_Cuspy_Laplace", dict(PARAMS, k_frame=1.0 / N, shape=[side, side], seed=seed, nchunk=nchunk))
        )
        seed += N
    return systems


def ragged_nopassing(systems, max_n=None):
    """No-passing versions of the Cuspy-Laplace chains of the ragged workload
    (the reference's no-passing minimiser exists for Cuspy + Laplace1d only)."""
    out = []
    for name, kw in systems:
        if name == "System_Cuspy_Laplace" and (max_n is None or kw["shape"][0] <= max_n):
            out.append(("System_Cuspy_Laplace_Nopassing", dict(kw, eta=0.0, dt=0.0)))
    return out


def rg_cpu_baseline(target_s=10.0):
    """The oracle on a sample of the ragged systems (one per thread): events of
    the same protocol, counted as sum N x time steps."""
    from oracle import pyoracle

    pyoracle.lib()
    cores = cpu_threads()
    systems = ragged_systems()
    rng = np.random.default_rng(1)
    pick = rng.choice(len(systems) - 8, size=cores, replace=False)
    cls = {"Cuspy": "Cuspy", "SemiSmooth": "SemiSmooth", "Smooth": "Smooth"}
    objs = []
    for i in pick:
        name, kw = systems[i]
        pot = name.split("_")[1]
        inter = "Quartic1d" if name.endswith("Quartic") else "Laplace1d"
        s = pyoracle._System(rank=1, potential=cls[pot], interactions=inter, **kw)
        s.minimise()
        objs.append((s, pot == "Smooth", int(np.prod(kw["shape"]))))

    def event(item):
        s, smooth, N = item
        i0 = s.inc
        if not smooth:
            s.eventDrivenStep(RG_EPS, False)
        s.eventDrivenStep(RG_EPS, True)
        s.minimise()
        return N * (s.inc - i0)

    work, t0 = 0, time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        while time.perf_counter() - t0 < target_s:
            work += sum(ex.map(event, objs))
    dt = time.perf_counter() - t0
    return {
        "value": work / dt,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{cores} randomly chosen systems of the workload (sizes {sorted(o[2] for o in objs)}), "
        f"one per thread, events of the same protocol for {dt:.1f} s",
    }


def qs_params(seed=0):
    return dict(PARAMS, k_frame=1.0 / QS_N, shape=[QS_N], seed=seed)


def qs_cpu_baseline(target_s=10.0):
    """The oracle running the examples' protocol on replicas of the quasistatic
    workload, one per thread, for about `target_s`; sum N x time steps."""
    from oracle import pyoracle

    pyoracle.lib()
    cores = cpu_threads()
    systems = [pyoracle.Line1d.System_Cuspy_Laplace(**qs_params(r)) for r in range(cores)]
    for s in systems:
        s.u_frame = 0.0
        s.minimise()

    def pair(s):
        i0 = s.inc
        s.eventDrivenStep(QS_EPS, False)
        s.eventDrivenStep(QS_EPS, True)
        s.minimise()
        return QS_N * (s.inc - i0)

    work, pairs, t0 = 0, 0, time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        while time.perf_counter() - t0 < target_s:
            work += sum(ex.map(pair, systems))
            pairs += 1
    dt = time.perf_counter() - t0
    return {
        "value": work / dt,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{cores} replicas (seeds 0..{cores - 1}), one per thread, {pairs} event pairs each of the "
        f"same protocol ({dt:.1f} s)",
    }


def qs_plain_path(pairs=5):
    """One system through the reference's API, the examples' Python loop."""
    import torch

    import frictionqpotspringblock_amd as fq

    s = fq.Line1d.System_Cuspy_Laplace(**qs_params(0))
    s.u_frame = 0.0
    s.minimise()
    torch.cuda.synchronize()
    i0, t0 = s.inc, time.perf_counter()
    for _ in range(pairs):
        s.eventDrivenStep(QS_EPS, False)
        s.eventDrivenStep(QS_EPS, True)
        s.minimise()
    torch.cuda.synchronize()
    return QS_N * (s.inc - i0) / (time.perf_counter() - t0)


def run_quasistatic(args):
    import torch

    from frictionqpotspringblock_amd.batch import Batch

    torch.cuda.init()
    tune = {}
    best = None
    if not args.no_autotune:
        # launch shapes tried on a copy of the workload (so that the timed
        # events are the same whatever is tried): event 0 and warm-up pairs (a
        # fifth of the timed pairs, at most 20: later events have larger
        # avalanches, with more well crossings per step, and the shapes rank
        # differently there), then two event pairs per shape, compared by
        # kernel time per particle-step
        trial = Batch(QS_R, seed_stride=1, potential="Cuspy", interactions="Laplace1d", **qs_params(0))
        trial.quasistatic(QS_EPS, 1)
        event = 1
        warm = min(20, args.steps // 5)
        if warm:
            trial.set_launch(*QS_SHAPES[-1])
            trial.quasistatic(QS_EPS, 2 * warm, first_event=event)
            event += 2 * warm
        for ppt, threads in QS_SHAPES:
            trial.set_launch(ppt, threads)
            i0 = trial.inc.copy()
            trial.quasistatic(QS_EPS, 4, first_event=event)
            event += 4
            work = float(QS_N * (trial.inc - i0).sum())
            tune[f"{ppt}x{threads}"] = trial.quasistatic_stats[0] * 1e6 / work
        best = min(tune, key=tune.get)
        del trial
    batch = Batch(QS_R, seed_stride=1, potential="Cuspy", interactions="Laplace1d", **qs_params(0))
    if best is not None:
        batch.set_launch(*[int(x) for x in best.split("x")])
    batch.quasistatic(QS_EPS, 1)  # event 0: the initial quench
    event = 1
    if args.warmup:
        batch.quasistatic(QS_EPS, 2 * args.warmup, first_event=event)
        event += 2 * args.warmup
    inc0 = batch.inc.copy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    q = batch.quasistatic(QS_EPS, 2 * args.steps, first_event=event)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms, launches = batch.quasistatic_stats
    steps = batch.inc - inc0
    work = float(QS_N * steps.sum())
    flop = FLOP_PER_PARTICLE_STEP_MIN * work
    achieved = flop / (ms * 1e-3) / 1e12
    ppt, threads = batch.launch
    plain = None if args.no_plain else qs_plain_path()
    cpu = None if args.no_cpu else cpu_leg("quasistatic", args)
    return {
        "metric": METRIC,
        "value": work / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Weibull yield landscapes from pcg32, replica r with seed r)",
        "config": {
            "workload": f"quasistatic: {QS_R} Line1d Cuspy-Laplace replicas x N={QS_N}, {args.steps} event pairs "
            "(elastic step, kick + minimise) each of the examples' protocol, device-side driver",
            "replicas": QS_R,
            "N": QS_N,
            "event_pairs": args.steps,
            "time_steps_per_replica_mean": float(steps.mean()),
            "time_steps_per_replica_max": int(steps.max()),
            "avalanche_size_mean": float(np.abs(q["S"][:, 1::2]).mean()),
            "launch": f"{ppt}x{threads}",
            "launch_tuning_ns_per_particle_step": tune,
            "driver_launches": launches,
            "plain_one_system_python_loop": plain,
        },
        "roofline": {
            "kernel": f"k_quasi_chain<Cuspy, Laplace1d, {ppt} x {threads}>",
            "bound": "valu",
            "achieved": achieved,
            "peak": 78.6,
            "unit": "TFLOP/s",
            "frac": achieved / 78.6,
            "traffic": None,
            "flop_per_particle_step": FLOP_PER_PARTICLE_STEP_MIN,
            "kernel_ms_total": ms,
        },
        "cpu_baseline": cpu,
    }


C0_EVENTS = 201  # event 0 and 100 event pairs of the examples' protocol


def c0_cpu_baseline(target_s=10.0):
    """The reference's CPU configuration on the oracle, whole: one system
    (N = 1024, seed 0), event 0 and 100 event pairs of the examples'
    protocol on one thread (the configuration is one chain: it does not
    divide over threads)."""
    from oracle import pyoracle
    from tests import protocols as P

    pyoracle.lib()
    s = pyoracle.Line1d.System_Cuspy_Laplace(**qs_params(0))
    t0 = time.perf_counter()
    P.quasistatic(s, C0_EVENTS, eps=QS_EPS)
    dt = time.perf_counter() - t0
    return {
        "value": QS_N * s.inc / dt,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"the whole configuration: one chain of {QS_N}, {C0_EVENTS} events ({s.inc} time steps), one "
        f"thread ({dt:.2f} s)",
        "seconds": dt,
    }


def run_c0(args):
    """The reference's CPU configuration (one Line1d Cuspy-Laplace chain,
    N = 1024, seed 0, Weibull(2, 1) increments in chunks of 200, the
    examples' protocol with eps = 1e-3: event 0 and 100 event pairs) on the
    GPU: through the reference's API (the examples' Python loop of
    eventDrivenStep / minimise calls, one launch sequence per call) and as
    one call of the device-side driver."""
    import torch

    import frictionqpotspringblock_amd as fq
    from tests import protocols as P

    torch.cuda.init()
    warm = fq.Line1d.System_Cuspy_Laplace(**qs_params(0))
    P.quasistatic(warm, 5, eps=QS_EPS)
    s = fq.Line1d.System_Cuspy_Laplace(**qs_params(0))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, S_loop, inc_loop = P.quasistatic(s, C0_EVENTS, eps=QS_EPS)
    torch.cuda.synchronize()
    wall_loop = time.perf_counter() - t0
    d = fq.Line1d.System_Cuspy_Laplace(**qs_params(0))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    q = d.quasistatic(QS_EPS, C0_EVENTS)
    torch.cuda.synchronize()
    wall_dev = time.perf_counter() - t0
    assert np.array_equal(q["S"], S_loop) and np.array_equal(q["inc"], inc_loop)
    steps = int(inc_loop[-1])
    cpu = None if args.no_cpu else cpu_leg("c0", args)
    return {
        "metric": METRIC,
        "value": QS_N * steps / wall_dev,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": wall_dev * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Weibull yield landscape from pcg32 seed 0)",
        "seconds": wall_dev,
        "config": {
            "workload": f"the reference's CPU configuration: one Line1d Cuspy-Laplace chain of {QS_N} (seed 0), "
            f"{C0_EVENTS} events of the examples' protocol (eps {QS_EPS}), {steps} time steps",
            "time_steps": steps,
            "avalanche_sizes_sum": int(np.abs(q["S"]).sum()),
            "device_driver_seconds": wall_dev,
            "python_loop_seconds": wall_loop,
        },
        "plain_value": QS_N * steps / wall_loop,
        "plain": "the examples' Python loop through the reference's API (System.eventDrivenStep / minimise)",
        "roofline": None,
        "roofline_note": "one 1024-particle chain is one workgroup on one CU: latency-bound, no roofline applies",
        "cpu_baseline": cpu,
    }


# FLOP per particle-step counted on the reference's expressions (the Verlet
# step with two correctors and the force assembly): Cuspy + Laplace1d 40;
# Quartic1d adds 9 (two cubes and their sum), SemiSmooth 10 (its three
# branches' parabolas), Smooth 16 plus 20 for the sine; Laplace2d adds 2 (four
# neighbours); a no-passing sweep (detail.h:1711-1741) 24 per particle.
POT_FLOP = {0: 0, 1: 10, 2: 36}
INT_FLOP = {1: 0, 2: 9, 5: 2}


def rg_groups(sets):
    """Every launch group with work since the last reset of its set's
    timers: its key, path, size, FLOP per particle-step, time and
    particle-steps."""
    groups = []
    for rs in sets:
        for g, (ms, ps) in enumerate(rs.group_times()):
            if ms <= 0.0 or ps <= 0.0:
                continue
            info = rs.group_info(g)
            rank, pot, inter, ext, minim = (int(x) for x in info["key"].split("-")[:5])
            flop = 24.0 if minim == 1 else 40.0 + POT_FLOP.get(pot, 0) + INT_FLOP.get(inter, 0)
            groups.append({"key": info["key"], "path": info["path"], "N": info["N"], "systems": len(info["systems"]),
                           "block": info["block"], "ms": ms, "particle_steps": ps, "flop_per_particle_step": flop,
                           "tflops": flop * ps / (ms * 1e-3) / 1e12})
    return groups


def rg_group_roofline(groups, total_flop, wall):
    """The ragged run's roofline from its own launch groups, each timed on
    its own (groups one after the other, events on the group's stream around
    its part of each operation): the dominant group -- the one with the most
    FLOP -- against the FP64 peak (the longest one is named too: after many
    events it can be a few small resident systems in a large avalanche, a
    latency-bound workgroup each); and the FLOP of the timed (concurrent) run
    over its wall clock."""
    top = max(groups, key=lambda x: x["flop_per_particle_step"] * x["particle_steps"])
    longest = max(groups, key=lambda x: x["ms"])
    return {
        "kernel": f"launch group {top['key']} ({top['path']}, {top['systems']} systems up to N = {top['N']}), "
        "the one with the most FLOP, timed alone",
        "bound": "valu",
        "achieved": top["tflops"],
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": top["tflops"] / FP64_PEAK_TFLOPS,
        "traffic": None,
        "group_ms": top["ms"],
        "flop_per_particle_step": top["flop_per_particle_step"],
        "longest_group": {k: longest[k] for k in ("key", "path", "N", "systems", "ms", "tflops")},
        "all_groups_over_wall_tflops": total_flop / wall / 1e12,
        "groups": sorted(groups, key=lambda x: -x["ms"])[:8],
    }


def rg_tile_roofline(replicas=8, n=65536, steps=2048):
    """The ragged workload's dominant kernel on its own: the blocked tiled
    path (k_tile_block, K steps per launch, with its decision kernel) on
    Cuspy-Laplace chains of the workload's largest size, flowSteps at the
    workload's parameters, timed with HIP events on the batch's stream."""
    from frictionqpotspringblock_amd.batch import Batch

    b = Batch(replicas, seed_stride=n, potential="Cuspy", interactions="Laplace1d",
              **dict(PARAMS, k_frame=1.0 / n, shape=[n]))
    b.flowSteps(256, V_FRAME)
    ms = b.flowSteps(steps, V_FRAME, timed=True)
    ps = replicas * n * steps
    achieved = FLOP_PER_PARTICLE_STEP * ps / (ms * 1e-3) / 1e12
    return {
        "kernel": f"k_tile_block<Cuspy, Laplace1d> + k_tile_block_decide ({replicas} chains x {n}, "
        "blocks of 128 steps over tiles of 1024 particles)",
        "bound": "valu",
        "achieved": achieved,
        "peak": 78.6,
        "unit": "TFLOP/s",
        "frac": achieved / 78.6,
        "traffic": None,
        "flop_per_particle_step": FLOP_PER_PARTICLE_STEP,
        "particle_steps_per_s": ps / (ms * 1e-3),
        "ms_per_block": ms / (steps / 128),
    }


def run_ragged(args):
    import torch

    from frictionqpotspringblock_amd.ragged import Ragged

    torch.cuda.init()
    systems = ragged_systems()
    rg = Ragged(systems)
    sizes = np.array(rg.sizes, dtype=float)
    smooth = np.array(["Smooth" in n for n, _ in systems])
    tuned = {}
    if not args.no_autotune:
        # launch shape / block length of the larger groups, from trial events
        # on a copy of the set (this set's events do not depend on it)
        t0 = time.perf_counter()
        res = rg.autotune(RG_EPS, events=1, smooth=smooth, min_particles=RG_TUNE_MIN_PARTICLES,
                          margin=RG_TUNE_MARGIN)
        tuned = {r["key"]: {"path": r["path"], "choice": r["choice"],
                            "ns_per_particle_step": {k: round(v, 4) for k, v in
                                                     (r.get("trials_ns_per_particle_step") or {}).items()}}
                 for r in res.values()}
        tuned["seconds"] = time.perf_counter() - t0
    nps = ragged_nopassing(systems)
    rgn = Ragged(nps)
    nsizes = np.array(rgn.sizes, dtype=float)

    def events(first, k):
        """k event pairs (elastic step, kick + minimise) from protocol event
        `first` on the systems with maxUniformDisplacement, k kicks + minimise
        on the Smooth ones (whose maxUniformDisplacement raises)"""
        rg.quasistatic(RG_EPS, 2 * k, first_event=first, active=~smooth)
        rg.quasistatic(RG_EPS, k, first_event=(first + 1) // 2, kicks_only=True, active=smooth)

    rg.quasistatic(RG_EPS, 1)  # event 0: the initial quench
    rgn.minimise()
    if args.warmup:
        events(1, args.warmup)
    first = 1 + 2 * args.warmup
    inc0 = np.array(rg.inc, dtype=float)
    rg.reset_group_times()
    rgn.reset_group_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    events(first, args.steps)
    torch.cuda.synchronize()
    t_inertial = time.perf_counter() - t0
    work = float(np.sum(sizes * (np.array(rg.inc, dtype=float) - inc0)))
    for _ in range(args.steps):
        rgn.eventDrivenStep(RG_EPS, False)
        rgn.eventDrivenStep(RG_EPS, True)
        rgn.minimise()
        work += float(np.sum(nsizes * rgn.last_steps))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_flop = sum(x["flop_per_particle_step"] * x["particle_steps"] for x in rg_groups([rg, rgn]))
    # each group's own time (the concurrent groups share the GPU, so their
    # event times overlap): one more event pair with the groups one after
    # the other, untimed by the wall clock
    for s in (rg, rgn):
        s.serial = True
        s.reset_group_times()
    events(first + 2 * args.steps, 1)
    rgn.eventDrivenStep(RG_EPS, False)
    rgn.eventDrivenStep(RG_EPS, True)
    rgn.minimise()
    roof = rg_group_roofline(rg_groups([rg, rgn]), total_flop, wall)
    for s in (rg, rgn):
        s.serial = False
    roof["synthetic_tile_leg"] = rg_tile_roofline()
    cpu = None if args.no_cpu else cpu_leg("ragged", args)
    return {
        "metric": METRIC,
        "value": work / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Weibull yield landscapes from pcg32, disjoint seeds per system)",
        "config": {
            "workload": f"ragged: {len(systems)} systems ({int(sizes.sum())} particles, {rg.groups} launch groups), "
            f"{args.steps} events each (device-side quasistatic driver), then {args.steps} no-passing events on "
            f"{len(nps)} Cuspy-Laplace chains",
            "systems": len(systems),
            "particles": int(sizes.sum()),
            "groups": rg.groups,
            "nopassing_systems": len(nps),
            "inertial_seconds": t_inertial,
            "particle_steps": work,
            "autotune": tuned,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }


def run_flow(args):
    blk = args.block or BLOCK
    import torch

    from frictionqpotspringblock_amd.batch import Batch

    torch.cuda.init()
    kw = dict(PARAMS)
    batch = Batch(R, seed_stride=1, potential="Cuspy", interactions="Laplace1d", **kw)
    tuned = {}
    if not args.no_autotune:
        tuned = batch.autotune(V_FRAME, steps=blk)
        # the profiled shape (its HBM counters are in FLOW_TRAFFIC_BYTES)
        # unless another one is clearly faster: trial times within 2% are noise
        if (4, 1024) in tuned and tuned[(4, 1024)] <= 1.02 * min(tuned.values()):
            batch.set_launch(4, 1024)
    for _ in range(args.warmup):
        batch.flowSteps(blk, V_FRAME, block=blk)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev_ms = batch.flowSteps(args.steps * blk, V_FRAME, block=blk, timed=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # every replica took every step, and its state is finite
    inc = batch.inc
    assert np.all(inc == inc[0]) and inc[0] >= blk * args.steps, inc[:4]
    u = batch.u
    assert np.all(np.isfinite(u[:: max(1, R // 16)])) and np.all(np.isfinite(batch.u_frame))

    particle_steps = R * N * blk * args.steps
    value = particle_steps / wall
    kernel_ms = ev_ms / args.steps  # average duration of one resident launch
    flop = FLOP_PER_PARTICLE_STEP * R * N * blk
    achieved = flop / (kernel_ms * 1e-3) / 1e12

    plain = None if args.no_plain else plain_path()
    cpu = None if args.no_cpu else cpu_leg("flow", args)
    ppt, threads = batch.launch
    profiled = (ppt, threads) == (4, 1024) and blk == BLOCK
    return {
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
            f"{blk}-step resident launches, v_frame={V_FRAME}",
            "replicas": R,
            "N": N,
            "flow_steps": blk * args.steps,
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
            "traffic": FLOW_TRAFFIC_BYTES if profiled else None,
            "traffic_source": "profiles/k_flow_cuspy_laplace_flow_R1024_N4096.md (rocprofv3 --pmc FETCH_SIZE, "
            "WRITE_SIZE of the 4 x 1024 launch of 1000 steps)" if profiled else None,
            "kernel": f"k_flow<Cuspy, Laplace1d, {ppt} x {threads}>",
            "kernel_ms": kernel_ms,
            "hbm_gbs": FLOW_TRAFFIC_BYTES / (kernel_ms * 1e-3) / 1e9 if profiled else None,
        },
        "cpu_baseline": cpu,
    }


def measure_peaks():
    """The HBM and FP64 matrix-core peaks measured on this GPU, run after the
    timed region: a device-to-device copy of 1 GiB (read plus written bytes
    per second, best of 10) and a float64 matrix product of two 8192 x 8192
    matrices (torch.matmul, i.e. the vendor BLAS on the FP64 matrix cores;
    2 n^3 FLOP, best of 5). Both timed with events on torch's stream."""
    import torch

    def best_ms(fn, reps):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for a, b in ev:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        return min(a.elapsed_time(b) for a, b in ev)

    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    ms = best_ms(lambda: y.copy_(x), 10)
    copy_gbs = 2 * x.numel() * 8 / (ms * 1e-3) / 1e9
    del x, y
    n = 8192
    g = torch.Generator(device="cuda").manual_seed(0)
    a = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g)
    b = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g)
    c = torch.empty_like(a)
    for _ in range(2):
        torch.matmul(a, b, out=c)
    ms = best_ms(lambda: torch.matmul(a, b, out=c), 5)
    gemm_tf = 2.0 * n**3 / (ms * 1e-3) / 1e12
    del a, b, c
    torch.cuda.empty_cache()
    return {
        "hbm_copy_GBs": copy_gbs,
        "fp64_matmul_TFs": gemm_tf,
        "published": {"hbm_GBs": 8000.0, "fp64_TFs": 78.6},
        "how": "1 GiB device copy (read + write bytes), best of 10; float64 torch.matmul 8192^3, best of 5",
    }


CPU_LEGS = {
    "flow": lambda t: cpu_baseline(target_s=t),
    "longrange": lambda t: lr_cpu_baseline(target_s=t),
    "adjoint": lambda t: ad_cpu_baseline(target_s=t),
    "ragged": lambda t: rg_cpu_baseline(target_s=t),
    "quasistatic": lambda t: qs_cpu_baseline(target_s=t),
    "c0": lambda t: c0_cpu_baseline(target_s=t),
}

# the other configurations, measured after the headline in the default run
# (name, function, timed bench steps, warmup steps)
EXTRAS = [
    ("c0", "run_c0", 1, 0),
    ("quasistatic", "run_quasistatic", 20, 1),
    ("ragged", "run_ragged", 10, 1),
    ("longrange", "run_longrange", 20, 1),
    ("adjoint", "run_adjoint", 2, 1),
]


def extras(args):
    """Every other configuration at a reduced length (the same workloads; their
    own --workload runs are the full lengths), each with its value, roofline
    and CPU baseline."""
    out = {}
    keep = ("value", "unit", "steps", "warmup", "ms_per_step", "seconds", "config", "plain_value", "plain",
            "batched_over_plain", "roofline", "roofline_note", "cpu_baseline")
    for name, fn, steps, warmup in EXTRAS:
        a = argparse.Namespace(**vars(args))
        a.steps, a.warmup = steps, warmup
        t0 = time.perf_counter()
        try:
            line = globals()[fn](a)
        except Exception as e:  # one configuration failing does not hide the others
            out[name] = {"error": f"{type(e).__name__}: {e}"}
            continue
        rec = {k: line[k] for k in keep if k in line}
        rec["seconds_total"] = time.perf_counter() - t0
        out[name] = rec
        print(f"[bench] {name}: {rec.get('value', 0):.3e} particle-steps/s ({rec['seconds_total']:.1f} s)",
              file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument(
        "--workload", choices=["flow", "longrange", "adjoint", "ragged", "quasistatic", "c0", "lrqs"],
        default="flow"
    )
    ap.add_argument(
        "--steps", type=int, default=None,
        help="timed bench steps (flow: resident blocks of 1000 flowSteps, default 100; "
        "longrange: blocks of 100 timeSteps, default 200; adjoint: forward + backward passes, default 5; "
        "ragged: quasistatic events, default 50; quasistatic: event pairs, default 100; c0: ignored)",
    )
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--block", type=int, default=None, help="flow: steps per resident launch (default 1000)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=4.0, help="CPU work per oracle build of a CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain one-system path")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-peaks", action="store_true", help="skip the copy / matrix-product peak measurement")
    ap.add_argument("--no-extras", action="store_true",
                    help="flow: skip the other configurations measured after the headline")
    ap.add_argument("--cpu-leg", default=None, help=argparse.SUPPRESS)
    ap.add_argument(
        "--copy-probe", action="store_true",
        help="only run a plain device copy of 1 GiB three times (calibration of HBM byte counters) and exit",
    )
    args = ap.parse_args()
    if args.cpu_leg:
        # child process of cpu_leg: one CPU baseline with the oracle build
        # named by FQO_LIB (or the parity build)
        print(json.dumps(CPU_LEGS[args.cpu_leg](args.cpu_seconds)), flush=True)
        return
    if args.copy_probe:
        import torch

        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": x.numel() * 8, "copy_probe_bytes_written": y.numel() * 8}))
        return
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: replicas are independent and would shard across GPUs "
            "without communication; run with --gpus 1",
            file=sys.stderr,
        )
        sys.exit(2)
    if args.steps is None:
        args.steps = {"flow": 100, "longrange": 200, "adjoint": 5, "ragged": 50, "quasistatic": 100,
                      "c0": 1, "lrqs": 5}[args.workload]
    run = {
        "flow": run_flow,
        "longrange": run_longrange,
        "adjoint": run_adjoint,
        "ragged": run_ragged,
        "quasistatic": run_quasistatic,
        "c0": run_c0,
        "lrqs": run_lr_quasistatic,
    }[args.workload]
    line = run(args)
    if args.workload == "flow" and not args.no_extras:
        line["workloads"] = extras(args)
    if not args.no_peaks:
        line["peaks_measured"] = measure_peaks()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
