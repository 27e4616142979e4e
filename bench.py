"""Benchmark: batched VMC walker moves on one MI355X.

Workload (plain FP64): periodic cubic cell L = 18 bohr, 128 ions, 512 electrons (256 per spin),
256 real orbitals per spin on a 64^3 periodic B-spline table, J1 + J2 B-spline Jastrows,
1024 walkers (one RNG stream per walker), delay rank 32, drift-diffusion tau = 0.3.
One step = one sweep (every electron of every walker proposes one move) + evaluateGL + the
kinetic local energy.  Inputs are synthetic (seeded) and resident in HBM before timing.

Prints one JSON line with `value` = electron moves per second of the whole job.

`--workload uneven`: 1024 walkers over four systems of 128 / 256 / 512 / 1024 electrons (cells at
the same electron density, 48^3 / 64^3 / 64^3 / 80^3 tables, a quarter of each system's walkers
with complex orbitals at a seeded twist), FP64, advanced together in one captured graph per
sub-step, delay rank and kernel forms chosen by the autotuner.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "electron moves/sec/GPU + effective samples/sec, learned+mixed vs plain FP64"
HBM_PEAK_GBS = 8000.0
GRAPH = True  # replay each sweep from a captured HIP graph (--eager: launch kernel by kernel)
LAUNCH = "auto"  # plain sweep's launch form: auto (timed trial), single, partitioned or window
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def autocorr_kappa(x):
    """Integrated autocorrelation time kappa = 1 + 2 sum rho(i), truncated at the first rho <= 0
    (the estimator of tests/scripts/check_scalars.py in the reference)."""
    x = np.asarray(x, dtype=float)
    n = len(x)
    if n < 3:
        return 1.0
    d = x - x.mean()
    var = np.dot(d, d) / n
    if var == 0:
        return 1.0
    kappa = 1.0
    for i in range(1, n):
        rho = np.dot(d[:-i], d[i:]) / (n - i) / var
        if rho <= 0:
            break
        kappa += 2 * rho
    return kappa


def segment_bytes_per_move(n, N, n_ion, d, sc=8, sd=8, s=8):
    """Algorithmic HBM bytes of one proposed move in the fused move kernel (per walker) at delay d:
    spline 64 n sc; inverse-row update (2 d n + d^2 + n) sd; accept (d n + 6 n) sd;
    J2 13 N s; J1 3 N_I s (sc, sd: bytes per table / determinant element; s: FP64 positions and
    Jastrow accumulators)."""
    return 64 * n * sc + (2 * d * n + d * d + n) * sd + (d * n + 6 * n) * sd + 13 * N * s + 3 * n_ion * s


def segment_flops_per_move(n, N, n_ion, d):
    return 608 * n + (4 * d * n + 2 * d * d) + 8 * n + 2 * d * n + 40 * N + 40 * n_ion


def flush_bytes(n, k, s=8):  # s: bytes per determinant element
    """Compulsory bytes of one rank-k flush of one n x n inverse in a single pass: read and write
    Ainv, read U and V (n x k each) and Binv (k x k)."""
    return (2 * n * n + 2 * n * k + k * k) * s


def flush_flops(n, k):
    return 4 * k * n * n + 2 * n * k * k


def measured_traffic(kernel_prefix="qmi::k_sweep_segment"):
    """HBM bytes per dispatch of the dominant kernel from the committed rocprofv3 counter summary
    (read and write requests by size, TCC_EA0_RDREQ_{32,64,128}B and TCC_EA0_WRREQ{,_64B})."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_config1_fp64.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    for name, v in d["per_dispatch"].items():
        if name.startswith(kernel_prefix):
            return v["read_bytes"] + v["write_bytes"], f"profiles/hbm_traffic_config1_fp64.md ({name})"
    return None, None


def coulomb_cost(N, n_ion, nk, pair_flops=40, k_flops=30):
    """Floating-point operations per walker of the Ewald energy: each electron-electron and
    electron-ion pair costs ~40 (minimum image 15, distance 6, spline 19) and each (k, electron)
    term ~30 (phase 5, sincos ~20, accumulate 4); the rest is O(nk)."""
    pairs = N * (N - 1) // 2 + N * n_ion
    return dict(pairs=pairs, k_terms=nk * N, flops=pairs * pair_flops + nk * N * k_flops)


def bench_nlpp(ctx, system, log, rcut=1.5, reps=3):
    """Non-local pseudopotential energy of every walker (s, p channels, 12-point rule, random
    rotations): device time per evaluation and virtual-move ratios per second."""
    x = np.linspace(0.0, 2.0, 201)
    v = np.array([-1.5 * np.exp(-1.2 * x * x), 0.8 * np.exp(-0.9 * x * x)])
    # second derivatives of these Gaussians (the splines interpolate them)
    v2 = np.array([-1.5 * np.exp(-1.2 * x * x) * (4 * 1.44 * x * x - 2.4),
                   0.8 * np.exp(-0.9 * x * x) * (4 * 0.81 * x * x - 1.8)])
    g = (1 + 5 ** 0.5) / 2
    q = np.array([[0, 1, g], [0, -1, g], [0, 1, -g], [0, -1, -g], [1, g, 0], [-1, g, 0], [1, -g, 0], [-1, -g, 0],
                  [g, 0, 1], [-g, 0, 1], [g, 0, -1], [-g, 0, -1]], dtype=float)
    q /= np.linalg.norm(q, axis=1)[:, None]
    ctx.nlpp_setup([0, 1], x[1] - x[0], v, v2, rcut, q, np.full(12, 1.0 / 12))
    rng = np.random.default_rng(5)
    rot = np.array([np.linalg.qr(rng.normal(size=(3, 3)))[0] for _ in range(ctx.nw)])
    ctx.timing_enable(True)
    ctx.timing_reset()
    for _ in range(reps):
        ctx.mw_evaluateNLPP(rot)
    tm = ctx.timing().get("nlpp", dict(ms=0.0, calls=0))
    ctx.timing_enable(False)
    # electron-ion pairs within the cutoff (host count over the current positions)
    R = ctx.walkers_get()
    ions, L = system.ions, system.L
    npairs = 0
    for a in ions:
        d = R - a
        d -= L * np.round(d / L)
        npairs += int(np.sum(np.sum(d * d, axis=2) < rcut * rcut))
    ms = tm["ms"] / max(1, tm["calls"])
    out = dict(ms_per_evaluation=ms, rcut=rcut, channels=[0, 1], points=12, pairs_per_walker=npairs / ctx.nw,
               ratios_per_s=npairs * 12 / (ms * 1e-3) if ms > 0 else 0.0)
    log("[bench] nlpp: " + json.dumps(out))
    return out


def bench_dmc(ctx, steps, log, tau=0.01, with_nlpp=False, nonlocal_moves=0, untimed=5):
    """DMC on the context's walkers (local energy kinetic + Ewald Coulomb, plus the non-local
    pseudopotential when one is set on the context, optionally with T-moves; fluctuating
    population, feedback toward the starting count): electron moves per second over `steps` timed
    steps after `untimed` ones, with the per-step device time of each part."""
    nw = ctx.nw
    ctx.dmc_init(tau=tau, capacity=4 * nw, seed=2024, warmup_steps=200, nonlocal_moves=nonlocal_moves)
    ctx.dmc_run(untimed)
    ctx.timing_enable(True)
    ctx.timing_reset()
    t1 = time.perf_counter()
    st = ctx.dmc_run(steps)
    dt = time.perf_counter() - t1
    tm = ctx.timing()
    ctx.timing_enable(False)
    moves = float(np.sum(st[:, 3])) * ctx.N
    per = {k: v["ms"] / steps for k, v in tm.items()}
    out = dict(moves_per_s=moves / dt, ms_per_step=dt * 1e3 / steps, tau=tau, steps=steps,
               local_energy="kinetic + Ewald Coulomb" + (" + non-local pseudopotential" if with_nlpp else ""),
               walkers_mean=float(np.mean(st[:, 3])), walkers_min=int(np.min(st[:, 4])),
               walkers_max=int(np.max(st[:, 4])), energy_last=float(st[-1, 0]),
               device_ms_per_step=per)
    if nonlocal_moves:
        out["nonlocal_moves"] = "T-moves v0"
        out["tmoves_accepted_per_step"] = float(np.mean(st[:, 10]))
    log("[bench] dmc: " + json.dumps(out))
    return out


def cpu_baseline(system, cfg, nwalk=16, nthreads=None):
    """The CPU FP64 oracle (a from-scratch restatement of the reference's CPU batched path) on a
    bounded sample: nwalk walkers of the same system, one crowd per thread, one sweep."""
    from oracle import oracle as orc
    from qmcpack_amd.system import make_walkers
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    wpc = max(1, nwalk // nthreads)
    R = make_walkers(system, nwalk, cfg["seed"] + 1)
    o = orc.VMC(system, R, wpc, 123)
    o.evaluate_log()
    t0 = time.perf_counter()
    o.run(1, 1, cfg["tau"], 0, nthreads=nthreads)
    dt = time.perf_counter() - t0
    moves = nwalk * system.N
    return dict(value=moves / dt, unit="electron moves/s", cores=min(nthreads, (nwalk + wpc - 1) // wpc),
                kind="port",
                sample=f"{nwalk} walkers x {system.N} electrons, 1 sweep of the same system "
                       f"({dt:.1f} s), delay rank {system.delay_rank}, oracle in C++ FP64")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_small_config(nthreads):
    """The small CPU-runnable configuration (16 walkers x 128 electrons, 48^3 table, delay rank 1,
    tau 0.3) on the oracle, as specified (one crowd, 1 thread) and with crowds = threads: electron
    moves/s over 2 sweeps each."""
    from oracle import oracle as orc
    from qmcpack_amd.system import make_walkers, workload_system
    s, cfg = workload_system("small")
    R = make_walkers(s, cfg["nw"], cfg["seed"])
    out = {}
    for label, threads, wpc in (("one_crowd_1_thread", 1, cfg["nw"]), ("crowds_eq_threads", nthreads,
                                                                       max(1, cfg["nw"] // nthreads))):
        o = orc.VMC(s, R, wpc, cfg["seed"])
        o.evaluate_log()
        t0 = time.perf_counter()
        o.run(2, 1, cfg["tau"], 0, nthreads=threads)
        dt = time.perf_counter() - t0
        out[label] = dict(value=cfg["nw"] * s.N * 2 / dt, unit="electron moves/s", threads=threads,
                          crowds=(cfg["nw"] + wpc - 1) // wpc)
    return out


def chain_summary(E, dt, nw, N, steps):
    """moves/s, kappa (mean over walkers), effective samples/s and the mean local energy with its
    error bar sqrt(var kappa / (nw steps))."""
    kappa = float(np.mean([autocorr_kappa(E[:, w]) for w in range(nw)])) if steps >= 3 else 1.0
    emean = float(np.mean(E))
    eerr = float(np.sqrt(np.var(E) * kappa / E.size))
    return dict(moves_per_s=nw * N * steps / dt, ms_per_sweep=dt / steps * 1e3, kappa=kappa,
                effective_samples_per_s=nw * steps / kappa / dt, energy=emean, energy_err=eerr)


def pick_launch(ctx, tau, seed):
    """Launch form of the plain sweep, timed over 3 untimed sweeps each: one captured graph per
    sub-step with all walkers in one launch per move; eager launches with the walkers in two
    partitions on two streams (their launches overlap, which hides the tail of each move's launch:
    the workgroups dispatched last on a compute unit finish last); or one launch per delay window
    (k_sweep_segment: each workgroup runs its walker's moves of the window back to back, so only
    the window's end waits for the slowest).  Same arithmetic in every form."""
    forms = {"single": ("graph" if GRAPH else "eager", 1, 0),
             "partitioned": ("eager", 2, 0),
             "window": ("graph" if GRAPH else "eager", 1, 1),
             "window-partitioned": ("eager", 2, 1)}
    names = list(forms) if LAUNCH == "auto" else [LAUNCH]
    best = None
    for name in names:
        mode, P, mk = forms[name]
        ctx.tune_set(move_kernel=mk)
        ctx.tune_partitions(P)
        ctx.graph_enable(mode == "graph")
        ctx.vmc_run(1, 1, tau, 0, seed + 3)
        t1 = time.perf_counter()
        ctx.vmc_run(3, 1, tau, 0, seed + 4)
        dt = time.perf_counter() - t1
        if best is None or dt < best[0]:
            best = (dt, name)
    name = best[1]
    mode, P, mk = forms[name]
    ctx.tune_set(move_kernel=mk)
    ctx.tune_partitions(P)
    ctx.graph_enable(mode == "graph")
    return {"single": f"{mode}, one launch of all walkers per move",
            "partitioned": "eager, 2 walker partitions on 2 streams, one launch per move each",
            "window": f"{mode}, one launch per delay window (k_sweep_segment)",
            "window-partitioned": "eager, 2 walker partitions on 2 streams, one launch per delay window each "
                                  "(k_sweep_segment)"}[name]


def run_variant(qmi, system, R, cfg, precision, learned, steps, warmup, train_sweeps, lr, log):
    """One context: optional training of the learned proposal, warmup, then `steps` timed sweeps."""
    nw, N, tau, seed = R.shape[0], system.N, cfg["tau"], cfg["seed"]
    ctx = qmi.Context(0, precision=precision)
    ctx.system_create(system)
    ctx.walkers_load(R, 1)
    ctx.mw_evaluateLog()
    launch = "graph" if GRAPH else "eager"
    if not learned:
        launch = pick_launch(ctx, tau, seed)
    extra = {}
    if learned:
        ctx.proposal_config(learned=True, train=True, lr=lr, init_seed=7)
        t1 = time.perf_counter()
        ctx.vmc_run(train_sweeps, 1, tau, 0, seed + 11, energies=False)
        st = ctx.proposal_stats()
        extra = dict(train_samples=int(st["moves"]), train_steps=int(st["train_steps"]),
                     train_s=time.perf_counter() - t1)
        ctx.proposal_config(learned=True, train=False)
    if warmup > 0:
        ctx.vmc_run(warmup, 1, tau, 0, seed)
    if learned:
        ctx.proposal_stats()
    t1 = time.perf_counter()
    E, _ = ctx.vmc_run(steps, 1, tau, 0, seed)
    dt = time.perf_counter() - t1
    out = chain_summary(E, dt, nw, N, steps)
    if learned:
        st = ctx.proposal_stats()
        out.update(extra, mean_acceptance=st["acceptance"], mean_objective=st["objective"])
    else:
        acc = ctx.vmc_stats()
        out.update(mean_acceptance=acc["accepted"] / max(1, acc["accepted"] + acc["rejected"]))
    out["launch"] = launch if not learned else "eager"
    log(f"[bench] {precision}{' + learned' if learned else ''}: " + json.dumps(out))
    return ctx, out


def bench_uneven(args, log):
    """Uneven batching: eight populations (four systems, real and complex orbitals) in one batch."""
    from qmcpack_amd import qmi
    from qmcpack_amd.system import UNEVEN, make_walkers, uneven_populations
    t0 = time.perf_counter()
    pops = uneven_populations(**UNEVEN)
    log(f"[bench] uneven populations built in {time.perf_counter() - t0:.1f}s: "
        + ", ".join(f"{lab} x {n}" for _, n, lab in pops))
    tau = UNEVEN["tau"]
    ctxs = []
    for i, (s, n, lab) in enumerate(pops):
        c = qmi.Context(0)
        c.system_create(s)
        c.walkers_load(make_walkers(s, n, UNEVEN["seed"] + i), 1)
        c.mw_evaluateLog()
        ctxs.append(c)
    b = qmi.Batch(ctxs)
    seeds = [UNEVEN["seed"] + 100 * i for i in range(len(ctxs))]
    b.vmc_run(1, 1, tau, 0, seeds)
    t1 = time.perf_counter()
    if args.no_autotune:
        choice = np.array([[c.tune_get()[x] for x in ("delay_rank", "move_kernel", "flush_variant")] for c in ctxs])
        tms = np.full(len(ctxs) + 2, np.nan)
    else:
        choice, tms = b.autotune(tau, trials=2)
    tune_s = time.perf_counter() - t1
    tuned = {lab: dict(delay_rank=int(ch[0]), move_kernel=("per delay window" if ch[1] else "per move"),
                       flush=("three LDS passes" if ch[2] else "auto (matrix cores for n <= 512)"),
                       substep_ms=None if np.isnan(tms[i]) else float(tms[i]))
             for i, ((_, _, lab), ch) in enumerate(zip(pops, choice))}
    log(f"[bench] autotune ({tune_s:.1f}s): " + json.dumps(tuned) +
        f"; chained {tms[-2]:.2f} ms, concurrent {tms[-1]:.2f} ms -> {'concurrent' if b.concurrent() else 'chained'}")
    tune_times = None if args.no_autotune else dict(chained_ms=float(tms[-2]), concurrent_ms=float(tms[-1]))
    if args.warmup > 0:
        b.vmc_run(args.warmup, 1, tau, 0, seeds)
    t1 = time.perf_counter()
    E = b.vmc_run(args.steps, 1, tau, 0, seeds)
    dt = time.perf_counter() - t1
    moves = sum(n * s.N for s, n, _ in pops) * args.steps
    value = moves / dt
    per_pop = {lab: dict(walkers=n, electrons=s.N, energy=float(np.mean(e)),
                         moves_per_s_share=n * s.N * args.steps / dt)
               for (s, n, lab), e in zip(pops, E)}
    # per-kernel device time: one eager sweep per context with HIP events
    kernels, dominant = {}, None
    for (s, n, lab), c, ch in zip(pops, ctxs, choice):
        c.timing_enable(True)
        c.timing_reset()
        c.graph_enable(False)
        c.vmc_run(1, 1, tau, 0, seeds[0])
        tm = c.timing()
        c.timing_enable(False)
        for name, v in tm.items():
            kernels[f"{lab}: {name}"] = dict(ms_per_call=v["ms"] / max(1, v["calls"]), calls_per_step=v["calls"])
            if name in ("move_one", "sweep_segment") and (dominant is None or v["ms"] > dominant[2]["ms"]):
                dominant = (lab, name, v, s, n, int(ch[0]), ctxs.index(c))
    log("[bench] kernels: " + json.dumps(kernels))
    lab, kname, v, s, n, k, ci = dominant
    sz = 16 if ctxs[ci].cplx else 8
    nn = max(s.n_up, s.n_dn)
    sweep_bytes = n * (s.N // k) * sum(segment_bytes_per_move(nn, s.N, len(s.ions), d, sz, sz) for d in range(k))
    bytes_launch = sweep_bytes / max(1.0, v["calls"])
    avg_ms = v["ms"] / max(1, v["calls"])
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "electron moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "uneven batching VMC sweeps: 1024 walkers over four systems of 128 / 256 / 512 / 1024 "
                        "electrons at one electron density (L = 11.34 / 14.29 / 18 / 22.68 bohr, N/4 ions, "
                        "48^3 / 64^3 / 64^3 / 80^3 tables), a quarter of each system's walkers with complex "
                        "orbitals at a seeded twist, FP64, one captured graph per sub-step, autotuned",
            "populations": {lab: dict(walkers=n, electrons=s.N, grid=s.M[0], complex=s.kpoints is not None)
                            for s, n, lab in pops},
            "autotune": tuned,
            "branches": "concurrent" if b.concurrent() else "chained",
            "branch_substep_ms": tune_times,
            "autotune_s": tune_s,
        },
        "populations": per_pop,
        "kernels": kernels,
        "roofline": {
            "kernel": f"{kname} of the {lab} population (largest device time)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": avg_ms,
        },
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        nthreads = min(16, os.cpu_count() or 1)
        tot_moves, tot_t, parts = 0, 0.0, []
        for s, n, lab in pops:
            if s.kpoints is not None:
                continue
            nwalk = 16 if s.N <= 512 else 8
            o = orc.VMC(s, make_walkers(s, nwalk, 7), 1, 123)
            o.evaluate_log()
            t1 = time.perf_counter()
            o.run(1, 1, tau, 0, nthreads=nthreads)
            d = time.perf_counter() - t1
            tot_moves += nwalk * s.N
            tot_t += d
            parts.append(f"{nwalk} x {s.N}e ({d:.1f} s)")
        out["cpu_baseline"] = dict(value=tot_moves / tot_t, unit="electron moves/s", cores=nthreads, kind="port",
                                   sample="one sweep of " + ", ".join(parts) + " (real-orbital systems, delay rank "
                                          "of the system, oracle in C++ FP64, one walker per thread)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed sweeps (the workload's 50)")
    ap.add_argument("--warmup", type=int, default=100,
                    help="untimed equilibration sweeps from the uniform start; with fewer than ~100 the "
                         "energies of the variants still drift and their comparison is biased")
    ap.add_argument("--walkers", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-walkers", type=int, default=1024)
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "fp32", "bf16", "t32", "tbf16"], default="fp64",
                    help="storage precision of the headline run: fp64 (the plain form), fp32 (FP32 table "
                         "and determinant engine), bf16 (BF16 table, FP32 determinant engine), t32 / tbf16 "
                         "(FP32 / BF16 table, FP64 determinant engine)")
    ap.add_argument("--variants", default="fp32,bf16,t32,learned:fp32,learned:bf16,learned:t32",
                    help="comma-separated extra runs (precision, or learned:<precision>); empty for none")
    ap.add_argument("--train-sweeps", type=int, default=2,
                    help="training sweeps of the learned proposal (2 x 1024 x 512 = 1.05e6 moves)")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--no-nlpp", dest="nlpp", action="store_false", help="skip the pseudopotential energy leg")
    ap.add_argument("--dmc-steps", type=int, default=20, help="timed DMC steps on the workload (0: skip)")
    ap.add_argument("--dmc-tmove-steps", type=int, default=5,
                    help="timed DMC steps with T-moves after the DMC leg (needs the pseudopotential leg; 0: skip)")
    ap.add_argument("--no-autotune", action="store_true", help="uneven workload: keep the default knobs")
    ap.add_argument("--eager", action="store_true",
                    help="launch the sweeps kernel by kernel instead of replaying a captured graph")
    ap.add_argument("--launch", choices=["auto", "single", "partitioned", "window", "window-partitioned"], default="auto",
                    help="plain sweeps: all walkers in one launch per move (graph unless --eager), two "
                         "walker partitions on two streams (eager), one launch per delay window, or the "
                         "fastest in a timed trial")
    ap.add_argument("--workload", choices=["plain", "uneven"], default="plain",
                    help="plain: the 1024 x 512-electron FP64 workload with variants; uneven: 1024 walkers over "
                         "four systems of 128-1024 electrons, a quarter complex, in one captured graph")
    args = ap.parse_args()
    global GRAPH, LAUNCH
    GRAPH = not args.eager
    LAUNCH = args.launch
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs (walker sharding, one "
                                   "rank per GPU) are not part of it", "n_gpus": args.gpus}))
        sys.exit(2)

    from qmcpack_amd import qmi
    from qmcpack_amd.system import workload_system, make_walkers

    def log(*a):
        if not args.quiet:
            print(*a, file=sys.stderr, flush=True)

    if args.workload == "uneven":
        if "--steps" not in sys.argv:
            args.steps = 20
        if "--warmup" not in sys.argv:
            args.warmup = 2
        print(json.dumps(bench_uneven(args, log)))
        return

    t0 = time.perf_counter()
    system, cfg = workload_system("production")
    nw = args.walkers or cfg["nw"]
    R = make_walkers(system, nw, cfg["seed"])
    log(f"[bench] system built in {time.perf_counter() - t0:.1f}s: N={system.N} n={system.n_up} "
        f"M={system.M[0]} norb={system.norb} ions={len(system.ions)} k={system.delay_rank} nw={nw}")
    sc = {"fp64": 8, "fp32": 4, "bf16": 2, "t32": 4, "tbf16": 2}[args.precision]
    sd = 4 if args.precision in ("fp32", "bf16") else 8
    ctx, head = run_variant(qmi, system, R, cfg, args.precision, False, args.steps, args.warmup, 0, args.lr, log)
    value = head["moves_per_s"]
    tau, seed = cfg["tau"], cfg["seed"]

    # per-kernel device time (HIP events on the library stream; eager launches of all walkers), a
    # separate short pass
    ctx.graph_enable(False)
    ctx.tune_partitions(1)
    ctx.timing_enable(True)
    ctx.timing_reset()
    nprof = 2
    ctx.vmc_run(nprof, 1, tau, 0, seed)
    tm = ctx.timing()
    ctx.timing_reset()
    # periodic Coulomb energy (Ewald) of every walker once per step, as the Hamiltonian would be
    # evaluated; timed apart from the sweep (the sweep's local energy is the kinetic part)
    zion = system.N / len(system.ions)
    ctx.coulomb_setup(zion)
    for _ in range(3):
        ctx.mw_evaluatePotential()
    tc = ctx.timing().get("coulomb", dict(ms=0.0, calls=0))
    coul = coulomb_cost(system.N, len(system.ions), ctx.coulomb_nk())
    ctx.timing_enable(False)
    nlpp = bench_nlpp(ctx, system, log) if args.nlpp else None
    dmc = bench_dmc(ctx, args.dmc_steps, log, with_nlpp=nlpp is not None) if args.dmc_steps > 0 else None
    dmc_tm = (bench_dmc(ctx, args.dmc_tmove_steps, log, with_nlpp=True, nonlocal_moves=1, untimed=2)
              if nlpp is not None and args.dmc_tmove_steps > 0 else None)
    del ctx
    n, N, nion, k = system.n_up, system.N, len(system.ions), system.delay_rank
    # the fused move kernel runs either once per move (move_one) or once per delay window of k
    # moves (sweep_segment); either way one sweep performs N moves per walker at delays
    # d = 0..k-1 repeated N/k times.
    kname = "move_one" if "move_one" in tm else "sweep_segment"
    seg = tm.get(kname, dict(ms=0.0, calls=0))
    sweep_bytes = nw * (N // k) * sum(segment_bytes_per_move(n, N, nion, d, sc, sd) for d in range(k))
    sweep_flops = nw * (N // k) * sum(segment_flops_per_move(n, N, nion, d) for d in range(k))
    launches_per_sweep = seg["calls"] / nprof
    bytes_launch = sweep_bytes / max(1.0, launches_per_sweep)
    flops_launch = sweep_flops / max(1.0, launches_per_sweep)
    avg_ms = seg["ms"] / max(1, seg["calls"])
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    fl = tm.get("flush", dict(ms=0.0, calls=0))
    fl_ms = fl["ms"] / max(1, fl["calls"])
    kernels = {name: dict(ms_per_call=v["ms"] / max(1, v["calls"]), calls_per_step=v["calls"] / nprof)
               for name, v in tm.items()}
    log("[bench] kernels: " + json.dumps(kernels))

    variants = {}
    for v in [x for x in args.variants.split(",") if x]:
        learned = v.startswith("learned:")
        prec = v.split(":", 1)[1] if learned else v
        c2, variants[v] = run_variant(qmi, system, R, cfg, prec, learned, args.steps, args.warmup,
                                      args.train_sweeps, args.lr, log)
        del c2
    for v, r in variants.items():
        r["effective_samples_vs_plain"] = r["effective_samples_per_s"] / head["effective_samples_per_s"]
        d = abs(r["energy"] - head["energy"])
        r["energy_diff_sigmas"] = d / np.sqrt(r["energy_err"] ** 2 + head["energy_err"] ** 2)

    traffic, traffic_src = (measured_traffic("qmi::k_move_one" if kname == "move_one" else "qmi::k_sweep_segment")
                            if nw == 1024 and args.precision == "fp64" else (None, None))
    workload = {"fp64": "plain FP64", "fp32": "mixed precision (FP32 table, FP32 determinant engine)",
                "bf16": "mixed precision (BF16 table, FP32 determinant engine)",
                "t32": "mixed precision (FP32 table, FP64 determinant engine)",
                "tbf16": "mixed precision (BF16 table, FP64 determinant engine)"}[args.precision]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "electron moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_sweep"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "bf16": "bf16+f32", "t32": "f32+f64", "tbf16": "bf16+f64"}[args.precision],
        "data": "synthetic",
        "config": {
            "workload": f"{workload} VMC sweeps: 1024 walkers x 512 electrons (256 per spin), 256 orbitals "
                        "on a 64^3 periodic B-spline table, 128 ions, L=18 bohr, J1+J2 B-spline Jastrows, "
                        "delay rank 32, tau 0.3, one RNG stream per walker",
            "walkers": nw, "electrons": N, "orbitals_per_spin": n, "grid": 64, "ions": nion,
            "delay_rank": k, "precision": args.precision, "launch": head["launch"],
        },
        "effective_samples_per_s": head["effective_samples_per_s"],
        "kappa": head["kappa"],
        "energy": [head["energy"], head["energy_err"]],
        "acceptance": head["mean_acceptance"],
        "variants": variants,
        "kernels": kernels,
        "roofline": {
            "kernel": ("k_move_one" if kname == "move_one" else "k_sweep_segment")
                      + " (fused move: inverse row, drift, Jastrows, spline VGL, ratio, MH, accept)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": avg_ms,
            "flops_per_launch": flops_launch,
        },
        "flush_roofline": {
            "kernel": "k_flush_mfma (T = Ainv U^T - W, T' = T Binv, Ainv -= T' V on the matrix cores)",
            "avg_ms": fl_ms,
            "achieved_gbs": (nw * flush_bytes(n, k, sd)) / (fl_ms * 1e-3) / 1e9 if fl_ms > 0 else 0.0,
            "achieved_tflops": (nw * flush_flops(n, k)) / (fl_ms * 1e-3) / 1e12 if fl_ms > 0 else 0.0,
            "mfma_peak_tflops": FP32_PEAK_TFLOPS if sd == 4 else FP64_PEAK_TFLOPS,
        },
        "dmc": dmc,
        "dmc_tmoves": dmc_tm,
        "nlpp": nlpp,
        "coulomb_roofline": {
            "kernel": "k_coulomb (Ewald: short-range pair sums from an LDS spline, long-range rho_k)",
            "avg_ms": tc["ms"] / max(1, tc["calls"]),
            "walkers": nw,
            "k_vectors": coul["k_terms"] // N,
            "flops_per_walker": coul["flops"],
            "achieved_tflops": nw * coul["flops"] / (tc["ms"] / max(1, tc["calls"]) * 1e-3) / 1e12
                               if tc["ms"] > 0 else 0.0,
            "fp64_peak_tflops": FP64_PEAK_TFLOPS,
        },
    }
    if not args.no_cpu:
        t1 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(system, cfg, nwalk=args.cpu_walkers)
        out["cpu_baseline"]["cpu_model"] = cpu_model()
        out["cpu_baseline"]["small_config"] = cpu_small_config(out["cpu_baseline"]["cores"])
        log(f"[bench] cpu baseline {out['cpu_baseline']['value']:.3e} moves/s "
            f"({time.perf_counter() - t1:.1f}s); small config " + json.dumps(out["cpu_baseline"]["small_config"]))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
