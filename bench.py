"""Throughput of the per-step particle hot path on one MI355X.

Workload (the value): 1,000,000 beads of 3 types in a 2,000 A periodic box, 6
tabulated WCA+DH pair tables (1,024 knots), a 256^3 FP32 membrane grid sampled
tricubically, Langevin (BAOAB, gamma = 1/ps), cutoff 30 A + 5 A skin, rebuild period 20.

A step is SimManager's per-step sequence (rebuild policy, forces, PMF, integrator)
over the whole system. Inputs are resident in HBM before the timed region; timing
uses HIP events on the stream the library launches on; the value is the median of
REPEATS timed runs of `--steps` steps. The same JSON line carries the replica batch
(512 BD replicas, aggregate particle-steps/s, against the same replicas as separate
Patches), force matching (gradient evaluations/s on 64 BD snapshots x 100k beads), a
measured copy bandwidth beside the 8 TB/s peak, and the oracle's CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "mars-md_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E
REPEATS = 5  # timed runs per figure; the median is reported
# Algorithmic bytes per unit of the two companion workloads: a batch particle-step
# (28 B state + (144 B rebuild + 8 B x 9.0 half pairs) / 20 at density 1e-4, the
# accounting of algorithmic_bytes) and one force-matching gradient evaluation (64
# snapshots x 100k beads x (16 B positions and types + 12 B target forces + 144 B
# per-snapshot rebuild + 8 B x 9.0 half pairs)).
FM_BYTES_PER_EVAL = 64 * 100_000 * (16.0 + 12.0 + 144.0 + 8.0 * 9.0)


def median(xs):
    xs = sorted(xs)
    k = len(xs)
    return xs[k // 2] if k % 2 else 0.5 * (xs[k // 2 - 1] + xs[k // 2])


def copy_bandwidth(torch, n: int = 1 << 28, reps: int = REPEATS) -> dict:
    """Measured streaming bandwidth: y = x over n float32 (4n bytes read, 4n written),
    timed with events, median of `reps` copies."""
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        e0.record()
        y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) * 1e-3)
    t = median(times)
    del x, y
    return {"GBps": 8.0 * n / t / 1e9, "bytes_per_copy": 8 * n, "peak_GBps": HBM_PEAK / 1e9,
            "frac_of_peak": 8.0 * n / t / HBM_PEAK, "repeats": reps}


def algorithmic_bytes(spec) -> dict:
    """Bytes a step of this method must move (per particle-step), as accounted for
    the workload: state S_state, grid S_field, and the rebuild share
    (S_rebuild + 8 P_half) / K. Returns the figures and the per-launch bytes of the
    fused step kernel (S_state + S_field per particle)."""
    import numpy as np

    n = spec.n
    langevin = spec.params.integrator == 1
    s_state = 52.0 if langevin else 28.0
    s_topo = 0.0
    if spec.bonds is not None:
        s_topo += 8.0 * len(spec.bonds) / n
    if spec.angles is not None:
        s_topo += 12.0 * len(spec.angles) / n
    s_field = sum(g.values.size * 4 for g in spec.grids) / n
    L = np.prod(spec.box.size)
    rho = n / L
    p_half = 0.5 * rho * 4.0 / 3.0 * np.pi * spec.params.pairlist_cutoff ** 3
    s_rebuild = 144.0
    K = spec.params.rebuild_period
    b_alg = s_state + s_topo + s_field + (s_rebuild + 8.0 * p_half) / K
    return {"s_state": s_state, "s_topo": s_topo, "s_field": s_field, "p_half": p_half,
            "b_alg_per_particle_step": b_alg, "step_kernel_bytes_per_launch": (s_state + s_topo + s_field) * n}


def cpu_baseline(spec, steps: int):
    """The oracle (C++ restatement of the reference functors) on the same workload: one
    untimed step (the pairlist build), then `steps` timed steps on one thread and
    `steps` more with OpenMP over the host's CPU share (per-particle loops split, pair
    forces into per-thread buffers summed in thread order). All timed steps lie before
    the first rebuild check (step 21), so no rebuild is timed."""
    from oracle import pyoracle

    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    threads = max(1, min(threads, 16))
    steps = max(1, min(steps, 9))
    sim = pyoracle.Sim(spec)
    sim.run(1)
    t0 = time.perf_counter()
    sim.run(steps, first_step=1)
    dt1 = time.perf_counter() - t0
    sim.set_threads(threads)
    t0 = time.perf_counter()
    sim.run(steps, first_step=1 + steps)
    dtn = time.perf_counter() - t0
    host = f"host {platform.processor() or platform.machine()} nproc={os.cpu_count()}"
    return {"value": spec.n * steps / dtn, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"steps {2 + steps}..{1 + 2 * steps} of {spec.n} beads (same workload) with {threads} OpenMP "
                      f"threads, after the step-1 pairlist build; {host}",
            "single_thread": {"value": spec.n * steps / dt1, "unit": "particle-steps/s", "cores": 1,
                              "sample": f"steps 2..{1 + steps}, one thread"}}


def copy_probe(torch, n: int = 1 << 28, reps: int = 3):
    """A plain streaming kernel of known size: y = -x over n float32 (4n bytes read and
    4n written per launch). Profiled with the same counters as the step kernel, it gives
    the factor between counter readings and bytes moved."""
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe": {"elements": n, "bytes_read_per_launch": 4 * n,
                                     "bytes_written_per_launch": 4 * n, "launches": reps}}))


def timed_runs(torch, run, reps: int = REPEATS, stream=None):
    """Milliseconds of each of `reps` calls of run(rep), each bracketed by HIP events
    on `stream` after a device synchronisation."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = []
    for rep in range(reps):
        torch.cuda.synchronize()
        e0.record(stream)
        run(rep)
        e1.record(stream)
        torch.cuda.synchronize()
        out.append(e0.elapsed_time(e1))
    return out


def batch_bytes_per_particle_step(specs) -> float:
    """Algorithmic bytes per particle-step of the replica batch (the per-replica
    accounting of algorithmic_bytes, weighted by replica size)."""
    total = sum(s.n for s in specs)
    return sum(algorithmic_bytes(s)["b_alg_per_particle_step"] * s.n for s in specs) / total


def bench_batch(torch, steps: int, warmup: int, sequential: bool = True) -> dict:
    """Aggregate particle-steps/s of 512 independent BD replicas (N log-uniform in
    [500, 20000], density 1e-4, config-0 table) stepped together at the autotuner's
    launch (median of REPEATS runs of K = max(steps, 100) steps: whole runs of five
    rebuild periods, so that the rebuild each run starts with on an idle GPU weighs as
    it does in a long run), and of the same replicas stepped one after another as
    separate Patches."""
    import marsmd
    from marsmd import configs
    from marsmd.batch import Batch

    specs = configs.config3()
    total = sum(s.n for s in specs)
    K = max(steps, 100)
    b = Batch(specs)
    # warmup: the cell configuration per size bucket (timed on the walk), then the launch
    # configuration (timed on real steps), both the autotuners' real steps
    cell_steps = 40
    cells = b.autotune_cells(cell_steps, cache=False)
    tune_steps = max(warmup, 160)
    choice = b.autotune(tune_steps, first_step=cell_steps, cache=False)
    step = cell_steps + tune_steps

    def run(rep):
        nonlocal step
        b.run(K, first_step=step)
        step += K

    times = timed_runs(torch, run)
    ms = median(times)
    # the same steps with and without hipGraph replay of the runs between host decisions,
    # at the workgroups per CU kept
    by_graph = {}
    for graph in (False, True):
        b.set_launch(choice["blocks_per_cu"], graph)
        by_graph[graph] = median(timed_runs(torch, run, reps=3)) / K
    b.set_launch(choice["blocks_per_cu"], choice["use_graph"])
    # step-kernel timing in a pass of its own (its events add dispatch gaps)
    b.set_kernel_timing(True)
    b.run(K, first_step=step)
    step += K
    kms, launches = b.kernel_timing()
    b.set_kernel_timing(False)
    value = total * K / (ms / 1e3)
    b_alg = batch_bytes_per_particle_step(specs)
    out = {"metric": "aggregate particle-steps/s, 512 replicas", "value": value, "unit": "particle-steps/s",
           "steps": K, "repeats": REPEATS, "ms_per_step": ms / K, "ms_per_step_runs": [t / K for t in times],
           "warmup": cell_steps + tune_steps,
           "config": {"workload": "bd-512-replicas-500..20000-beads-density1e-4", "replicas": len(specs),
                      "beads": total, "integrator": "brownian", "cutoff_A": 30.0, "skin_A": 5.0,
                      "rebuild_period": 20},
           "launch": {"blocks_per_cu": choice["blocks_per_cu"], "use_graph": choice["use_graph"],
                      "tuned_ms_per_step": choice["ms_per_step"]},
           "ms_per_step_eager": by_graph[False], "ms_per_step_graph": by_graph[True],
           "cells": {"per_size_bucket": {f"2^{bk}": {"home_target": v[0], "fine_offset": v[1]}
                                         for bk, v in cells["cells"].items()},
                     "walk_ms": cells["walk_ms"]},
           "step_kernel_avg_us": kms / max(launches, 1) * 1e3,
           "roofline": {"bound": "hbm", "b_alg_per_particle_step": b_alg, "achieved": value * b_alg / 1e9,
                        "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": value * b_alg / HBM_PEAK}}
    del b
    if sequential:
        # the same replicas as separate Patches, one after another on the same stream
        patches = [marsmd.Patch(s) for s in specs]
        for p in patches:
            p.run(warmup)
        Kseq = min(K, 40)
        first = [warmup]

        def run_seq(rep):
            for p in patches:
                p.run(Kseq, first_step=first[0])
            first[0] += Kseq

        tseq = median(timed_runs(torch, run_seq, reps=3))
        seq = total * Kseq / (tseq / 1e3)
        out["sequential_patches"] = {"value": seq, "steps": Kseq, "repeats": 3, "ms_per_step": tseq / Kseq}
        out["speedup_over_sequential"] = value / seq
    return out


def bench_fm(torch, evals: int, warmup: int) -> dict:
    """Gradient evaluations per second of force matching: forces of 64 BD snapshots of
    100k beads for 10 tables x 512 knots, MSE loss, FP64 adjoint, Adam step (median of
    REPEATS runs of `evals` evaluations)."""
    from marsmd import configs
    from marsmd.fm import ForceMatching, force_matching_loss

    n, L, S = 100_000, 1000.0, 64
    frames, type_ids = configs.bd_snapshots(n, L, S)
    w = configs.config4(n=n, L=L, snapshots=S, frames=frames)
    assert (w.type_ids == type_ids).all()
    fm = ForceMatching(w.positions, w.type_ids, w.box, w.tables, w.term_matrix, w.cutoff)
    target = fm.forces(fm.knots(w.target_tables))[..., :3].clone()
    knots = fm.knots().clone().requires_grad_(True)
    opt = torch.optim.Adam([knots], lr=0.02)
    last = [None]

    def one():
        opt.zero_grad()
        loss = force_matching_loss(knots, fm, target)
        loss.backward()
        opt.step()
        last[0] = loss
        return loss

    first = one().item()
    for _ in range(max(warmup, 2)):
        one()
    K = max(evals, 10)
    times = timed_runs(torch, lambda rep: [one() for _ in range(K)])
    ms = median(times)
    value = K / (ms / 1e3)
    return {"metric": "force-matching gradient evaluations/s (64 BD snapshots x 100k beads, 10 tables x 512 knots)",
            "value": value, "unit": "grad evals/s", "evals": K, "repeats": REPEATS, "ms_per_eval": ms / K,
            "ms_per_eval_runs": [t / K for t in times], "dtype": "f32+f64",
            "config": {"workload": w.name, "snapshots": S, "beads": n, "tables": len(w.tables),
                       "knots_per_table": 512, "pairs_total": fm.num_pairs()},
            "roofline": {"bound": "hbm", "bytes_per_eval": FM_BYTES_PER_EVAL, "achieved": value * FM_BYTES_PER_EVAL / 1e9,
                         "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": value * FM_BYTES_PER_EVAL / HBM_PEAK},
            "loss_first": first, "loss_last": last[0].item()}


def valu_issue(torch, choice, spec):
    """The fused step kernel's VALU issue share (rocprofv3 SQ counters, profiles/) at the
    launch shape kept, or None when no profile of that shape is there."""
    sq = ROOT / "profiles" / "k_step_langevin_1M_grid256_sq.json"
    if not sq.exists() or choice["split_pmf"]:
        return None
    rec = json.loads(sq.read_text())
    # the persistent step kernel's threads: workgroups per CU (0: all that fit, 4 at its
    # 128-VGPR bound) x CUs x 256
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    threads = (choice["blocks_per_cu"] or 4) * cus * 256
    if rec.get("workload") != spec.name or str(threads) not in rec["by_threads"]:
        return None
    return {"valu_share": rec["by_threads"][str(threads)]["valu_share"], "source": str(sq.relative_to(ROOT))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=9)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="leave the replica batch out of the line")
    ap.add_argument("--no-fm", action="store_true", help="leave force matching out of the line")
    ap.add_argument("--workload", choices=["langevin", "batch", "fm"], default="langevin",
                    help="langevin: the per-step path at 1M beads (with the batch and force matching in the "
                         "same line); batch: only the 512 BD replicas; fm: only force matching")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a streaming kernel of known size (reads and writes 1 GiB each), the "
                         "scale for hardware byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_probe:
        copy_probe(torch)
        return
    if args.workload == "fm":
        print(json.dumps(dict(bench_fm(torch, max(args.steps // 10, 10), max(args.warmup // 4, 2)), n_gpus=1,
                              higher_is_better=True, scaling="weak", vs_baseline=None, data="synthetic (seeded BD frames)")))
        return
    if args.workload == "batch":
        print(json.dumps(dict(bench_batch(torch, args.steps, args.warmup), n_gpus=1, higher_is_better=True,
                              scaling="weak", vs_baseline=None, dtype="f32", data="synthetic (seeded)")))
        return

    import marsmd
    from marsmd import configs

    copy_bw = copy_bandwidth(torch)
    spec = configs.config2(n=args.n, grid_n=args.grid)
    patch = marsmd.Patch(spec)
    stream = torch.cuda.current_stream()
    # warmup: the autotuners' real steps (the walk's cell configuration timed on
    # re-walks of the first build's positions, then launch candidates timed on the step
    # kernels), then the configuration they keep
    warmup = max(args.warmup, 120)
    cell_steps = 10
    cells = patch.autotune_cells(cell_steps)
    choice = patch.autotune(warmup - cell_steps, first_step=cell_steps)
    torch.cuda.synchronize()

    # timed region (the value): REPEATS runs of `steps` steps, each one run() call
    # bracketed by HIP events on the stream the library launches on, nothing
    # instrumented inside; the median run is the value
    step = [warmup]

    def run(rep):
        patch.run(args.steps, first_step=step[0])
        step[0] += args.steps

    times = timed_runs(torch, run, stream=stream)
    total_ms = median(times)
    # the step kernel's duration: as many further steps with the library bracketing
    # every step's kernels (the fused step kernel, and the PMF kernel when split) with
    # HIP events on the same stream (the events add dispatch gaps, which is why this
    # pass is not the value's)
    patch.set_kernel_timing(True)
    patch.run(args.steps, first_step=step[0])
    kernel_ms, launches = patch.kernel_timing()
    patch.set_kernel_timing(False)
    ms_step = total_ms / args.steps
    value = spec.n * args.steps / (total_ms / 1e3)
    acc = algorithmic_bytes(spec)
    t_kernel = kernel_ms / max(launches, 1) * 1e-3
    traffic, traffic_src = None, None
    pmc = ROOT / "profiles" / "k_step_langevin_1M_grid256_pmc.json"
    if pmc.exists():
        rec = json.loads(pmc.read_text())
        if rec.get("workload") == spec.name:
            traffic, traffic_src = rec["hbm_bytes_per_launch"], str(pmc.relative_to(ROOT))
    # the step kernel is issue-bound on VALU work rather than on HBM: quote the measured
    # VALU share of its issue cycles (rocprofv3 SQ counters) for the launch shape kept
    valu = valu_issue(torch, choice, spec)
    achieved = acc["step_kernel_bytes_per_launch"] / t_kernel
    del patch
    torch.cuda.synchronize()
    out = {
        "metric": "particle-steps/s at 1M beads",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warmup,
        "ms_per_step": ms_step,
        "repeats": REPEATS,
        "ms_per_step_runs": [t / args.steps for t in times],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded)",
        "config": {"workload": spec.name, "beads": spec.n, "box_A": spec.box.size[0], "types": len(spec.types),
                   "pair_tables": len(spec.nb_tables), "grid": f"{args.grid}^3 tricubic", "integrator": "langevin",
                   "cutoff_A": spec.params.cutoff, "skin_A": spec.params.pairlist_cutoff - spec.params.cutoff,
                   "rebuild_period": spec.params.rebuild_period},
        "hbm_roofline_frac_whole_step": value * acc["b_alg_per_particle_step"] / HBM_PEAK,
        "b_alg_per_particle_step": acc["b_alg_per_particle_step"],
        "copy_bandwidth": copy_bw,
        "launch": {"blocks_per_cu": choice["blocks_per_cu"], "split_pmf": choice["split_pmf"],
                   "tuned_ms_per_step": choice["ms_per_step"]},
        "cells": cells,
        "roofline": {"bound": "hbm",
                     "limiter": "valu",
                     "kernel": ("k_pmf_force + k_step (grid force, then pair force + BAOAB)" if choice["split_pmf"]
                                else "k_step (fused pair force + PMF + BAOAB)"),
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK,
                     "frac_of_measured_copy": achieved / (copy_bw["GBps"] * 1e9),
                     "traffic": traffic, "traffic_unit": "B/launch",
                     "traffic_source": traffic_src,
                     "avg_launch_us": t_kernel * 1e6, "launches": launches,
                     "other_ms_per_step": ms_step - kernel_ms / max(launches, 1),
                     "valu_frac": None if valu is None else valu["valu_share"],
                     "issue": valu},
    }
    if not args.no_batch:
        out["batch"] = bench_batch(torch, args.steps, args.warmup)
        torch.cuda.synchronize()
    if not args.no_fm:
        out["fm"] = bench_fm(torch, max(args.steps // 2, 10), max(args.warmup // 4, 2))
        torch.cuda.synchronize()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(spec, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
