"""Throughput of the per-step particle hot path on one MI355X.

Workload (default): 1,000,000 beads of 3 types in a 2,000 A periodic box, 6 tabulated
WCA+DH pair tables (1,024 knots), a 256^3 FP32 membrane grid sampled tricubically,
Langevin (BAOAB, gamma = 1/ps), cutoff 30 A + 5 A skin, rebuild period 20.

A step is SimManager's per-step sequence (rebuild policy, forces, PMF, integrator)
over the whole system. Inputs are resident in HBM before the timed region; timing
uses HIP events on the stream the library launches on. Prints one JSON line.
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


def bench_batch(args, torch):
    """Aggregate particle-steps/s of 512 independent BD replicas (N log-uniform in
    [500, 20000], density 1e-4, config-0 table) stepped together, and of the same
    replicas stepped one after another as separate Patches."""
    import marsmd
    from marsmd import configs
    from marsmd.batch import Batch

    specs = configs.config3()
    total = sum(s.n for s in specs)
    K = args.steps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # the default launch (eager, every workgroup that fits), then the autotuned one on a
    # fresh batch: the autotuner's steps are the second batch's warmup
    b = Batch(specs)
    b.run(args.warmup)
    torch.cuda.synchronize()
    e0.record()
    b.run(K, first_step=args.warmup)
    e1.record()
    torch.cuda.synchronize()
    ms_default = e0.elapsed_time(e1)
    del b
    b = Batch(specs)
    tune_steps = max(args.warmup, 160)
    choice = b.autotune(tune_steps, cache=False)
    torch.cuda.synchronize()
    e0.record()
    b.run(K, first_step=tune_steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    # step-kernel timing in a pass of its own (its events add dispatch gaps)
    b.set_kernel_timing(True)
    b.run(K, first_step=tune_steps + K)
    kms, launches = b.kernel_timing()
    b.set_kernel_timing(False)
    # the same replicas as separate Patches, one after another on the same stream
    patches = [marsmd.Patch(s) for s in specs]
    for p in patches:
        p.run(args.warmup)
    torch.cuda.synchronize()
    Kseq = min(K, 40)
    e0.record()
    for p in patches:
        p.run(Kseq, first_step=args.warmup)
    e1.record()
    torch.cuda.synchronize()
    ms_seq = e0.elapsed_time(e1)
    value = total * K / (ms / 1e3)
    seq = total * Kseq / (ms_seq / 1e3)
    print(json.dumps({
        "metric": "aggregate particle-steps/s, 512 replicas", "value": value, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": K, "warmup": tune_steps, "ms_per_step": ms / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded)",
        "config": {"workload": "bd-512-replicas-500..20000-beads-density1e-4", "replicas": len(specs),
                   "beads": total, "integrator": "brownian", "cutoff_A": 30.0, "skin_A": 5.0, "rebuild_period": 20},
        "launch": {"blocks_per_cu": choice["blocks_per_cu"], "use_graph": choice["use_graph"],
                   "tuned_ms_per_step": choice["ms_per_step"]},
        "default_launch_ms_per_step": ms_default / K,
        "step_kernel_avg_us": kms / max(launches, 1) * 1e3,
        "sequential_patches": {"value": seq, "steps": Kseq, "ms_per_step": ms_seq / Kseq},
        "speedup_over_sequential": value / seq}))


def bd_frames(n: int, L: float, snapshots: int, seed: int = 42, warmup: int = 100, stride: int = 10):
    """Brownian-dynamics frames of the 4-type WCA+DH system (the force-matching
    workload's snapshots), in global-id order, from the GPU Patch."""
    import numpy as np

    import marsmd
    from marsmd import configs, rng
    from marsmd.system import RunParams, SystemSpec
    from marsmd.core import ParticleType
    from marsmd import _capi as api

    tables, tm = configs.fm_tables()
    type_ids = np.minimum((rng.uniform(seed, rng.STREAM_POSITIONS ^ 0x4, n)[:, 0] * 4).astype(np.int32), 3)
    q = [1.0, -1.0, 0.5, -0.5]
    spec = SystemSpec(box=marsmd.PeriodicBox((L, L, L)),
                      types=[ParticleType(mass=1.0, charge=qi, diffusion=(150.0,) * 3) for qi in q],
                      positions=configs.uniform_positions(n, L, seed), type_ids=type_ids,
                      params=RunParams(cutoff=30.0, pairlist_cutoff=35.0, rebuild_period=20, timestep=1e-4,
                                       kT=configs.KT, integrator=api.INTEGRATOR_BROWNIAN),
                      nb_tables=tables, term_matrix=tm, seed=seed, name="fm-bd")
    patch = marsmd.Patch(spec)
    patch.run(warmup)
    out = np.zeros((snapshots, n, 3), np.float32)
    step = warmup
    for s in range(snapshots):
        patch.run(stride, first_step=step)
        step += stride
        h = patch.copy_particles_to_host()
        out[s, h["id"]] = h["pos"][:, :3]
    return out, type_ids


def bench_fm(args, torch):
    """Gradient evaluations per second of force matching: forces of 64 snapshots of
    100k beads for 10 tables x 512 knots, MSE loss, FP64 adjoint, Adam step."""
    from marsmd import configs
    from marsmd.fm import ForceMatching, force_matching_loss

    n, L, S = 100_000, 1000.0, 64
    frames, type_ids = bd_frames(n, L, S)
    w = configs.config4(n=n, L=L, snapshots=S, frames=frames)
    assert (w.type_ids == type_ids).all()
    fm = ForceMatching(w.positions, w.type_ids, w.box, w.tables, w.term_matrix, w.cutoff)
    target = fm.forces(fm.knots(w.target_tables))[..., :3].clone()
    knots = fm.knots().clone().requires_grad_(True)
    opt = torch.optim.Adam([knots], lr=0.02)

    def one():
        opt.zero_grad()
        loss = force_matching_loss(knots, fm, target)
        loss.backward()
        opt.step()
        return loss

    first = one().item()
    for _ in range(max(args.warmup // 4, 2)):
        one()
    K = max(args.steps // 10, 10)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(K):
        loss = one()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    print(json.dumps({
        "metric": "force-matching gradient evaluations/s (64 BD snapshots x 100k beads, 10 tables x 512 knots)",
        "value": K / (ms / 1e3), "unit": "grad evals/s", "n_gpus": 1, "steps": K, "warmup": max(args.warmup // 4, 2),
        "ms_per_step": ms / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded BD frames)",
        "config": {"workload": w.name, "snapshots": S, "beads": n, "tables": len(w.tables), "knots_per_table": 512,
                   "pairs_total": fm.num_pairs()},
        "loss_first": first, "loss_last": loss.item()}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=9)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["langevin", "batch", "fm"], default="langevin",
                    help="langevin: the per-step path at 1M beads; batch: 512 BD replicas (500..20000 beads) "
                         "stepped together; fm: force-matching gradient evaluations (64 BD snapshots x 100k "
                         "beads, 10 tables x 512 knots)")
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
        bench_fm(args, torch)
        return
    if args.workload == "batch":
        bench_batch(args, torch)
        return

    import marsmd
    from marsmd import configs

    spec = configs.config2(n=args.n, grid_n=args.grid)
    patch = marsmd.Patch(spec)
    stream = torch.cuda.current_stream()
    # warmup: the autotuner's real steps (launch candidates timed on the step kernels),
    # then the configuration it keeps
    warmup = max(args.warmup, 60)
    choice = patch.autotune(warmup)
    torch.cuda.synchronize()

    # timed region (the value): `steps` steps in one run() call, bracketed by HIP
    # events on the stream the library launches on, nothing instrumented inside
    first = warmup
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    patch.run(args.steps, first_step=first)
    e1.record(stream)
    torch.cuda.synchronize()
    total_ms = e0.elapsed_time(e1)
    # the step kernel's duration: as many further steps with the library bracketing
    # every step's kernels (the fused step kernel, and the PMF kernel when split) with
    # HIP events on the same stream (the events add dispatch gaps, which is why this
    # pass is not the value's)
    patch.set_kernel_timing(True)
    patch.run(args.steps, first_step=first + args.steps)
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
    valu = None
    sq = ROOT / "profiles" / "k_step_langevin_1M_grid256_sq.json"
    if sq.exists() and not choice["split_pmf"]:
        rec = json.loads(sq.read_text())
        # the persistent step kernel's threads: workgroups per CU (0: all that fit, 4 at its
        # 128-VGPR bound) x CUs x 256
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        threads = (choice["blocks_per_cu"] or 4) * cus * 256
        if rec.get("workload") == spec.name and threads is not None and str(threads) in rec["by_threads"]:
            valu = {"valu_share": rec["by_threads"][str(threads)]["valu_share"], "source": str(sq.relative_to(ROOT))}
    achieved = acc["step_kernel_bytes_per_launch"] / t_kernel
    out = {
        "metric": "particle-steps/s at 1M beads",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warmup,
        "ms_per_step": ms_step,
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
        "launch": {"blocks_per_cu": choice["blocks_per_cu"], "split_pmf": choice["split_pmf"],
                   "tuned_ms_per_step": choice["ms_per_step"]},
        "roofline": {"bound": "hbm",
                     "kernel": ("k_pmf_force + k_step (grid force, then pair force + BAOAB)" if choice["split_pmf"]
                                else "k_step (fused pair force + PMF + BAOAB)"),
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_unit": "B/launch",
                     "traffic_source": traffic_src,
                     "avg_launch_us": t_kernel * 1e6, "launches": launches,
                     "other_ms_per_step": ms_step - kernel_ms / max(launches, 1),
                     "issue": valu},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(spec, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
