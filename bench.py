"""Benchmark of the RDME step on one MI355X: voxel-steps/s on the 256^3 large-cell workload.

Workload: cell_simulation_amd.configs.large_cell -- 256^3 lattice, 16 one-byte slots per voxel, 4
regions, 20 species, 30 reactions, ~2M particles (seed 1234), dt = 10 us. Inputs are resident in HBM
before the timed region. One "step" = one full RDME step (diffusion, reaction, compaction) of the whole
lattice.

Prints ONE JSON line. `roofline` prices the step kernel against HBM: algorithmic bytes per voxel-step
= 2*S (read and write every slot) + 1 (region label) = 33 B at S=16, divided by the kernel's average
duration measured with HIP events on the simulation's stream. `cpu_baseline` times the CPU restatement of the
reference shaders on every host core this process may use (hop draws, reactions and compaction in parallel,
claims in the canonical order; bit-exact against the serial oracle) on a bounded sample of the same workload,
and the 64^3 spherical cell for its full 1000 steps. `configs` times the other single-cell workloads on the
GPU (64^3 for its 1000 steps, 128^3 gene expression for its 10000 steps), `batched` the 512-cell ensemble as
stated (empty lattices, genes off) and started near its stationary state, against the same cells' plain
launches, and `precision` the 512^3 precision-stress workload with its FP32-vs-FP64 audit.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def calibrate_copy(mib: int, steps: int) -> int:
    """Device-to-device copy of `mib` MiB, `steps` times: reads and writes exactly mib MiB per launch."""
    import torch

    n = mib << 20
    src = torch.randint(0, 255, (n,), dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read_per_launch": n, "bytes_written_per_launch": n,
                      "launches": steps}), flush=True)
    return 0


def ensemble_leg(args, populated: bool = False) -> dict:
    """Replicate cells (gene-expression network, edges log-uniform 32..96, rates x LogNormal(0.2), S=8,
    seed 99): one batched Ensemble (autotuned, hipGraph blocks) against the same cells stepped one
    plain launch sequence per cell. Values are replicate-voxel-steps per second. `populated`: each cell starts
    near its stationary state (genes on with probability k_on / (k_on + k_off), mRNA and protein at their
    stationary means) instead of empty with its genes off."""
    from cell_simulation_amd import Ensemble, configs, default_itime_schedule

    log(f"ensemble: building {args.ensemble_cells} cells (populated={populated})")
    cells, seeds = configs.ensemble_cells(args.ensemble_cells, seed=99, populated=populated)
    particles = sum(int(c.species_counts()[1:].sum()) for c in cells)
    ens = Ensemble(cells, seeds)
    vox = ens.info()["total_voxels"]
    ens.run(2)  # warm-up (first launches, page-in)
    ens.synchronize()
    tune_ms = ens.autotune(3)
    log(f"ensemble: {vox} voxels per step, autotuned tiling {ens.tiling} ({tune_ms:.3f} ms/step in trials)")
    ens_ms = ens.run_timed(args.ensemble_steps, args.graph_block)
    batched_value = vox * args.ensemble_steps / (ens_ms / 1e3)
    c = ens.counters()
    tiling = ens.tiling
    ens.close()

    if args.plain_steps <= 0:  # the batched leg alone (profiling runs)
        for sim in cells:
            sim.close()
        return {"populated": populated, "particles": particles, "voxels_per_step": vox,
                "unit": "replicate-voxel-steps/s", "batched_value": batched_value,
                "batched_ms_per_step": ens_ms / args.ensemble_steps, "batched_steps": args.ensemble_steps,
                "graph_block": args.graph_block, "tiling": tiling, "plain_value": None, "speedup": None,
                "cme_events": int(c[5]), "lattice_reactions_fired": int(c[3]), "exceptions": int(c[1])}
    log(f"plain: preparing {len(cells)} simulations")
    for sim in cells:
        sim.prepare_for_gpu(0)
    its = [default_itime_schedule(s, args.plain_steps + 1) for s in seeds]
    for sim, it in zip(cells, its):  # warm-up step
        sim.step(0, int(it[0]), 0)
    for sim in cells:
        sim.synchronize()
    t0 = time.perf_counter()
    for k in range(1, args.plain_steps + 1):
        for sim, it in zip(cells, its):
            sim.step(k, int(it[k]), 0)
    for sim in cells:
        sim.synchronize()
    plain_s = time.perf_counter() - t0
    plain_value = vox * args.plain_steps / plain_s
    for sim in cells:
        sim.close()
    b_alg = 2 * 8 + 1
    start = ("started near the stationary state (genes on with probability k_on/(k_on+k_off), mRNA and protein at "
             "their stationary means)" if populated else "started as stated: empty lattices, genes off")
    return {
        "workload": f"{len(cells)} replicate cells, gene-expression network (well-mixed gene switching, lattice "
                    f"mRNA/protein), edges log-uniform 32..96, rates x LogNormal(sigma 0.2), S=8, dt=8us (every replicate's "
                    f"hop vector sums below 1), seed 99, {start}",
        "populated": populated, "particles": particles,
        "voxels_per_step": vox, "unit": "replicate-voxel-steps/s",
        "batched_value": batched_value, "batched_ms_per_step": ens_ms / args.ensemble_steps,
        "batched_steps": args.ensemble_steps, "graph_block": args.graph_block, "tiling": tiling,
        "plain_value": plain_value, "plain_ms_per_step": plain_s * 1e3 / args.plain_steps,
        "plain_steps": args.plain_steps, "plain_path": "one cs_sim per cell, its launches per step (own stream)",
        "speedup": batched_value / plain_value,
        "bytes_per_voxel_step": b_alg,
        # the batched kernel skips the (tile, plane) pieces of a cell that are empty and were empty, so it moves
        # fewer bytes than b_alg per voxel-step: this dense-equivalent rate is not an HBM figure
        "batched_dense_equivalent_gbs": batched_value * b_alg / 1e9,
        "batched_hbm": _ensemble_hbm(args, ens_ms / args.ensemble_steps, populated),
        "cme_events": int(c[5]), "lattice_reactions_fired": int(c[3]), "exceptions": int(c[1]),
    }


def _ensemble_hbm(args, ms_per_step: float, populated: bool):
    """HBM bytes the batched step moved (rocprofv3 FETCH_SIZE / WRITE_SIZE per step, scaled by the copy
    calibration, from the traffic JSON of this start state) over this run's time per step, against the 8 TB/s
    peak."""
    path = args.ensemble_traffic_json
    if populated:
        path = path.replace("_traffic.json", "_populated_traffic.json")
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        tj = json.load(f)
    if tj.get("cells") != args.ensemble_cells or bool(tj.get("populated", False)) != populated:
        return None
    gbs = tj["bytes_per_step"] / (ms_per_step * 1e-3) / 1e9
    return {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "traffic_per_step": tj["bytes_per_step"], "source": os.path.relpath(path, ROOT)}


def precision_leg(steps: int) -> dict:
    """512^3 precision-stress workload (12 lattice + 40 well-mixed species, 120 reactions, ~10M particles, S=8):
    `steps` steps with the device's f64 audit of every well-mixed event (compensated f32 total vs plain f64
    total; choice under f64 vs the f32 choice)."""
    from cell_simulation_amd import configs, default_itime_schedule

    log("precision: building the 512^3 workload")
    sim = configs.precision_stress()
    sim.prepare_for_gpu(0)
    sim.set_cme_audit(True)
    its = default_itime_schedule(2024, steps)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(0, its)
    sim.synchronize()
    wall = time.perf_counter() - t0
    a = sim.cme_audit()
    c = sim.counters()
    inf = sim.info()
    return {"workload": "precision_stress 512^3 S=8, 12 lattice + 40 well-mixed species, 120 reactions (40 lattice "
                        "reactions with rate constants log-uniform 1e-6..1e1 /s; 80 well-mixed reactions whose "
                        "propensities at the initial counts are log-uniform 1e-6..1e4 /s), ~10M particles, dt=5us, "
                        "seed 2024",
            "steps": steps, "ms_per_step": wall * 1e3 / steps,
            "voxel_steps_per_s": inf["n_voxels"] * steps / wall,
            "cme_events": a["events"], "max_rel_propensity_error_f32_kahan_vs_f64": a["max_rel_error"],
            "selection_divergences": a["divergences"],
            "selection_divergence_rate": a["divergences"] / a["events"] if a["events"] else None,
            "lattice_reactions_fired": int(c[3]), "near_threshold": int(c[2]), "exceptions": int(c[1]),
            "tight_voxel_steps": int(c[0]), "failed_claims": int(c[8]), "slot_s_aliases": int(c[9])}


def host_cores() -> tuple[int, str]:
    """The cores this process may run on, and the CPU model (from /proc/cpuinfo)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model


def cpu_leg(args, N: int, S: int) -> dict:
    """The CPU restatement on every core this process may use: a bounded sample of the 256^3 workload (one
    warm-up step, then args.cpu_steps timed), and the 64^3 spherical cell for its full 1000 steps."""
    from cell_simulation_amd import configs, default_itime_schedule
    from oracle.cs_oracle import Oracle

    cores, model = host_cores()
    log(f"cpu baseline: {cores} threads on {model}, {args.cpu_steps} steps of {args.edge}^3")
    orc = Oracle.from_simulation(configs.large_cell(seed=1234, edge=args.edge))
    ci = default_itime_schedule(1234, args.cpu_steps + 1)
    orc.run_mt(ci[:1], cores)  # warm-up step (page-in), not timed
    t1 = time.perf_counter()
    orc.run_mt(ci[1:], cores)
    ct = time.perf_counter() - t1
    del orc
    out = {"value": N * args.cpu_steps / ct, "unit": "voxel-steps/s", "cores": cores, "kind": "port",
           "sample": f"{args.cpu_steps} step(s) of the {args.edge}^3 S={S} large-cell workload after one warm-up "
                     f"step, C restatement of the reference shaders on {cores} threads (hop draws, reactions and "
                     "compaction in parallel, claims in canonical order; bit-exact against the serial oracle)",
           "cpu_model": model}
    if args.cpu_cfg1:
        log("cpu baseline: 64^3 spherical cell, 1000 steps")
        orc = Oracle.from_simulation(configs.spherical_cell(seed=42))
        ci = default_itime_schedule(42, 1000)
        t1 = time.perf_counter()
        orc.run_mt(ci, cores)
        ct = time.perf_counter() - t1
        out["spherical_cell_64"] = {"value": 64 ** 3 * 1000 / ct, "unit": "voxel-steps/s", "steps": 1000,
                                    "seconds": ct}
    return out


def kernel_variants(args, steps: int) -> dict:
    """The 256^3 step on each step kernel the library can take, `steps` timed steps each (HIP events around the
    kernel, after 5 warm-up steps): the wave-march kernel with its tables in LDS (the default), the same kernel with
    only the hop tables in LDS and the reaction tables in global memory (the form it takes when the tables pass
    24 KB, for instance with hundreds of species and reactions), with every table in global memory (when even the
    hop tables pass 24 KB), the block-shared plane-march kernel and the one-lane-per-voxel kernel."""
    from cell_simulation_amd import configs, default_itime_schedule

    variants = [("wave_lds_tables", {}), ("wave_split_tables", {"CS_WAVE_GLOBAL_TABLES": "1"}),
                ("wave_global_tables", {"CS_WAVE_GLOBAL_TABLES": "2"}),
                ("tiled", {"CS_STEP_KERNEL": "tiled"}), ("one_lane_per_voxel", {"CS_STEP_KERNEL": "simple"})]
    out = {}
    for name, env in variants:
        sim = configs.large_cell(seed=1234, edge=args.edge)
        info = sim.info()
        N, S = info["n_voxels"], info["max_particles_site"]
        os.environ.update(env)
        try:
            sim.prepare_for_gpu(0)
        finally:
            for k in env:
                os.environ.pop(k, None)
        its = default_itime_schedule(1234, 5 + steps)
        sim.run(0, its[:5])
        ms = sim.run_timed(its[5:]) / steps
        gbs = (2 * S + 1) * N / (ms * 1e-3) / 1e9
        out[name] = {"ms_per_step": ms, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS}
        log(f"kernel {name}: {ms:.4f} ms/step")
        sim.close()
    return out


def single_cell_configs(args) -> dict:
    """The 64^3 spherical cell for its 1000 steps and the 128^3 gene-expression cell for its 10000 steps (as
    stated, and started near its stationary state), plain one-launch-per-step path, wall time of the run."""
    from cell_simulation_amd import configs, default_itime_schedule

    out = {}
    runs = [("spherical_cell_64", lambda: configs.spherical_cell(seed=42), 42, 1000),
            ("gene_expression_128", lambda: configs.gene_expression(seed=7), 7, 10000),
            ("gene_expression_128_populated", lambda: configs.gene_expression(seed=7, populated=True), 7, 10000)]
    for name, build, seed, steps in runs:
        sim = build()
        n = sim.info()["n_voxels"]
        sim.prepare_for_gpu(0)
        its = default_itime_schedule(seed, steps + 1)
        sim.step(0, int(its[0]), 0)  # warm-up
        sim.synchronize()
        t0 = time.perf_counter()
        sim.run(1, its[1:])
        sim.synchronize()
        wall = time.perf_counter() - t0
        c = sim.counters()
        out[name] = {"voxels": n, "steps": steps, "value": n * steps / wall, "unit": "voxel-steps/s",
                     "ms_per_step": wall * 1e3 / steps, "particles": int(sim.species_counts()[1:].sum()),
                     "fired": int(c[3]), "cme_events": int(c[5]), "exceptions": int(c[1])}
        log(f"{name}: {wall * 1e3 / steps:.4f} ms/step")
        sim.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed steps (the 256^3 workload's run length: 2000)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--edge", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="timed steps of the CPU baseline on the 256^3 workload, after one warm-up step (0 = skip)")
    ap.add_argument("--cpu-cfg1", type=int, default=1, help="also time the CPU baseline on the 64^3 cell (1/0)")
    ap.add_argument("--kernel-variants", type=int, default=50, metavar="STEPS",
                    help="time the 256^3 step on every step kernel for this many steps (0: skip)")
    ap.add_argument("--configs", type=int, default=1,
                    help="also time the 64^3 and 128^3 single-cell workloads on the GPU (1/0)")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "large_cell_256_traffic.json"),
                    help="per-launch HBM bytes measured with rocprofv3 --pmc (FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--ensemble-cells", type=int, default=512,
                    help="replicate cells of the batched-vs-plain comparison (0 = skip it)")
    ap.add_argument("--ensemble-steps", type=int, default=5000,
                    help="timed steps of the batched ensemble (the workload's run length: 5000 steps in 100-step graphs)")
    ap.add_argument("--plain-steps", type=int, default=100, help="timed steps of the per-cell plain launches")
    ap.add_argument("--populated", type=int, default=1,
                    help="also run the batched leg with the cells started near their stationary state (1/0)")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the batched-ensemble leg and print its record (for profiling it)")
    ap.add_argument("--ensemble-traffic-json", default=os.path.join(ROOT, "profiles", "ensemble_512_traffic.json"),
                    help="HBM bytes per batched step measured with rocprofv3 --pmc (FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--graph-block", type=int, default=100, help="steps per captured hipGraph")
    ap.add_argument("--precision", type=int, default=1000, metavar="STEPS",
                    help="also run the 512^3 precision-stress workload for STEPS steps (its run length: 1000) with "
                         "the f64 audit of the well-mixed propensities and report it (0 = skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="instead of the step: copy a MIB-MiB device buffer --steps times (a kernel of known "
                         "bytes for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy, args.steps)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: the RDME step is benchmarked per device and this build "
              "has no multi-GPU path (slab decomposition is a design note only).", file=sys.stderr)
        return 2

    from cell_simulation_amd import configs, default_itime_schedule

    if args.ensemble_only:
        print(json.dumps({"batched": ensemble_leg(args, populated=bool(args.populated))}), flush=True)
        return 0

    log(f"building large_cell edge={args.edge}")
    sim = configs.large_cell(seed=1234, edge=args.edge)
    info = sim.info()
    N, S = info["n_voxels"], info["max_particles_site"]
    sim.prepare_for_gpu(0)
    log("device ready")
    its = default_itime_schedule(1234, args.warmup + args.steps)
    if args.warmup:
        sim.run(0, its[: args.warmup])
    sim.synchronize()
    # the state the timed steps start from, written back before each of the two passes below so that both run
    # the same steps on the same data (the write also re-plans the kernel's work items from this state)
    start_state = sim.lattice()
    sim.write_lattice(start_state)
    sim.reset_counters()

    # Timed region: K steps enqueued back to back, bracketed by stream synchronisation.
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(args.warmup, its[args.warmup:])
    sim.synchronize()
    wall = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {wall:.4f} s")
    counters = sim.counters()

    # Kernel duration with HIP events around each launch, on the simulation's stream: a second pass over the
    # same steps (same start state, same itimes), so the events time exactly the work the wall clock timed.
    sim.write_lattice(start_state)
    kernel_ms = sim.run_timed(its[args.warmup:])
    avg_kernel_s = kernel_ms / 1e3 / args.steps
    del start_state

    b_alg = 2 * S + 1
    achieved_gbs = b_alg * N / avg_kernel_s / 1e9
    kernel = {"simple": "rdme_step_kernel", "tiled": "rdme_step_tiled"}.get(os.environ.get("CS_STEP_KERNEL", ""),
                                                                          "rdme_step_wave")
    traffic = None
    if args.traffic_json and os.path.exists(args.traffic_json):
        with open(args.traffic_json) as f:
            tj = json.load(f)
        # measured bytes apply only to the kernel and lattice they were measured on
        if tj.get("edge") == args.edge and str(tj.get("kernel", "")).startswith(kernel):
            traffic = tj.get("bytes_per_launch")

    value = N * args.steps / wall
    sim.close()
    cpu = cpu_leg(args, N, S) if args.cpu_steps > 0 else None
    single = single_cell_configs(args) if args.configs else None
    variants = kernel_variants(args, args.kernel_variants) if args.kernel_variants > 0 else None

    batched = None
    if args.ensemble_cells > 0:
        batched = ensemble_leg(args)
        if args.populated:
            batched["populated_run"] = ensemble_leg(args, populated=True)
    precision = precision_leg(args.precision) if args.precision > 0 else None

    out = {
        "metric": "voxel-steps/s (RDME diff+react, 256^3) batched vs plain; % of HBM peak",
        "value": value,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"large_cell {args.edge}^3 S={S}: 4 regions, 20 species, 30 reactions, "
                               "~2M particles, dt=10us, seed 1234, plain one-launch-per-step path",
                   "voxels": N, "slots_per_voxel": S, "path": "plain"},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel, "bytes_per_voxel_step": b_alg,
                     "avg_kernel_us": avg_kernel_s * 1e6},
        "cpu_baseline": cpu,
        "diagnostics": {"tight_voxel_steps": int(counters[0]), "exceptions": int(counters[1]),
                        "near_threshold": int(counters[2]), "fired": int(counters[3]),
                        "failed_claims": int(counters[8]), "slot_s_aliases": int(counters[9])},
        "configs": single,
        "kernel_variants": variants,
        "batched": batched,
        "precision": precision,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
