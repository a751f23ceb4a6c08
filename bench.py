"""Offspring-pipeline throughput on one MI355X.

Workload (single population, the configuration the metric is quoted on):
N = 100,000 diploid (H = 200,000 haplotype rows), 50,176 allele slots
(W = 784 uint64 words per row), synthetic steady-state standing variation with
column frequencies from the neutral 1/f spectrum, Poisson(5) crossovers per
gamete, one additive trait with N(0, 0.1^2) effects on new mutations,
environmental noise N(0, 1), Gaussian stabilising fitness exp(-z^2/2).  The
mutation rate is 0.01 per gamete: the rate at which ~50k segregating sites are
the neutral equilibrium for N = 1e5 (4 N mu sum 1/i ~ 5e4); the literal
Poisson(10) would add 2e6 new columns per generation (DESIGN.md has the
arithmetic).

A step is one generation of the pipeline (draws, recombination + mutation,
free space, traits, fitness, selection) over resident HBM state, replayed as
a captured hipGraph.  `value` = offspring genomes (haplotype rows) per second.
The dominant kernel (recombination) is timed with HIP events on the
library's stream in a second pass; the CPU baseline is the oracle (a C++
restatement of the reference engine) timed for one generation of the same
population with the reference's thread model (contiguous offspring batches,
one per thread) on the host's CPU share, and on one core; its output is
compared bit for bit.

The rest of the metric rides on the same line:
  mfma_vs_plain          the precision workload (200,000 haplotypes x 65,536
                         sites x 16 traits, MVN effects): exact int8 MFMA and
                         split-BF16 MFMA contraction against the plain FP32
                         kernel, with the error of each against FP64;
  batched_vs_sequential  the replicate workload (256 replicates, N
                         log-uniform on [1e3, 5e4]): one persistent launch per
                         phase and a graph per generation, against each
                         replicate's own graph in turn;
  large_population       the footprint of the N = 5e5, 16-trait population.
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

N_IND = 100_000
ALLELES = 50_176          # W = 784 words
SEGREGATING = 46_000      # columns with standing variation; the rest start free
MU, RHO = 0.01, 5.0
PEAK_HBM_GBPS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload_config(api):
    return api.make_config(N_IND, MU, RHO, seed=42, traits=1, fitness=api.FIT_GAUSSIAN_FIXED,
                           neutral_p=0.0, mean=0.0, sigma=0.1, fitness_scale=1.0, env_sigma=1.0,
                           quantum_log2=-24, max_alleles=ALLELES + 8 * 512)


def standing_variation(seed=42):
    """Per-slot frequencies from P(f) ~ 1/f on [1/H, 1 - 1/H] (inverse
    transform f = a (b/a)^u), uniform positions k 2^-32, neutral effects."""
    h = 2 * N_IND
    rng = np.random.default_rng(seed)
    a, b = 1.0 / h, 1.0 - 1.0 / h
    freqs = np.zeros(ALLELES)
    freqs[:SEGREGATING] = a * (b / a) ** rng.uniform(0, 1, SEGREGATING)
    rng.shuffle(freqs)
    positions = np.floor(rng.uniform(0, 1, ALLELES) * 2**32) / 2**32
    effects = np.zeros((ALLELES, 1))
    return freqs, positions, effects


def traffic_from_profiles():
    """HBM bytes per recombination launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated by a copy kernel), if any."""
    p = os.path.join(ROOT, "profiles", "k_recombine_N1e5_W784_rho5_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return float(json.load(f)["bytes_per_launch"])
    except Exception:
        return None


def host_cpu():
    """CPU model, logical CPUs and the CPU share this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    return model, nproc, share


def mfma_vs_plain(api):
    """The precision workload: 200,000 x 65,536 x 16, f ~ U(0,1), MVN effects."""
    from clotho_amd.bench_traits import effects_mvn, errors

    h, s, t = 200_000, 65_536, 16
    freqs = np.random.default_rng(42).uniform(0, 1, s)
    eff = effects_mvn(s, t)
    z64, _ = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_EXACT_FP64, seed=42)
    z32, ms32 = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_PLAIN_FP32, seed=42,
                                             repeats=3)
    zi, msi = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_EXACT_I8, seed=42,
                                           repeats=10)
    zb, msb = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_MFMA_BF16, seed=42,
                                           split_terms=3, repeats=10)
    n32, ni, nb = errors(z32, z64)[0], errors(zi, z64)[0], errors(zb, z64)[0]
    flops = 2.0 * h * s * t
    return {
        "workload": "H=200000 x S=65536 x T=16, genotypes Bernoulli(f), f~U(0,1) seed 42, "
                    "effects MVN(unit var, corr 0.3) seed 43; error = per trait "
                    "max|z-z*|/max|z*| against an FP64 sum",
        "plain_fp32_ms": ms32, "plain_fp32_err_max": float(n32.max()),
        "exact_i8_ms": msi, "exact_i8_err_per_trait": [float(x) for x in ni],
        "exact_i8_speedup": ms32 / msi,
        "exact_i8_exec_tops": 4 * flops / (msi * 1e-3) / 1e12,
        "bf16x3_ms": msb, "bf16x3_err_per_trait": [float(x) for x in nb],
        "bf16x3_speedup": ms32 / msb,
        "bf16x3_alg_tflops": flops / (msb * 1e-3) / 1e12,
        "bf16x3_exec_tflops": 3 * flops / (msb * 1e-3) / 1e12,
        "bf16x3_exec_frac_of_2500": 3 * flops / (msb * 1e-3) / 1e12 / 2500.0,
    }


def batched_vs_sequential(api, generations):
    """The replicate workload, both modes from the same founders."""
    from clotho_amd.bench_batch import MU_SCALE, CAPACITY, replicate_configs

    cfgs, N, _, _ = replicate_configs(api, 256)
    ms = {}
    shapes = {}
    grid_pct = None
    phases = None
    rec_bytes = None
    for mode in ("graph", "sequential"):
        b = api.Batch(cfgs)
        b.run(1, mode=mode)  # untimed first generation (and the tuner's gametes)
        if mode == "graph":
            grid_pct = b.autotune()
        ms[mode] = b.run(generations - 1, mode=mode)
        shapes[mode] = [b.replicate(r).shape().as_dict() for r in range(len(cfgs))]
        if mode == "graph":
            # per-phase kernel time of the batched launches after the timed
            # run (eager, with timing events), and the recombination's
            # algorithmic bytes at those widths: two parent rows read and one
            # child row written per offspring haplotype of every replicate
            words = np.array([s["words"] for s in shapes[mode]], dtype=np.float64)
            rec_bytes = float((24.0 * words * 2.0 * N).sum())
            b.set_timing(True)
            b.run(3, mode="eager")
            n, ph = b.timings()
            b.set_timing(False)
            phases = {k: round(float(ph[i]) / max(n, 1), 4) for i, k in
                      enumerate(["generation", "draw", "recombine", "traits", "free_space",
                                 "fitness"])}
        b.close()
    generations -= 1
    genomes = 2.0 * float(N.sum()) * generations
    rec_gbps = rec_bytes / (phases["recombine"] * 1e-3) / 1e9 if phases and phases["recombine"] else None
    return {
        "workload": f"256 replicates, N log-uniform [1e3, 5e4], rho U[0.5, 5], "
                    f"mu U[0.5, 5] x {MU_SCALE}, {CAPACITY} slots each, 1 trait, "
                    f"generations 2..{generations + 1} from empty founders timed",
        "batched_ms_per_generation": ms["graph"] / generations,
        "sequential_ms_per_generation": ms["sequential"] / generations,
        "batched_genomes_per_s": genomes / (ms["graph"] * 1e-3),
        "sequential_genomes_per_s": genomes / (ms["sequential"] * 1e-3),
        "speedup": ms["sequential"] / ms["graph"],
        "shapes_equal": shapes["graph"] == shapes["sequential"],
        "batched_grid_percent_of_default": grid_pct,
        "batched_phase_ms_after_run": phases,
        "batched_recombine_alg_bytes": rec_bytes,
        "batched_recombine_alg_gbps": rec_gbps,
        "batched_recombine_frac_of_8000": rec_gbps / 8000.0 if rec_gbps else None,
    }


def large_population_footprint(api):
    """Device bytes of the N = 5e5, 16-trait population: two probe contexts
    give bytes per allele slot, the bench capacity is allocated and read
    back, and the free HBM gives the largest capacity that fits."""
    from clotho_amd.bench_large import config as large_config

    probes = []
    for slots in (1 << 16, 1 << 17):
        p = api.Simulation(large_config(api, slots))
        probes.append((slots, p.device_bytes()))
        p.close()
    (sa, ba), (sb, bb) = probes
    per_slot = (bb - ba) / (sb - sa)
    fixed = ba - per_slot * sa
    slots = 262_144 + 8 * 4096
    p = api.Simulation(large_config(api, slots))
    dev = p.device_bytes()
    p.close()
    free, total = api.device_memory(0)
    return {"haplotypes": 1_000_000, "traits": 16, "slots": slots, "device_bytes": dev,
            "device_gib": dev / 2**30, "bytes_per_slot": per_slot, "fixed_bytes": fixed,
            "hbm_total_gib": total / 2**30,
            "max_slots_in_free_hbm": int((free * 0.97 - fixed) / per_slot) // 512 * 512}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity leg")
    ap.add_argument("--eager", action="store_true", help="time eager launches instead of a graph")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the precision, replicate and footprint legs")
    ap.add_argument("--batch-generations", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus above 1 is not supported "
              "(the multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from clotho_amd import api

    cfg = workload_config(api)
    h = 2 * N_IND
    copy_gbps = api.stream_copy_gbps(1 << 30, 20)
    log(f"[bench] measured device copy bandwidth {copy_gbps:.0f} GB/s")

    sim = api.Simulation(cfg)
    freqs, positions, effects = standing_variation()
    t0 = time.perf_counter()
    sim.set_synthetic(freqs, positions, effects, seed=42)
    s0 = sim.shape()
    log(f"[bench] synthetic population H={h} W={s0.words} alleles={s0.alleles} "
        f"free={s0.free_count} ({time.perf_counter() - t0:.2f}s)")

    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle

        rows = sim.rows()
        pos, eff, _ = sim.alleles()
        orc = oracle.OracleSim(cfg)
        orc.set_population(rows, pos, eff, None)
        # the reference's thread model: T contiguous offspring batches, one
        # per thread (examples/QTLSim/engine_parallel_pipeline.hpp:203-226),
        # on the CPU share of this process; then the same generation on one core
        model, nproc, share = host_cpu()
        threads = max(1, min(share, 16))
        orc_mt = oracle.OracleSim(cfg)
        orc_mt.set_population(rows, pos, eff, None)
        del rows
        orc_mt.set_threads(threads)
        log(f"[bench] oracle loaded; timing one generation on {threads} threads, then one core")
        t0 = time.perf_counter()
        orc_mt.step()
        t_mt = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.step()
        t_cpu = time.perf_counter() - t0
        mt_equal = bool(np.array_equal(orc_mt.rows(), orc.rows()))
        orc_mt.close()
        sim.run(1)
        g_rows = sim.rows()
        parity = bool(np.array_equal(g_rows, orc.rows()))
        del g_rows
        fx, ls = sim.free_space()
        ofx, ols = orc.free_space()
        parity = parity and np.array_equal(fx, ofx) and np.array_equal(ls, ols)
        parity = parity and np.array_equal(sim.traits(), orc.traits())
        parity = parity and np.array_equal(sim.fitness(), orc.fitness())
        top, bot, _, _ = sim.gametes()
        otop, obot, _, _ = orc.gametes()
        parity = parity and np.array_equal(top, otop) and np.array_equal(bot, obot)
        cpu = {"value": h / t_mt, "unit": "offspring genomes/s", "cores": threads,
               "kind": "port",
               "sample": f"one full generation (H={h}, W={s0.words}) of the oracle in {threads} "
                         f"contiguous offspring batches on {threads} threads, {t_mt:.2f} s; the "
                         f"same generation on one thread {t_cpu:.1f} s (phases crossover/mutate/"
                         f"phenotype/fixed/fitness/finalize/draws = "
                         + "/".join(f"{x:.2f}" for x in orc.phase_seconds()) + " s)",
               "seconds": t_mt, "single_thread_value": h / t_cpu,
               "single_thread_seconds": t_cpu, "threads_equal_single": mt_equal,
               "cpu_model": model, "nproc": nproc, "cpu_share": share}
        orc.close()
        log(f"[bench] cpu generation {t_mt:.2f}s on {threads} threads, {t_cpu:.2f}s on one; "
            f"parity={parity}")

    graph = not args.eager
    sim.run(args.warmup, graph=graph)
    chunks = sim.autotune()
    shape = sim.launch_shape()
    log(f"[bench] autotuned launch shapes: {shape}")
    sim.run(1, graph=graph)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(args.steps, graph=graph)
    sim.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    s1 = sim.shape()
    log(f"[bench] {args.steps} generations in {elapsed * 1e3:.2f} ms; W={s1.words} "
        f"alleles={s1.alleles} M={s1.new_mutations} free={s1.free_count}")

    # per-phase kernel time on the library stream (eager pass with HIP events)
    sim.set_timing(True)
    sim.run(args.steps, graph=False)
    ngen, ms = sim.timings()
    sim.set_timing(False)
    rec_ms = ms[2] / max(ngen, 1)
    words = s1.words
    alg_bytes = 24.0 * words * h
    achieved = alg_bytes / (rec_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles()
    phases = {k: round(float(ms[i]) / max(ngen, 1), 4) for i, k in
              enumerate(["generation", "draw", "recombine", "traits", "free_space", "fitness"])}
    log(f"[bench] per-generation ms (eager): {phases}")

    # per-period population statistics (column and row margins, all rows)
    _, _, stats_ms = sim.allele_counts(timed=True)
    _, _, stats_ms = sim.allele_counts(timed=True)
    log(f"[bench] population statistics (all rows): {stats_ms:.3f} ms")

    sim.close()
    extras = {}
    if not args.no_extras:
        t0 = time.perf_counter()
        extras["mfma_vs_plain"] = mfma_vs_plain(api)
        log(f"[bench] precision workload: {extras['mfma_vs_plain']} "
            f"({time.perf_counter() - t0:.1f}s)")
        t0 = time.perf_counter()
        extras["batched_vs_sequential"] = batched_vs_sequential(api, args.batch_generations)
        log(f"[bench] replicate workload: {extras['batched_vs_sequential']} "
            f"({time.perf_counter() - t0:.1f}s)")
        t0 = time.perf_counter()
        extras["large_population"] = large_population_footprint(api)
        log(f"[bench] large population footprint: {extras['large_population']} "
            f"({time.perf_counter() - t0:.1f}s)")

    ms_per_step = elapsed / args.steps * 1e3
    value = h * args.steps / elapsed
    out = {
        "metric": "offspring genomes/s/GPU",
        "value": value,
        "unit": "offspring genomes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "single population N=1e5 (H=2e5), W=784 (50,176 slots), "
                               "neutral 1/f standing variation, rho=5, mu=0.01/gamete, "
                               "1 trait N(0,0.1^2), env N(0,1), fitness exp(-z^2/2)",
                   "haplotypes": h, "words": words, "rho": RHO, "mu": MU,
                   "graph": graph, "row_chunks": chunks, "launch_shape": shape},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBPS, "traffic": traffic,
                     "kernel": "k_recombine", "kernel_ms": rec_ms,
                     "alg_bytes_per_launch": alg_bytes,
                     # bytes the kernel actually moves (hardware counters) per
                     # second: siblings share one load of their parents' rows
                     "traffic_gbps": traffic / (rec_ms * 1e-3) / 1e9 if traffic else None,
                     "traffic_frac": (traffic / (rec_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS
                                      if traffic else None)},
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
        "hbm_copy_gbps_measured": copy_gbps,
        "phase_ms": phases,
        "diploid_per_s": value / 2,
        "population_stats": {"ms": stats_ms,
                             "gbps": 8.0 * words * h / (stats_ms * 1e-3) / 1e9 if stats_ms else None},
        **extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
