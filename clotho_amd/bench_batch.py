"""Replicate-batch throughput: 256 independent populations of uneven size and
parameters advanced together, against the same populations advanced one
replicate after another.

Workload: R = 256 replicates from empty founders; N_r log-uniform on
[1e3, 5e4]; rho_r ~ U[0.5, 5] crossovers per gamete; one additive trait with
N(0, 0.1^2) effects, environmental noise N(0, 1), fitness exp(-z^2/2);
mu_r ~ U[0.5, 5] x 1e-3 new mutations per gamete, capacity 16,384 slots per
replicate (the unscaled rates put ~1e6 new columns per generation into the
largest replicates, far beyond any dense bitset; at 1e-3 the largest
replicate holds ~7e3 segregating sites after 500 generations).  Draws of
N_r, rho_r and mu_r use numpy's default_rng(42).

Two batches with identical configurations run the same number of generations
from the same founders: one in batched mode (each kernel of the generation
launched once for all replicates as a persistent grid, the generation
captured as one hipGraph), one in sequential mode (each replicate's own
captured generation graph, replicates one after another on one stream).  The
replicates of the two runs are then compared bit for bit (shapes of every
replicate; rows, free lists, traits and fitness of a sample).  One JSON line
on stdout: offspring genomes per second of both modes and their ratio.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

MU_SCALE = 1e-3
CAPACITY = 16384


def replicate_configs(api, R=256, seed=42):
    rng = np.random.default_rng(seed)
    N = np.exp(rng.uniform(np.log(1e3), np.log(5e4), R)).astype(np.int64)
    rho = rng.uniform(0.5, 5.0, R)
    mu = rng.uniform(0.5, 5.0, R) * MU_SCALE
    cfgs = [api.make_config(int(N[r]), float(mu[r]), float(rho[r]), seed=42, replicate=r,
                            traits=1, fitness=api.FIT_GAUSSIAN_FIXED, neutral_p=0.0, mean=0.0,
                            sigma=0.1, fitness_scale=1.0, env_sigma=1.0, quantum_log2=-24,
                            max_alleles=CAPACITY) for r in range(R)]
    return cfgs, N, rho, mu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--replicates", type=int, default=256)
    ap.add_argument("--generations", type=int, default=500)
    ap.add_argument("--sample", type=int, default=12, help="replicates compared in full")
    ap.add_argument("--batched-only", action="store_true",
                    help="run the batched mode alone (no comparison; for profiling)")
    args = ap.parse_args()

    from clotho_amd import api

    cfgs, N, rho, mu = replicate_configs(api, args.replicates)
    H = 2 * N
    genomes_per_gen = float(H.sum())

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    results = {}
    batches = {}
    for mode in (("graph",) if args.batched_only else ("graph", "sequential")):
        t0 = time.perf_counter()
        b = api.Batch(cfgs)
        log(f"[batch] {mode}: {len(cfgs)} replicates allocated ({time.perf_counter() - t0:.1f}s)")
        ms = 0.0
        done = 0
        while done < args.generations:
            step = min(100, args.generations - done)
            ms += b.run(step, mode=mode)
            done += step
            log(f"[batch] {mode}: {done} generations, {ms:.1f} ms")
        results[mode] = ms
        batches[mode] = b

    if args.batched_only:
        b = batches["graph"]
        b.set_timing(True)
        b.run(5, mode="eager")
        n, ph = b.timings()
        print(json.dumps({"batched_graph_ms": results["graph"], "generations": args.generations,
                          "phase_ms": [round(float(x) / max(n, 1), 4) for x in ph]}), flush=True)
        b.close()
        return
    # bit-exact agreement of the two runs
    bg, bs = batches["graph"], batches["sequential"]
    shapes_equal = True
    words = []
    for r in range(len(cfgs)):
        a, c = bg.replicate(r).shape(), bs.replicate(r).shape()
        words.append(a.words)
        for key in ("generation", "words", "alleles", "fixed_count", "lost_count", "free_count"):
            if getattr(a, key) != getattr(c, key):
                shapes_equal = False
    order = np.argsort(H)
    sample = sorted(set(order[: args.sample // 2].tolist() + order[-(args.sample - args.sample // 2):].tolist()))
    full_equal = True
    for r in sample:
        x, y = bg.replicate(r), bs.replicate(r)
        full_equal = full_equal and np.array_equal(x.rows(), y.rows())
        fx, lx = x.free_space()
        fy, ly = y.free_space()
        full_equal = full_equal and np.array_equal(fx, fy) and np.array_equal(lx, ly)
        full_equal = full_equal and np.array_equal(x.traits(), y.traits())
        full_equal = full_equal and np.array_equal(x.fitness(), y.fitness())

    # per-phase time of the batched kernels (eager launches, a few generations)
    bg.set_timing(True)
    ngen = 5
    bg.run(ngen, mode="eager")
    n, ph = bg.timings()
    phases = {k: round(float(ph[i]) / max(n, 1), 4) for i, k in
              enumerate(["generation", "draw", "recombine", "traits", "free_space", "fitness"])}
    words = np.asarray(words)
    bytes_per_gen = float((24.0 * words * H).sum())
    for b in batches.values():
        b.close()

    g_rate = genomes_per_gen * args.generations / (results["graph"] * 1e-3)
    s_rate = genomes_per_gen * args.generations / (results["sequential"] * 1e-3)
    out = {
        "metric": "offspring genomes/s/GPU (replicate batch)",
        "replicates": len(cfgs),
        "generations": args.generations,
        "haplotypes_total": int(H.sum()),
        "batched_graph": {"value": g_rate, "ms": results["graph"],
                          "ms_per_generation": results["graph"] / args.generations},
        "sequential_graphs": {"value": s_rate, "ms": results["sequential"],
                              "ms_per_generation": results["sequential"] / args.generations},
        "speedup": g_rate / s_rate,
        "parity": {"shapes_equal_all": shapes_equal, "full_state_equal_sample": full_equal,
                   "sample": [int(x) for x in sample]},
        "final_words": {"min": int(words.min()), "median": float(np.median(words)),
                        "max": int(words.max())},
        "batched_phase_ms_final": phases,
        "recombine_alg_bytes_final_gen": bytes_per_gen,
        "config": {"N": "log-uniform [1e3, 5e4]", "rho": "U[0.5, 5]",
                   "mu": f"U[0.5, 5] x {MU_SCALE}", "capacity_slots": CAPACITY, "traits": 1,
                   "fitness": "exp(-z^2/2), env N(0,1), effects N(0, 0.1^2)", "seed": 42},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
