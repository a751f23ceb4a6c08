"""Large-population throughput: N = 5e5 diploid (H = 1e6 haplotype rows),
16 traits, multivariate stabilising fitness exp(-|z|^2/2), Poisson(10)
crossovers per gamete, fixed and lost columns compacted away after every
generation, and the HBM footprint of the population.

Standing variation: `--slots` columns (default 262,144, W = 4,096 words per
row) with frequencies from the neutral 1/f spectrum on [1/H, 1 - 1/H]
(seed 42, drawn on the device), positions uniform, effects multivariate
normal with unit variances and pairwise correlation 0.3 (seed 43), scaled by
EFFECT_SD so that |z| is of order one.  New mutations: 0.01 per gamete with
independent N(0, EFFECT_SD^2) effects per trait (the reference's
TraitSpaceGenerator2 draws traits independently); the literal 20 per gamete
would add 2e7 columns per generation, far beyond any dense bitset.

Reports offspring genomes per second over timed hipGraph generations, the
per-phase kernel time (the last phase includes fitness, commit and the
compaction), the algorithmic bandwidth of the recombination phase, the
context's device bytes, and the largest slot capacity that fits the free
HBM (footprint model from two probe contexts, then verified by allocating
it).  One JSON line on stdout.
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

N_IND = 500_000
TRAITS = 16
RHO = 10.0
MU = 0.01
EFFECT_SD = 0.002


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def config(api, slots_cap, compact=True):
    return api.make_config(N_IND, MU, RHO, seed=42, traits=TRAITS,
                           fitness=api.FIT_MV_STABILIZING, neutral_p=0.0, mean=0.0,
                           sigma=EFFECT_SD, fitness_scale=1.0, env_sigma=0.0, quantum_log2=-24,
                           max_alleles=slots_cap, compact=compact)


def standing_variation(slots, seed=42):
    h = 2 * N_IND
    rng = np.random.default_rng(seed)
    a, b = 1.0 / h, 1.0 - 1.0 / h
    freqs = a * (b / a) ** rng.uniform(0, 1, slots)
    positions = np.floor(rng.uniform(0, 1, slots) * 2**32) / 2**32
    cov = np.full((TRAITS, TRAITS), 0.3) + 0.7 * np.eye(TRAITS)
    z = np.random.default_rng(43).standard_normal((slots, TRAITS)) @ np.linalg.cholesky(cov).T
    effects = np.round(z * EFFECT_SD * 2**24) / 2**24
    return freqs, positions, effects


def compaction_spans(api, slots, generations):
    """Where a compaction after each generation removes columns, and the
    input words each output tile of the compaction's row pass draws on: a
    tile of tw words whose kept bits come from at most 64 input words takes
    the one-plan path, and 64 + m words with m * rows_per_step <= 64 the
    same path with the m extra words packed for a step's rows in one pass
    (counted per rows per step); otherwise up to 128 words take the two-plan
    path and more the per-row path.  The generations run without
    compact_every and are compacted explicitly after their free-space lists
    are read, which is the same sequence of states."""
    sim = api.Simulation(config(api, slots + 8 * 4096, compact=False))
    freqs, positions, effects = standing_variation(slots)
    sim.set_synthetic(freqs, positions, effects, seed=42)
    for g in range(generations):
        sim.run(1, graph=False)
        s = sim.shape()
        fixed, lost = sim.free_space()
        keep = np.ones(s.words * 64, bool)
        keep[s.alleles:] = False
        keep[np.asarray(fixed, np.int64)] = False
        keep[np.asarray(lost, np.int64)] = False
        cnt = keep.reshape(s.words, 64).sum(1)
        dst = np.concatenate([[0], np.cumsum(cnt)[:-1]])
        kept = int(cnt.sum())
        w2 = (kept + 511) // 512 * 8
        rec = {"generation": g + 1, "alleles": s.alleles, "words": s.words,
               "removed": s.alleles - kept, "removed_by_eighth":
               [round(float(x.mean()), 4) for x in np.array_split(~keep[:s.alleles], 8)]}
        for tw in (64, 60, 56, 48):
            paths = [0, 0, 0]
            spill = {4: 0, 8: 0, 16: 0}
            widest = 0
            for t in range((w2 + tw - 1) // tw):
                wa = int(np.searchsorted(dst, t * tw * 64, side="right")) - 1
                wb = int(np.searchsorted(dst, (t + 1) * tw * 64, side="left"))
                widest = max(widest, wb - wa)
                paths[0 if wb - wa <= 64 else 1 if wb - wa <= 128 else 2] += 1
                for rows in spill:
                    spill[rows] += int(0 < wb - wa - 64 and (wb - wa - 64) * rows <= 64)
            rec[f"tiles_{tw}"] = {"one_plan": paths[0], "over_64": paths[1],
                                  "one_pass_extra_by_rows_per_step": spill,
                                  "per_row": paths[2], "widest_span": widest}
        print(json.dumps(rec), flush=True)
        sim.compact()
    sim.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--slots", type=int, default=262_144)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-max", action="store_true", help="skip the largest-capacity allocation")
    ap.add_argument("--compact-tile-words", type=int, default=0,
                    help="compaction tile width after autotuning (48, 56, 60 or 64; 0 keeps the tuned one)")
    ap.add_argument("--traits-variant", type=int, default=0,
                    help="traits kernel shape after autotuning (1..traits_variants; 0 keeps the tuned one)")
    ap.add_argument("--compact-rows", type=int, default=0,
                    help="compaction rows per wave step after autotuning (4, 8 or 16; 0 keeps the tuned one)")
    ap.add_argument("--spans", type=int, default=0,
                    help="instead of the benchmark: G generations compacted one by one, printing "
                         "where columns are removed and how many input words each compaction "
                         "tile draws on")
    args = ap.parse_args()

    from clotho_amd import api

    if args.spans:
        compaction_spans(api, args.slots, args.spans)
        return
    free0, total = api.device_memory(0)
    h = 2 * N_IND
    cap = args.slots + 8 * 4096
    sim = api.Simulation(config(api, cap))
    dev_bytes = sim.device_bytes()
    freqs, positions, effects = standing_variation(args.slots)
    t0 = time.perf_counter()
    sim.set_synthetic(freqs, positions, effects, seed=42)
    s0 = sim.shape()
    log(f"[large] H={h} W={s0.words} slots={s0.alleles} free={s0.free_count} "
        f"device bytes {dev_bytes / 2**30:.1f} GiB ({time.perf_counter() - t0:.1f}s)")
    sim.run(args.warmup, graph=True)
    chunks = sim.autotune()
    if args.compact_tile_words:
        sim.set_launch_shape(compact_tile_words=args.compact_tile_words)
    if args.compact_rows:
        sim.set_launch_shape(compact_rows=args.compact_rows)
    if args.traits_variant:
        sim.set_launch_shape(traits_variant=args.traits_variant)
    shape = sim.launch_shape()
    log(f"[large] launch shape {shape}")
    sim.run(1, graph=True)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(args.steps, graph=True)
    sim.sync()
    elapsed = time.perf_counter() - t0
    s1 = sim.shape()
    log(f"[large] {args.steps} generations in {elapsed * 1e3:.1f} ms; W={s1.words} "
        f"alleles={s1.alleles} M={s1.new_mutations} chunks={chunks}")
    sim.set_timing(True)
    sim.run(3, graph=False)
    ngen, ms = sim.timings()
    sim.set_timing(False)
    phases = {k: round(float(ms[i]) / max(ngen, 1), 3) for i, k in
              enumerate(["generation", "draw", "recombine", "traits", "free_space",
                         "fitness_commit_compact"])}
    log(f"[large] per-generation ms (eager): {phases}")
    words = sim.shape().words
    rec_gbps = 24.0 * words * h / (phases["recombine"] * 1e-3) / 1e9
    sim.close()
    del sim

    # footprint model and the largest capacity that fits the free HBM
    probes = []
    for s_cap in (1 << 16, 1 << 17):
        p = api.Simulation(config(api, s_cap))
        probes.append((s_cap, p.device_bytes()))
        p.close()
    (s_a, b_a), (s_b, b_b) = probes
    per_slot = (b_b - b_a) / (s_b - s_a)
    fixed = b_a - per_slot * s_a
    free1, _ = api.device_memory(0)
    s_max = int((free1 * 0.97 - fixed) / per_slot) // 512 * 512
    max_ok = None
    max_bytes = None
    if not args.no_max:
        try:
            big = api.Simulation(config(api, s_max))
            max_bytes = big.device_bytes()
            big.close()
            max_ok = True
        except Exception as e:  # allocation failure is a result, not an error
            log(f"[large] allocation at {s_max} slots failed: {e}")
            max_ok = False
    value = h * args.steps / elapsed
    out = {
        "metric": "offspring genomes/s/GPU (large population)",
        "value": value,
        "unit": "offspring genomes/s",
        "steps": args.steps,
        "ms_per_step": elapsed / args.steps * 1e3,
        "config": {"N": N_IND, "haplotypes": h, "traits": TRAITS, "rho": RHO, "mu": MU,
                   "standing_slots": args.slots, "words_final": words,
                   "fitness": "exp(-|z|^2/2), 16 traits", "compaction": "every generation",
                   "effects": f"MVN(rho=0.3) x {EFFECT_SD}", "row_chunks": chunks,
                   "launch_shape": shape},
        "phase_ms": phases,
        "recombine_alg_gbps": rec_gbps,
        "hbm": {"device_bytes": dev_bytes, "device_gib": dev_bytes / 2**30,
                "bytes_per_slot": per_slot, "fixed_bytes": fixed,
                "total_gib": total / 2**30, "free_gib_at_start": free0 / 2**30,
                "max_slots": s_max, "max_words": s_max // 64, "max_alloc_ok": max_ok,
                "max_alloc_gib": (max_bytes / 2**30) if max_bytes else None},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
