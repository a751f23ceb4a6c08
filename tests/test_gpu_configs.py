"""The benchmark configurations at their own workloads, on the GPU.

* Config 1 at its stated rates (Poisson(10) mutations and Poisson(5)
  crossovers per gamete), at a reduced population of N = 4,000: ~8e4 new
  mutations per generation, far more than the free slots, so most of them
  append columns (the mutation pre-draw of examples/QTLSim/engine_parallel_pipeline.hpp:327-384
  at M >> F).  Bit-exact against the oracle for 25 generations.
* Config 3: the 256 replicates of the batch benchmark (N log-uniform on
  [1e3, 5e4]), batched graphs against sequential per-replicate graphs for
  every replicate's shape, and eight replicates against the oracle in full.
* Config 4: N = 5e5 (H = 1e6), 16 traits with correlated effects,
  multivariate stabilising fitness and compaction after every generation.
  Compacted and uncompacted runs give identical traits, fitness and gametes;
  a 2,000-row sample carries the same alleles in both, and its traits equal
  the exact sums recomputed with numpy from the downloaded rows and effects.
"""
import numpy as np
import pytest

import oracle
from clotho_amd import api
from clotho_amd.bench_batch import replicate_configs
from clotho_amd.bench_large import standing_variation
from tests.parity_util import compare_gametes, compare_state

pytestmark = pytest.mark.gpu


def test_config1_stated_rates_reduced_population():
    cfg = api.make_config(4000, 10.0, 5.0, seed=42, traits=1, neutral_p=0.0, sigma=0.1,
                          fitness=api.FIT_GAUSSIAN_FIXED, fitness_scale=1.0, env_sigma=1.0,
                          max_alleles=1 << 19)
    sim = api.Simulation(cfg)
    orc = oracle.OracleSim(cfg)
    appended = 0
    for g in range(25):
        free_before = orc.shape().free_count
        sim.run(1, graph=True)
        orc.step()
        s = orc.shape()
        appended += max(0, s.new_mutations - free_before)
        if (g + 1) % 5 == 0:
            print(f"generation {g + 1}: alleles {s.alleles} W {s.words} M {s.new_mutations}",
                  flush=True)
            compare_gametes(sim, orc, f"generation {g + 1}")
            compare_state(sim, orc, f"generation {g + 1}")
    s = sim.shape()
    # ~8e4 mutations per generation, of which ~4e5 appended new columns in all
    assert s.new_mutations > 70_000 and appended > 300_000
    sim.close()


def test_config3_replicate_batch_workload():
    cfgs, N, rho, mu = replicate_configs(api, 256)
    gens = 5
    batched = api.Batch(cfgs)
    batched.run(gens, mode="graph")
    print(f"batched graphs: {gens} generations of {len(cfgs)} replicates", flush=True)
    seq = api.Batch(cfgs)
    seq.run(gens, mode="sequential")
    print(f"sequential graphs: {gens} generations", flush=True)
    for r in range(len(cfgs)):
        a, b = batched.replicate(r).shape(), seq.replicate(r).shape()
        for key in ("generation", "n_individuals", "words", "alleles", "new_mutations",
                    "fixed_count", "lost_count", "free_count", "error"):
            assert getattr(a, key) == getattr(b, key), f"replicate {r}: {key}"
    seq.close()
    order = np.argsort(N)
    sample = sorted(set(order[:3].tolist() + order[-3:].tolist() + order[[64, 192]].tolist()))
    assert len(sample) == 8
    for r in sample:
        print(f"replicate {r}: N={N[r]} rho={rho[r]:.3f} mu={mu[r]:.5f}", flush=True)
        orc = oracle.OracleSim(cfgs[r])
        for _ in range(gens):
            orc.step()
        compare_gametes(batched.replicate(r), orc, f"replicate {r}")
        compare_state(batched.replicate(r), orc, f"replicate {r}")
        orc.close()
    batched.close()


def _carried(rows, positions, ages, exclude=()):
    """Per row: sorted (position key, age) pairs of the alleles it carries,
    leaving out the slots in `exclude`."""
    bits = np.unpackbits(rows.view(np.uint8), axis=1, bitorder="little").astype(bool)
    bits = bits[:, : positions.size]
    bits[:, np.asarray(exclude, np.int64)] = False
    keys = (positions * 2.0**32).astype(np.uint64) << np.uint64(32)
    tags = keys | ages.astype(np.uint64)
    out = []
    for r in range(rows.shape[0]):
        cols = np.flatnonzero(bits[r])
        out.append(np.sort(tags[cols]))
    return out


def test_config4_large_population_compaction():
    slots, gens, T = 32_768, 2, 16
    freqs, positions, effects = standing_variation(slots)

    def context(compact):
        cfg = api.make_config(500_000, 0.01, 10.0, seed=42, traits=T,
                              fitness=api.FIT_MV_STABILIZING, neutral_p=0.0, sigma=0.002,
                              fitness_scale=1.0, quantum_log2=-24, max_alleles=2 * slots,
                              compact=compact)
        sim = api.Simulation(cfg)
        sim.set_synthetic(freqs, positions, effects, seed=42)
        sim.run(gens)
        return sim

    full = context(False)
    print("uncompacted run done", flush=True)
    comp = context(True)
    print("compacted run done", flush=True)
    sf, sc = full.shape(), comp.shape()
    assert sf.haplotypes == sc.haplotypes == 1_000_000
    assert sc.free_count == 0 and sc.alleles == sf.alleles - sf.free_count
    assert sc.words < sf.words
    for name, a, b in zip(("top", "bottom", "events", "mutations"), full.gametes(),
                          comp.gametes()):
        assert np.array_equal(a, b), f"gamete {name} differs"
    tf = full.traits()
    assert np.array_equal(tf, comp.traits()), "traits differ between compacted and full runs"
    assert np.array_equal(full.fitness(), comp.fitness()), "fitness differs"
    assert np.count_nonzero(tf) > 0.9 * tf.size

    rng = np.random.default_rng(5)
    sample = np.sort(rng.choice(sf.haplotypes, 2000, replace=False))
    pf, ef, af = full.alleles()
    fixed_f, _ = full.free_space()
    rows_f = full.rows()[sample]
    pc, ec, ac = comp.alleles()
    rows_c = comp.rows()[sample]
    full.close()
    comp.close()
    # exact sums: effects are integers k 2^-24 and every partial sum stays
    # below 2^53, so float64 products of 0/1 and k are exact
    k = np.rint(ef * 2.0**24)
    bits = np.unpackbits(rows_f.view(np.uint8), axis=1, bitorder="little")[:, : pf.size]
    want = (bits.astype(np.float64) @ k) * 2.0**-24
    assert np.array_equal(tf[sample], want), "sampled traits differ from the exact sums"
    # the compacted rows carry the same (position, age) alleles; the full
    # rows also carry the fixed columns, which the compaction removed
    pairs = zip(_carried(rows_f, pf, af, fixed_f), _carried(rows_c, pc, ac))
    for r, (a, b) in enumerate(pairs):
        assert np.array_equal(a, b), f"sample row {sample[r]}: carried alleles differ"
