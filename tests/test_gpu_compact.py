"""Column compaction on the GPU: bit-exact against the oracle's compaction,
and, at the bench shape, the slot-free invariant (compaction leaves traits and
fitness unchanged and keeps exactly the variable columns)."""
import numpy as np
import pytest

import oracle
from clotho_amd import api
from tests.parity_util import compare_gametes, compare_state, run_pair

pytestmark = pytest.mark.gpu


def test_compact_every_generation_matches_oracle():
    cfg = api.make_config(120, 3.0, 2.0, seed=21, traits=2, neutral_p=0.3, sigma=0.3,
                          fitness=api.FIT_MV_STABILIZING, fitness_scale=2.0, env_sigma=0.5,
                          max_alleles=1 << 14, compact=True)
    run_pair(cfg, 12)


def test_compact_graph_replay_multi_tile():
    # ~5k new columns per generation: widths of several 64-word tiles, with
    # packed words straddling tile boundaries
    cfg = api.make_config(300, 8.0, 3.0, seed=4, traits=1, neutral_p=0.5, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 16, compact=True)
    sim, orc = run_pair(cfg, 8, check_every=2, graph=True)
    assert sim.shape().words > 128


def test_explicit_compact_after_upload():
    rng = np.random.default_rng(8)
    n, alleles = 200, 1500
    words = ((alleles + 511) // 512) * 8
    freqs = rng.uniform(0, 1, alleles) ** 4
    bits = rng.uniform(0, 1, (2 * n, words * 64)) < np.pad(freqs, (0, words * 64 - alleles))[None, :]
    bits[:, alleles:] = False
    bits[:, [3, 700, 1499]] = True    # fixed columns
    bits[:, [0, 64, 65, 1000]] = False  # lost columns
    rows = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64)
    pos = np.floor(rng.uniform(0, 1, alleles) * 2**32) / 2**32
    eff = np.round(rng.normal(0, 0.2, (alleles, 1)) * 2**24) / 2**24
    cfg = api.make_config(n, 1.0, 2.0, seed=3, traits=1, neutral_p=0.2, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 13)
    sim, orc = api.Simulation(cfg), oracle.OracleSim(cfg)
    sim.set_population(rows, pos, eff)
    orc.set_population(rows, pos, eff)
    sim.compact()
    orc.compact()
    compare_state(sim, orc, "after compact")
    sim.compact()  # nothing left to remove
    orc.compact()
    compare_state(sim, orc, "after second compact")
    for g in range(3):
        sim.run(1)
        orc.step()
        compare_gametes(sim, orc, f"generation {g + 1}")
        compare_state(sim, orc, f"generation {g + 1}")


def test_compaction_invariant_at_bench_shape():
    import bench

    h = 2 * bench.N_IND
    freqs, positions, effects = bench.standing_variation()
    effects = np.round(np.random.default_rng(9).normal(0, 0.1, effects.shape) * 2**24) / 2**24
    sims = []
    for compact in (False, True):
        cfg = api.make_config(bench.N_IND, bench.MU, bench.RHO, seed=42, traits=1,
                              fitness=api.FIT_GAUSSIAN_FIXED, neutral_p=0.0, sigma=0.1,
                              env_sigma=1.0, max_alleles=bench.ALLELES + 8 * 512, compact=compact)
        s = api.Simulation(cfg)
        s.set_synthetic(freqs, positions, effects, seed=42)
        sims.append(s)
    a, b = sims
    for g in range(3):
        a.run(1, graph=True)
        b.run(1, graph=True)
        sa, sb = a.shape(), b.shape()
        print(f"generation {g + 1}: words {sa.words} -> {sb.words}, kept {sb.alleles}", flush=True)
        assert sb.alleles == sa.variable_count and sb.free_count == 0
        assert np.array_equal(a.traits(), b.traits())
        assert np.array_equal(a.fitness(), b.fitness())
        ga, gb = a.gametes(), b.gametes()
        for x, y in zip(ga, gb):
            assert np.array_equal(x, y)
    # the variable columns of the uncompacted run and all columns of the
    # compacted one are the same alleles with the same carrier counts (slot
    # order differs: the uncompacted run reuses free slots), over a row sample
    ra, rb = a.rows()[:4000], b.rows()[:4000]
    fa, la = a.free_space()
    pa, ea, aa = a.alleles()
    pb, eb, ab = b.alleles()
    free = np.zeros(pa.size, bool)
    free[fa] = True
    free[la] = True
    ca = np.unpackbits(ra.view(np.uint8), axis=1, bitorder="little")[:, : pa.size].sum(0)
    cb = np.unpackbits(rb.view(np.uint8), axis=1, bitorder="little")[:, : pb.size].sum(0)
    ta = sorted(zip(pa[~free], aa[~free], ea[~free, 0], ca[~free]))
    tb = sorted(zip(pb, ab, eb[:, 0], cb))
    assert ta == tb
    assert a.rows().shape[0] == h


@pytest.mark.parametrize("rows_per_step", [4, 8, 16])
@pytest.mark.parametrize("tile_words", [64, 60, 56, 48])
@pytest.mark.parametrize("removed", [0.01, 0.04, 0.5])
def test_compact_row_paths(removed, tile_words, rows_per_step):
    """Four tiles of input words.  With 1% or 4% of the columns removed an
    output tile of 48-60 words draws on at most 64 input words (plan in
    registers, rows in groups) and a tile of 64 words on 1-4 more, packed for
    all of a step's rows at once; with half removed the narrower tiles draw
    on 97-122 input words (two plans) and 64-word tiles on 127-130 (two
    plans, or the plan read per row).  Every tile width and step gives the
    oracle's rows."""
    rng = np.random.default_rng(int(removed * 100) + 11)
    n, alleles = 150, 16000
    words = ((alleles + 511) // 512) * 8
    freqs = rng.uniform(0.05, 0.95, alleles)
    gone = rng.uniform(0, 1, alleles) < removed
    bits = rng.uniform(0, 1, (2 * n, words * 64)) < np.pad(freqs, (0, words * 64 - alleles))[None, :]
    bits[:, alleles:] = False
    lost = np.flatnonzero(gone)
    bits[:, lost[::2]] = False   # lost columns
    bits[:, lost[1::2]] = True   # fixed columns
    rows = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64)
    pos = np.floor(rng.uniform(0, 1, alleles) * 2**32) / 2**32
    eff = np.round(rng.normal(0, 0.2, (alleles, 1)) * 2**24) / 2**24
    cfg = api.make_config(n, 1.0, 2.0, seed=5, traits=1, neutral_p=0.2, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 15)
    sim, orc = api.Simulation(cfg), oracle.OracleSim(cfg)
    sim.set_population(rows, pos, eff)
    orc.set_population(rows, pos, eff)
    sim.set_launch_shape(compact_tile_words=tile_words, compact_rows=rows_per_step)
    assert sim.launch_shape()["compact_tile_words"] == tile_words
    assert sim.launch_shape()["compact_rows"] == rows_per_step
    sim.compact()
    orc.compact()
    compare_state(sim, orc, f"after compact ({removed:.0%} removed, {tile_words}-word tiles)")
    sim.run(2)
    orc.step()
    orc.step()
    compare_state(sim, orc, "two generations later")
