"""The generic operator families on the GPU, bit for bit against the oracle:
given mate pairs (Crossover::update), host-directed mutations
(MutationGenerator), fitness of the current population
(Fitness<phenotype_evaluator>::update), per-gamete recombination of given
rows (common_crossover's per-gamete operator), and the reference's base
engine loop built from them (examples/QTLSim/qtlsim_engine.hpp:123-160)."""
import numpy as np
import pytest

import oracle
from clotho_amd import api
from tests.parity_util import compare_gametes, compare_state

pytestmark = pytest.mark.gpu


def _pair(cfg):
    return api.Simulation(cfg), oracle.OracleSim(cfg)


def _mutations(rng, h, n, traits, neutral_p=0.3, sigma=0.2):
    rows = rng.integers(0, h, n).astype(np.uint32)
    pos = np.floor(rng.uniform(0, 1, n) * 2**32) / 2**32
    eff = np.round(rng.normal(0, sigma, (n, traits)) * 2**24) / 2**24
    eff[rng.uniform(0, 1, (n, traits)) < neutral_p] = 0.0
    return rows, pos, eff


def test_given_parents_match_oracle():
    cfg = api.make_config(300, 1.0, 3.0, seed=17, traits=1, neutral_p=0.4, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, env_sigma=0.5, max_alleles=1 << 14)
    sim, orc = _pair(cfg)
    rng = np.random.default_rng(17)
    for g in range(6):
        if g % 2 == 1:  # every other generation takes given pairs
            pairs = rng.integers(0, 300, 600).astype(np.uint32)
            sim.set_parents(pairs)
            orc.set_parents(pairs)
        sim.run(1, graph=(g >= 3))
        orc.step()
        compare_gametes(sim, orc, f"generation {g + 1}")
        compare_state(sim, orc, f"generation {g + 1}")
        if g % 2 == 1:
            assert np.array_equal(sim.gametes()[0] >> 1, pairs)


def test_given_parents_rejects_bad_input():
    cfg = api.make_config(50, 1.0, 1.0, max_alleles=1 << 12)
    sim = api.Simulation(cfg)
    with pytest.raises(api.ClothoError, match="gamete parents"):
        sim.set_parents(np.zeros(99, np.uint32))
    with pytest.raises(api.ClothoError, match="outside"):
        sim.set_parents(np.full(100, 50, np.uint32))


def test_host_mutations_match_oracle():
    """Host-mutation contexts draw no mutations; clotho_add_mutations reuses
    fixed and lost slots in order, clears reused fixed columns, appends and
    widens the rows past the free slots."""
    cfg = api.make_config(200, 2.0, 2.0, seed=23, traits=2, neutral_p=0.3, sigma=0.2,
                          fitness=api.FIT_MV_STABILIZING, fitness_scale=2.0, env_sigma=0.4,
                          max_alleles=1 << 14, host_mutations=True)
    sim, orc = _pair(cfg)
    rng = np.random.default_rng(23)
    sizes = [300, 50, 1200, 0, 700]
    reused = 0
    for g, n in enumerate(sizes):
        sim.run(1)
        orc.step()
        assert sim.shape().new_mutations == 0
        reused += min(n, sim.shape().free_count)
        rows, pos, eff = _mutations(rng, 400, n, 2)
        assert np.array_equal(sim.add_mutations(rows, pos, eff), orc.add_mutations(rows, pos, eff))
        sim.evaluate_fitness()
        orc.evaluate_fitness()
        print(f"generation {g + 1}: {n} mutations, shape {sim.shape().as_dict()}", flush=True)
        compare_state(sim, orc, f"generation {g + 1}")
    # free slots were reused and the rows widened
    assert reused > 0 and sim.shape().words > 8


def test_add_mutations_after_upload_reuses_fixed_columns():
    rng = np.random.default_rng(5)
    n_ind, alleles = 64, 700
    words = ((alleles + 511) // 512) * 8
    bits = rng.uniform(0, 1, (2 * n_ind, words * 64)) < 0.3
    bits[:, alleles:] = False
    bits[:, [2, 100, 699]] = True      # fixed
    bits[:, [0, 1, 64, 500]] = False   # lost
    rows = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64)
    pos = np.floor(rng.uniform(0, 1, alleles) * 2**32) / 2**32
    eff = np.round(rng.normal(0, 0.3, (alleles, 1)) * 2**24) / 2**24
    cfg = api.make_config(n_ind, 1.0, 2.0, seed=5, traits=1, fitness=api.FIT_GAUSSIAN_FIXED,
                          max_alleles=1 << 12, host_mutations=True)
    sim, orc = _pair(cfg)
    sim.set_population(rows, pos, eff)
    orc.set_population(rows, pos, eff)
    m_rows, m_pos, m_eff = _mutations(rng, 2 * n_ind, 5, 1, neutral_p=0.0)
    slots = sim.add_mutations(m_rows, m_pos, m_eff)
    assert list(slots) == [2, 100, 699, 0, 1]
    assert np.array_equal(slots, orc.add_mutations(m_rows, m_pos, m_eff))
    compare_state(sim, orc, "after add_mutations")
    for g in range(3):
        sim.run(1)
        orc.step()
        compare_state(sim, orc, f"generation {g + 1}")


def test_crossover_rows_match_oracle():
    cfg = api.make_config(100, 3.0, 2.0, seed=8, traits=1, max_alleles=1 << 13)
    sim = api.Simulation(cfg)
    sim.run(4)
    s = sim.shape()
    pos, _, _ = sim.alleles()
    rows = sim.rows()
    rng = np.random.default_rng(8)
    G, stride = 64, 12
    top = rows[rng.integers(0, s.haplotypes, G)]
    bot = rows[rng.integers(0, s.haplotypes, G)]
    nev = rng.integers(0, stride + 1, G).astype(np.uint32)
    nev[:4] = 0
    keys = rng.integers(0, 2**32, (G, stride), dtype=np.uint64).astype(np.uint32)
    keys[4, :3] = (pos[[5, 9, 13]] * 2**32).astype(np.uint32)  # events at allele positions: ties
    nev[4] = 3
    out = sim.crossover_rows(top, bot, nev, keys, out_words=s.words + 8)
    assert np.all(out[:, s.words:] == 0)
    for g in range(G):
        ev = keys[g, : nev[g]].astype(np.float64) / 2**32
        want = oracle.crossover(top[g], bot[g], s.words, pos, ev)
        assert np.array_equal(out[g, : s.words], want), f"gamete {g}"
    assert np.array_equal(out[:4, : s.words], top[:4])


def test_base_engine_loop_matches_oracle():
    """qtlsim_engine's simulate (examples/QTLSim/qtlsim_engine.hpp:123-160):
    M ~ Poisson(2 N mu) on the host, mate pairs, crossover, M host mutations
    at uniform rows, traits, phenotype, fitness; both sides fed the same draws."""
    cfg = api.make_config(250, 1.5, 2.5, seed=41, traits=1, neutral_p=0.5, sigma=0.25,
                          fitness=api.FIT_GAUSSIAN, env_sigma=0.0, max_alleles=1 << 14,
                          host_mutations=True)
    sim, orc = _pair(cfg)
    rng = np.random.default_rng(41)
    for g in range(8):
        m = int(rng.poisson(cfg.mu * 2 * 250))
        pairs = rng.integers(0, 250, 500).astype(np.uint32)
        sim.set_parents(pairs)
        orc.set_parents(pairs)
        sim.run(1)
        orc.step()
        rows, pos, eff = _mutations(rng, 500, m, 1, neutral_p=0.5, sigma=0.25)
        sim.add_mutations(rows, pos, eff)
        orc.add_mutations(rows, pos, eff)
        sim.evaluate_fitness()
        orc.evaluate_fitness()
        compare_gametes(sim, orc, f"generation {g + 1}")
        compare_state(sim, orc, f"generation {g + 1}")


def _read_vectors(path, dtypes):
    out = []
    with open(path, "rb") as f:
        for dt in dtypes:
            n = int(np.frombuffer(f.read(8), np.uint64)[0])
            out.append(np.frombuffer(f.read(n * np.dtype(dt).itemsize), dt).copy())
    return out


def test_generic_families_through_cpp_adapters(tmp_path):
    """The C++ adapters of the generic families (clotho_amd/host/generic_demo.cpp)
    in the base engine's order, replayed through the oracle with the draws
    the adapters made: rows, traits, phenotypes, fitness and free space after
    every step; single gametes of common_crossover's per-gamete operator
    (immediate and batched) against the oracle's crossover; their free space
    through free_space_accumulator_mt's thread buffers; and the
    free_space_evaluator layout (per word: fixed slots, then lost slots)."""
    import subprocess

    from clotho_amd import build

    gens = 6
    cmd = [build.GENERIC_DEMO, str(tmp_path), str(gens), "99", "population_growth.size=120",
           "random_number.seed=5", "mutation.mu=1.5", "recombination.rho=3",
           "neutral.p=0.4", "traits.sigma=0.3", "traits.N=2", "fitness.name=mv_stabilizing",
           "fitness.params.scale=2", "alleles.capacity=16384"]
    subprocess.run(cmd, check=True, timeout=120)
    raw = open(tmp_path / "config.bin", "rb").read()
    cfg = api.Config.from_buffer_copy(raw[8:8 + C_SIZE()])
    assert cfg.flags == api.FLAG_HOST_MUTATIONS and cfg.n_individuals == 120
    T = cfg.trait_count
    orc = oracle.OracleSim(cfg)
    for g in range(gens):
        pairs, mrows, mpos, meff, slots, fixed, lost = _read_vectors(
            tmp_path / f"gen{g}.bin",
            [np.uint32, np.uint32, np.float64, np.float64, np.uint32, np.uint32, np.uint32])
        orc.set_parents(pairs)
        orc.step()
        got = orc.add_mutations(mrows, mpos, meff.reshape(-1, T))
        assert np.array_equal(got, slots), f"generation {g}: slots"
        orc.evaluate_fitness()
        ofx, ols = orc.free_space()
        assert np.array_equal(fixed, ofx) and np.array_equal(lost, ols), f"generation {g}"
    rows, traits, fitness, pheno, res, meta = _read_vectors(
        tmp_path / "final.bin",
        [np.uint64, np.float64, np.float64, np.float64, np.uint32, np.uint32])
    fo, lo, fr, M, W, A = (int(x) for x in meta)
    assert np.array_equal(rows.reshape(-1, W), orc.rows())
    assert np.array_equal(traits, orc.traits().reshape(-1))
    assert np.array_equal(fitness, orc.fitness())
    t = orc.traits()
    assert np.array_equal(pheno.reshape(-1, T), t[0::2] + t[1::2])
    # evaluator layout, from the oracle's lists plus the never-used slots [A, M)
    ofx, ols = orc.free_space()
    lost_all = np.concatenate([ols, np.arange(A, M, dtype=np.uint32)])
    want = []
    for w in range(0, (M + 63) // 64):
        want += [int(x) for x in ofx if x // 64 == w] + [int(x) for x in lost_all if x // 64 == w]
    assert fr == len(want) and list(res[:fr]) == want
    assert list(res[M:fo]) == [int(x) for x in ofx]
    assert list(res[2 * M:lo]) == [int(x) for x in lost_all]
    # single gametes
    picked, swaps, nev, keys, offspring, gfx, gls = _read_vectors(
        tmp_path / "gametes.bin",
        [np.uint32, np.uint32, np.uint32, np.uint32, np.uint64, np.uint32, np.uint32])
    G = picked.size // 2
    OW = offspring.size // G
    offspring = offspring.reshape(G, OW)
    rows2 = rows.reshape(-1, W)
    pos, _, _ = orc.alleles()
    k0 = 0
    for g in range(G):
        p0, p1 = rows2[picked[2 * g]], rows2[picked[2 * g + 1]]
        top, bot = (p1, p0) if swaps[g] else (p0, p1)
        ev = keys[k0:k0 + nev[g]].astype(np.float64) / 2**32
        k0 += nev[g]
        want = oracle.crossover(top, bot, OW, pos, ev)
        assert np.array_equal(offspring[g], want), f"gamete {g}"
    ofx2, ols2 = oracle.free_space(offspring, A)
    assert np.array_equal(gfx, ofx2) and np.array_equal(gls, ols2)


def C_SIZE():
    import ctypes

    return ctypes.sizeof(api.Config)
