"""Bit-exact parity of the HIP pipeline against the CPU oracle (through the C ABI)."""
import numpy as np
import pytest

from clotho_amd import api
from tests.parity_util import compare_state, run_pair

pytestmark = pytest.mark.gpu


def test_config0_first_generation():
    cfg = api.make_config(1000, 1.0, 1.0, seed=42, traits=1, neutral_p=1.0, max_alleles=1 << 15)
    run_pair(cfg, 1)


def test_config0_hundred_generations():
    cfg = api.make_config(1000, 1.0, 1.0, seed=42, traits=1, neutral_p=1.0, max_alleles=1 << 15)
    run_pair(cfg, 100, check_every=10)


def test_selection_gaussian_single_trait():
    cfg = api.make_config(400, 2.0, 3.0, seed=7, traits=1, neutral_p=0.3, sigma=0.1,
                          fitness=api.FIT_GAUSSIAN, max_alleles=1 << 14)
    run_pair(cfg, 30, check_every=3)


def test_stabilizing_fixed_sigma_env_noise():
    cfg = api.make_config(300, 4.0, 5.0, seed=11, traits=1, neutral_p=0.0, sigma=0.1,
                          fitness=api.FIT_GAUSSIAN_FIXED, env_sigma=1.0, max_alleles=1 << 14)
    run_pair(cfg, 20, check_every=4)


def test_multitrait_quadratic_and_mv():
    cfg = api.make_config(200, 3.0, 2.0, seed=5, traits=3, neutral_p=0.2, sigma=0.5,
                          fitness=api.FIT_MV_STABILIZING, fitness_scale=2.0, max_alleles=1 << 14)
    run_pair(cfg, 15, check_every=5)
    cfg = api.make_config(200, 3.0, 2.0, seed=6, traits=2, neutral_p=0.2, sigma=0.5,
                          fitness=api.FIT_QUADRATIC, fitness_scale=3.0, max_alleles=1 << 14)
    run_pair(cfg, 15, check_every=5)


def test_graph_replay_matches_eager():
    cfg = api.make_config(500, 2.0, 2.0, seed=3, traits=1, neutral_p=0.5, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 14)
    run_pair(cfg, 12, check_every=3, graph=True)


def _uploaded_population(n, alleles, traits, seed):
    rng = np.random.default_rng(seed)
    words = ((alleles + 511) // 512) * 8
    freqs = rng.uniform(0, 1, alleles) ** 3
    bits = (rng.uniform(0, 1, (2 * n, words * 64)) < np.pad(freqs, (0, words * 64 - alleles))[None, :])
    bits[:, rng.integers(0, alleles, 3)] = 1  # fixed columns
    rows = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64)
    pos = np.floor(rng.uniform(0, 1, alleles) * 2**32) / 2**32
    eff = np.round(rng.normal(0, 0.2, (alleles, traits)) * 2**24) / 2**24
    fit = rng.uniform(0.1, 1.0, n)
    return rows, pos, eff, fit


def test_uploaded_population_with_fixed_columns():
    cfg = api.make_config(150, 3.0, 4.0, seed=13, traits=2, neutral_p=0.4, sigma=0.2,
                          fitness=api.FIT_MV_STABILIZING, fitness_scale=3.0, max_alleles=1 << 13)
    run_pair(cfg, 6, check_every=1, population=_uploaded_population(150, 900, 2, 1))


def test_many_events_and_strand_bias_edges():
    for bias in (0.0, 1.0):
        cfg = api.make_config(120, 1.5, 25.0, seed=17, traits=1, neutral_p=0.5, sigma=0.3,
                              bias=bias, fitness=api.FIT_QUADRATIC, fitness_scale=2.0,
                              max_alleles=1 << 13)
        run_pair(cfg, 8, check_every=2)


def test_no_mutation_no_recombination():
    rows, pos, eff, fit = _uploaded_population(100, 600, 1, 2)
    for mu, rho in ((0.0, 3.0), (2.0, 0.0)):
        cfg = api.make_config(100, mu, rho, seed=19, traits=1, neutral_p=0.0, sigma=0.1,
                              fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 12)
        run_pair(cfg, 5, check_every=1, population=(rows, pos, eff, fit))


def test_single_individual():
    cfg = api.make_config(1, 2.0, 2.0, seed=23, traits=1, neutral_p=0.5, sigma=0.3,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 12)
    run_pair(cfg, 10, check_every=5)


def test_capacity_overflow_is_an_error():
    cfg = api.make_config(500, 5.0, 1.0, seed=29, traits=1, neutral_p=1.0, max_alleles=2048)
    sim = api.Simulation(cfg)
    with pytest.raises(api.ClothoError, match="capacity"):
        sim.run(3)


def test_synthetic_population_bits():
    import oracle

    cfg = api.make_config(64, 1.0, 1.0, seed=31, traits=1, max_alleles=1 << 12)
    sim = api.Simulation(cfg)
    rng = np.random.default_rng(4)
    freqs = rng.uniform(0, 1, 700)
    pos = np.floor(rng.uniform(0, 1, 700) * 2**32) / 2**32
    sim.set_synthetic(freqs, pos, np.zeros((700, 1)), seed=77)
    rows = sim.rows()
    thr = np.floor(freqs * 2**32)
    for r in (0, 5, 127):
        bits = np.unpackbits(rows[r].view(np.uint8), bitorder="little")
        for j in range(0, 700, 37):
            d = oracle.draw(77, 0, 7, 0, r, j // 4)
            assert bits[j] == (d[j % 4] < thr[j])
        assert not bits[700:].any()


def test_autotune_keeps_parity():
    rows, pos, eff, fit = _uploaded_population(300, 2000, 1, 5)
    cfg = api.make_config(300, 2.0, 5.0, seed=37, traits=1, neutral_p=0.2, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, env_sigma=0.5, max_alleles=1 << 13)
    sim = api.Simulation(cfg)
    orc = __import__("oracle").OracleSim(cfg)
    sim.set_population(rows, pos, eff, fit)
    orc.set_population(rows, pos, eff, fit)
    assert sim.autotune() >= 1  # before any generation: zero gametes
    compare_state(sim, orc, "after autotune")
    for g in range(4):
        sim.run(1, graph=(g % 2 == 1))
        orc.step()
        assert sim.autotune() >= 1
        compare_state(sim, orc, f"autotuned generation {g + 1}")


@pytest.mark.parametrize("traits", [7, 16, 29])
def test_many_traits(traits):
    # 2, 4 and 8 MFMA N tiles (four traits per tile), several K groups per row
    cfg = api.make_config(160, 3.0, 2.0, seed=60 + traits, traits=traits, neutral_p=0.2,
                          sigma=0.3, fitness=api.FIT_MV_STABILIZING, fitness_scale=3.0,
                          max_alleles=1 << 13)
    run_pair(cfg, 5, check_every=1, population=_uploaded_population(160, 2500, traits, traits))


def _dense_population(n, alleles, traits, seed):
    """Rows of density 1/4 (AND of two random words) over `alleles` columns,
    three of them fixed; generated word-wise so that large n stays cheap."""
    rng = np.random.default_rng(seed)
    words = ((alleles + 511) // 512) * 8
    draw = lambda: rng.integers(0, np.iinfo(np.int64).max, size=(2 * n, words), dtype=np.int64).view(np.uint64)
    rows = draw() & draw()
    cols = np.arange(words * 64)
    live = np.packbits((cols < alleles).astype(np.uint8), bitorder="little").view(np.uint64)
    rows &= live[None, :]
    for c in rng.integers(0, alleles, 3):
        rows[:, c // 64] |= np.uint64(1) << np.uint64(c % 64)
    pos = np.floor(rng.uniform(0, 1, alleles) * 2**32) / 2**32
    eff = np.round(rng.normal(0, 0.2, (alleles, traits)) * 2**24) / 2**24
    fit = rng.uniform(0.1, 1.0, n)
    return rows, pos, eff, fit


@pytest.mark.parametrize("n", [20001, 125000])
def test_trait_row_grids(n):
    # up to four traits: 40,002 rows give waves of four and of five 16-row
    # tiles on the even grid; 250,000 rows exceed one round of resident
    # blocks and take the fixed four-tile grid
    cfg = api.make_config(n, 0.01, 1.0, seed=71, traits=2, neutral_p=0.2, sigma=0.3,
                          fitness=api.FIT_GAUSSIAN_FIXED, env_sigma=1.0, max_alleles=1 << 13)
    print(f"trait row grids: n={n}", flush=True)
    run_pair(cfg, 2, population=_dense_population(n, 700, 2, 71))


def test_clustered_positions():
    # every standing position inside two 2^-8 buckets: most keys of a word
    # share an event's bucket, so the lookup's walk past the compared keys runs
    rows, pos, eff, fit = _uploaded_population(200, 1500, 1, 83)
    rng = np.random.default_rng(83)
    pos = np.floor((0.5 + rng.uniform(-2.0**-8, 2.0**-8, pos.size)) * 2**32) / 2**32
    cfg = api.make_config(200, 1.0, 25.0, seed=83, traits=1, neutral_p=0.3, sigma=0.3,
                          fitness=api.FIT_GAUSSIAN_FIXED, env_sigma=1.0, max_alleles=1 << 13)
    run_pair(cfg, 6, population=(rows, pos, eff, fit))


@pytest.mark.parametrize("n", [300000, 530000])
def test_selection_block_totals(n):
    # up to 2048 blocks of 256 individuals the gamete and scatter blocks scan
    # the selection and parent-count block totals themselves (n = 300,000:
    # 1,172 blocks); beyond, one-block scans run first (n = 530,000)
    cfg = api.make_config(n, 0.001, 1.0, seed=89, traits=1, neutral_p=0.2, sigma=0.3,
                          fitness=api.FIT_GAUSSIAN_FIXED, env_sigma=1.0, max_alleles=1 << 13)
    print(f"selection block totals: n={n}", flush=True)
    run_pair(cfg, 2, population=_dense_population(n, 300, 1, 89))
