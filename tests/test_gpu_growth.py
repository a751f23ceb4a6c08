"""Population growth: offspring population sizes that change between
generations (the pN of the reference's population_growth models,
examples/QTLSim/engine_parallel_pipeline.hpp:181-197), bit-exact against the
oracle through the C ABI."""
import numpy as np
import pytest

import oracle
from clotho_amd import api
from tests.parity_util import compare_gametes, compare_state

pytestmark = pytest.mark.gpu


def _run_schedule(cfg, sizes, graph=False, check_parents=True):
    sim = api.Simulation(cfg)
    orc = oracle.OracleSim(cfg)
    n_par = cfg.n_individuals
    for g, n in enumerate(sizes):
        sim.set_population_size(n)
        orc.set_population_size(n)
        sim.run(1, graph=graph)
        orc.step()
        where = f"generation {g + 1} ({n_par} -> {n} individuals)"
        assert sim.shape().n_individuals == n, where
        compare_gametes(sim, orc, where)
        compare_state(sim, orc, where)
        if check_parents:
            top, bottom, _, _ = sim.gametes()
            assert top.size == 2 * n and int(top.max()) < 2 * n_par, where
        n_par = n
    return sim, orc


def test_grow_and_shrink():
    cfg = api.make_config(200, 2.0, 3.0, seed=41, traits=1, neutral_p=0.3, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, env_sigma=0.5, max_alleles=1 << 14,
                          max_individuals=600)
    _run_schedule(cfg, [200, 400, 600, 100, 37, 37, 300, 1, 5])


def test_size_dependent_fitness_constants():
    # gaussian and quadratic scale with sqrt(4 N mu) of the evaluated population
    for kind, seed in ((api.FIT_GAUSSIAN, 43), (api.FIT_QUADRATIC, 44)):
        cfg = api.make_config(150, 3.0, 2.0, seed=seed, traits=1, neutral_p=0.2, sigma=0.4,
                              fitness=kind, fitness_scale=2.0, max_alleles=1 << 14,
                              max_individuals=500)
        _run_schedule(cfg, [300, 450, 90, 90, 210])


def test_graph_runs_after_size_changes():
    cfg = api.make_config(250, 2.0, 4.0, seed=45, traits=2, neutral_p=0.3, sigma=0.3,
                          fitness=api.FIT_MV_STABILIZING, fitness_scale=2.0, max_alleles=1 << 14,
                          max_individuals=400)
    _run_schedule(cfg, [250, 250, 400, 400, 400, 120, 120, 120], graph=True)


def test_growth_with_compaction():
    cfg = api.make_config(180, 2.5, 3.0, seed=47, traits=1, neutral_p=0.4, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 14, compact=True,
                          max_individuals=360)
    _run_schedule(cfg, [360, 90, 270, 270, 45])


def test_linear_growth_schedule():
    # linear_population_growth: pN' = A pN + B, truncated to an integer
    A, B = 1.5, 20.0
    n, sizes = 60, []
    for _ in range(6):
        n = int(A * n + B)
        sizes.append(n)
    cfg = api.make_config(60, 1.5, 2.0, seed=49, traits=1, neutral_p=0.5, sigma=0.3,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 14,
                          max_individuals=max(sizes))
    _run_schedule(cfg, sizes)


def test_size_beyond_capacity_is_an_error():
    cfg = api.make_config(100, 1.0, 1.0, seed=51, max_alleles=1 << 12, max_individuals=150)
    sim = api.Simulation(cfg)
    sim.set_population_size(150)
    with pytest.raises(api.ClothoError, match="capacity"):
        sim.set_population_size(151)
    with pytest.raises(api.ClothoError):
        sim.set_population_size(0)
    sim.run(2)
    assert sim.shape().n_individuals == 150
