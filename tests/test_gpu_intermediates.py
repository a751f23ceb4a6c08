"""Intermediate generations (generations.intermediates of the reference's
batched pipeline engine, examples/QTLSim/engine_parallel_pipeline_batched.hpp):
k generations between free-space analyses, bit-exact against the oracle
through the C ABI, eager, graph-replayed and batched."""
import numpy as np
import pytest

from clotho_amd import api
from tests.parity_util import run_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [1, 3])
def test_intermediates_match_oracle(k):
    cfg = api.make_config(200, 3.0, 2.5, seed=70 + k, traits=2, neutral_p=0.3, sigma=0.3,
                          fitness=api.FIT_MV_STABILIZING, fitness_scale=2.0, env_sigma=0.5,
                          max_alleles=1 << 14, intermediates=k)
    sim, orc = run_pair(cfg, 4 * (k + 1) + 1)
    assert sim.shape().generation == orc.shape().generation


def test_intermediates_graph_and_upload():
    from tests.test_gpu_parity import _uploaded_population

    cfg = api.make_config(150, 2.0, 3.0, seed=77, traits=1, neutral_p=0.4, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 13, intermediates=2)
    run_pair(cfg, 9, check_every=1, graph=True, population=_uploaded_population(150, 900, 1, 3))


def test_intermediates_in_a_batch():
    cfgs = [api.make_config(n, 2.0, 2.0, seed=80 + r, replicate=r, traits=1, neutral_p=0.5,
                            sigma=0.2, fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 13,
                            intermediates=r % 3)
            for r, n in enumerate([120, 300, 75, 200])]
    batch = api.Batch(cfgs)
    batch.run(7, mode="graph")
    for r, cfg in enumerate(cfgs):
        lone = api.Simulation(cfg)
        lone.run(7)
        rep = batch.replicate(r)
        assert np.array_equal(rep.rows(), lone.rows()), r
        assert np.array_equal(rep.fitness(), lone.fitness()), r
        assert rep.shape().as_dict() == lone.shape().as_dict(), r


def test_intermediates_with_compaction_is_an_error():
    cfg = api.make_config(100, 1.0, 1.0, seed=1, max_alleles=1 << 12, intermediates=2,
                          compact=True)
    with pytest.raises(api.ClothoError):
        api.Simulation(cfg)
