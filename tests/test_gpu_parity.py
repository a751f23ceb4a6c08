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
