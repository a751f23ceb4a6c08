"""The C++ adapters (Engine<..., mi355x_pipeline>, common_crossover,
free_space_accumulator_mt, GeneralFitness, fitness_toolkit) driven like the
reference's qtlsim_main: same results as the oracle for the same keys."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from clotho_amd import api, build

pytestmark = pytest.mark.gpu


def fnv1a(rows):
    h = 1469598103934665603
    for b in rows.astype("<u8").tobytes():
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


@pytest.mark.parametrize("graph", [False, True])
def test_engine_adapter_matches_oracle(graph):
    keys = ["population_growth.params.B=200", "mutation.mu=2", "recombination.rho=3",
            "neutral.p=0.3", "traits.sigma=0.2", "fitness.name=gaussian_fixed",
            "fitness.params.scale=1.5", "random_number.seed=9"]
    args = [build.DEMO, *keys, "--generations", "12"] + (["--graph"] if graph else [])
    out = json.loads(subprocess.run(args, check=True, capture_output=True, text=True).stdout)
    cfg = api.make_config(200, 2.0, 3.0, seed=9, traits=1, neutral_p=0.3, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, fitness_scale=1.5, max_alleles=32768)
    orc = oracle.OracleSim(cfg)
    for _ in range(12):
        orc.step()
    s = orc.shape()
    assert out["generation"] == s.generation and out["alleles"] == s.alleles
    assert out["free"] == s.free_count and out["fixed"] == s.fixed_count
    assert out["rows_fnv1a"] == fnv1a(orc.rows())
