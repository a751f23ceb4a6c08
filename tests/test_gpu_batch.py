"""Replicate batches: every replicate of a batched run equals a lone context
with the same configuration, bit for bit, in all three run modes (batched
graph, batched eager, sequential per-replicate graphs), and small replicates
equal the CPU oracle."""
import numpy as np
import pytest

import oracle
from clotho_amd import api
from tests.parity_util import compare_gametes, compare_state

pytestmark = pytest.mark.gpu


def _configs(traits=1, wide=True):
    specs = [  # n, mu, rho, fitness, neutral_p, env
        (300, 2.0, 1.5, api.FIT_GAUSSIAN_FIXED, 0.2, 1.0),
        (1000, 0.7, 4.0, api.FIT_GAUSSIAN, 0.5, 0.0),
        (37, 3.0, 0.5, api.FIT_CONSTANT, 1.0, 0.0),
        (2500, 0.4, 2.5, api.FIT_GAUSSIAN_FIXED, 0.0, 0.5),
        (1, 1.0, 1.0, api.FIT_GAUSSIAN_FIXED, 0.0, 1.0),
    ]
    if wide:  # more than 64 crossover events per gamete: the wide recombination kernel
        specs.append((150, 1.5, 60.0, api.FIT_GAUSSIAN_FIXED, 0.3, 1.0))
    cfgs = []
    for r, (n, mu, rho, fit, neutral, env) in enumerate(specs):
        kind = fit
        if traits > 1 and fit in (api.FIT_GAUSSIAN, api.FIT_GAUSSIAN_FIXED):
            kind = api.FIT_MV_STABILIZING
        cfgs.append(api.make_config(n, mu, rho, seed=1000 + r, replicate=r, traits=traits,
                                    fitness=kind, neutral_p=neutral, sigma=0.2, env_sigma=env,
                                    fitness_scale=1.5, max_alleles=1 << 16))
    return cfgs


def _lone(cfg, generations):
    sim = api.Simulation(cfg)
    sim.run(generations)
    return sim


def test_batch_graph_matches_lone_contexts():
    cfgs = _configs()
    batch = api.Batch(cfgs)
    batch.run(5, mode="graph")
    batch.run(4, mode="graph")  # odd then even parity
    for r, cfg in enumerate(cfgs):
        print(f"replicate {r}: n={cfg.n_individuals} rho={cfg.rho}", flush=True)
        lone = _lone(cfg, 9)
        compare_state(batch.replicate(r), lone, f"replicate {r}")
        compare_gametes(batch.replicate(r), lone, f"replicate {r}")
    batch.close()


def test_batch_modes_agree():
    cfgs = _configs(wide=False)
    runs = {}
    for mode in ("graph", "eager", "sequential"):
        b = api.Batch(cfgs)
        b.run(3, mode=mode)
        b.run(3, mode="eager" if mode == "graph" else mode)
        runs[mode] = b
    for r in range(len(cfgs)):
        compare_state(runs["eager"].replicate(r), runs["graph"].replicate(r), f"eager/graph {r}")
        compare_state(runs["sequential"].replicate(r), runs["graph"].replicate(r),
                      f"sequential/graph {r}")
    for b in runs.values():
        b.close()


def test_batch_replicates_match_oracle():
    cfgs = _configs(wide=False)[:3]
    batch = api.Batch(cfgs)
    orcs = [oracle.OracleSim(c) for c in cfgs]
    for step in range(2):
        batch.run(5, mode="graph")
        for r, o in enumerate(orcs):
            for _ in range(5):
                o.step()
            compare_state(batch.replicate(r), o, f"replicate {r} generation {5 * (step + 1)}")
    batch.close()


@pytest.mark.parametrize("traits", [3, 6, 20])
def test_batch_multitrait(traits):
    # 1, 2 and 8 digit-plane tiles: the batched packing visits the live
    # 16-word groups of every tile, and the batched trait kernels of up to 8,
    # 16 and 32 traits run
    cfgs = _configs(traits=traits, wide=False)
    batch = api.Batch(cfgs)
    batch.run(6, mode="graph")
    for r, cfg in enumerate(cfgs):
        print(f"traits {traits} replicate {r}", flush=True)
        compare_state(batch.replicate(r), _lone(cfg, 6), f"replicate {r}")
    batch.close()


def test_batch_uploaded_population():
    cfgs = _configs(wide=False)[:2]
    rng = np.random.default_rng(5)
    pops = []
    for cfg in cfgs:
        n, alleles = cfg.n_individuals, 700
        words = ((alleles + 511) // 512) * 8
        bits = rng.uniform(0, 1, (2 * n, words * 64)) < 0.3
        bits[:, alleles:] = False
        bits[:, 5] = True  # a fixed column
        rows = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64)
        pos = np.floor(rng.uniform(0, 1, alleles) * 2**32) / 2**32
        eff = np.round(rng.normal(0, 0.2, (alleles, 1)) * 2**24) / 2**24
        pops.append((rows, pos, eff))
    batch = api.Batch(cfgs)
    lones = []
    for r, cfg in enumerate(cfgs):
        batch.replicate(r).set_population(*pops[r])
        lone = api.Simulation(cfg)
        lone.set_population(*pops[r])
        lones.append(lone)
    batch.run(4, mode="graph")
    for r, lone in enumerate(lones):
        lone.run(4)
        compare_state(batch.replicate(r), lone, f"uploaded replicate {r}")
    batch.close()


def test_batch_capacity_error_names_replicate():
    cfgs = _configs(wide=False)[:3]
    cfgs[2] = api.make_config(500, 5.0, 1.0, seed=9, replicate=2, traits=1, max_alleles=1024)
    batch = api.Batch(cfgs)
    with pytest.raises(api.ClothoError, match="replicate 2"):
        batch.run(20, mode="graph")
    batch.close()
