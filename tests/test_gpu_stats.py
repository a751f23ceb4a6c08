"""Population statistics on the GPU: column and row margins of the current
population (all rows or a row sample) equal the oracle's set-bit walk and a
numpy brute force; the C++ allele_frequency adapter reports the reference's
stats fields."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from clotho_amd import api, build as hip_build

pytestmark = pytest.mark.gpu


def _bits(rows, alleles):
    return np.unpackbits(rows.view(np.uint8), axis=1, bitorder="little")[:, :alleles]


@pytest.mark.parametrize("n,mu,gens", [(1000, 1.0, 20), (75, 3.0, 9), (500, 0.5, 1)])
def test_counts_match_oracle_and_brute_force(n, mu, gens):
    cfg = api.make_config(n, mu, 1.0, seed=13, traits=1, neutral_p=1.0, max_alleles=1 << 15)
    sim = api.Simulation(cfg)
    sim.run(gens)
    s = sim.shape()
    print(f"n={n} words={s.words} alleles={s.alleles}", flush=True)
    rows = sim.rows()
    cols, rc = sim.allele_counts()
    ocols, orc = oracle.allele_counts(rows, s.alleles)
    assert np.array_equal(cols, ocols) and np.array_equal(rc, orc)
    b = _bits(rows, s.alleles)
    assert np.array_equal(cols, b.sum(0)) and np.array_equal(rc, b.sum(1))
    sample = api.sample_rows(13, s.generation, 0, s.haplotypes, max(1, s.haplotypes // 3))
    cols, rc = sim.allele_counts(sample)
    ocols, orc = oracle.allele_counts(rows, s.alleles, sample)
    assert np.array_equal(cols, ocols) and np.array_equal(rc, orc)


def test_counts_before_any_generation():
    sim = api.Simulation(api.make_config(10, 1.0, 1.0, max_alleles=1024))
    cols, rc = sim.allele_counts()
    assert cols.size == 0 and rc.size == 20 and not rc.any()


def test_counts_at_bench_shape():
    import bench

    cfg = bench.workload_config(api)
    sim = api.Simulation(cfg)
    f, p, e = bench.standing_variation()
    sim.set_synthetic(f, p, e, seed=42)
    sim.run(2, graph=True)
    s = sim.shape()
    cols, rc, ms = sim.allele_counts(timed=True)
    print(f"bench shape: {ms:.3f} ms for H={s.haplotypes} W={s.words}", flush=True)
    assert int(cols.sum()) == int(rc.sum())  # both count every set bit once
    rows = sim.rows()
    pick = np.arange(0, s.haplotypes, 97)
    assert np.array_equal(rc[pick], _bits(rows[pick], s.alleles).sum(1))
    sample = api.sample_rows(5, s.generation, 0, s.haplotypes, 3000)
    gcols, grc = sim.allele_counts(sample)
    ocols, orc = oracle.allele_counts(rows, s.alleles, sample)
    assert np.array_equal(gcols, ocols) and np.array_equal(grc, orc)


def test_engine_demo_statistics():
    demo = hip_build.DEMO
    args = [demo, "population_growth.params.B=400", "mutation.mu=2", "recombination.rho=1.5",
            "random_number.seed=9", "--generations", "12"]
    for k in (0, 200):
        out = json.loads(subprocess.check_output(args + ["--stats", str(k)], text=True))
        cfg = api.make_config(400, 2.0, 1.5, seed=9, traits=1, neutral_p=1.0, max_alleles=32768)
        orc = oracle.OracleSim(cfg)
        for _ in range(12):
            orc.step()
        s = orc.shape()
        sample = None if k == 0 else api.sample_rows(9, s.generation, 0, s.haplotypes, k)
        cols, rc = oracle.allele_counts(orc.rows(), s.alleles, sample)
        st = out["allele_frequency"]["stats"]
        for name, v in (("sequences_per_allele", cols), ("alleles_per_sequence", rc)):
            d = st[name]
            assert d["total"] == v.size and d["min"] == v.min() and d["max"] == v.max()
            assert d["mean"] == pytest.approx(v.mean(), rel=1e-12)
            assert d["median"] == pytest.approx(np.median(v), rel=1e-12)
            assert d["variance"] == pytest.approx(v.var(), rel=1e-9)


def test_pairwise_differences_match_oracle():
    """Device pairwise differences against the oracle on a simulated
    population: all rows, an unordered sample, rows spanning several 64-word
    chunks, the empty cases, and an out-of-range row."""
    cfg = api.make_config(200, 6.0, 2.0, seed=31, traits=1, neutral_p=0.5, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 16)
    sim = api.Simulation(cfg)
    sim.run(12)
    rows = sim.rows()
    assert rows.shape[1] > 64
    d = sim.pairwise_differences(np.arange(rows.shape[0]))
    assert np.array_equal(d, oracle.pairwise_differences(rows))
    sample = np.random.default_rng(2).permutation(rows.shape[0])[:77].astype(np.uint32)
    d, ms = sim.pairwise_differences(sample, timed=True)
    assert np.array_equal(d, oracle.pairwise_differences(rows, sample)) and ms > 0
    assert sim.pairwise_differences(np.array([3], np.uint32)).size == 0
    assert sim.pairwise_differences(np.array([], np.uint32)).size == 0
    with pytest.raises(api.ClothoError, match="out of range"):
        sim.pairwise_differences(np.array([0, rows.shape[0]], np.uint32))
