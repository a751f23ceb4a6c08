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


def test_engine_demo_log_files(tmp_path):
    """The demo's .config/.status/.performance records (the reference
    driver's loop) carry the statistics, free space, alleles and per-generation
    counts of the oracle's run."""
    prefix = str(tmp_path / "run")
    args = [build.DEMO, "population_growth.params.B=300", "mutation.mu=2",
            "recombination.rho=1.5", "random_number.seed=4", "log.period=3",
            "log.sampling.0.size=100", "--generations", "6", "--log-prefix", prefix]
    subprocess.check_call(args, stdout=subprocess.DEVNULL)
    # the configuration block as the reference driver writes it (QS/qtlsim_main.cpp:120-121)
    cfg_log = json.load(open(prefix + ".0.config.json"))["configuration"]
    assert cfg_log["mutation"]["mu"] == "2" and cfg_log["log"]["period"] == "3"
    cfg = api.make_config(300, 2.0, 1.5, seed=4, traits=1, neutral_p=1.0, max_alleles=32768)
    orc = oracle.OracleSim(cfg)
    free_counts = []
    for k in range(2):
        for _ in range(3):
            orc.step()
            free_counts.append(orc.shape().free_count)
        st = json.load(open(f"{prefix}.{k}.status.json"))
        s = orc.shape()
        cols, rc = oracle.allele_counts(orc.rows(), s.alleles)
        af = st["allele_frequency"]
        assert [int(x) for x in af["distribution"]] == cols.tolist()
        fd = np.bincount(cols, minlength=s.haplotypes + 1)
        assert [int(x) for x in af["frequency_distribution"]] == fd.tolist()
        assert float(af["stats"]["alleles_per_sequence"]["mean"]) == pytest.approx(rc.mean())
        sample = api.sample_rows(4, s.generation, 0, s.haplotypes, 100)
        scols, src = oracle.allele_counts(orc.rows(), s.alleles, sample)
        sa = st["sample"]["0"]["allele_frequency"]
        assert [int(x) for x in sa["distribution"]] == scols.tolist()
        pd = st["population_data"]
        assert int(pd["free_space"]["free_size"]) == s.free_count
        fx, ls = orc.free_space()
        assert [int(x) for x in pd["free_space"]["free"]] == fx.tolist() + ls.tolist()
        pos, _, age = orc.alleles()
        assert [float(a["position"]) for a in pd["allele_space"]] == pos.tolist()
        assert [int(a["age"]) for a in pd["allele_space"]] == age.tolist()
        assert [float(x) for x in pd["child"]["fitness"]] == orc.fitness().tolist()
    perf = json.load(open(prefix + ".0.performance.json"))
    assert len(perf["performance"]["simulate"]) == 6
    assert len(perf["performance"]["stats"]) == 2
    assert [int(x) for x in perf["memory"]["free_count"]] == free_counts


@pytest.mark.parametrize("graph", [False, True])
def test_engine_demo_population_growth(tmp_path, graph):
    """A JSON configuration with predefined population growth through the C++
    driver: each generation's offspring count is sizes[g mod 4]
    (predefined_population_growth) and the run matches the oracle driven with
    the same sizes.  With --graph the runs of equal size replay graphs."""
    sizes = [300, 500, 200, 200]
    conf = {"configuration": {
        "mutation": {"mu": "2"}, "recombination": {"rho": "1.5"},
        "random_number": {"seed": "6"}, "neutral": {"p": "1"},
        "fitness": {"name": "constant"}, "alleles": {"capacity": "32768"},
        "population_growth": {"name": "predefined", "params": {"sizes": [str(v) for v in sizes]}},
        "generations": {"size": "6"}}}
    path = tmp_path / "growth.json"
    path.write_text(json.dumps(conf))
    args = [build.DEMO, "-c", str(path)] + (["--graph"] if graph else [])
    out = json.loads(subprocess.run(args, check=True, capture_output=True, text=True).stdout)
    cfg = api.make_config(300, 2.0, 1.5, seed=6, traits=1, neutral_p=1.0, max_alleles=32768,
                          max_individuals=500)
    orc = oracle.OracleSim(cfg)
    for g in range(1, 7):
        orc.set_population_size(sizes[g % 4])
        orc.step()
    s = orc.shape()
    assert s.n_individuals == 200
    assert out["generation"] == s.generation and out["alleles"] == s.alleles
    assert out["free"] == s.free_count and out["fixed"] == s.fixed_count
    assert out["rows_fnv1a"] == fnv1a(orc.rows())


@pytest.mark.parametrize("graph", [False, True])
def test_engine_demo_intermediates(graph):
    """generations.intermediates = 2 through the C++ driver: each simulate()
    is one batch of three generations (the batched engine's simulate), and
    three of them match nine oracle steps with the same setting."""
    args = [build.DEMO, "population_growth.params.B=250", "mutation.mu=2",
            "recombination.rho=2", "random_number.seed=8", "neutral.p=0.4",
            "fitness.name=gaussian_fixed", "generations.intermediates=2",
            "--generations", "3"] + (["--graph"] if graph else [])
    out = json.loads(subprocess.run(args, check=True, capture_output=True, text=True).stdout)
    cfg = api.make_config(250, 2.0, 2.0, seed=8, traits=1, neutral_p=0.4,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=32768, intermediates=2)
    orc = oracle.OracleSim(cfg)
    for _ in range(9):
        orc.step()
    s = orc.shape()
    assert out["generation"] == s.generation and out["alleles"] == s.alleles
    assert out["free"] == s.free_count and out["fixed"] == s.fixed_count
    assert out["rows_fnv1a"] == fnv1a(orc.rows())


def test_engine_demo_pairwise_sample(tmp_path):
    """log.sampling_size entries with pairwise = true add the sample's
    pairwise-difference statistics to each .status record (the reference's
    sample_log_params); they match the oracle on the same sample rows."""
    prefix = str(tmp_path / "pw")
    args = [build.DEMO, "population_growth.params.B=150", "mutation.mu=3",
            "recombination.rho=2", "random_number.seed=5", "log.period=2",
            'log.sampling_size=[{"size": "60", "pairwise": "true"}, {"size": "20"}]',
            "--generations", "2", "--log-prefix", prefix]
    subprocess.check_call(args, stdout=subprocess.DEVNULL)
    cfg = api.make_config(150, 3.0, 2.0, seed=5, traits=1, neutral_p=1.0, max_alleles=32768)
    orc = oracle.OracleSim(cfg)
    orc.step()
    orc.step()
    st = json.load(open(f"{prefix}.0.status.json"))
    s = orc.shape()
    sample = np.unique(api.sample_rows(5, s.generation, 0, s.haplotypes, 60))
    d = np.sort(oracle.pairwise_differences(orc.rows(), sample)).astype(np.float64)
    pd = st["sample"]["0"]["pairwise_difference"]
    assert float(pd["min"]) == d[0] and float(pd["max"]) == d[-1]
    assert float(pd["mean"]) == pytest.approx(d.mean())
    assert float(pd["median"]) == float(np.median(d))
    assert float(pd["variance"]) == pytest.approx(d.var())
    assert "pairwise_difference" not in st["sample"]["1"]
