"""The configuration front end of the C++ driver (include/clotho_amd/config.hpp,
population_growth.hpp, engine_demo): JSON configuration files in the
reference's layout, the population-growth toolkit and the fitness toolkit
registry.  CPU only: --registry and --print-config never touch a GPU."""
import json
import os
import subprocess

import pytest

from clotho_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def demo(*args, check=True):
    if not os.path.exists(build.DEMO):
        build.build_host_demo()
    r = subprocess.run([build.DEMO, *args], capture_output=True, text=True)
    if check:
        assert r.returncode == 0, r.stderr
    return r


def registry(*args):
    return json.loads(demo("--registry", *args).stdout)


def test_linear_growth_schedule_closed_forms():
    # pN' = A pN + B truncated; founders growth(0, 0) = B
    d = registry("population_growth.params.A=1.5", "population_growth.params.B=20",
                 "--generations", "6", "--schedule", "6")
    assert d["schedule"] == [20, 50, 95, 162, 263, 414, 641]
    assert d["n"] == 20 and d["max_individuals"] == 641
    # A = 2, B = 1: 2^(g+1) - 1
    d = registry("population_growth.params.A=2", "population_growth.params.B=1",
                 "--generations", "10", "--schedule", "10")
    assert d["schedule"] == [2 ** (g + 1) - 1 for g in range(11)]
    # A = 0: constant B, the capacity is the size
    d = registry("population_growth.params.A=0", "population_growth.params.B=777",
                 "--schedule", "4")
    assert d["schedule"] == [777] * 5 and d["n"] == d["max_individuals"] == 777
    # A = 0.5 shrinks toward 2 B: 2B (1 - 2^-(g+1)) truncated at every step
    d = registry("population_growth.params.A=0.5", "population_growth.params.B=100",
                 "--schedule", "5")
    assert d["schedule"] == [100, 150, 175, 187, 193, 196]
    # capacity over the default 10 generations: 198, 199, 199, 199, 199 follow
    assert d["max_individuals"] == 199


def test_predefined_growth_cycles_by_generation():
    d = registry("population_growth.name=predefined", "population_growth.params.sizes=[100,300,50]",
                 "--generations", "7", "--schedule", "7")
    assert d["schedule"] == [100, 300, 50, 100, 300, 50, 100, 300]
    assert d["n"] == 100 and d["max_individuals"] == 300
    # one size: constant
    d = registry("population_growth.name=predefined", "population_growth.params.sizes=[64]",
                 "--schedule", "3")
    assert d["schedule"] == [64] * 4


def test_json_configuration_file(tmp_path):
    # the reference's layout: a "configuration" block, every leaf a string
    cfg = {"configuration": {
        "mutation": {"mu": "0.5"},
        "recombination": {"rho": "2.5", "sequence_bias": {"p": "0.25"}},
        "random_number": {"seed": "1234"},
        "fitness": {"name": "gaussian"},
        "population_growth": {"name": "predefined", "params": {"sizes": ["200", "400"]}},
        "generations": {"size": "5"},
        "log": {"period": "2", "sampling_size": [{"size": "50", "pairwise": "false"}]},
    }}
    path = tmp_path / "run.json"
    path.write_text(json.dumps(cfg))
    d = registry("-c", str(path), "--schedule", "5")
    assert d["n"] == 200 and d["max_individuals"] == 400
    assert d["mu"] == 0.5 and d["rho"] == 2.5 and d["seed"] == 1234
    assert d["fitness_kind"] == 1  # CLOTHO_FIT_GAUSSIAN
    assert d["schedule"] == [200, 400, 200, 400, 200, 400]
    # key=value after -c overrides; unquoted numbers read alike
    path.write_text(json.dumps({"configuration": {"mutation": {"mu": 0.125},
                                                  "population_growth": {"name": "linear",
                                                                        "params": {"A": 0, "B": 90}}}}))
    d = registry("-c", str(path), "mutation.mu=0.25")
    assert d["mu"] == 0.25 and d["n"] == 90


def test_print_config_round_trip(tmp_path):
    out = json.loads(demo("--print-config", "population_growth.name=predefined",
                          "population_growth.params.sizes=[30,60]", "fitness.name=quadratic",
                          "fitness.params.scale=3").stdout)
    c = out["configuration"]
    assert c["population_growth"]["params"]["sizes"] == ["30", "60"]
    tk = c["population_growth"]["toolkit"]
    assert tk["linear"]["params"] == {"A": "1", "B": "0"}
    assert tk["predefined"]["params"]["sizes"] == ["10000"]
    assert set(c["fitness"]["toolkit"]) == {"constant", "gaussian", "gaussian_fixed",
                                            "mv_stabilizing", "quadratic"}
    assert c["fitness"]["toolkit"]["quadratic"]["params"]["scale"] == "1"
    # the printed configuration reads back to the same run
    path = tmp_path / "printed.json"
    path.write_text(json.dumps(out))
    a = registry("-c", str(path), "--schedule", "4")
    b = registry("population_growth.name=predefined", "population_growth.params.sizes=[30,60]",
                 "fitness.name=quadratic", "fitness.params.scale=3", "--schedule", "4")
    assert a == b


def test_bad_configurations_are_errors(tmp_path):
    path = tmp_path / "bad.json"
    path.write_text('{"configuration": {"mutation": {"mu": }}}')
    r = demo("--registry", "-c", str(path), check=False)
    assert r.returncode != 0 and "JSON" in r.stderr
    r = demo("--registry", "population_growth.name=unknown", check=False)
    assert r.returncode != 0 and "founder size" in r.stderr
    r = demo("--registry", "-c", str(tmp_path / "missing.json"), check=False)
    assert r.returncode != 0 and "cannot read" in r.stderr
