"""CPU checks of the C ABI library and the C++ adapters: the library loads
without a GPU, exports every entry point include/clotho_amd.h declares,
rejects invalid configurations before touching a device, and the adapter
driver parses the reference's configuration keys."""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from clotho_amd import api, build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "clotho_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(clotho_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    L = api.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(api.EXPORTS) == set(names)


def test_invalid_configs_rejected_before_device():
    L = api.lib()
    ctx = C.c_void_p()
    bad = [
        api.make_config(0, 1.0, 1.0),                       # N = 0
        api.make_config(10, 1.0, 1.0, bias=1.5),            # bias outside [0, 1]
        api.make_config(10, 1.0, 1.0, neutral_p=-0.1),
        api.make_config(10, 100.0, 1.0),                     # mu beyond the Poisson table
        api.make_config(10, 1.0, 1.0, traits=33),            # trait capacity
        api.make_config(10, 0.0, 1.0, fitness=api.FIT_GAUSSIAN),  # sigma = 0
        api.make_config(10, 1.0, 1.0, quantum_log2=3),
        api.make_config(10, 1.0, 1.0, max_alleles=0),
        api.make_config(10, 1.0, 1.0, intermediates=2, compact=True),  # batch consumes the list
        api.make_config(10, 1.0, 1.0, max_individuals=(1 << 28) + 1),
        # int8 digit sums of a row stay within int32 only up to 2^24 slots
        api.make_config(10, 1.0, 1.0, max_alleles=(1 << 24) + 1),
    ]
    for cfg in bad:
        assert L.clotho_create(C.addressof(cfg), 0, C.byref(ctx)) == -1


def test_poisson_tables_agree_with_oracle():
    for lam in (0.0, 0.01, 1.0, 5.0, 10.0, 20.0, 40.0):
        assert api.poisson_table(lam).tobytes() == oracle.poisson_table(lam).tobytes()


def test_engine_adapter_registry_and_config():
    demo = build.DEMO
    if not os.path.exists(demo):
        build.build_host_demo()
    out = subprocess.run([demo, "--registry", "fitness.name=quadratic",
                          "configuration.population_growth.params.B=250", "mutation.mu=0.5"],
                         check=True, capture_output=True, text=True).stdout
    d = json.loads(out)
    assert set(d["fitness_tools"]) == {"constant", "gaussian", "quadratic", "gaussian_fixed",
                                       "mv_stabilizing"}
    assert d["n"] == 250 and d["mu"] == 0.5 and d["fitness_kind"] == api.FIT_QUADRATIC
    # a model without a positive founder size is rejected
    r = subprocess.run([demo, "--registry", "population_growth.params.B=0"],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "founder size" in r.stderr


def test_exact_contraction_rejects_unquantisable_effects():
    """CLOTHO_TRAITS_EXACT_I8 quantises effects to k 2^-q with the largest q
    keeping |k| <= 2^30 - 1; non-finite effects, and magnitudes that would
    need q outside [-60, 60], are rejected before any device work."""
    rows = np.zeros((4, 8), np.uint64)
    rows[:, 0] = 0b1011
    good = np.full((512, 2), 0.25)
    bad_sets = [
        np.where(np.arange(1024).reshape(512, 2) == 7, np.inf, good),
        np.where(np.arange(1024).reshape(512, 2) == 9, -np.inf, good),
        np.where(np.arange(1024).reshape(512, 2) == 3, np.nan, good),
        np.full((512, 2), 2.0**95),     # q = -66
        np.full((512, 2), 2.0**-40),    # q = 69
    ]
    for eff in bad_sets:
        with pytest.raises(api.ClothoError, match="invalid argument"):
            api.trait_contract(rows, eff, api.TRAITS_EXACT_I8)
        freqs = np.full(512, 0.5)
        with pytest.raises(api.ClothoError, match="invalid argument"):
            api.trait_contract_synthetic(4, freqs, eff, api.TRAITS_EXACT_I8)


def test_mutation_allocator_draws_poisson():
    """mutation_allocator::allocate(N) draws M ~ Poisson(mu N) from the caller's
    engine (include/clotho/data_spaces/mutation/mutation_allocator.hpp:41-47):
    mean and variance both mu N, and a chi-square fit of the counts."""
    import scipy.stats as st

    gen = build.GENERIC_DEMO
    if not os.path.exists(gen):
        build.build_host_demo()
    k, n, mu = 4000, 500, 0.02  # lambda = 10
    out = subprocess.run([gen, "--allocator-sample", str(k), str(n), f"mutation.mu={mu}"],
                         check=True, capture_output=True, text=True).stdout
    x = np.array([int(v) for v in out.split()])
    lam = mu * n
    assert x.size == k
    assert abs(x.mean() - lam) < 4 * np.sqrt(lam / k)
    assert abs(x.var() - lam) < 4 * lam * np.sqrt(2.0 / k) + 0.5
    obs = np.bincount(np.minimum(x, 20), minlength=21)[3:21]
    exp = st.poisson.pmf(np.arange(3, 21), lam) * k
    exp[-1] = st.poisson.sf(19, lam) * k
    assert st.chisquare(obs, exp * obs.sum() / exp.sum()).pvalue > 1e-3
