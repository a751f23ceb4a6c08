"""The launch-shape autotuner leaves every visible buffer as it was and the
shapes it picks keep the results bit-exact: the recombination's row chunks,
the traits pass's kernel shape (1, 5 and 16 traits), the compaction's rows per
step and row chunks, and the persistent grid of replicate batches."""
import numpy as np
import pytest

import oracle
from clotho_amd import api
from tests.parity_util import compare_gametes, compare_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("traits,compact", [(1, False), (5, False), (16, True)])
def test_autotuned_shapes_keep_parity(traits, compact):
    kind = api.FIT_GAUSSIAN_FIXED if traits == 1 else api.FIT_MV_STABILIZING
    cfg = api.make_config(700, 2.0, 3.0, seed=90 + traits, traits=traits, neutral_p=0.2,
                          sigma=0.2, fitness=kind, fitness_scale=3.0, env_sigma=0.5,
                          max_alleles=1 << 14, compact=compact)
    sim, orc = api.Simulation(cfg), oracle.OracleSim(cfg)
    seen = set()
    for g in range(5):
        sim.run(1, graph=(g % 2 == 1))
        orc.step()
        sim.autotune()
        shape = sim.launch_shape()
        print(f"generation {g + 1}: {shape}", flush=True)
        assert 1 <= shape["traits_variant"] <= shape["traits_variants"]
        assert shape["compact_rows"] in (4, 8, 16) and shape["compact_chunks"] >= 1
        assert shape["compact_tile_words"] == 64  # not tuned: stays as set
        seen.add(shape["traits_variant"])
        compare_state(sim, orc, f"after autotune {g + 1}")
    for g in range(3):  # the tuned shapes, eager and replayed
        sim.run(1, graph=(g == 1))
        orc.step()
        compare_gametes(sim, orc, f"tuned generation {g + 1}")
        compare_state(sim, orc, f"tuned generation {g + 1}")


@pytest.mark.parametrize("traits,variant,rows,tile_words",
                         [(3, 1, 4, 64), (3, 2, 16, 60), (3, 3, 8, 56), (6, 2, 8, 48),
                          (12, 2, 4, 56), (20, 2, 16, 64)])
def test_each_shape_is_exact(traits, variant, rows, tile_words):
    """Every shape the tuner can pick computes the same generation: traits
    kernels of each trait range, compaction rows per step, tile widths and
    chunk counts."""
    cfg = api.make_config(900, 3.0, 2.0, seed=71 + traits, traits=traits, neutral_p=0.1,
                          sigma=0.3, fitness=api.FIT_MV_STABILIZING, fitness_scale=2.0,
                          max_alleles=1 << 16, compact=True)
    sim, orc = api.Simulation(cfg), oracle.OracleSim(cfg)
    sim.set_launch_shape(traits_variant=variant, compact_rows=rows, compact_chunks=3,
                         compact_tile_words=tile_words, recombine_chunks=5)
    with pytest.raises(api.ClothoError):
        sim.set_launch_shape(traits_variant=sim.launch_shape()["traits_variants"] + 1)
    with pytest.raises(api.ClothoError):
        sim.set_launch_shape(compact_tile_words=40)
    for g in range(4):
        sim.run(1, graph=(g >= 2))
        orc.step()
        compare_state(sim, orc, f"shape {variant}/{rows} generation {g + 1}")


def test_batch_autotune_keeps_parity():
    cfgs = [api.make_config(n, mu, rho, seed=500 + r, replicate=r, traits=1,
                            fitness=api.FIT_GAUSSIAN_FIXED, neutral_p=0.3, sigma=0.2,
                            env_sigma=1.0, max_alleles=1 << 14)
            for r, (n, mu, rho) in enumerate([(400, 1.0, 2.0), (1200, 0.5, 4.0), (90, 2.0, 1.0)])]
    batch = api.Batch(cfgs)
    batch.run(3, mode="graph")
    pct = batch.autotune()
    assert pct in (50, 100, 200)
    batch.run(3, mode="graph")
    batch.run(2, mode="eager")
    for r, cfg in enumerate(cfgs):
        orc = oracle.OracleSim(cfg)
        for _ in range(8):
            orc.step()
        compare_gametes(batch.replicate(r), orc, f"replicate {r}")
        compare_state(batch.replicate(r), orc, f"replicate {r}")
    batch.close()
