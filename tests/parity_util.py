"""Shared helpers of the GPU parity tests: run the HIP library and the CPU
oracle on the same configuration and compare every output bit for bit."""
import numpy as np

import oracle
from clotho_amd import api


def compare_state(sim, orc, where=""):
    gs, os_ = sim.shape(), orc.shape()
    for key in ("generation", "n_individuals", "haplotypes", "words", "alleles", "new_mutations", "fixed_count",
                "lost_count", "free_count", "variable_count", "all_neutral", "fitness_constant"):
        assert getattr(gs, key) == getattr(os_, key), f"{where}: shape.{key} {getattr(gs, key)} != {getattr(os_, key)}"
    assert gs.error == 0
    g_rows, o_rows = sim.rows(), orc.rows()
    assert g_rows.shape == o_rows.shape
    if not np.array_equal(g_rows, o_rows):
        bad = np.argwhere(g_rows != o_rows)
        r, w = bad[0]
        raise AssertionError(f"{where}: {len(bad)} row words differ, first at row {r} word {w}: "
                             f"{int(g_rows[r, w]):#x} != {int(o_rows[r, w]):#x}")
    gp, ge, ga = sim.alleles()
    op, oe, oa = orc.alleles()
    assert np.array_equal(gp, op), f"{where}: positions differ"
    assert np.array_equal(ge, oe), f"{where}: effects differ"
    assert np.array_equal(ga, oa), f"{where}: ages differ"
    gf, gl = sim.free_space()
    of, ol = orc.free_space()
    assert np.array_equal(gf, of), f"{where}: fixed lists differ"
    assert np.array_equal(gl, ol), f"{where}: lost lists differ"
    gt, ot = sim.traits(), orc.traits()
    assert np.array_equal(gt, ot), f"{where}: traits differ (max |d| {np.max(np.abs(gt - ot)) if gt.size else 0})"
    gfit, ofit = sim.fitness(), orc.fitness()
    assert np.array_equal(gfit, ofit), f"{where}: fitness differs"


def compare_gametes(sim, orc, where=""):
    for name, a, b in zip(("top", "bottom", "events", "mutations"), sim.gametes(), orc.gametes()):
        assert np.array_equal(a, b), f"{where}: gamete {name} differs"


def run_pair(cfg, generations, check_every=1, graph=False, population=None):
    sim = api.Simulation(cfg)
    orc = oracle.OracleSim(cfg)
    if population is not None:
        rows, pos, eff, fit = population
        sim.set_population(rows, pos, eff, fit)
        orc.set_population(rows, pos, eff, fit)
        compare_state(sim, orc, "after upload")
    for g in range(generations):
        sim.run(1, graph=graph)
        orc.step()
        if (g + 1) % check_every == 0 or g + 1 == generations:
            compare_gametes(sim, orc, f"generation {g + 1}")
            compare_state(sim, orc, f"generation {g + 1}")
    return sim, orc
