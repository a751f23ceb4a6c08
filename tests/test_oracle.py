"""Pins of the CPU oracle: known-answer vectors, closed forms, brute force on
tiny inputs, and the reference's own hand-built unit-test constructions.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong tie rule or index, swapped strands) fails it.  Reference files
are cited relative to the reference root.
"""
import math

import numpy as np
import pytest
import scipy.stats as st

import oracle
from clotho_amd import api


# ----------------------------------------------------------------- Philox
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32 with 10 rounds.
    kats = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, out in kats:
        assert list(oracle.philox(ctr, key)) == out


def test_stream_counter_layout():
    # draw(seed, gen, stream, rep, index, sub) = philox(ctr=(index, sub, gen, stream|rep<<8), key=seed)
    seed = 0x0123456789ABCDEF
    got = oracle.draw(seed, 17, 3, 5, 1000, 2)
    want = oracle.philox([1000, 2, 17, 3 | (5 << 8)], [seed & 0xFFFFFFFF, seed >> 32])
    assert list(got) == list(want)


def test_uniform_streams_are_uniform():
    # position keys (32-bit) and 53-bit uniforms of 4096 draws: KS test
    keys = np.array([oracle.draw(42, 1, 3, 0, k, 0)[0] for k in range(4096)], np.float64)
    assert st.kstest(keys / 2**32, "uniform").pvalue > 1e-3


# ------------------------------------------------------ deterministic math
def test_det_log_exp_sqrt_recip_against_libm():
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.logspace(-300, 300, 3001), rng.uniform(1e-6, 10, 2000)])
    for x in xs:
        if x == 1.0:
            continue
        assert abs(oracle.det_log(x) - math.log(x)) <= 4e-16 * abs(math.log(x)) + 3e-17
        assert abs(oracle.det_sqrt(x) - math.sqrt(x)) <= 4e-16 * math.sqrt(x)
        assert abs(oracle.det_recip(x) - 1 / x) <= 4e-16 / x
    for y in np.linspace(-700, 700, 4001):
        assert abs(oracle.det_exp(y) - math.exp(y)) <= 4e-16 * math.exp(y)
    assert oracle.det_exp(0.0) == 1.0
    assert oracle.det_exp(-800.0) == 0.0
    assert oracle.det_log(1.0) == 0.0


def test_normal_ppf_matches_published_accuracy():
    # Acklam's approximation: relative error below 1.15e-9 everywhere.
    ps = np.concatenate([np.logspace(-16, -1.5, 400), np.linspace(0.02, 0.98, 961),
                         1 - np.logspace(-16, -1.5, 400)])
    for p in ps:
        want = st.norm.ppf(p)
        got = oracle.normal_ppf(p)
        assert abs(got - want) <= 1.2e-9 * max(abs(want), 1e-2), p
    assert oracle.normal_ppf(0.5) == 0.0


def test_poisson_table_matches_scipy():
    for lam in (0.0, 0.5, 1.0, 5.0, 10.0, 20.0, 40.0):
        cdf = oracle.poisson_table(lam)
        assert cdf[-1] == 1.0
        ks = np.arange(len(cdf) - 1)
        assert np.allclose(cdf[:-1], st.poisson.cdf(ks, lam), rtol=1e-12, atol=1e-15)
        if lam > 0:
            assert 1.0 - st.poisson.cdf(len(cdf) - 2, lam) < 1e-15
    with pytest.raises(ValueError):
        oracle.poisson_table(100.0)


# ------------------------------------------------------------ recombination
def brute_crossover(top, bottom, wc, positions, events):
    """child[j] = odd(#{e <= pos_j}) ? bottom[j] : top[j] (zero-filled)."""
    out = np.zeros(wc, np.uint64)
    ev = np.asarray(events, np.float64)
    for w in range(wc):
        word = 0
        for j in range(64):
            s = 64 * w + j
            t = int(top[w]) >> j & 1 if w < len(top) else 0
            b = int(bottom[w]) >> j & 1 if w < len(bottom) else 0
            if t == b:
                bit = t
            else:
                cnt = int(np.sum(ev <= positions[s])) if s < len(positions) else 0
                bit = b if (cnt & 1) else t
            word |= bit << j
        out[w] = word
    return out


def test_crossover_reference_construction():
    """UT/data_spaces/crossover_test.cpp:53-110 rebuilt with its event injected.

    Parent 1 = rows 2 (top) and 3 (bottom); the test's mt19937(1234) stream
    produces one event at 0.562272, injected here.  Expected: 77 set
    (homozygous), 33 unset, 88 set, 99 unset, 22 set, 15 set."""
    words = 8
    positions = np.zeros(512, np.float64)
    positions[88], positions[99], positions[22], positions[15] = 0.75, 0.85, 0.5622, 0.562273
    rows = np.zeros((4, words), np.uint64)

    def setbit(r, j):
        rows[r, j // 64] |= np.uint64(1 << (j % 64))

    for r, j in [(0, 28), (0, 110), (1, 28), (2, 22), (2, 77), (2, 99), (3, 77), (3, 88), (3, 15)]:
        setbit(r, j)
    child = oracle.crossover(rows[2], rows[3], words, positions, [0.562272])

    def bit(j):
        return int(child[j // 64]) >> (j % 64) & 1

    assert bit(77) == 1 and bit(33) == 0 and bit(88) == 1
    assert bit(99) == 0 and bit(22) == 1 and bit(15) == 1
    # offspring 0 from parent 0 with no event: a copy of the top strand
    c0 = oracle.crossover(rows[0], rows[1], words, positions, [])
    assert int(c0[0]) >> 28 & 1 == 1 and np.array_equal(c0, rows[0])
    # swapping strands (sequence_bias) mirrors the result on het sites
    sw = oracle.crossover(rows[3], rows[2], words, positions, [0.562272])
    assert int(sw[99 // 64]) >> (99 % 64) & 1 == 1 and int(sw[0]) >> 22 & 1 == 0


def test_crossover_tie_rule_is_upper_bound():
    # region_classifier uses upper_bound: an event equal to a position counts
    # (include/clotho/classifiers/region_classifier.hpp:73).
    positions = np.array([0.25, 0.5, 0.75] + [0.0] * 509)
    top = np.array([0b001] + [0] * 7, np.uint64)
    bottom = np.array([0b110] + [0] * 7, np.uint64)
    child = oracle.crossover(top, bottom, 8, positions, [0.5])
    assert int(child[0]) & 0b111 == 0b111  # slot 1 (pos 0.5 == event) takes bottom
    child = oracle.crossover(top, bottom, 8, positions, [0.5000001])
    assert int(child[0]) & 0b111 == 0b101


@pytest.mark.parametrize("seed", range(6))
def test_crossover_brute_force(seed):
    rng = np.random.default_rng(seed)
    wp = 8 * int(rng.integers(1, 3))
    wc = wp + 8 * int(rng.integers(0, 2))
    positions = np.floor(rng.uniform(0, 1, wc * 64) * 2**32) / 2**32
    top = rng.integers(0, 2**63, wp, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, wp, dtype=np.uint64)
    bottom = rng.integers(0, 2**63, wp, dtype=np.uint64)
    ne = int(rng.integers(0, 9))
    events = list(positions[rng.integers(0, wp * 64, ne // 2)]) + list(rng.uniform(0, 1, ne - ne // 2))
    got = oracle.crossover(top, bottom, wc, positions, events)
    want = brute_crossover(top, bottom, wc, positions, events) if ne else np.concatenate(
        [top, np.zeros(wc - wp, np.uint64)])
    assert np.array_equal(got, want)


# --------------------------------------------------------------- free space
def test_free_space_reference_construction():
    """UT/data_spaces/free_space_test.cpp:30-66: 11 individuals (22 rows) x 200
    alleles, column 101 set in every row, column 33 in row 2 -> fixed 1,
    lost 198, free 199."""
    rows = np.zeros((22, 8), np.uint64)
    rows[:, 101 // 64] |= np.uint64(1 << (101 % 64))
    rows[2, 0] |= np.uint64(1 << 33)
    fixed, lost = oracle.free_space(rows, 200)
    assert list(fixed) == [101]
    assert len(lost) == 198 and len(fixed) + len(lost) == 199
    assert 33 not in lost and 101 not in lost and list(lost) == sorted(lost)
    assert lost.max() < 200  # the < M filter drops padding columns


@pytest.mark.parametrize("seed", range(4))
def test_free_space_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    h, w = int(rng.integers(1, 12)), 8 * int(rng.integers(1, 3))
    m = int(rng.integers(1, w * 64 + 1))
    dens = rng.uniform(0, 1, w * 64)
    bits = (rng.uniform(0, 1, (h, w * 64)) < dens).astype(np.uint8)
    bits[:, m:] = 0
    bits[:, rng.integers(0, m, 5)] = 1  # some fixed columns
    rows = np.packbits(bits, axis=1, bitorder="little").view(np.uint64)
    fixed, lost = oracle.free_space(rows, m)
    col_and = bits.all(axis=0)
    col_or = bits.any(axis=0)
    assert list(fixed) == [j for j in range(m) if col_and[j]]
    assert list(lost) == [j for j in range(m) if not col_or[j]]


# ---------------------------------------------------------------- phenotype
def test_phenotype_naive_sum():
    # UT/data_spaces/row_grouped_trait_accumulator_test.cpp:131-168 style:
    # compare with a naive per-bit sum (exact for dyadic effects).
    rng = np.random.default_rng(9)
    h, w, t = 7, 16, 3
    bits = (rng.uniform(0, 1, (h, w * 64)) < 0.3).astype(np.uint8)
    rows = np.packbits(bits, axis=1, bitorder="little").view(np.uint64)
    eff = np.round(rng.normal(0, 1, (w * 64, t)) * 2**20) / 2**20
    got = oracle.phenotype(rows, eff)
    assert np.array_equal(got, bits.astype(np.float64) @ eff)


# ---------------------------------------------------------------- selection
def test_selection_uniform_when_constant():
    n = 1000
    par = oracle.select(np.full(n, 0.5), 42, 3)
    for i in range(0, n, 97):
        d = oracle.draw(42, 3, 0, 0, i, 0)
        k0 = ((int(d[1]) << 32) | int(d[0])) >> 11
        k1 = ((int(d[3]) << 32) | int(d[2])) >> 11
        assert par[2 * i] == (k0 * n) >> 53 and par[2 * i + 1] == (k1 * n) >> 53


def test_selection_proportional_chi_square():
    n = 50
    rng = np.random.default_rng(4)
    f = rng.uniform(0, 1, n)
    f[7] = 0.0
    counts = np.zeros(n)
    for gen in range(1, 81):
        par = oracle.select(f, 11, gen)
        counts += np.bincount(par, minlength=n)
    assert counts[7] == 0
    exp = counts.sum() * f / f.sum()
    mask = exp > 0
    chi = ((counts[mask] - exp[mask]) ** 2 / exp[mask]).sum()
    assert st.chi2.sf(chi, mask.sum() - 1) > 1e-4


def test_selection_inverse_cdf_definition():
    # idx = min{i : floor(k T / 2^53) < C[i]} on the quantised cumulative table
    n = 37
    rng = np.random.default_rng(5)
    f = rng.uniform(0.1, 2.0, n)
    par = oracle.select(f, 99, 2)
    L = max(1, math.ceil(math.log2(n)))
    sc = 52 - L - (math.frexp(f.max())[1])
    q = np.floor(np.ldexp(f, sc))
    C = np.cumsum(q)
    T = int(C[-1])
    for i in range(n):
        d = oracle.draw(99, 2, 0, 0, i, 0)
        k = ((int(d[1]) << 32) | int(d[0])) >> 11
        target = (k * T) >> 53
        assert par[2 * i] == int(np.searchsorted(C, target, side="right"))


# ------------------------------------------------------------------ fitness
def test_fitness_closed_forms():
    n, mu = 100, 0.01
    cfg = api.make_config(n, mu, 1.0, fitness=api.FIT_GAUSSIAN)
    sigma = math.sqrt(4 * n * mu)
    for x in (0.0, 0.3, -1.7, 5.0):
        want = math.exp(-x * x / (2 * sigma * sigma)) / (sigma * math.sqrt(2 * math.pi))
        assert abs(oracle.fitness_value(cfg, n, [x]) - want) <= 1e-15 * max(want, 1e-300)
    cfg = api.make_config(n, mu, 1.0, fitness=api.FIT_QUADRATIC, fitness_scale=2.0)
    s = 2.0 * math.sqrt(4 * n * mu)
    assert oracle.fitness_value(cfg, n, [0.0]) == 1.0
    assert oracle.fitness_value(cfg, n, [s]) == 0.0 and oracle.fitness_value(cfg, n, [-3 * s]) == 0.0
    assert abs(oracle.fitness_value(cfg, n, [0.5 * s]) - 0.75) < 1e-15
    cfg = api.make_config(n, mu, 1.0, traits=3, fitness=api.FIT_MV_STABILIZING, fitness_scale=1.0)
    z = [0.5, -1.0, 2.0]
    assert abs(oracle.fitness_value(cfg, n, z) - math.exp(-sum(v * v for v in z) / 2)) < 1e-16
    cfg = api.make_config(n, mu, 1.0, fitness=api.FIT_CONSTANT)
    assert oracle.fitness_value(cfg, n, [123.0]) == 1.0


# ----------------------------------------------------- whole-generation oracle
def test_generation_invariants():
    """One generation: child rows are recombinations of the recorded parents,
    new mutations land on previously free columns, one per slot, in the order
    [fixed asc, lost asc, appended]; free lists equal the column AND / ~OR."""
    cfg = api.make_config(60, 3.0, 2.0, seed=21, traits=1, neutral_p=0.5, sigma=0.3,
                          fitness=api.FIT_GAUSSIAN_FIXED)
    sim = oracle.OracleSim(cfg)
    for _ in range(5):
        sim.step()
    parent = sim.rows()
    pfix, plost = sim.free_space()
    pos_before, _, _ = sim.alleles()
    a_before = sim.shape().alleles
    sim.step()
    child = sim.rows()
    shape = sim.shape()
    pos, eff, age = sim.alleles()
    top, bot, nev, nmut = sim.gametes()
    assert shape.new_mutations == nmut.sum()
    # remove fixed columns from the parent
    pclear = parent.copy()
    for j in pfix:
        pclear[:, j // 64] &= ~np.uint64(1 << (j % 64))
    free = list(pfix) + list(plost)
    slots = [int(free[k]) if k < len(free) else a_before + k - len(free) for k in range(int(nmut.sum()))]
    assert len(set(slots)) == len(slots)
    assert shape.alleles == a_before + max(0, nmut.sum() - len(free))
    off = np.concatenate([[0], np.cumsum(nmut)]).astype(np.int64)
    mutated = np.zeros_like(child)
    for o in range(len(nmut)):
        for k in range(off[o], off[o + 1]):
            s = slots[k]
            mutated[o, s // 64] |= np.uint64(1 << (s % 64))
            assert age[s] == shape.generation - 1
    recomb = child ^ mutated
    # mutated bits were clear after recombination
    assert not np.any(recomb & mutated)
    bits_p = np.unpackbits(pclear.view(np.uint8), axis=1, bitorder="little")
    bits_c = np.unpackbits(recomb.view(np.uint8), axis=1, bitorder="little")
    for o in range(len(top)):
        t, b = bits_p[top[o]], bits_p[bot[o]]
        homo = t == b
        w = min(len(t), len(bits_c[o]))
        assert np.array_equal(bits_c[o][:w][homo[:w]], t[:w][homo[:w]])
        if nev[o] == 0:
            assert np.array_equal(bits_c[o][:w], t[:w])
        assert top[o] // 2 == bot[o] // 2 and top[o] != bot[o]
    fx, ls = sim.free_space()
    cb = np.unpackbits(child.view(np.uint8), axis=1, bitorder="little")[:, : shape.alleles]
    assert list(fx) == list(np.nonzero(cb.all(axis=0))[0])
    assert list(ls) == list(np.nonzero(~cb.any(axis=0))[0])
    # traits equal the naive per-bit sum (effects are dyadic, sums exact)
    assert np.array_equal(sim.traits(), cb.astype(np.float64) @ eff)


def test_threads_do_not_change_results():
    cfg = api.make_config(200, 2.0, 2.0, seed=8, traits=1, neutral_p=0.3, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, env_sigma=0.5)
    a, b = oracle.OracleSim(cfg), oracle.OracleSim(cfg)
    b.set_threads(4)
    for _ in range(8):
        a.step()
        b.step()
    assert np.array_equal(a.rows(), b.rows())
    assert np.array_equal(a.fitness(), b.fitness())
    assert np.array_equal(a.traits(), b.traits())


def test_config0_dynamics_plausible():
    """Config 0 (neutral, mu = rho = 1, N = 1000): segregating sites after g
    generations grow like theta * sum 1/i with theta = 4 N mu = 4000 per
    gamete-pair... bounded checks only (parity is the GPU tests' job)."""
    cfg = api.make_config(1000, 1.0, 1.0, seed=42, traits=1, neutral_p=1.0)
    sim = oracle.OracleSim(cfg)
    for _ in range(40):
        sim.step()
    s = sim.shape()
    assert s.all_neutral == 1 and s.fitness_constant == 1
    assert 1500 <= s.new_mutations <= 2500
    assert 8000 < s.variable_count < 20000
    assert s.words == ((s.alleles + 511) // 512) * 8


# ---------------------------------------------------------------- compaction
def _genome_alleles(sim, exclude_free=True):
    """Per haplotype: sorted tuples (position, effects..., age) of its set
    bits, skipping the fixed/lost slots when exclude_free."""
    rows = sim.rows()
    pos, eff, age = sim.alleles()
    free = set()
    if exclude_free:
        fx, ls = sim.free_space()
        free = set(int(j) for j in fx) | set(int(j) for j in ls)
    bits = np.unpackbits(rows.view(np.uint8), axis=1, bitorder="little")[:, : pos.size]
    out = []
    for r in range(rows.shape[0]):
        slots = [j for j in np.flatnonzero(bits[r]) if j not in free]
        out.append(sorted((pos[j], *eff[j].tolist(), int(age[j])) for j in slots))
    return out


def test_compaction_preserves_population_as_allele_multisets():
    """Every draw is keyed by gamete or mutation index, never by slot, so a run
    that compacts after every generation carries, per genome, the same alleles
    (position, effect, age) as the uncompacted run minus its fixed/lost
    columns, and identical traits, fitness and gametes."""
    base = dict(seed=21, traits=2, neutral_p=0.3, sigma=0.3, fitness=api.FIT_MV_STABILIZING,
                fitness_scale=2.0, env_sigma=0.5, max_alleles=1 << 14)
    a = oracle.OracleSim(api.make_config(120, 3.0, 2.0, **base))
    b = oracle.OracleSim(api.make_config(120, 3.0, 2.0, compact=True, **base))
    for g in range(12):
        a.step()
        b.step()
        sa, sb = a.shape(), b.shape()
        assert sb.free_count == 0 and sb.alleles == sa.variable_count
        assert sb.words == ((sb.alleles + 511) // 512) * 8
        assert np.array_equal(a.traits(), b.traits())
        assert np.array_equal(a.fitness(), b.fitness())
        for x, y in zip(a.gametes(), b.gametes()):
            assert np.array_equal(x, y)
        assert _genome_alleles(a) == _genome_alleles(b, exclude_free=False), f"generation {g}"


def test_compaction_hand_built_case():
    """Five slots, slot 1 fixed (all rows) and slot 3 lost (no row): the kept
    slots 0, 2, 4 become 0, 1, 2 with their positions, effects and bits."""
    cfg = api.make_config(2, 0.0, 0.0, seed=1, traits=1, neutral_p=0.0, max_alleles=512)
    o = oracle.OracleSim(cfg)
    # rows (H = 4), bits over slots 0..4
    bits = np.array([[1, 1, 0, 0, 1],
                     [0, 1, 1, 0, 0],
                     [1, 1, 1, 0, 0],
                     [0, 1, 0, 0, 1]], np.uint8)
    rows = np.zeros((4, 8), np.uint64)
    for r in range(4):
        for j in range(5):
            if bits[r, j]:
                rows[r, 0] |= np.uint64(1) << np.uint64(j)
    pos = np.array([0.5, 0.25, 0.75, 0.125, 0.0625])
    eff = np.array([[1.0], [2.0], [3.0], [4.0], [5.0]])
    o.set_population(rows, pos, eff)
    fx, ls = o.free_space()
    assert fx.tolist() == [1] and ls.tolist() == [3]
    traits_before = o.traits().copy()
    o.compact()
    s = o.shape()
    assert s.alleles == 3 and s.words == 8 and s.free_count == 0
    p2, e2, _ = o.alleles()
    assert p2.tolist() == [0.5, 0.75, 0.0625] and e2[:, 0].tolist() == [1.0, 3.0, 5.0]
    want = [0b101, 0b010, 0b011, 0b100]
    assert o.rows()[:, 0].tolist() == want and not o.rows()[:, 1:].any()
    assert np.array_equal(o.traits(), traits_before)


# ---------------------------------------------------------------- statistics
def test_allele_counts_brute_force():
    rng = np.random.default_rng(12)
    for h, alleles in ((1, 1), (7, 70), (130, 1000), (64, 512)):
        words = ((alleles + 511) // 512) * 8
        bits = rng.uniform(0, 1, (h, words * 64)) < rng.uniform(0, 1, words * 64)[None, :]
        bits[:, alleles:] = False
        rows = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64)
        cols, rc = oracle.allele_counts(rows, alleles)
        assert np.array_equal(cols, bits[:, :alleles].sum(0))
        assert np.array_equal(rc, bits.sum(1))
        sample = rng.choice(h, size=max(1, h // 3), replace=False).astype(np.uint32)
        cols, rc = oracle.allele_counts(rows, alleles, sample)
        assert np.array_equal(cols, bits[sample, :alleles].sum(0))
        assert np.array_equal(rc, bits[sample].sum(1))


def test_sample_rows_properties():
    """The product's host sampler (random_sample_generator semantics): k >=
    population gives every row in order; otherwise k distinct in-range rows,
    deterministic in (seed, generation, replicate), roughly uniform."""
    assert api.sample_rows(1, 1, 0, 10, 10).tolist() == list(range(10))
    assert api.sample_rows(1, 1, 0, 10, 25).tolist() == list(range(10))
    s = api.sample_rows(42, 5, 3, 1000, 300)
    assert s.size == 300 and len(set(s.tolist())) == 300 and s.max() < 1000
    assert np.array_equal(s, api.sample_rows(42, 5, 3, 1000, 300))
    assert not np.array_equal(s, api.sample_rows(42, 6, 3, 1000, 300))
    hits = np.zeros(50, np.int64)
    for g in range(400):
        hits[api.sample_rows(7, g, 0, 50, 10)] += 1
    # each row is drawn with probability 10/50 per sample: 80 expected of 400
    assert hits.min() > 45 and hits.max() < 120


def test_oracle_population_size_changes():
    """Population growth in the oracle: an explicit size equal to N changes
    nothing; after a size change the child has 2 n rows and n fitness values,
    its gametes draw parents from the previous population only, and bottom is
    always the other haplotype of the top row's individual."""
    cfg = api.make_config(100, 1.5, 2.0, seed=9, traits=1, neutral_p=0.5, sigma=0.3,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 13,
                          max_individuals=300)
    a, b = oracle.OracleSim(cfg), oracle.OracleSim(cfg)
    b.set_population_size(100)
    for _ in range(3):
        a.step()
        b.step()
    assert np.array_equal(a.rows(), b.rows())
    assert np.array_equal(a.fitness(), b.fitness())
    n_par = 100
    for n in (250, 40, 40, 1, 300):
        a.set_population_size(n)
        a.step()
        s = a.shape()
        assert s.n_individuals == n and s.haplotypes == 2 * n
        top, bottom, _, _ = a.gametes()
        assert top.size == 2 * n and int(top.max()) < 2 * n_par
        assert np.array_equal(top ^ 1, bottom)
        assert a.rows().shape[0] == 2 * n and a.fitness().size == n
        # every slot of the allele table is fixed, lost or segregating in the child
        assert s.free_count + s.variable_count == s.alleles
        n_par = n
    with pytest.raises(ValueError):
        a.set_population_size(0)


def test_oracle_intermediate_generations():
    """k intermediate generations: generation numbers start at k + 1; between
    analyses the free list only loses the slots the new mutations took (from
    the front, fixed before lost) and new alleles beyond it are appended; after
    the last generation of a batch the lists are the free space of the
    current rows (or_free_space); an upload starts a batch."""
    k = 2
    cfg = api.make_config(120, 3.0, 2.0, seed=12, traits=1, neutral_p=0.5, sigma=0.3,
                          fitness=api.FIT_GAUSSIAN_FIXED, max_alleles=1 << 14, intermediates=k)
    orc = oracle.OracleSim(cfg)
    assert orc.shape().generation == k + 1
    for step in range(3 * (k + 1)):
        before = orc.shape()
        fx0, ls0 = orc.free_space()
        free0 = np.concatenate([fx0, ls0])
        orc.step()
        s = orc.shape()
        assert s.generation == before.generation + 1
        m, F = s.new_mutations, before.free_count
        assert s.alleles == before.alleles + max(0, m - F)
        pos, _, ages = orc.alleles()
        new_slots = np.flatnonzero(ages == before.generation)
        expect = np.concatenate([free0[:min(m, F)], before.alleles + np.arange(max(0, m - F))])
        assert sorted(new_slots.tolist()) == sorted(expect.tolist())
        fx, ls = orc.free_space()
        if (step + 1) % (k + 1):  # intermediate: the unused tail of the old lists
            assert np.array_equal(np.concatenate([fx, ls]), free0[min(m, F):])
        else:  # end of a batch: the free space of the current rows
            rows = orc.rows()
            ofx, ols = oracle.free_space(rows, s.alleles)
            assert np.array_equal(fx, ofx) and np.array_equal(ls, ols)
    with pytest.raises(ValueError):  # no compaction with intermediate generations
        oracle.OracleSim(api.make_config(100, 1.0, 1.0, intermediates=1, compact=True,
                                         max_alleles=1 << 12))


def test_oracle_pairwise_differences_against_numpy():
    """popcount(row a xor row b) for the pairs a < b of a sample, packed row by
    row, against numpy's bit unpacking; empty and single-row samples give no pairs."""
    rng = np.random.default_rng(21)
    rows = rng.integers(0, 2**63, (40, 24), dtype=np.uint64) ^ rng.integers(0, 2**63, (40, 24),
                                                                            dtype=np.uint64)
    bits = np.unpackbits(rows.view(np.uint8), axis=1)
    for sample in (np.arange(40), rng.permutation(40)[:17], np.array([5, 3]), np.array([7]),
                   np.array([], dtype=np.int64)):
        d = oracle.pairwise_differences(rows, sample)
        ref = [int((bits[a] != bits[b]).sum()) for i, a in enumerate(sample) for b in sample[i + 1:]]
        assert d.tolist() == ref


# ------------------------------------------- new-allele and noise laws
def test_new_allele_effect_law():
    """TraitSpaceGenerator2 (include/clotho/data_spaces/phenotype_evaluator/
    trait_space_generator.hpp:118-129) with weight_distribution_helper
    (include/clotho/data_spaces/generators/weight_distribution_helper.hpp:24-33):
    per allele and trait, independently, Bernoulli(neutral.p) gives 0, else
    Normal(traits.mean, traits.sigma).  About 1e5 new alleles of one oracle
    generation are checked against that law: the neutral fraction per trait
    and jointly (independence), a KS test of the non-neutral values against
    N(0.3, 0.7^2) (a sigma^2-for-sigma slip, a dropped mean or a shared
    draw fails), and the 2^q quantisation grid."""
    mean, sigma, p, q = 0.3, 0.7, 0.4, -24
    cfg = api.make_config(2000, 25.0, 0.0, seed=7, traits=2, neutral_p=p, mean=mean,
                          sigma=sigma, quantum_log2=q, max_alleles=1 << 17)
    orc = oracle.OracleSim(cfg)
    orc.step()
    pos, eff, age = orc.alleles()
    m = orc.shape().new_mutations
    assert m == eff.shape[0] and m > 95_000
    neutral = eff == 0.0
    sd = math.sqrt(m * p * (1 - p))
    for t in range(2):
        assert abs(neutral[:, t].sum() - p * m) <= 4 * sd
    both = np.logical_and(neutral[:, 0], neutral[:, 1]).sum()
    assert abs(both - p * p * m) <= 4 * math.sqrt(m * p * p * (1 - p * p))
    vals = eff[~neutral]
    assert st.kstest(vals, "norm", args=(mean, sigma)).pvalue > 1e-3
    assert abs(vals.mean() - mean) <= 4 * sigma / math.sqrt(vals.size)
    assert abs(vals.std() - sigma) <= 4 * sigma / math.sqrt(2 * vals.size)
    k = eff * 2.0**-q
    assert np.array_equal(k, np.round(k)) and np.abs(k).max() <= 2**30 - 1
    assert st.kstest(pos, "uniform").pvalue > 1e-3
    assert np.all(age == 1)


def test_environmental_noise_law():
    """Environmental noise N(0, env_sigma^2) per individual and trait, added to
    the diploid phenotype before fitness (no reference counterpart).  The
    draws are KS-tested against N(0, 1.7^2), and an all-neutral population
    with fixed-sigma Gaussian fitness exp(-z^2/(2 s^2)) must score exactly
    those draws (traits are zero, so z is the noise alone)."""
    env, s = 1.7, 2.0
    cfg = api.make_config(20_000, 0.5, 1.0, seed=11, traits=1, neutral_p=1.0,
                          fitness=api.FIT_GAUSSIAN_FIXED, fitness_scale=s, env_sigma=env,
                          max_alleles=1 << 16)
    noise = oracle.env_noise(cfg, 1, 20_000)
    assert st.kstest(noise, "norm", args=(0.0, env)).pvalue > 1e-3
    assert abs(noise.std() - env) <= 4 * env / math.sqrt(2 * noise.size)
    orc = oracle.OracleSim(cfg)
    orc.step()  # generation 1
    assert np.all(orc.traits() == 0.0)
    want = np.exp(-(noise * noise) / (2 * s * s))
    np.testing.assert_allclose(orc.fitness(), want, rtol=1e-14, atol=0)
    # a different trait index or generation is a different draw
    other = oracle.env_noise(cfg, 2, 100)
    assert not np.array_equal(other, noise[:100])


# --------------------------------------------- generic operator families
def test_oracle_given_parents_equal_the_drawn_ones():
    """Crossover<RNG,GS>::update(parents, mate_pairs, offspring)
    (include/clotho/data_spaces/crossover/crossover.hpp:49-72): gamete 2i
    recombines pair i's first individual, 2i+1 its second.  Giving the
    oracle exactly the pairs its own selection draws must reproduce the
    undirected generation bit for bit; other pairs must be used as given."""
    cfg = api.make_config(150, 1.5, 2.0, seed=31, traits=1, neutral_p=0.3, sigma=0.2,
                          fitness=api.FIT_GAUSSIAN_FIXED, env_sigma=0.5, max_alleles=1 << 14)
    a, b = oracle.OracleSim(cfg), oracle.OracleSim(cfg)
    for _ in range(3):
        a.step()
        b.step()
    # generation 4: b takes a's selection as given pairs
    drawn = oracle.select(b.fitness(), cfg.seed, b.shape().generation)
    b.set_parents(drawn)
    a.step()
    b.step()
    assert np.array_equal(a.rows(), b.rows())
    assert np.array_equal(a.gametes()[0] >> 1, drawn)
    # arbitrary pairs: every gamete's strands come from its given individual
    rng = np.random.default_rng(3)
    pairs = rng.integers(0, 150, 300).astype(np.uint32)
    b.set_parents(pairs)
    b.step()
    top, bot, _, _ = b.gametes()
    assert np.array_equal(top >> 1, pairs) and np.array_equal(bot >> 1, pairs)
    assert np.all((top ^ bot) == 1)
    with pytest.raises(ValueError):
        b.set_parents(np.zeros(299, np.uint32))
    with pytest.raises(ValueError):
        b.set_parents(np.full(300, 150, np.uint32))


def test_oracle_add_mutations_reference_known_answer():
    """The reference's mutation unit test (unittest/data_spaces/mutation_test.cpp:39-66):
    20 empty genomes over 200 allele slots have 200 free slots; 15 mutations
    leave 185, and the free-space analysis of the mutated population finds
    185.  Also the closed form: each mutation sets exactly its bit, slots
    follow the free list in order, popcount grows by 15."""
    cfg = api.make_config(10, 0.0, 0.0, seed=1, traits=1, neutral_p=0.0, max_alleles=1 << 12,
                          host_mutations=True)
    orc = oracle.OracleSim(cfg)
    words = 8
    orc.set_population(np.zeros((20, words), np.uint64), np.arange(200) / 256.0,
                       np.zeros((200, 1)))
    s = orc.shape()
    assert s.free_count == 200 and s.lost_count == 200 and s.alleles == 200
    rng = np.random.default_rng(1234)
    rows = rng.integers(0, 20, 15).astype(np.uint32)
    pos = np.floor(rng.uniform(0, 1, 15) * 2**32) / 2**32
    eff = np.round(rng.normal(0, 1, (15, 1)) * 2**24) / 2**24
    slots = orc.add_mutations(rows, pos, eff)
    assert list(slots) == list(range(15))  # the lost slots, ascending
    s = orc.shape()
    assert s.free_count == 185 and s.alleles == 200
    got = orc.rows()
    bits = np.unpackbits(got.view(np.uint8), axis=1, bitorder="little")
    assert bits.sum() == 15
    for k in range(15):
        assert bits[rows[k], slots[k]] == 1
    p, e, age = orc.alleles()
    assert np.array_equal(p[slots], pos) and np.array_equal(e[slots, 0], eff[:, 0])
    want = np.zeros(20)
    np.add.at(want, rows, eff[:, 0])
    assert np.array_equal(orc.traits()[:, 0], want)


def test_oracle_add_mutations_fixed_reuse_and_widening():
    """A reused fixed slot leaves every row before the new allele is set; more
    mutations than free slots append columns and widen the rows."""
    cfg = api.make_config(4, 0.0, 0.0, seed=1, traits=2, neutral_p=0.0, max_alleles=1 << 12,
                          host_mutations=True)
    orc = oracle.OracleSim(cfg)
    rows = np.zeros((8, 8), np.uint64)
    rows[:, 0] = 0b101  # slot 0 and 2 fixed, slot 1 lost
    rows[0, 0] |= 0b1000  # slot 3 segregating
    eff = np.full((4, 2), 0.5)
    orc.set_population(rows, np.array([0.1, 0.2, 0.3, 0.4]), eff)
    fx, ls = orc.free_space()
    assert list(fx) == [0, 2] and list(ls) == [1]
    n = 600  # 3 free slots, 597 appended: 601 slots need 16 words
    mrows = (np.arange(n) % 8).astype(np.uint32)
    pos = (np.arange(n) + 1) / 1024.0
    meff = np.full((n, 2), 0.25)
    slots = orc.add_mutations(mrows, pos, meff)
    assert list(slots[:3]) == [0, 2, 1] and list(slots[3:6]) == [4, 5, 6]
    s = orc.shape()
    assert s.words == 16 and s.alleles == 601
    got = orc.rows()
    bits = np.unpackbits(got.view(np.uint8), axis=1, bitorder="little")[:, :601].astype(int)
    want = np.zeros((8, 601), int)
    want[0, 3] = 1
    want[mrows, slots] = 1
    assert np.array_equal(bits, want)
    # traits: 0.25 per new allele, 0.5 for the one old segregating allele
    t = orc.traits()
    assert np.array_equal(t[:, 0], want.sum(1) * 0.25 + np.where(np.arange(8) == 0, 0.25, 0.0))


def test_oracle_evaluate_fitness_closed_form():
    """Fitness<phenotype_evaluator>::update: f_i = op(trait(2i) + trait(2i+1)
    + noise of the producing generation), with the configured metric."""
    cfg = api.make_config(64, 1.0, 1.0, seed=9, traits=2, neutral_p=0.2, sigma=0.4,
                          fitness=api.FIT_MV_STABILIZING, fitness_scale=1.5, env_sigma=0.3,
                          max_alleles=1 << 12)
    orc = oracle.OracleSim(cfg)
    for _ in range(3):
        orc.step()
    before = orc.fitness().copy()
    orc.evaluate_fitness()
    after = orc.fitness()
    gen = orc.shape().generation - 1
    t = orc.traits().reshape(64, 2, 2).sum(axis=1)
    noise = np.stack([oracle.env_noise(cfg, gen, 64, trait=j) for j in range(2)], axis=1)
    want = [oracle.fitness_value(cfg, 64, t[i] + noise[i]) for i in range(64)]
    assert np.array_equal(after, np.array(want))
    assert np.array_equal(after, before)  # the step computed the same thing
