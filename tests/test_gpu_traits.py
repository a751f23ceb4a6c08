"""Trait contraction on the GPU (plain FP32 walk, FP64 walk, split-BF16 MFMA)
against an FP64 reference of the same op (numpy matmul of the 0/1 genotype
matrix with the effects)."""
import numpy as np
import pytest

from clotho_amd import api

pytestmark = pytest.mark.gpu


def random_case(h, sites, t, density, seed):
    rng = np.random.default_rng(seed)
    w = ((sites + 511) // 512) * 8
    bits = (rng.uniform(0, 1, (h, w * 64)) < density).astype(np.uint8)
    bits[:, sites:] = 0
    rows = np.packbits(bits, axis=1, bitorder="little").view(np.uint64)
    eff = rng.normal(0, 1, (sites, t))
    ref = bits[:, :sites].astype(np.float64) @ eff
    return rows, eff, ref


def normwise(z, ref):
    return np.max(np.abs(z - ref), axis=0) / np.max(np.abs(ref), axis=0)


@pytest.mark.parametrize("h,sites,t", [(1000, 1024, 16), (333, 700, 5), (48, 64, 1)])
def test_contraction_modes(h, sites, t):
    rows, eff, ref = random_case(h, sites, t, 0.5, seed=h + t)
    z64, _ = api.trait_contract(rows, eff, api.TRAITS_EXACT_FP64)
    assert np.max(normwise(z64, ref)) < 1e-13
    z32, _ = api.trait_contract(rows, eff, api.TRAITS_PLAIN_FP32)
    assert np.max(normwise(z32, ref)) < 1e-5
    zw, _ = api.trait_contract(rows, eff, api.TRAITS_WALK_FP32)
    assert np.max(normwise(zw, ref)) < 1e-5
    tol = {1: 1e-2, 2: 3e-5, 3: 1e-6}
    for terms in (1, 2, 3):
        zm, _ = api.trait_contract(rows, eff, api.TRAITS_MFMA_BF16, split_terms=terms)
        err = np.max(normwise(zm, ref))
        assert err < tol[terms], (terms, err)


def test_contraction_bf16_exact_on_small_integers():
    # integer effects below 2^8 are exact in one BF16 term and sums stay exact in FP32
    rng = np.random.default_rng(3)
    rows, _, _ = random_case(256, 512, 16, 0.3, seed=3)
    bits = np.unpackbits(rows.view(np.uint8), axis=1, bitorder="little")[:, :512]
    eff = rng.integers(-100, 100, (512, 16)).astype(np.float64)
    zm, _ = api.trait_contract(rows, eff, api.TRAITS_MFMA_BF16, split_terms=1)
    assert np.array_equal(zm, bits.astype(np.float64) @ eff)


def test_synthetic_contraction_matches_numpy_fp64():
    """Device-drawn genotypes (the precision workload's generator) rebuilt on
    the CPU: the FP64 walk equals numpy float64 and 3-term split BF16 stays
    below 1e-6 normwise."""
    import oracle

    h, s, t = 400, 3000, 16
    freqs = np.random.default_rng(42).uniform(0, 1, s)
    rng = np.random.default_rng(43)
    eff = rng.standard_normal((s, t))
    rows = oracle.synth_rows(42, freqs, 0, h)
    bits = np.unpackbits(rows.view(np.uint8), axis=1, bitorder="little")[:, :s]
    ref = bits.astype(np.float64) @ eff
    z64, _ = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_EXACT_FP64, seed=42)
    assert np.max(normwise(z64, ref)) < 1e-13
    zm, _ = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_MFMA_BF16, seed=42,
                                         split_terms=3)
    assert np.max(normwise(zm, ref)) < 1e-6


@pytest.mark.parametrize("h,sites,t", [(1000, 1024, 16), (333, 700, 5), (48, 64, 1), (700, 3000, 9)])
def test_contraction_exact_i8(h, sites, t):
    """The pipeline's exact path: effects quantised to k 2^-q with the largest
    q keeping |k| below 2^30; the result equals the float64 sum of the
    quantised effects bit for bit (integer sums below 2^53)."""
    rows, eff, ref = random_case(h, sites, t, 0.4, seed=7 * h + t)
    bits = np.unpackbits(rows.view(np.uint8), axis=1, bitorder="little")[:, :sites]
    q = 29 - int(np.floor(np.log2(np.max(np.abs(eff)))))
    while np.max(np.abs(eff)) * 2.0**q > 2**30 - 1:
        q -= 1
    k = np.round(eff * 2.0**q)  # ties: numpy rounds half to even, llround away from zero
    ties = np.abs(eff * 2.0**q - np.trunc(eff * 2.0**q)) == 0.5
    k[ties] = np.trunc(eff * 2.0**q)[ties] + np.sign(eff[ties])
    exact = (bits.astype(np.int64) @ k.astype(np.int64)).astype(np.float64) * 2.0**-q
    zi, _ = api.trait_contract(rows, eff, api.TRAITS_EXACT_I8)
    assert np.array_equal(zi, exact)
    # against the unquantised float64 product: at most half a quantum per set bit
    bound = 0.5 * 2.0**-q * bits.sum(axis=1, keepdims=True) + 1e-12 * np.abs(ref)
    assert np.all(np.abs(zi - ref) <= bound)


def test_synthetic_contraction_exact_i8():
    import oracle

    h, s, t = 400, 3000, 16
    freqs = np.random.default_rng(42).uniform(0, 1, s)
    eff = np.random.default_rng(43).standard_normal((s, t))
    rows = oracle.synth_rows(42, freqs, 0, h)
    zr, _ = api.trait_contract(rows, eff, api.TRAITS_EXACT_I8)
    zs, _ = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_EXACT_I8, seed=42)
    assert np.array_equal(zr, zs)
