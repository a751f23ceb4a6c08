"""clotho_trait_contract_exact_noalloc: the exact contraction over
caller-owned device buffers with its digit planes in the library's fixed
device array.  Checked bit for bit against clotho_trait_contract's exact
mode (which allocates its own workspace) and numpy's exact sums, on two calls
in a row, the second smaller (the array keeps the first call's planes), and
on a non-default stream."""
import ctypes as C

import numpy as np
import pytest

from clotho_amd import api

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


def _device(hip, arr):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), max(arr.nbytes, 8)) == 0
    assert hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
    return p


def _case(rng, h, sites, traits):
    words = ((sites + 511) // 512) * 8
    bits = rng.uniform(0, 1, (h, words * 64)) < 0.3
    bits[:, sites:] = False
    rows = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64)
    k = rng.integers(-(2**30 - 1), 2**30, (sites, traits)).astype(np.int32)
    k[0, 0] = 2**30 - 1  # the allocating entry point's quantum is then 2^-24
    return np.ascontiguousarray(rows), k, words


def _noalloc(hip, rows, k, words, stream):
    h, t = rows.shape[0], k.shape[1]
    d_rows, d_k = _device(hip, rows), _device(hip, k)
    out = np.zeros((h, t), np.int64)
    d_out = _device(hip, out)
    rc = api.lib().clotho_trait_contract_exact_noalloc(d_rows, h, words, d_k, k.shape[0], t,
                                                       d_out, stream)
    assert rc == 0
    assert hip.hipStreamSynchronize(stream) == 0
    assert hip.hipMemcpy(out.ctypes.data, d_out, out.nbytes, D2H) == 0
    for p in (d_rows, d_k, d_out):
        hip.hipFree(p)
    return out


def _allocating(rows, k):
    eff = k.astype(np.float64) * 2.0**-24
    z, _ = api.trait_contract(rows, eff, api.TRAITS_EXACT_I8)
    return z * 2.0**24


def test_noalloc_matches_allocating_entry_point():
    hip = _hip()
    rng = np.random.default_rng(77)
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    for h, sites, traits, st in ((3000, 70_000, 16, None), (500, 1000, 3, stream)):
        print(f"case H={h} sites={sites} traits={traits}", flush=True)
        rows, k, words = _case(rng, h, sites, traits)
        got = _noalloc(hip, rows, k, words, st)
        want = _allocating(rows, k)
        assert np.array_equal(got.astype(np.float64), want)
        kf = k.astype(np.float64)
        for r0 in range(0, h, 250):  # integer sums below 2^53: exact in float64
            bits = np.unpackbits(rows[r0:r0 + 250].view(np.uint8), axis=1,
                                 bitorder="little")[:, :sites]
            exact = bits.astype(np.float64) @ kf
            assert np.array_equal(got[r0:r0 + 250].astype(np.float64), exact)
    hip.hipStreamDestroy(stream)


def test_noalloc_rejects_sizes_beyond_its_array():
    L = api.lib()
    dummy = C.c_void_p(8)  # never dereferenced: the size checks come first
    assert L.clotho_trait_contract_exact_noalloc(dummy, 4, 2056, dummy, 1000, 3, dummy,
                                                 None) == api_err("capacity")
    assert L.clotho_trait_contract_exact_noalloc(dummy, 4, 16, dummy, 1000, 17, dummy,
                                                 None) == api_err("capacity")
    assert L.clotho_trait_contract_exact_noalloc(dummy, 4, 12, dummy, 700, 3, dummy,
                                                 None) == api_err("arg")


def api_err(kind):
    return {"arg": -1, "capacity": -3}[kind]
