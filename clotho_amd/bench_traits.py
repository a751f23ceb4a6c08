"""Trait contraction at the precision-configuration shape: 200,000 haplotypes
x 65,536 sites x 16 traits, genotype bits Bernoulli(f_j) with f_j ~ U(0,1)
(seed 42, drawn on the device), effects from a multivariate normal with unit
variances and pairwise correlation 0.3 (seed 43).

Reports kernel time of the plain FP32 set-bit walk, the FP64 walk, the
pipeline's exact int8 path (effects quantised to k 2^-q, int8 digit planes
on int8 MFMA, exact integer sums) and the split-BF16 MFMA contraction (1, 2,
3 BF16 terms, FP32 accumulation per K chunk, FP64 chunk reduction), the MFMA/plain speedup, and per-trait error
against an FP64 reference: normwise max_i |z_i - z*_i| / max_i |z*_i| and the
median elementwise relative error.  "plain" is the dense bit x effect FP32
kernel (lane = haplotype); "walk" is the sparse set-bit walk.  The FP64 reference is the GPU FP64 walk
over all rows (FP64 sums; the GPU tests pin it against numpy float64 on rows
rebuilt from the same generator).  One JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def effects_mvn(sites, traits, rho=0.3, seed=43):
    cov = np.full((traits, traits), rho) + (1 - rho) * np.eye(traits)
    L = np.linalg.cholesky(cov)
    z = np.random.default_rng(seed).standard_normal((sites, traits))
    return z @ L.T


def errors(z, ref):
    norm = np.max(np.abs(z - ref), axis=0) / np.max(np.abs(ref), axis=0)
    rel = np.abs(z - ref) / np.maximum(np.abs(ref), 1e-300)
    return norm, np.median(rel, axis=0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--haplotypes", type=int, default=200_000)
    ap.add_argument("--sites", type=int, default=65_536)
    ap.add_argument("--traits", type=int, default=16)
    ap.add_argument("--repeats", type=int, default=10)
    ap.add_argument("--modes", default="fp64,plain,walk,i8,bf16",
                    help="comma-separated subset to run (fp64 is the error reference)")
    args = ap.parse_args()
    modes = set(args.modes.split(","))

    from clotho_amd import api

    h, s, t = args.haplotypes, args.sites, args.traits
    freqs = np.random.default_rng(42).uniform(0, 1, s)
    eff = effects_mvn(s, t)
    out = {"workload": f"trait contraction H={h} S={s} T={t}, f~U(0,1), MVN effects rho=0.3"}

    if "fp64" not in modes:  # timing only: no error reference
        for name, mode, terms in (("exact_i8", api.TRAITS_EXACT_I8, 3),
                                  ("mfma_bf16x3", api.TRAITS_MFMA_BF16, 3)):
            if name.split("_")[-1].startswith("i8") and "i8" not in modes:
                continue
            if name.startswith("mfma") and "bf16" not in modes:
                continue
            _, ms = api.trait_contract_synthetic(h, freqs, eff, mode, seed=42, split_terms=terms,
                                                 repeats=args.repeats)
            out[name + "_ms"] = ms
        print(json.dumps(out), flush=True)
        return
    z64, ms64 = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_EXACT_FP64, seed=42,
                                             repeats=2)
    out["fp64_walk_ms"] = ms64

    z32, ms32 = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_PLAIN_FP32, seed=42,
                                             repeats=max(1, args.repeats // 5))
    n32, m32 = errors(z32, z64)
    out["plain_fp32_ms"] = ms32
    out["plain_fp32_err_normwise_max"] = float(n32.max())
    zw, msw = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_WALK_FP32, seed=42,
                                           repeats=1)
    out["walk_fp32_ms"] = msw
    out["walk_fp32_err_normwise_max"] = float(errors(zw, z64)[0].max())
    zi, msi = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_EXACT_I8, seed=42,
                                           repeats=args.repeats)
    ni, mi = errors(zi, z64)
    out["exact_i8_ms"] = msi
    out["exact_i8_speedup_vs_plain"] = ms32 / msi
    out["exact_i8_err_normwise_max"] = float(ni.max())
    out["exact_i8_err_median_rel_max"] = float(mi.max())
    # four int8 digit planes per trait: 2 H S (4 T) integer operations executed
    out["exact_i8_exec_tops"] = 2.0 * h * s * 4 * t / (msi * 1e-3) / 1e12
    out["exact_i8_row_gbps"] = h * (((s + 511) // 512) * 64) / (msi * 1e-3) / 1e9
    for terms in (1, 2, 3):
        zm, msm = api.trait_contract_synthetic(h, freqs, eff, api.TRAITS_MFMA_BF16, seed=42,
                                               split_terms=terms, repeats=args.repeats)
        nm, mm = errors(zm, z64)
        key = f"mfma_bf16x{terms}"
        out[key + "_ms"] = msm
        out[key + "_speedup_vs_plain"] = ms32 / msm
        out[key + "_err_normwise_per_trait"] = [float(x) for x in nm]
        out[key + "_err_median_rel_per_trait"] = [float(x) for x in mm]
        out[key + "_err_normwise_max"] = float(nm.max())
        flops = 2.0 * h * s * t
        out[key + "_alg_tflops"] = flops / (msm * 1e-3) / 1e12
        out[key + "_exec_tflops"] = terms * flops / (msm * 1e-3) / 1e12
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
