"""Measures how far Scheme II's componentwise accuracy falls short of its truncation bound, per
exponent spread phi of the reference generator (utils/matrix_gen.hpp:15-34).

The bound gives K_param(S, K) bits relative to K 2^(E_i + F_j) (E_i, F_j: step 1's row and column
exponents); what a result delivers is -log10 max |C - C_exact| / (|A||B|)_ij.  The difference, in
decimal digits, grows with the spread of the inputs' exponents, since |A||B| falls below
K 2^(E_i + F_j) when few products reach the row/column maxima.  The library's digits -> S rule
(scheme2_slices_for_accuracy, ozablas_amd/csrc/crt_host.cpp) adds, per phi, the worst shortfall
printed here plus a margin of 0.8-1.05 digits for problems larger than these (the maximum over
more entries reaches further into the tail).

Everything runs on the CPU oracle against its FP128 product: S = 10, three seeds per shape.
Run from the repository root after building the oracle: python3 tests/golden/scheme2_shortfall.py"""
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
from oracle_lib import Oracle  # noqa: E402

S = 10
SHAPES = ((1024, 256), (256, 64), (256, 4096))  # (M = N, K)


def shortfall(orc, phi, MN, K, seeds=3):
    worst = math.inf
    for seed in range(seeds):
        A, B = orc.generate(MN, K, phi, 42 + seed), orc.generate(K, MN, phi, 43 + seed)
        exact, absab = orc.gemm_fp128(A, B), orc.gemm_abs(A, B)
        C = orc.s2_gemm(A, B, S)
        cw = np.max(np.abs(C - exact) / np.where(absab > 0, absab, 1.0))
        worst = min(worst, -math.log10(cw))
    return orc.kparam(S, K) * math.log10(2) - worst


def main():
    orc = Oracle()
    for phi in (0.0, 0.5, 1.0, 1.5, 2.0, 3.0, 4.0):
        vals = [shortfall(orc, phi, MN, K) for MN, K in SHAPES]
        cells = "  ".join(f"M=N={MN} K={K}: {v:5.2f}" for (MN, K), v in zip(SHAPES, vals))
        print(f"phi {phi:3.1f}  worst {max(vals):5.2f}  ({cells})", flush=True)


if __name__ == "__main__":
    main()
