"""Writes the committed golden fixtures (tests/golden/golden.npz and generator_sha256.json).

Inputs come from the reference's own generator (utils/matrix_gen.hpp) compiled from the reference
sources into oracle/_ref/libozref.so by `make -C oracle ref`; stage outputs of the tiny cases are
the oracle's, which the GPU test suite checks bit for bit against the reference's own stage
kernels.  Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import Oracle, RefBuild  # noqa: E402

GEN_CASES = [(256, 256, 0.5, 42), (256, 256, 0.5, 43), (64, 64, 4.0, 7), (100, 37, 0.0, 1), (33, 65, 2.0, 1234)]


def main():
    ref = RefBuild()
    orc = Oracle()
    sha = {}
    for rows, cols, phi, seed in GEN_CASES:
        X = ref.generate(rows, cols, phi, seed)
        sha[f"{rows}x{cols}_phi{phi}_seed{seed}"] = hashlib.sha256(X.tobytes()).hexdigest()
    with open(os.path.join(HERE, "generator_sha256.json"), "w") as f:
        json.dump(sha, f, indent=1, sort_keys=True)

    out = {}
    # Tiny end-to-end cases: inputs from the reference generator, every stage's output.
    for name, (M, N, K, phi, S) in {"s2_9x7x11": (9, 7, 11, 1.0, 5), "s2_16x16x16_s9": (16, 16, 16, 2.0, 9),
                                    "s1_9x7x11": (9, 7, 11, 1.0, 4)}.items():
        A = ref.generate(M, K, phi, 42)
        B = ref.generate(K, N, phi, 43)
        out[f"{name}_A"], out[f"{name}_B"] = A, B
        if name.startswith("s2"):
            kp = orc.kparam(S, K)
            shA, shB = orc.row_shifts(A, kp), orc.col_shifts(B, kp)
            As, Bs = orc.slice_s2(A, shA, S, True), orc.slice_s2(B, shB, S, False)
            Ct = orc.s2_products(As, Bs)
            C = orc.recon_s2(Ct, shA, shB)
            assert np.array_equal(C.view(np.uint64), orc.s2_gemm(A, B, S).view(np.uint64))
        else:
            beta = orc.beta(K)
            shA, shB = orc.row_shifts(A, 0), orc.col_shifts(B, 0)
            As, Bs = orc.slice_s1(A, shA, S, beta, True), orc.slice_s1(B, shB, S, beta, False)
            Ct = orc.s1_groups(As, Bs)
            C = orc.recon_s1(Ct, shA, shB, beta)
            assert np.array_equal(C.view(np.uint64), orc.s1_gemm(A, B, S).view(np.uint64))
        out.update({f"{name}_shA": shA, f"{name}_shB": shB, f"{name}_As": As, f"{name}_Bs": Bs,
                    f"{name}_Ct": Ct, f"{name}_C": C, f"{name}_fp128": ref.gemm_fp128(A, B)})
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", sorted(out)[:4], "...", len(out), "arrays")


if __name__ == "__main__":
    main()
