"""Pins the oracle (oracle/ozoracle.c) before it is trusted as the checker of the GPU path.

Each oracle function is checked against something independent of its own formula: the reference's
own sources compiled into oracle/_ref, closed forms evaluated with Python big integers and exact
rationals, brute force over tiny inputs, and invariants a slip (dropped term, wrong sign or index,
transposed operand) would break.  CPU only."""
import hashlib
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODULI = [251, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191, 181, 179, 173, 167, 163, 157, 151, 149]


def big(limbs):
    return sum(int(v) << (64 * i) for i, v in enumerate(limbs))


# ----------------------------------------------------------------------------- generator
def test_generator_matches_reference_fixture(orc):
    with open(os.path.join(GOLDEN, "generator_sha256.json")) as f:
        sha = json.load(f)
    for key, digest in sha.items():
        shape, phi, seed = key.split("_")
        rows, cols = map(int, shape.split("x"))
        X = orc.generate(rows, cols, float(phi[3:]), int(seed[4:]))
        assert hashlib.sha256(X.tobytes()).hexdigest() == digest, key


def test_generator_matches_reference_build(orc, ref):
    for rows, cols, phi, seed in [(17, 23, 1.0, 5), (200, 3, 3.0, 99), (1, 1, 0.25, 0)]:
        a, b = orc.generate(rows, cols, phi, seed), ref.generate(rows, cols, phi, seed)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_generator_distribution_sanity(orc):
    X = orc.generate(300, 300, 0.0, 3)  # phi = 0: (U - 0.5) with U in (0, 1]
    assert X.min() > -0.5 and X.max() <= 0.5
    assert abs(X.mean()) < 0.01 and abs(X.std() - math.sqrt(1 / 12)) < 0.01


# ----------------------------------------------------------------------------- CRT tables
def test_moduli_are_distinct_primes(orc):
    m = [int(v) for v in orc.moduli()]
    assert m == MODULI
    for p in m:
        assert all(p % d for d in range(2, int(p ** 0.5) + 1))


@pytest.mark.parametrize("S", range(1, 21))
def test_crt_tables_closed_forms(orc, S):
    M, Mh, part, inv = orc.crt_tables(S)
    prod = math.prod(MODULI[:S])
    assert big(M) == prod and big(Mh) == prod // 2
    for s in range(S):
        P = big(part[s])
        assert P * MODULI[s] == prod
        assert (P * int(inv[s])) % MODULI[s] == 1 and 0 < int(inv[s]) < MODULI[s]


def test_crt_tables_match_reference(orc, ref):
    for S in range(1, 21):
        for a, b in zip(orc.crt_tables(S), ref.crt_tables(S)):
            if a.ndim == 2:
                np.testing.assert_array_equal(a[:S], b[:S])
            elif a.size == 20:
                np.testing.assert_array_equal(a[:S], b[:S])
            else:
                np.testing.assert_array_equal(a, b)


def test_to_double_256_known_answers(orc):
    assert orc.to_double_256([0, 0, 0, 0]) == 0.0
    assert orc.to_double_256([2 ** 53 + 1, 0, 0, 0]) == float(2 ** 53)  # ties to even
    assert orc.to_double_256([2 ** 64 - 1, 0, 0, 0]) == 2.0 ** 64
    assert orc.to_double_256([0, 1, 0, 0]) == 2.0 ** 64
    assert orc.to_double_256([5, 0, 3, 0]) == 3.0 * 2.0 ** 128  # low limb below half an ulp
    # double rounding: limb 1 rounds up to 2^64 exactly, adding it to ~2^64 then rounds again
    v = orc.to_double_256([2 ** 64 - 1, 2 ** 53 - 1, 0, 0])
    assert v == (float(2 ** 64) + float(2 ** 53 - 1) * 2.0 ** 64)


# ----------------------------------------------------------------------------- host parameters
@pytest.mark.parametrize("S", range(2, 21))
def test_kparam_is_largest_k_with_exact_products(orc, S):
    """k = floor(0.5 log2((M/2 - 1)/K)) is the largest k with K * 4^k <= M/2 - 1 (the bound that
    keeps every dot product of k-bit integers inside the CRT range)."""
    prod = math.prod(MODULI[:S])
    for K in [1, 2, 3, 7, 64, 100, 255, 256, 1000, 1024, 4096, 5000, 8192, 16384, 65536, 131072]:
        k = orc.kparam(S, K)
        bound = Fraction(prod, 2) - 1
        assert K * Fraction(4) ** k <= bound < K * Fraction(4) ** (k + 1), (S, K, k)


def test_kparam_and_beta_match_reference(orc, ref):
    for S in range(1, 21):
        for K in [1, 5, 64, 256, 1024, 4096, 8192, 65536]:
            assert orc.kparam(S, K) == ref.lib.ref_scheme2_kparam(S, K)
    for K in list(range(1, 3000, 7)) + [2 ** 17, 2 ** 17 + 1, 2 ** 20]:
        assert orc.beta(K) == ref.lib.ref_scheme1_beta(K)


def test_beta_closed_form(orc):
    for K in list(range(1, 5000)) + [2 ** e for e in range(13, 30)] + [2 ** e + 1 for e in range(13, 30)]:
        b = max([1] + [bb for bb in range(1, 8) if K <= 2 ** (31 - 2 * bb)])
        assert orc.beta(K) == b, K


# ----------------------------------------------------------------------------- step 1
def special(rng, rows, cols):
    X = (rng.random((rows, cols)) - 0.5) * np.exp(3.0 * rng.standard_normal((rows, cols)))
    X[0, :] = 0.0
    X[1, 0] = 5e-310
    X[1, 1:] = 0.0
    X[2, 0] = 3e-320
    X[2, 1:] = -0.0
    X[3, 3 % cols] = 1e300
    return X


def test_shifts_brute_force(orc, rng):
    X = special(rng, 40, 50)
    for off in (0, 11, 54):
        want_r = [off - max([-1023] + [math.frexp(abs(v))[1] for v in row if abs(v) > 0]) for row in X]
        want_c = [off - max([-1023] + [math.frexp(abs(v))[1] for v in col if abs(v) > 0]) for col in X.T]
        np.testing.assert_array_equal(orc.row_shifts(X, off), want_r)
        np.testing.assert_array_equal(orc.col_shifts(X, off), want_c)


# ----------------------------------------------------------------------------- step 2, Scheme II
def exact_rint_ldexp(a, sh):
    """round-half-even(a * 2^sh) as a Python int, computed exactly."""
    q = Fraction(a) * (Fraction(2) ** sh)
    return round(q)  # Fraction.__round__ rounds half to even


@pytest.mark.parametrize("S,K", [(4, 256), (12, 300), (20, 50), (20, 1)])
def test_scheme2_residues_exact(orc, rng, S, K):
    A = special(rng, 12, K if K > 4 else 5)
    kp = orc.kparam(S, A.shape[1])
    sh = orc.row_shifts(A, kp)
    R = orc.slice_s2(A, sh, S, True)
    for i in range(A.shape[0]):
        for j in range(A.shape[1]):
            t = exact_rint_ldexp(float(A[i, j]), int(sh[i]))
            assert abs(t) <= 2 ** kp
            for s in range(S):
                r = int(R[s, i, j])
                assert (t - r) % MODULI[s] == 0 and abs(r) <= (MODULI[s] - 1) // 2, (i, j, s)


def test_scheme2_columns_use_column_shifts(orc, rng):
    B = special(rng, 30, 9)
    sh = orc.col_shifts(B, 20)
    Rc = orc.slice_s2(B, sh, 6, False)
    Rr = orc.slice_s2(np.ascontiguousarray(B.T), sh, 6, True)
    np.testing.assert_array_equal(Rc, Rr.transpose(0, 2, 1))


def test_symmetric_mod(orc):
    for m in MODULI:
        for a in [0, 1, -1, m, -m, (m - 1) // 2, (m + 1) // 2, -(m + 1) // 2, 2 ** 62 + 12345, -(2 ** 62) - 7]:
            r = orc.symmetric_mod(a, m)
            assert (a - r) % m == 0 and abs(r) <= (m - 1) // 2


# ----------------------------------------------------------------------------- step 2, Scheme I
@pytest.mark.parametrize("S", [1, 2, 5, 8, 10])
def test_scheme1_slices_invariants(orc, rng, S):
    A = special(rng, 10, 64)
    beta = orc.beta(64)
    sh = orc.row_shifts(A, 0)
    Sl = orc.slice_s1(A, sh, S, beta, True)
    assert np.abs(Sl.astype(int)).max() <= 2 ** (beta - 1)
    for i in range(A.shape[0]):
        E = -int(sh[i])
        if E < -1000 or E >= 970 + beta:
            continue  # all-zero / tiny rows; huge rows overflow sigma (see the next test)
        for j in range(A.shape[1]):
            acc = sum(Fraction(int(Sl[s, i, j])) * Fraction(2) ** (E + 1 - beta - beta * s) for s in range(S))
            # remainder after S slices is at most half the last slice unit
            assert abs(Fraction(float(A[i, j])) - acc) <= Fraction(2) ** (E - beta * S), (i, j)


def test_scheme1_sigma_overflow_quirk(orc):
    """Reference behaviour kept as is: when a row maximum reaches 2^(970+beta), sigma =
    0.75 * 2^(E+54-beta) overflows to inf (step2_slicing.hpp:63) and every slice of that row is 0."""
    A = np.array([[1e300, 1.0], [3.0, -2.0]])
    beta = orc.beta(2)
    sh = orc.row_shifts(A, 0)
    Sl = orc.slice_s1(A, sh, 4, beta, True)
    assert np.all(Sl[:, 0, :] == 0) and np.any(Sl[:, 1, :] != 0)


# ----------------------------------------------------------------------------- step 3
def test_igemm_matches_int64_matmul(orc, rng):
    A = rng.integers(-128, 128, size=(37, 300), dtype=np.int64).astype(np.int8)
    B = rng.integers(-128, 128, size=(300, 29), dtype=np.int64).astype(np.int8)
    C, bad = orc.igemm(A, B)
    assert bad == 0
    np.testing.assert_array_equal(C, A.astype(np.int64) @ B.astype(np.int64))


def test_s1_groups_are_antidiagonal_sums(orc, rng):
    S = 4
    As = rng.integers(-64, 65, size=(S, 5, 20), dtype=np.int64).astype(np.int8)
    Bs = rng.integers(-64, 65, size=(S, 20, 6), dtype=np.int64).astype(np.int8)
    T = orc.s1_groups(As, Bs)
    for g in range(2, S + 2):
        want = sum(As[s - 1].astype(np.int64) @ Bs[g - s - 1].astype(np.int64) for s in range(1, g))
        np.testing.assert_array_equal(T[g - 2], want)


# ----------------------------------------------------------------------------- step 4
@pytest.mark.parametrize("S", [3, 7, 8, 12, 20])
def test_crt_reconstructs_exact_integer_product(orc, rng, S):
    """With valid inputs the CRT value is the exact integer sum of t_a * t_b; below 2^64 the FP64
    result is that integer correctly rounded, scaled by 2^-(shift_A + shift_B)."""
    M, N, K = 6, 5, 40
    A = special(rng, M, K)
    B = np.ascontiguousarray(special(rng, N, K).T)
    kp = orc.kparam(S, K)
    shA, shB = orc.row_shifts(A, kp), orc.col_shifts(B, kp)
    C = orc.recon_s2(orc.s2_products(orc.slice_s2(A, shA, S, True), orc.slice_s2(B, shB, S, False)), shA, shB)
    for i in range(M):
        for j in range(N):
            X = sum(exact_rint_ldexp(float(A[i, k]), int(shA[i])) * exact_rint_ldexp(float(B[k, j]), int(shB[j]))
                    for k in range(K))
            assert abs(X) < math.prod(MODULI[:S]) // 2
            try:
                want = math.ldexp(float(X), -(int(shA[i]) + int(shB[j])))
            except OverflowError:  # C's ldexp saturates to +-inf
                want = math.copysign(math.inf, X)
            if abs(X) < 2 ** 64 or math.isinf(want):
                assert C[i, j] == want, (i, j)
            else:
                assert abs(C[i, j] - want) <= abs(want) * 2.0 ** -52


def test_crt_digits_definition(orc, rng):
    Ct = rng.integers(-(2 ** 31), 2 ** 31 - 1, size=(6, 4, 5), dtype=np.int64).astype(np.int32)
    Ct[0, 0, :3] = [-(2 ** 31), 2 ** 31 - 1, 0]
    D = orc.crt_digits(Ct)
    for s in range(6):  # Euclidean residue (Python's % on non-negative moduli)
        np.testing.assert_array_equal(D[s], Ct[s].astype(np.int64) % MODULI[s])


def test_crt_basis_reconstructs(orc):
    """x = sum_s (x mod m_s) Q_s mod M with Q_s = inv_s * M/m_s: the identity the gfx950
    reconstruction uses in place of the reference's sum of c_s * M/m_s."""
    for S in (1, 3, 8, 17, 20):
        M, Mh, part, inv = orc.crt_tables(S)
        Mi = sum(int(M[i]) << (64 * i) for i in range(4))
        Q = [int(inv[s]) * (Mi // MODULI[s]) for s in range(S)]
        assert all(q < Mi for q in Q)
        for x in [0, 1, Mi // 2, Mi - 1, (Mi * 7) // 13, 123456789 % Mi]:
            assert sum((x % MODULI[s]) * Q[s] for s in range(S)) % Mi == x


def test_recon_s1_exact_power_of_two_sum(orc, rng):
    """Scheme I reconstruction is sum_g T_g * 2^(e_i + e_j + 2 - 2beta) * 2^(2beta - beta g) in
    ascending g; with small T_g every partial sum is exact, so it equals the rational value."""
    S, M, N, beta = 5, 4, 3, 7
    T = rng.integers(-100, 100, size=(S, M, N), dtype=np.int64).astype(np.int32)
    shA = rng.integers(-5, 5, size=M).astype(np.int32)
    shB = rng.integers(-5, 5, size=N).astype(np.int32)
    C = orc.recon_s1(T, shA, shB, beta)
    for i in range(M):
        for j in range(N):
            e = -int(shA[i]) - int(shB[j]) + 2 - 2 * beta
            want = sum(Fraction(int(T[g - 2, i, j])) * Fraction(2) ** (e + 2 * beta - beta * g) for g in range(2, S + 2))
            assert Fraction(C[i, j]) == want


# ----------------------------------------------------------------------------- end-to-end oracles
def test_fp128_gemm_exact_rational(orc, rng):
    A = (rng.random((5, 30)) - 0.5) * np.exp(4 * rng.standard_normal((5, 30)))
    B = (rng.random((30, 4)) - 0.5) * np.exp(4 * rng.standard_normal((30, 4)))
    C = orc.gemm_fp128(A, B)
    for i in range(5):
        for j in range(4):
            exact = sum(Fraction(float(A[i, k])) * Fraction(float(B[k, j])) for k in range(30))
            assert abs(Fraction(C[i, j]) - exact) <= abs(exact) * Fraction(2) ** -52 + Fraction(2) ** -1074


U = Fraction(1, 2 ** 53)


def gamma(n):
    """gamma_n = n u / (1 - n u), the classical bound on n roundings (Higham, Accuracy and
    Stability of Numerical Algorithms, Lemma 3.1)."""
    return n * U / (1 - n * U)


def wide_pair(rng, M, K, N, phi):
    A = (rng.random((M, K)) - 0.5) * np.exp(phi * rng.standard_normal((M, K)))
    B = (rng.random((K, N)) - 0.5) * np.exp(phi * rng.standard_normal((K, N)))
    return A, B


def test_gemm_abs_exact_rational(orc, rng):
    """|A||B| (the denominator of every componentwise error) against the exact rational sum:
    K roundings, so within gamma_K of it (Higham Lemma 3.4 with nonnegative terms).  A dropped
    absolute value, a transposed operand or a skipped term breaks the bound."""
    for (M, K, N), phi in (((5, 30, 4), 4.0), ((3, 257, 6), 1.0), ((1, 1, 1), 2.0)):
        A, B = wide_pair(rng, M, K, N, phi)
        C = orc.gemm_abs(A, B)
        for i in range(M):
            for j in range(N):
                exact = sum(abs(Fraction(float(A[i, k]))) * abs(Fraction(float(B[k, j]))) for k in range(K))
                assert abs(Fraction(C[i, j]) - exact) <= gamma(K) * exact, (i, j)
    # the same products in the same order, so the transposed product is the same bits
    A, B = wide_pair(rng, 20, 50, 13, 2.0)
    np.testing.assert_array_equal(orc.gemm_abs(A, B), orc.gemm_abs(B.T.copy(), A.T.copy()).T)
    # and it bounds the signed product
    assert np.all(orc.gemm_abs(A, B) >= np.abs(orc.gemm_fp128(A, B)) * (1 - 2.0 ** -50))


def test_gemm_dd_exact_rational(orc, rng):
    """The double-double product (Ogita-Rump-Oishi Dot2 with an error-free TwoProduct and TwoSum,
    rounded once) against the exact sum: |C - x| <= u|x| + gamma_K^2 (|A||B|) (Ogita, Rump,
    Oishi, SIAM J. Sci. Comput. 26 (2005), Prop. 3.5 / 4.5; K^2 2^-106 order, not 2^-106)."""
    for (M, K, N), phi in (((4, 40, 3), 4.0), ((2, 600, 2), 1.0), ((3, 1, 5), 2.0)):
        A, B = wide_pair(rng, M, K, N, phi)
        C = orc.gemm_dd(A, B)
        for i in range(M):
            for j in range(N):
                prods = [Fraction(float(A[i, k])) * Fraction(float(B[k, j])) for k in range(K)]
                exact = sum(prods)
                absab = sum(abs(p) for p in prods)
                assert abs(Fraction(C[i, j]) - exact) <= U * abs(exact) + gamma(K) ** 2 * absab, (i, j)


def test_fp128_and_dd_match_reference(orc, ref, rng):
    A = (rng.random((40, 70)) - 0.5) * np.exp(2 * rng.standard_normal((40, 70)))
    B = (rng.random((70, 30)) - 0.5) * np.exp(2 * rng.standard_normal((70, 30)))
    c128 = orc.gemm_fp128(A, B)
    np.testing.assert_array_equal(c128, ref.gemm_fp128(A, B))
    dd = orc.gemm_dd(A, B)
    assert np.max(np.abs(dd - c128) / orc.gemm_abs(A, B)) <= 2.0 ** -52


def test_error_metrics_match_reference(orc, ref, rng):
    e = rng.standard_normal(1000)
    a = e + rng.standard_normal(1000) * 1e-10
    e[0] = 0.0
    np.testing.assert_array_equal(orc.errors(e, a), ref.errors(e, a))


def test_golden_tiny_cases(orc):
    g = np.load(os.path.join(GOLDEN, "golden.npz"), allow_pickle=False)
    for name, S in (("s2_9x7x11", 5), ("s2_16x16x16_s9", 9)):
        A, B = g[f"{name}_A"], g[f"{name}_B"]
        got = orc.s2_gemm(A, B, S)
        np.testing.assert_array_equal(got.view(np.uint64), g[f"{name}_C"].view(np.uint64))
        np.testing.assert_array_equal(orc.slice_s2(A, g[f"{name}_shA"], S, True), g[f"{name}_As"])
    A, B = g["s1_9x7x11_A"], g["s1_9x7x11_B"]
    np.testing.assert_array_equal(orc.s1_gemm(A, B, 4).view(np.uint64), g["s1_9x7x11_C"].view(np.uint64))


def test_c1_config_accuracy(orc):
    """M=N=K=256, S=4, phi=0.5, seeds 42/43: both schemes within their truncation bound of the FP128
    product (componentwise, relative to |A||B|)."""
    A, B = orc.generate(256, 256, 0.5, 42), orc.generate(256, 256, 0.5, 43)
    exact, absab = orc.gemm_fp128(A, B), orc.gemm_abs(A, B)
    for f in (orc.s1_gemm, orc.s2_gemm):
        C = f(A, B, 4)
        assert np.max(np.abs(C - exact) / absab) < 2.0 ** -8
    # more moduli -> smaller error, down to FP64 level
    errs = [np.max(np.abs(orc.s2_gemm(A, B, S) - exact) / absab) for S in (4, 8, 12, 16)]
    assert errs[0] > errs[1] > errs[2] and errs[3] < 1e-15


def test_crt_digit_quotient_off_by_one_only_on_multiples():
    """The GEMM epilogue's digit c = v - floor(v * fl(1/m)) * m (double arithmetic, v an INT32
    product) is corrected only by c == m -> 0: the floor is the exact quotient unless m divides v,
    and then it is the quotient or one less.  Checked here in IEEE double for every multiple of
    each modulus up to |v| < 2^28 and its neighbours."""
    mods = [251, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191, 181, 179, 173, 167, 163, 157, 151, 149]
    for m in mods:
        rcp = np.float64(1.0) / np.float64(m)
        k = np.arange(-(2**28) // m, (2**28) // m + 1, dtype=np.int64)
        for r in (0, 1, m - 1):
            v = k * m + r
            q = np.floor(v.astype(np.float64) * rcp).astype(np.int64)
            c = v - q * m
            if r == 0:
                assert np.all((c == 0) | (c == m)), m
            else:
                assert np.all(c == r), m
