"""The accuracy machinery behind every reported error and digit count: the GPU double-double GEMM
(oza_dd_gemm, the full-matrix accuracy reference of the benchmarks) and the error statistics
(oza_error_stats), against the oracle (itself pinned to exact rationals and the reference build
in test_oracle.py) and against numpy definitions.  All calls go through the C ABI."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def gamma(n):
    return n * U / (1 - n * U)


def dd_gemm(ex, A, B):
    from ozablas_amd import _lib as L

    M, K = A.shape
    N = B.shape[1]
    dA, dB = ex.to_device(np.ascontiguousarray(A)), ex.to_device(np.ascontiguousarray(B))
    hi, ab = ex.alloc(8 * M * N), ex.alloc(8 * M * N)
    L.check(L.lib.oza_dd_gemm(ex.ctx, M, N, K, dA.ptr, dB.ptr, hi.ptr, ab.ptr))
    ex.synchronize()
    return hi.download((M, N), np.float64), ab.download((M, N), np.float64)


def error_stats(ex, exact, actual, absab=None):
    from ozablas_amd import _lib as L

    de, da = ex.to_device(np.ascontiguousarray(exact)), ex.to_device(np.ascontiguousarray(actual))
    db = ex.to_device(np.ascontiguousarray(absab)) if absab is not None else None
    out = (C.c_double * 4)()
    L.check(L.lib.oza_error_stats(ex.ctx, exact.size, de.ptr, da.ptr, db.ptr if db else None, out))
    return np.array(out[:])


def wide_pair(rng, M, K, N, phi):
    A = (rng.random((M, K)) - 0.5) * np.exp(phi * rng.standard_normal((M, K)))
    B = (rng.random((K, N)) - 0.5) * np.exp(phi * rng.standard_normal((K, N)))
    return A, B


@pytest.mark.parametrize("M,K,N,phi", [(1, 1, 1, 1.0), (64, 100, 48, 1.0), (33, 300, 17, 4.0), (128, 1000, 96, 2.0),
                                       (20, 4097, 9, 0.5)])
def test_dd_gemm_matches_oracle(ex, orc, rng, M, K, N, phi):
    """Same Dot2 in the same k order, both TwoProducts exact in this range: the same values.
    |A||B| is accumulated with fma on the GPU and with separate roundings in the oracle, so the two
    agree within the K-rounding bound of each."""
    A, B = wide_pair(rng, M, K, N, phi)
    hi, ab = dd_gemm(ex, A, B)
    np.testing.assert_array_equal(hi, orc.gemm_dd(A, B))
    want = orc.gemm_abs(A, B)
    assert np.all(np.abs(ab - want) <= 2 * gamma(K) * want)
    # the signed product is inside its bound of the FP128 product
    c128 = orc.gemm_fp128(A, B)
    assert np.all(np.abs(hi - c128) <= 2 * U * np.abs(c128) + 2 * gamma(K) ** 2 * want + 2.0 ** -1070)


def test_dd_gemm_special_values(ex, orc, rng):
    """Zero rows and columns, subnormals and huge entries: within the Dot2 bound of FP128 (the
    TwoProducts may differ from the oracle's Dekker split where products underflow)."""
    from test_gpu_parity import special_matrix

    A = special_matrix(rng, 40, 70)
    B = special_matrix(rng, 70, 30)
    hi, ab = dd_gemm(ex, A, B)
    c128, want = orc.gemm_fp128(A, B), orc.gemm_abs(A, B)
    assert np.all(np.abs(ab - want) <= 2 * gamma(70) * want)
    assert np.all(np.abs(hi - c128) <= 2 * U * np.abs(c128) + 2 * gamma(70) ** 2 * want + 2.0 ** -1060)
    assert np.all(hi[1, :] == 0) and np.all(hi[:, 2] == 0) and np.all(ab[1, :] == 0)


@pytest.mark.parametrize("n", [1, 1000, 3_000_000])
def test_error_stats_match_oracle_and_numpy(ex, orc, rng, n):
    """max |d| and max relative (the reference's metric) equal the oracle's restatement of
    utils/matrix_compare.hpp; the componentwise maximum and the bit-difference count equal numpy's."""
    e = rng.standard_normal(n) * np.exp(3 * rng.standard_normal(n))
    a = e * (1 + rng.standard_normal(n) * 1e-12)
    keep = rng.random(n) < 0.3
    a[keep] = e[keep]  # identical entries
    if n > 5:
        e[0] = 0.0
        a[0] = 1e-300  # relative error where exact == 0 is |d|
        e[1] = a[1] = 0.0
        e[2], a[2] = 0.0, -0.0  # equal values, different bits
    absab = np.abs(e) * (1 + rng.random(n))
    if n > 5:
        absab[3] = 0.0  # componentwise where |A||B| == 0 is |d|
    got = error_stats(ex, e, a, absab)
    want = orc.errors(e, a)
    assert got[0] == want[0] and got[1] == want[1]
    d = np.abs(a - e)
    cw = np.where(absab > 0, d / np.where(absab > 0, absab, 1.0), d)
    assert got[2] == np.max(cw)
    assert got[3] == np.count_nonzero(e.view(np.uint64) != a.view(np.uint64))
    # without |A||B| the componentwise slot stays zero
    assert error_stats(ex, e, a)[2] == 0.0


def test_mfma_ceiling_probe(ex):
    """The power-ceiling probe reports a positive rate and a plausible shader clock, and no more
    than the nominal peak at its own clock (4 SIMDs x 65536 ops per 32 cycles per CU)."""
    from ozablas_amd import _lib as L

    cus = 256  # MI355X
    for zero in (0, 1):
        out = (C.c_double * 2)()
        L.check(L.lib.oza_mfma_ceiling(ex.ctx, zero, out))
        tops, ghz = out[0], out[1]
        assert 0.3 < ghz < 4.0, ghz
        assert 0 < tops <= cus * 4 * 65536 / 32 * ghz * 1e9 / 1e12 * 1.05, (tops, ghz)


@pytest.mark.parametrize("phi", [0.0, 1.0, 2.0, 4.0])
def test_dd_reference_stands_in_for_fp128(ex, orc, phi):
    """The precision protocol measures errors against the GPU double-double product instead of the
    reference's FP128 one: the max relative error (utils/matrix_compare.hpp:47) and the componentwise
    error of an emulated product come out the same against either, to 1 %."""
    from test_gpu_parity import gen_pair, run_pipeline

    M = N = 128
    K = 300
    A, B = gen_pair(orc, M, N, K, phi, 42)
    hi, ab = dd_gemm(ex, A, B)
    c128 = orc.gemm_fp128(A, B)
    for S in (5, 9, 13):  # emulation errors well above FP64 rounding, where both references agree
        C, ws = run_pipeline(ex, 2, A, B, S)
        ws.free()
        r_dd, r_128 = orc.errors(hi, C)[1], orc.errors(c128, C)[1]
        assert abs(r_dd - r_128) <= 0.01 * r_128, (S, r_dd, r_128)
        cw_dd = np.max(np.abs(C - hi) / ab)
        cw_128 = np.max(np.abs(C - c128) / orc.gemm_abs(A, B))
        assert abs(cw_dd - cw_128) <= 0.01 * cw_128, (S, cw_dd, cw_128)


@pytest.mark.parametrize("phi", [0.5, 1.0, 2.0, 4.0])
def test_slices_for_accuracy_on_the_gpu_pipeline(ex, orc, phi):
    """oza_scheme2_slices_for_accuracy's moduli count, run through the GPU pipeline, reaches the
    requested componentwise digits against the FP128 product (fresh seeds)."""
    import ozablas_amd as oz
    from test_gpu_parity import run_pipeline

    for K, d in [(96, 8), (700, 13), (3000, 10)]:
        try:
            S = oz.scheme2_slices_for_accuracy(K, d, phi)
        except ValueError:
            continue  # beyond 20 moduli
        A, B = orc.generate(192, K, phi, 3000 + K), orc.generate(K, 160, phi, 4000 + K)
        C, ws = run_pipeline(ex, 2, A, B, S)
        ws.free()
        exact, absab = orc.gemm_fp128(A, B), orc.gemm_abs(A, B)
        cw = np.max(np.abs(C - exact) / np.where(absab > 0, absab, 1.0))
        assert cw <= 10.0 ** -d, (phi, K, d, S, cw)
