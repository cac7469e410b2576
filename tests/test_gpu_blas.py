"""BLAS-style front end (oza_gemm): layouts, transposes, leading dimensions, alpha / beta.

With alpha = 1, beta = 0 every layout/transpose combination must give the bits of the plain
row-major pipeline on dense copies of op(A), op(B) (which is itself pinned to the oracle); with
general alpha / beta each entry must be fma(alpha, X, beta * C0) rounded once, computed here
exactly with rational arithmetic.  All calls go through the C ABI via ozablas_amd._lib."""
from fractions import Fraction

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oz():
    import ozablas_amd as oz

    return oz


def stored(X, layout, ld):
    """Flat buffer holding matrix X (r x c) in the given layout with leading dimension ld; the
    padding between rows/columns holds NaN, which must never be read."""
    r, c = X.shape
    if layout == 0:
        buf = np.full((r, ld), np.nan)
        buf[:, :c] = X
    else:
        buf = np.full((c, ld), np.nan)
        buf[:, :r] = X.T
    return buf.ravel()


def read_back(buf, layout, r, c, ld):
    if layout == 0:
        return buf.reshape(r, ld)[:, :c]
    return buf.reshape(c, ld)[:, :r].T


def plain(ex, scheme, S, A, B):
    oz = _oz()
    M, K = A.shape
    N = B.shape[1]
    ws = (oz.WorkspaceScheme2 if scheme == 2 else oz.WorkspaceScheme1)(ex, M, N, K, S)
    dA, dB, dC = ex.to_device(A), ex.to_device(B), ex.alloc(8 * M * N)
    (oz.ozaki_scheme2_gemm if scheme == 2 else oz.ozaki_scheme1_gemm)(ws, dA, dB, dC)
    ex.synchronize()
    out = dC.download((M, N), np.float64)
    ws.free()
    return out


def call(ex, scheme, S, layout, ta, tb, A_log, B_log, alpha, beta, C0, pad=3, workspace=None):
    """oza_gemm on logical op(A) = A_log (M x K), op(B) = B_log (K x N); returns C (M x N)."""
    oz = _oz()
    M, K = A_log.shape
    N = B_log.shape[1]
    A_st = A_log.T if ta else A_log  # what is stored
    B_st = B_log.T if tb else B_log
    lda = (A_st.shape[1] if layout == 0 else A_st.shape[0]) + pad
    ldb = (B_st.shape[1] if layout == 0 else B_st.shape[0]) + pad
    ldc = (N if layout == 0 else M) + pad
    dA = ex.to_device(stored(A_st, layout, lda))
    dB = ex.to_device(stored(B_st, layout, ldb))
    cbuf = stored(C0, layout, ldc)
    dC = ex.to_device(cbuf)
    oz.gemm(ex, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, transA=ta, transB=tb, layout=layout,
            scheme=scheme, slices=S, workspace=workspace)
    ex.synchronize()
    out = dC.download(cbuf.shape, np.float64)
    inside = np.zeros(cbuf.shape, bool)
    read_back(inside, layout, M, N, ldc)[...] = True
    assert np.isnan(out[~inside]).all()  # the padding of C is never written
    return read_back(out, layout, M, N, ldc)


def same_bits(x, y):
    return np.array_equal(np.ascontiguousarray(x).view(np.uint64), np.ascontiguousarray(y).view(np.uint64))


@pytest.mark.parametrize("scheme,S", [(2, 8), (1, 4)])
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_layouts_and_transposes_match_plain(ex, orc, scheme, S, layout, ta, tb):
    M, N, K = 37, 29, 70
    A = orc.generate(M, K, 1.0, 11)
    B = orc.generate(K, N, 1.0, 12)
    got = call(ex, scheme, S, layout, ta, tb, A, B, 1.0, 0.0, np.full((M, N), np.nan))
    ref = plain(ex, scheme, S, A, B)
    assert same_bits(got, ref)
    if scheme == 2:
        assert same_bits(got, orc.s2_gemm(A, B, S))
    else:
        assert same_bits(got, orc.s1_gemm(A, B, S))


def exact_fma(a, x, t):
    return float(Fraction(a) * Fraction(float(x)) + Fraction(float(t)))


@pytest.mark.parametrize("alpha,beta", [(0.75, -1.5), (-3.0, 1.0), (1.0, 0.5), (2.5, 0.0)])
def test_alpha_beta(ex, orc, rng, alpha, beta):
    M, N, K = 21, 18, 40
    A = orc.generate(M, K, 1.0, 21)
    B = orc.generate(K, N, 1.0, 22)
    C0 = rng.standard_normal((M, N))
    X = plain(ex, 2, 10, A, B)
    for layout in (0, 1):
        got = call(ex, 2, 10, layout, 0, 1, A, B, alpha, beta, C0)
        if beta == 0.0:
            exp = alpha * X
        else:
            t = beta * C0
            exp = np.array([[exact_fma(alpha, X[i, j], t[i, j]) for j in range(N)] for i in range(M)])
        assert same_bits(got, exp), (alpha, beta, layout)


def test_beta_zero_does_not_read_c_and_alpha_zero_scales(ex, orc):
    M, N, K = 16, 24, 33
    A = orc.generate(M, K, 1.0, 31)
    B = orc.generate(K, N, 1.0, 32)
    got = call(ex, 2, 6, 0, 0, 0, A, B, 1.0, 0.0, np.full((M, N), np.nan))
    assert not np.isnan(got).any()
    C0 = np.arange(M * N, dtype=np.float64).reshape(M, N)
    got = call(ex, 2, 6, 0, 0, 0, A, B, 0.0, -2.0, C0)
    assert same_bits(got, -2.0 * C0)
    got = call(ex, 2, 6, 1, 1, 0, A, B, 0.0, 0.0, np.full((M, N), np.nan))
    assert (got == 0.0).all()


def test_k_zero_and_explicit_workspace(ex, orc):
    oz = _oz()
    M, N = 9, 7
    C0 = np.ones((M, N))
    got = call(ex, 2, 4, 0, 0, 0, np.zeros((M, 0)), np.zeros((0, N)), 1.0, 3.0, C0)
    assert (got == 3.0).all()
    A = orc.generate(M, 50, 1.0, 41)
    B = orc.generate(50, N, 1.0, 42)
    wsb = oz.gemm_workspace_bytes(2, M, N, 50, 12)
    ws = ex.alloc(wsb)
    got = call(ex, 2, 12, 0, 0, 0, A, B, 1.0, 0.0, np.zeros((M, N)), workspace=ws)
    assert same_bits(got, plain(ex, 2, 12, A, B))


def test_invalid_leading_dimensions(ex):
    oz = _oz()
    buf = ex.alloc(8 * 1024)
    with pytest.raises(ValueError, match="lda"):
        oz.gemm(ex, 8, 8, 8, 1.0, buf, 7, buf, 8, 0.0, buf, 8)
    with pytest.raises(ValueError, match="ldb"):
        oz.gemm(ex, 8, 8, 8, 1.0, buf, 8, buf, 7, 0.0, buf, 8)
    with pytest.raises(ValueError, match="ldc"):
        oz.gemm(ex, 8, 9, 8, 1.0, buf, 8, buf, 9, 0.0, buf, 8)
    with pytest.raises(ValueError, match="ldb"):  # column-major: B is K x N, ldb >= K
        oz.gemm(ex, 8, 8, 9, 1.0, buf, 8, buf, 8, 0.0, buf, 8, layout=oz.COL_MAJOR)
    with pytest.raises(ValueError):
        oz.gemm(ex, 8, 8, 8, 1.0, buf, 8, buf, 8, 0.0, buf, 8, transA=2)
