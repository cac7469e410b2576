"""oza_gemm_noalloc: the BLAS-style front end with its workspace in the library's device scratch
array instead of an allocation.  Its results must carry the bits of oza_gemm with a caller
workspace, call after call on the shared array (a larger call, then a smaller one that finds the
larger call's planes still there), and a problem larger than the array must be rejected before
anything runs.  All calls go through the C ABI."""
import numpy as np
import pytest

from test_gpu_blas import read_back, stored

pytestmark = pytest.mark.gpu


def _run(ex, entry, scheme, S, layout, ta, tb, A_log, B_log, alpha, beta, C0, pad=3):
    """One call of oza_gemm (entry='ws', caller workspace) or oza_gemm_noalloc (entry='noalloc')
    on logical op(A) = A_log (M x K), op(B) = B_log (K x N); returns C (M x N)."""
    from ozablas_amd import _lib as L

    M, K = A_log.shape
    N = B_log.shape[1]
    A_st = A_log.T if ta else A_log
    B_st = B_log.T if tb else B_log
    lda = (A_st.shape[1] if layout == 0 else A_st.shape[0]) + pad
    ldb = (B_st.shape[1] if layout == 0 else B_st.shape[0]) + pad
    ldc = (N if layout == 0 else M) + pad
    dA = ex.to_device(stored(A_st, layout, lda))
    dB = ex.to_device(stored(B_st, layout, ldb))
    cbuf = stored(C0, layout, ldc)
    dC = ex.to_device(cbuf)
    if entry == "ws":
        ws = ex.alloc(L.gemm_workspace_bytes(scheme, M, N, K, S))
        L.check(L.lib.oza_gemm(ex.ctx, scheme, S, layout, ta, tb, M, N, K, alpha, dA.ptr, lda, dB.ptr, ldb, beta,
                               dC.ptr, ldc, ws.ptr, ws.nbytes))
    else:
        L.check(L.lib.oza_gemm_noalloc(ex.ctx, scheme, S, layout, ta, tb, M, N, K, alpha, dA.ptr, lda, dB.ptr, ldb,
                                       beta, dC.ptr, ldc))
    ex.synchronize()
    return read_back(dC.download(cbuf.shape, np.float64), layout, M, N, ldc)


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


# (scheme, S, layout, transA, transB, M, N, K, alpha, beta): a larger call, then a smaller one
CASES = [
    ((2, 12, 0, 0, 0, 700, 520, 300, 1.0, 0.0), (2, 7, 0, 0, 0, 129, 65, 77, 1.0, 0.0)),
    ((1, 6, 1, 1, 0, 600, 300, 257, -1.5, 0.25), (2, 20, 1, 0, 1, 40, 33, 16, 2.0, -0.5)),
]


@pytest.mark.parametrize("first,second", CASES)
def test_noalloc_matches_workspace_calls(ex, orc, first, second):
    rng = np.random.default_rng(7)
    for scheme, S, layout, ta, tb, M, N, K, alpha, beta in (first, second):
        A = orc.generate(M, K, 1.0, 42)
        B = orc.generate(K, N, 1.0, 43)
        C0 = rng.standard_normal((M, N))
        want = _run(ex, "ws", scheme, S, layout, ta, tb, A, B, alpha, beta, C0)
        got = _run(ex, "noalloc", scheme, S, layout, ta, tb, A, B, alpha, beta, C0)
        np.testing.assert_array_equal(bits(got), bits(want), err_msg=f"{scheme} {S} {M}x{N}x{K}")


def test_noalloc_rejects_problems_beyond_the_scratch(ex):
    from ozablas_amd import _lib as L

    M = N = K = 4096  # 17 moduli: about 850 MiB of workspace
    need = L.gemm_workspace_bytes(2, M, N, K, 17)
    assert need > 128 << 20
    # Null operands: were the size check to let the call through, oza_gemm's own null-pointer check
    # would still stop it before any launch.  The message tells the two rejections apart.
    st = L.lib.oza_gemm_noalloc(ex.ctx, 2, 17, 0, 0, 0, M, N, K, 1.0, None, K, None, N, 0.0, None, N)
    assert st == L.INVALID_ARGUMENT
    assert b"oza_gemm_noalloc holds" in L.lib.oza_last_error()
