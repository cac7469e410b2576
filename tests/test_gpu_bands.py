"""Scheme II in row bands (oza_ctx_set_bands): step 4 of band b runs on a second stream beside
step 3 of band b+1.  Whatever the band count, every result must carry the bits of the oracle and
of the unbanded pipeline, including ragged last bands, back-to-back calls that reuse the
workspace, the timed path and the BLAS front end with beta != 0.  All calls go through the C ABI."""
import numpy as np
import pytest

from test_gpu_blas import call
from test_gpu_parity import gen_pair, run_pipeline

pytestmark = pytest.mark.gpu


@pytest.fixture
def banded(ex):
    yield ex
    ex.set_bands(0)


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


# tiles of 256 rows: M = 1000 -> 4 tile rows, 513 -> 3, 2048 -> 8; K = 16 with 20 moduli takes the
# wide residue path (K_param 73)
@pytest.mark.parametrize("M,N,K,S", [(1000, 300, 257, 16), (513, 700, 64, 8), (2048, 256, 512, 20),
                                     (1024, 300, 16, 20)])
@pytest.mark.parametrize("nb", [1, 2, 3, 4, 8])
def test_banded_pipeline_bit_exact(banded, orc, M, N, K, S, nb):
    ex = banded
    A, B = gen_pair(orc, M, N, K, 1.0, 42)
    ex.set_bands(nb)
    got, ws = run_pipeline(ex, 2, A, B, S)
    ws.free()
    np.testing.assert_array_equal(bits(got), bits(orc.s2_gemm(A, B, S)))


@pytest.mark.parametrize("M,N,K,S", [(1000, 300, 257, 6), (1280, 512, 129, 10), (513, 700, 64, 3)])
@pytest.mark.parametrize("nb", [1, 2, 3, 4])
def test_banded_scheme1_bit_exact(banded, orc, M, N, K, S, nb):
    ex = banded
    A, B = gen_pair(orc, M, N, K, 2.0, 42)
    ex.set_bands(nb)
    got, ws = run_pipeline(ex, 1, A, B, S)
    ws.free()
    np.testing.assert_array_equal(bits(got), bits(orc.s1_gemm(A, B, S)))


def test_banded_back_to_back_calls(banded, orc):
    """Two calls on one workspace without a host synchronization in between: the second call's
    step 1 must not overwrite shifts or digits that the first call's side stream still reads."""
    import ozablas_amd as oz

    ex = banded
    ex.set_bands(4)
    M, N, K, S = 1024, 768, 300, 12
    A1, B1 = gen_pair(orc, M, N, K, 1.0, 42)
    A2, B2 = gen_pair(orc, M, N, K, 2.0, 7)
    ws = oz.WorkspaceScheme2(ex, M, N, K, S)
    d = [ex.to_device(X) for X in (A1, B1, A2, B2)]
    C1, C2 = ex.alloc(8 * M * N), ex.alloc(8 * M * N)
    oz.ozaki_scheme2_gemm(ws, d[0], d[1], C1)
    oz.ozaki_scheme2_gemm(ws, d[2], d[3], C2)
    ex.synchronize()
    np.testing.assert_array_equal(bits(C1.download((M, N), np.float64)), bits(orc.s2_gemm(A1, B1, S)))
    np.testing.assert_array_equal(bits(C2.download((M, N), np.float64)), bits(orc.s2_gemm(A2, B2, S)))
    ws.free()


def test_banded_timed_call(banded, orc):
    import ozablas_amd as oz

    ex = banded
    ex.set_bands(3)
    A, B = gen_pair(orc, 768, 512, 200, 1.0, 5)
    t = oz.OzaTimings()
    got, ws = run_pipeline(ex, 2, A, B, 10, t)
    ws.free()
    np.testing.assert_array_equal(bits(got), bits(orc.s2_gemm(A, B, 10)))
    assert t.step3_ms > 0 and t.step4_ms > 0 and t.total_ms >= t.step3_ms


@pytest.mark.parametrize("layout,ta,tb", [(0, 0, 0), (1, 1, 0), (1, 0, 1)])
def test_banded_blas_front_end(banded, orc, rng, layout, ta, tb):
    """alpha / beta and both layouts: banded and unbanded calls give the same bits."""
    ex = banded
    M, N, K, S = 900, 520, 130, 14
    A = orc.generate(M, K, 1.0, 3)
    B = orc.generate(K, N, 1.0, 4)
    C0 = rng.standard_normal((M, N))
    out = {}
    for nb in (1, 4):
        ex.set_bands(nb)
        out[nb] = call(ex, 2, S, layout, ta, tb, A, B, 0.75, -1.5, C0)
    np.testing.assert_array_equal(bits(out[4]), bits(out[1]))


def test_set_bands_rejects_out_of_range(banded):
    from ozablas_amd import _lib

    for bad in (-1, 9):
        with pytest.raises(ValueError):
            banded.set_bands(bad)
    assert _lib.lib.oza_ctx_set_bands(banded.ctx, 9) == _lib.INVALID_ARGUMENT


def test_bands_for_reports_the_band_count(banded):
    ex = banded
    ex.set_bands(0)
    assert ex.bands_for(8192, 8192, 8192, 17) == 4  # 8 tile rows x 32 x 17 = 17 rounds of 256 CUs per band
    assert ex.bands_for(256, 256, 256, 4) == 1
    ex.set_bands(3)
    assert ex.bands_for(1000, 300, 257, 16) == 3
    assert ex.bands_for(300, 300, 257, 16) == 2  # at most one band per 256-row tile
    assert ex.bands_for(1000, 300, 257, 6, scheme=1) == 3
    ex.set_bands(1)
    assert ex.bands_for(8192, 8192, 8192, 17) == 1
    ex.set_bands(0)
    assert ex.bands_for(8192, 8192, 8192, 10, scheme=1) == 4  # 8 tile rows x 32 x 55 pair products per band
    assert ex.bands_for(512, 512, 512, 3, scheme=1) == 1


def test_banded_call_in_a_hip_graph(banded, orc):
    """A banded call captured from the context stream (the side stream joins the capture through
    the fork/join events; the work counters are cleared by a kernel, not a memset) replays to the
    same bits several times, with eager calls on other inputs and host synchronizations between
    the replays."""
    import ctypes as C

    import ozablas_amd as oz

    hip = C.CDLL("libamdhip64.so")
    ex = banded
    ex.set_bands(4)
    M, N, K, S = 1024, 640, 200, 12
    A1, B1 = gen_pair(orc, M, N, K, 1.0, 21)
    A2, B2 = gen_pair(orc, M, N, K, 2.0, 22)
    want1, want2 = orc.s2_gemm(A1, B1, S), orc.s2_gemm(A2, B2, S)
    ws = oz.WorkspaceScheme2(ex, M, N, K, S)
    d = [ex.to_device(X) for X in (A1, B1, A2, B2)]
    C1, C2 = ex.alloc(8 * M * N), ex.alloc(8 * M * N)
    st = C.c_void_p(ex.stream())
    graph, gexec = C.c_void_p(), C.c_void_p()
    assert hip.hipStreamBeginCapture(st, 1) == 0  # thread-local capture mode
    oz.ozaki_scheme2_gemm(ws, d[0], d[1], C1)
    assert hip.hipStreamEndCapture(st, C.byref(graph)) == 0
    assert hip.hipGraphInstantiate(C.byref(gexec), graph, None, None, C.c_size_t(0)) == 0
    try:
        for _ in range(3):
            oz.ozaki_scheme2_gemm(ws, d[2], d[3], C1)  # eager call on other inputs into the same C
            oz.ozaki_scheme2_gemm(ws, d[2], d[3], C2)
            ex.synchronize()
            np.testing.assert_array_equal(bits(C1.download((M, N), np.float64)), bits(want2))
            assert hip.hipGraphLaunch(gexec, st) == 0
            ex.synchronize()
            np.testing.assert_array_equal(bits(C1.download((M, N), np.float64)), bits(want1))
            np.testing.assert_array_equal(bits(C2.download((M, N), np.float64)), bits(want2))
    finally:
        hip.hipGraphExecDestroy(gexec)
        hip.hipGraphDestroy(graph)
        ws.free()
