"""Bit-exact parity of every stage of the gfx950 pipeline against the oracle (and, where the
reference build is present, of the oracle against the reference's own GPU kernels).

All calls go through the C ABI (include/ozablas_hip.h) via ozablas_amd._lib."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (7, 13, 5), (33, 17, 64), (129, 257, 300), (256, 256, 256), (300, 200, 1001)]


def _lib():
    from ozablas_amd import _lib as L

    return L


def gen_pair(orc, M, N, K, phi, seed):
    return orc.generate(M, K, phi, seed), orc.generate(K, N, phi, seed + 1)


def special_matrix(rng, rows, cols):
    """Exponential-distribution values plus zero rows/columns, subnormals and huge values."""
    X = (rng.random((rows, cols)) - 0.5) * np.exp(2.0 * rng.standard_normal((rows, cols)))
    if rows > 2:
        X[1, :] = 0.0
    if cols > 2:
        X[:, 2] = 0.0
    if rows > 3 and cols > 3:
        X[3, 0] = 5e-310  # subnormal
        X[3, 1] = -2.5e-320
        X[0, 3] = 1.0e300
    return X


def stage_shifts(ex, scheme, A, B, S):
    L = _lib()
    M, K = A.shape
    N = B.shape[1]
    dA, dB = ex.to_device(A), ex.to_device(B)
    sa, sb = ex.alloc(4 * M), ex.alloc(4 * N)
    L.check(L.lib.oza_stage1_shifts(ex.ctx, scheme, M, N, K, S, dA.ptr, dB.ptr, sa.ptr, sb.ptr))
    return sa.download((M,), np.int32), sb.download((N,), np.int32)


def stage_slices(ex, scheme, A, B, S, shA, shB):
    L = _lib()
    M, K = A.shape
    N = B.shape[1]
    b = L.workspace_bytes(scheme, M, N, K, S)
    dA, dB = ex.to_device(A), ex.to_device(B)
    sa, sb = ex.to_device(shA.astype(np.int32)), ex.to_device(shB.astype(np.int32))
    As, Bs = ex.alloc(b[2]), ex.alloc(b[3])
    L.check(L.lib.oza_stage2_slices(ex.ctx, scheme, M, N, K, S, dA.ptr, dB.ptr, sa.ptr, sb.ptr, As.ptr, Bs.ptr))
    ap, bp = ex.alloc(S * M * K), ex.alloc(S * K * N)
    L.check(L.lib.oza_unpack_slices(ex.ctx, M, N, K, S, As.ptr, Bs.ptr, ap.ptr, bp.ptr))
    return ap.download((S, M, K), np.int8), bp.download((S, K, N), np.int8), As, Bs


# ------------------------------------------------------------------------------------------ step 1
@pytest.mark.parametrize("M,N,K", SHAPES)
@pytest.mark.parametrize("scheme,S", [(1, 4), (2, 4), (2, 16)])
def test_stage1_shifts(ex, orc, M, N, K, scheme, S):
    A, B = gen_pair(orc, M, N, K, 1.0, 42)
    sa, sb = stage_shifts(ex, scheme, A, B, S)
    off = orc.kparam(S, K) if scheme == 2 else 0
    np.testing.assert_array_equal(sa, orc.row_shifts(A, off))
    np.testing.assert_array_equal(sb, orc.col_shifts(B, off))


def test_stage1_special_values(ex, orc, rng):
    A = special_matrix(rng, 40, 70)
    B = special_matrix(rng, 70, 30)
    B[5, :] = np.inf
    B[6, 4] = np.nan
    B[:, 7] = 1e-320
    A[8, :] = np.nan
    sa, sb = stage_shifts(ex, 2, A, B, 8)
    off = orc.kparam(8, 70)
    np.testing.assert_array_equal(sa, orc.row_shifts(A, off))
    np.testing.assert_array_equal(sb, orc.col_shifts(B, off))


# ------------------------------------------------------------------------------------------ step 2
@pytest.mark.parametrize("M,N,K", SHAPES)
@pytest.mark.parametrize("S", [1, 4, 8, 13, 20])
def test_stage2_scheme2_residues(ex, orc, M, N, K, S):
    A, B = gen_pair(orc, M, N, K, 2.0, 7)
    kp = orc.kparam(S, K)
    shA, shB = orc.row_shifts(A, kp), orc.col_shifts(B, kp)
    ap, bp, _, _ = stage_slices(ex, 2, A, B, S, shA, shB)
    np.testing.assert_array_equal(ap, orc.slice_s2(A, shA, S, True))
    np.testing.assert_array_equal(bp, orc.slice_s2(B, shB, S, False))


@pytest.mark.parametrize("M,N,K", SHAPES)
@pytest.mark.parametrize("S", [1, 3, 7, 10])
def test_stage2_scheme1_slices(ex, orc, M, N, K, S):
    A, B = gen_pair(orc, M, N, K, 2.0, 9)
    beta = orc.beta(K)
    shA, shB = orc.row_shifts(A, 0), orc.col_shifts(B, 0)
    ap, bp, _, _ = stage_slices(ex, 1, A, B, S, shA, shB)
    np.testing.assert_array_equal(ap, orc.slice_s1(A, shA, S, beta, True))
    np.testing.assert_array_equal(bp, orc.slice_s1(B, shB, S, beta, False))


def test_stage2_special_values(ex, orc, rng):
    A = special_matrix(rng, 37, 90)
    B = special_matrix(rng, 90, 41)
    for scheme, S in ((2, 20), (2, 6), (1, 8)):
        off = orc.kparam(S, 90) if scheme == 2 else 0
        shA, shB = orc.row_shifts(A, off), orc.col_shifts(B, off)
        ap, bp, _, _ = stage_slices(ex, scheme, A, B, S, shA, shB)
        if scheme == 2:
            np.testing.assert_array_equal(ap, orc.slice_s2(A, shA, S, True))
            np.testing.assert_array_equal(bp, orc.slice_s2(B, shB, S, False))
        else:
            beta = orc.beta(90)
            np.testing.assert_array_equal(ap, orc.slice_s1(A, shA, S, beta, True))
            np.testing.assert_array_equal(bp, orc.slice_s1(B, shB, S, beta, False))


def test_stage2_tiny_K_wide_residues(ex, orc):
    """K = 1, S = 20 gives K_param = 75: |rint(x)| up to 2^75, the widest residue path."""
    A, B = gen_pair(orc, 50, 60, 1, 1.0, 3)
    S = 20
    kp = orc.kparam(S, 1)
    assert kp >= 73
    shA, shB = orc.row_shifts(A, kp), orc.col_shifts(B, kp)
    ap, bp, _, _ = stage_slices(ex, 2, A, B, S, shA, shB)
    np.testing.assert_array_equal(ap, orc.slice_s2(A, shA, S, True))
    np.testing.assert_array_equal(bp, orc.slice_s2(B, shB, S, False))


# ------------------------------------------------------------------------------------------ step 3
@pytest.mark.parametrize("M,N,K", SHAPES)
@pytest.mark.parametrize("S", [1, 5, 12])
def test_stage3_scheme2_products_and_digits(ex, orc, M, N, K, S):
    L = _lib()
    A, B = gen_pair(orc, M, N, K, 1.0, 11)
    kp = orc.kparam(S, K)
    shA, shB = orc.row_shifts(A, kp), orc.col_shifts(B, kp)
    ap, bp, As, Bs = stage_slices(ex, 2, A, B, S, shA, shB)
    want = orc.s2_products(ap, bp)
    cp = ex.alloc(4 * S * M * N)
    L.check(L.lib.oza_stage3_s2_products(ex.ctx, M, N, K, S, As.ptr, Bs.ptr, cp.ptr))
    np.testing.assert_array_equal(cp.download((S, M, N), np.int32), want)
    b = L.workspace_bytes(2, M, N, K, S)
    dg = ex.alloc(b[4])
    L.check(L.lib.oza_stage3_s2_digits(ex.ctx, M, N, K, S, As.ptr, Bs.ptr, dg.ptr))
    pl = ex.alloc(S * M * N)
    L.check(L.lib.oza_unpack_digits(ex.ctx, M, N, S, dg.ptr, pl.ptr))
    np.testing.assert_array_equal(pl.download((S, M, N), np.uint8), orc.crt_digits(want))


@pytest.mark.parametrize("M,N,K", SHAPES)
@pytest.mark.parametrize("S", [1, 4, 9])
def test_stage3_scheme1_groups(ex, orc, M, N, K, S):
    L = _lib()
    A, B = gen_pair(orc, M, N, K, 1.0, 13)
    shA, shB = orc.row_shifts(A, 0), orc.col_shifts(B, 0)
    ap, bp, As, Bs = stage_slices(ex, 1, A, B, S, shA, shB)
    T = ex.alloc(4 * S * M * N)
    L.check(L.lib.oza_stage3_s1_groups(ex.ctx, M, N, K, S, As.ptr, Bs.ptr, T.ptr))
    np.testing.assert_array_equal(T.download((S, M, N), np.int32), orc.s1_groups(ap, bp))


def frag_tile(P):
    """Row-major [S][R][Kp] INT8 planes -> the library's fragment-tiled operand layout: 1 KiB blocks
    of 32 rows x 32 K-bytes, lane l of a block holding row l%32, K-bytes 16*(l//32)..+15."""
    S, R, Kp = P.shape
    return np.ascontiguousarray(P.reshape(S, R // 32, 32, Kp // 32, 2, 16).transpose(0, 1, 3, 4, 2, 5)).reshape(S, R, Kp)


def test_igemm_extreme_operands(ex, orc, rng):
    """Synthetic product-layout operands at the residue extremes (+-125) over K = 8192 drive the
    INT32 accumulators to ~K*125^2; products must stay exact."""
    L = _lib()
    M, N, K, S = 130, 70, 8192, 2
    Mp, Np, Kp = 256, 256, 8192
    vals = np.array([-125, 125, -124, 0, 1, -1], np.int8)
    Ap = np.zeros((S, Mp, Kp), np.int8)
    Bt = np.zeros((S, Np, Kp), np.int8)
    Ap[:, :M, :K] = rng.choice(vals, size=(S, M, K), p=[0.4, 0.4, 0.05, 0.05, 0.05, 0.05])
    Bt[:, :N, :K] = rng.choice(vals, size=(S, N, K), p=[0.4, 0.4, 0.05, 0.05, 0.05, 0.05])
    Ap[0, 0, :K] = 125
    Bt[0, 0, :K] = 125
    Ap[1, 1, :K] = -125
    Bt[1, 1, :K] = 125
    As, Bs = ex.to_device(frag_tile(Ap)), ex.to_device(frag_tile(Bt))
    cp = ex.alloc(4 * S * M * N)
    L.check(L.lib.oza_stage3_s2_products(ex.ctx, M, N, K, S, As.ptr, Bs.ptr, cp.ptr))
    got = cp.download((S, M, N), np.int32)
    want = orc.s2_products(Ap[:, :M, :K], np.ascontiguousarray(Bt[:, :N, :K].transpose(0, 2, 1)))
    np.testing.assert_array_equal(got, want)
    assert got[0, 0, 0] == K * 125 * 125 and got[1, 1, 1] == -K * 125 * 125


# ------------------------------------------------------------------------------------------ step 4
@pytest.mark.parametrize("S", list(range(1, 21)))
def test_stage4_scheme2_all_moduli_counts(ex, orc, S):
    L = _lib()
    M, N, K = 65, 70, 300
    A, B = gen_pair(orc, M, N, K, 1.0, 100 + S)
    kp = orc.kparam(S, K)
    shA, shB = orc.row_shifts(A, kp), orc.col_shifts(B, kp)
    Ct = orc.s2_products(orc.slice_s2(A, shA, S, True), orc.slice_s2(B, shB, S, False))
    want = orc.recon_s2(Ct, shA, shB)
    dct, dsa, dsb, dC = ex.to_device(Ct), ex.to_device(shA), ex.to_device(shB), ex.alloc(8 * M * N)
    L.check(L.lib.oza_stage4_s2_products(ex.ctx, M, N, K, S, dct.ptr, dsa.ptr, dsb.ptr, dC.ptr))
    got = dC.download((M, N), np.float64)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("S", [1, 2, 7, 8, 9, 16, 19, 20])
def test_stage4_scheme2_synthetic_residues(ex, orc, rng, S):
    """Arbitrary INT32 planes: every residue class reachable, including the acc == floor(M/2)
    boundary that valid inputs never produce."""
    L = _lib()
    M, N = 48, 80
    Ct = rng.integers(-(2 ** 31), 2 ** 31 - 1, size=(S, M, N), dtype=np.int64).astype(np.int32)
    Ct[:, 0, 0] = 0
    # X = floor(M/2): choose residues of that integer
    Mp, Mh, _, _ = orc.crt_tables(S)
    half = sum(int(v) << (64 * i) for i, v in enumerate(Mh))
    Mfull = sum(int(v) << (64 * i) for i, v in enumerate(Mp))
    # |Y| tiny against M (Y = 1, -1, -12345, 98765): the quotient estimate cannot tell the sign of
    # X/M - q there, so the reconstruction's wrong-guess path runs
    tiny = [1, Mfull - 1, Mfull - 12345, 98765]
    for s, m in enumerate(orc.moduli()[:S]):
        Ct[s, 0, 1] = half % int(m)
        Ct[s, 0, 2] = (half + 1) % int(m)
        Ct[s, 0, 3] = (half - 1) % int(m)
        for j, x in enumerate(tiny):
            Ct[s, 0, 4 + j] = x % int(m)
    shA = rng.integers(-60, 60, size=M).astype(np.int32)
    shB = rng.integers(-60, 60, size=N).astype(np.int32)
    want = orc.recon_s2(Ct, shA, shB)
    dct, dsa, dsb, dC = ex.to_device(Ct), ex.to_device(shA), ex.to_device(shB), ex.alloc(8 * M * N)
    L.check(L.lib.oza_stage4_s2_products(ex.ctx, M, N, 100, S, dct.ptr, dsa.ptr, dsb.ptr, dC.ptr))
    got = dC.download((M, N), np.float64)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("S", [1, 4, 8])
def test_stage4_scheme1_groups(ex, orc, S):
    L = _lib()
    M, N, K = 50, 77, 200
    A, B = gen_pair(orc, M, N, K, 2.0, 5)
    beta = orc.beta(K)
    shA, shB = orc.row_shifts(A, 0), orc.col_shifts(B, 0)
    T = orc.s1_groups(orc.slice_s1(A, shA, S, beta, True), orc.slice_s1(B, shB, S, beta, False))
    want = orc.recon_s1(T, shA, shB, beta)
    dT, dsa, dsb, dC = ex.to_device(T), ex.to_device(shA), ex.to_device(shB), ex.alloc(8 * M * N)
    L.check(L.lib.oza_stage4_s1_groups(ex.ctx, M, N, K, S, dT.ptr, dsa.ptr, dsb.ptr, dC.ptr))
    np.testing.assert_array_equal(dC.download((M, N), np.float64).view(np.uint64), want.view(np.uint64))


# ------------------------------------------------------------------------------------------ pipelines
def run_pipeline(ex, scheme, A, B, S, timings=None):
    import ozablas_amd as oz

    M, K = A.shape
    N = B.shape[1]
    ws = (oz.WorkspaceScheme1 if scheme == 1 else oz.WorkspaceScheme2)(ex, M, N, K, S)
    dA, dB, dC = ex.to_device(A), ex.to_device(B), ex.alloc(8 * M * N)
    (oz.ozaki_scheme1_gemm if scheme == 1 else oz.ozaki_scheme2_gemm)(ws, dA, dB, dC, timings)
    ex.synchronize()
    return dC.download((M, N), np.float64), ws


@pytest.mark.parametrize("M,N,K", SHAPES)
@pytest.mark.parametrize("S", [2, 4, 7, 8, 12, 16, 20])
def test_scheme2_pipeline_bit_exact(ex, orc, M, N, K, S):
    A, B = gen_pair(orc, M, N, K, 1.0, 42)
    got, _ = run_pipeline(ex, 2, A, B, S)
    want = orc.s2_gemm(A, B, S)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("M,N,K", SHAPES)
@pytest.mark.parametrize("S", [1, 3, 6, 10])
def test_scheme1_pipeline_bit_exact(ex, orc, M, N, K, S):
    A, B = gen_pair(orc, M, N, K, 2.0, 42)
    got, _ = run_pipeline(ex, 1, A, B, S)
    want = orc.s1_gemm(A, B, S)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


# Work orders of the persistent GEMMs: the dynamic (per-XCD counters) order needs a grid that is a
# multiple of 8 and at least two items per workgroup (2000 x 1900: 8 x 8 tiles with ragged edges,
# 768 and 640 items on 256 workgroups); with fewer items per workgroup (192, 160, 120 and 96 items,
# grid = items) or a grid that is not a multiple of 8 (105, 36) the static order runs.
@pytest.mark.parametrize("scheme,M,N,K,S", [(1, 1024, 1024, 200, 6), (1, 2048, 512, 300, 10),
                                            (1, 1000, 700, 129, 3), (1, 1280, 768, 64, 8),
                                            (1, 2000, 1900, 64, 10),
                                            (2, 1024, 1024, 200, 12), (2, 1280, 768, 64, 7),
                                            (2, 2000, 1900, 300, 12)])
def test_pipeline_work_orders(ex, orc, scheme, M, N, K, S):
    A, B = gen_pair(orc, M, N, K, 2.0 if scheme == 1 else 1.0, 11)
    got, _ = run_pipeline(ex, scheme, A, B, S)
    want = orc.s1_gemm(A, B, S) if scheme == 1 else orc.s2_gemm(A, B, S)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


def test_pipelines_special_values(ex, orc, rng):
    A = special_matrix(rng, 70, 129)
    B = special_matrix(rng, 129, 90)
    for scheme, S in ((2, 8), (2, 18), (1, 7)):
        got, _ = run_pipeline(ex, scheme, A, B, S)
        want = orc.s2_gemm(A, B, S) if scheme == 2 else orc.s1_gemm(A, B, S)
        np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


def test_c1_config_against_fp128(ex, orc):
    """M=N=K=256, slices=4, phi=0.5, seeds 42/43 (the reference's CPU-scale case): bit-exact with
    the oracle for both schemes, and within the truncation bound of the FP128 product."""
    import ozablas_amd as oz

    A, B = oz.generate_matrix(256, 256, 0.5, 42), oz.generate_matrix(256, 256, 0.5, 43)
    exact = orc.gemm_fp128(A, B)
    absab = orc.gemm_abs(A, B)
    for scheme in (1, 2):
        t = oz.OzaTimings()
        got, _ = run_pipeline(ex, scheme, A, B, 4, t)
        want = orc.s2_gemm(A, B, 4) if scheme == 2 else orc.s1_gemm(A, B, 4)
        np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
        assert t.total_ms > 0.0
        # Scheme II truncates each entry to 2^-(K_param) of its row/column maximum: with
        # K_param = 11 at S = 4, K = 256 the error is well below 2^-9 |A||B| componentwise.
        cw = np.max(np.abs(got - exact) / absab)
        assert cw < 2.0 ** -8, (scheme, cw)


# ------------------------------------------------------------------------------------------ reference kernels
def test_reference_kernels_match_oracle(ex, orc, ref, rng):
    """The reference's own stage 1/2/4 kernels (compiled from its sources into oracle/_ref) agree
    bit for bit with the oracle's restatement on the same inputs."""
    assert ref.lib.ref_gpu_init() == 0
    M, N, K = 70, 90, 130
    A = special_matrix(rng, M, K)
    B = special_matrix(rng, K, N)
    dA, dB = ex.to_device(A), ex.to_device(B)
    for S in (3, 7, 8, 12, 20):
        kp = orc.kparam(S, K)
        sa, sb = ex.alloc(4 * M), ex.alloc(4 * N)
        assert ref.lib.ref_gpu_row_shifts(dA.ptr, M, K, kp, sa.ptr) == 0
        assert ref.lib.ref_gpu_col_shifts(dB.ptr, K, N, kp, sb.ptr) == 0
        shA, shB = sa.download((M,), np.int32), sb.download((N,), np.int32)
        np.testing.assert_array_equal(shA, orc.row_shifts(A, kp))
        np.testing.assert_array_equal(shB, orc.col_shifts(B, kp))
        oa, ob = ex.alloc(S * M * K), ex.alloc(S * K * N)
        assert ref.lib.ref_gpu_slice_s2_A(dA.ptr, sa.ptr, M, K, S, oa.ptr) == 0
        assert ref.lib.ref_gpu_slice_s2_B(dB.ptr, sb.ptr, K, N, S, ob.ptr) == 0
        ap, bp = oa.download((S, M, K), np.int8), ob.download((S, K, N), np.int8)
        np.testing.assert_array_equal(ap, orc.slice_s2(A, shA, S, True))
        np.testing.assert_array_equal(bp, orc.slice_s2(B, shB, S, False))
        Ct = orc.s2_products(ap, bp)
        dct, dC = ex.to_device(Ct), ex.alloc(8 * M * N)
        assert ref.lib.ref_gpu_recon_s2(dct.ptr, sa.ptr, sb.ptr, M, N, S, dC.ptr) == 0
        want = orc.recon_s2(Ct, shA, shB)
        np.testing.assert_array_equal(dC.download((M, N), np.float64).view(np.uint64), want.view(np.uint64))
    # Scheme I
    for S in (2, 5, 9):
        beta = orc.beta(K)
        sa, sb = ex.alloc(4 * M), ex.alloc(4 * N)
        assert ref.lib.ref_gpu_row_shifts(dA.ptr, M, K, 0, sa.ptr) == 0
        assert ref.lib.ref_gpu_col_shifts(dB.ptr, K, N, 0, sb.ptr) == 0
        shA, shB = sa.download((M,), np.int32), sb.download((N,), np.int32)
        oa, ob = ex.alloc(S * M * K), ex.alloc(S * K * N)
        assert ref.lib.ref_gpu_slice_s1_A(dA.ptr, sa.ptr, M, K, S, beta, oa.ptr) == 0
        assert ref.lib.ref_gpu_slice_s1_B(dB.ptr, sb.ptr, K, N, S, beta, ob.ptr) == 0
        ap, bp = oa.download((S, M, K), np.int8), ob.download((S, K, N), np.int8)
        np.testing.assert_array_equal(ap, orc.slice_s1(A, shA, S, beta, True))
        np.testing.assert_array_equal(bp, orc.slice_s1(B, shB, S, beta, False))
        T = orc.s1_groups(ap, bp)
        dC = ex.to_device(np.zeros((M, N)))
        for g in range(2, S + 2):
            dT = ex.to_device(np.ascontiguousarray(T[g - 2]))
            assert ref.lib.ref_gpu_recon_s1_group(dT.ptr, sa.ptr, sb.ptr, M, N, g, beta, dC.ptr) == 0
        want = orc.recon_s1(T, shA, shB, beta)
        np.testing.assert_array_equal(dC.download((M, N), np.float64).view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("M,N,K,S", [(96, 64, 8000, 5), (1, 1, 5000, 20), (200, 130, 20000, 12),
                                     (256, 256, 32768, 16)])
def test_single_call_split_k(ex, orc, M, N, K, S):
    """Single calls with few 256 x 256 tiles and a long K run their GEMM in K parts whose digits are
    summed mod m_s (the workspace holds the partial planes): bit-identical to the unsplit grouped
    GEMM of the same problem, and to the oracle (whole C, or sampled rows x columns with the full K:
    step 1 depends only on whole rows of A and columns of B, so a block of C is the oracle's product
    of its rows and columns)."""
    import ozablas_amd as oz

    b = oz.workspace_bytes(2, M, N, K, S)
    Mp, Np = -(-M // 256) * 256, -(-N // 256) * 256
    split = b[4] // (S * Mp * Np)
    assert split > 1, (M, N, K, S, split)
    A, B = gen_pair(orc, M, N, K, 1.0, 77)
    got, ws = run_pipeline(ex, 2, A, B, S)
    ws.free()
    g = oz.GroupScheme2(ex, [(M, N, K, S)], tiles=[256], splits=[1])
    dA, dB, dC = ex.to_device(A), ex.to_device(B), ex.alloc(8 * M * N)
    g.run([dA], [dB], [dC], use_graph=False)
    ex.synchronize()
    g.free()
    np.testing.assert_array_equal(got.view(np.uint64), dC.download((M, N), np.float64).view(np.uint64))
    if M * N * K <= 50_000_000:
        want = orc.s2_gemm(A, B, S)
        np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
    else:
        rng = np.random.default_rng(K)
        rows = np.unique(np.r_[0, M - 1, rng.choice(M, 8, replace=False)])
        cols = np.unique(np.r_[0, N - 1, rng.choice(N, 8, replace=False)])
        want = orc.s2_gemm(np.ascontiguousarray(A[rows]), np.ascontiguousarray(B[:, cols]), S)
        np.testing.assert_array_equal(got[np.ix_(rows, cols)].view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("S", [1, 2, 4, 6, 7, 8, 9, 16, 17, 19, 20])
def test_stage4_scheme2_digits_synthetic(ex, orc, rng, S):
    """The reconstruction from blocked digits (the pipeline's step 4) on arbitrary residues: every
    residue class reachable, the acc == floor(M/2) boundary, and |Y| tiny against M (the wrong-guess
    path), packed by oza_pack_digits, against the oracle's reconstruction of the same residues."""
    L = _lib()
    M, N = 70, 90  # ragged 32 x 32 blocks
    Ct = np.stack([rng.integers(0, int(m), size=(M, N)) for m in orc.moduli()[:S]]).astype(np.int32)
    Mp, Mh, _, _ = orc.crt_tables(S)
    half = sum(int(v) << (64 * i) for i, v in enumerate(Mh))
    Mfull = sum(int(v) << (64 * i) for i, v in enumerate(Mp))
    special = [0, half, half + 1, half - 1, 1, Mfull - 1, Mfull - 12345, 98765]
    for s, m in enumerate(orc.moduli()[:S]):
        for j, x in enumerate(special):
            Ct[s, 0, j] = x % int(m)
            Ct[s, M - 1, N - 1 - j] = x % int(m)
    shA = rng.integers(-60, 60, size=M).astype(np.int32)
    shB = rng.integers(-60, 60, size=N).astype(np.int32)
    want = orc.recon_s2(Ct, shA, shB)
    b = L.workspace_bytes(2, M, N, 1, S)
    dg = ex.alloc(b[4])
    planar = ex.to_device(Ct.astype(np.uint8))
    L.check(L.lib.oza_pack_digits(ex.ctx, M, N, S, planar.ptr, dg.ptr))
    back = ex.alloc(S * M * N)
    L.check(L.lib.oza_unpack_digits(ex.ctx, M, N, S, dg.ptr, back.ptr))
    np.testing.assert_array_equal(back.download((S, M, N), np.uint8), Ct.astype(np.uint8))
    dsa, dsb, dC = ex.to_device(shA), ex.to_device(shB), ex.alloc(8 * M * N)
    L.check(L.lib.oza_stage4_s2_digits(ex.ctx, M, N, 1, S, dg.ptr, dsa.ptr, dsb.ptr, dC.ptr))
    got = dC.download((M, N), np.float64)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
