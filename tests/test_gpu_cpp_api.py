"""The C++ drop-in API (include/ozablas/...) driven the way the reference's own drivers use it;
results compared bit for bit with the oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "test_api")


@pytest.mark.parametrize("M,N,K,S1,S2", [(256, 256, 256, 4, 4), (97, 130, 513, 6, 17)])
def test_cpp_api_driver(orc, tmp_path, M, N, K, S1, S2):
    assert os.path.exists(BIN), "tests/cpp/bin/test_api not built (run __graft_entry__.build())"
    r = subprocess.run([BIN, str(tmp_path), str(M), str(N), str(K), str(S1), str(S2)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    A, B = orc.generate(M, K, 1.0, 42), orc.generate(K, N, 1.0, 43)
    c2 = np.fromfile(tmp_path / "c_s2.bin", np.float64).reshape(M, N)
    c1 = np.fromfile(tmp_path / "c_s1.bin", np.float64).reshape(M, N)
    np.testing.assert_array_equal(c2.view(np.uint64), orc.s2_gemm(A, B, S2).view(np.uint64))
    np.testing.assert_array_equal(c1.view(np.uint64), orc.s1_gemm(A, B, S1).view(np.uint64))
    kp = orc.kparam(S2, K)
    shA, shB = orc.row_shifts(A, kp), orc.col_shifts(B, kp)
    ap = np.fromfile(tmp_path / "a_s2.bin", np.int8).reshape(S2, M, K)
    bp = np.fromfile(tmp_path / "b_s2.bin", np.int8).reshape(S2, K, N)
    np.testing.assert_array_equal(ap, orc.slice_s2(A, shA, S2, True))
    np.testing.assert_array_equal(bp, orc.slice_s2(B, shB, S2, False))
    # the reference's planar C_s (WorkspaceScheme2::unpack_products) and the digits C_s mod m_s
    ct = np.fromfile(tmp_path / "c_tc_s2.bin", np.int32).reshape(S2, M, N)
    want_ct = orc.s2_products(ap, bp)
    np.testing.assert_array_equal(ct, want_ct)
    dg = np.fromfile(tmp_path / "digits_s2.bin", np.uint8).reshape(S2, M, N)
    np.testing.assert_array_equal(dg, orc.crt_digits(want_ct))
    # Scheme I group products T_g (WorkspaceScheme1::unpack_products)
    shA1, shB1 = orc.row_shifts(A, 0), orc.col_shifts(B, 0)
    beta = orc.beta(K)
    a1, b1 = orc.slice_s1(A, shA1, S1, beta, True), orc.slice_s1(B, shB1, S1, beta, False)
    tp = np.fromfile(tmp_path / "t_s1.bin", np.int32).reshape(S1, M, N)
    np.testing.assert_array_equal(tp, orc.s1_groups(a1, b1))
