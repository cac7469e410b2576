"""numpy-facing bindings of the oracle (oracle/liboracle.so) and of the reference build
(oracle/_ref/libozref.so).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libozref.so")

_P, _I, _I64, _D, _SZ = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_size_t
U64x4 = C.c_uint64 * 4


def _ptr(a: np.ndarray):
    return a.ctypes.data


class Oracle:
    """CPU restatement of the reference pipeline (oracle/ozoracle.c)."""

    def __init__(self, path: str = ORACLE_PATH):
        self.lib = L = C.CDLL(path)
        protos = {
            "orc_generate_matrix": [_I64, _I64, _D, _I, _P],
            "orc_moduli": [_P],
            "orc_crt_tables": [_I, _P, _P, _P, _P],
            "orc_to_double_256": [_P],
            "orc_scheme1_beta": [_I64],
            "orc_scheme2_kparam": [_I, _I64],
            "orc_row_shifts": [_P, _I64, _I64, C.c_int32, _P],
            "orc_col_shifts": [_P, _I64, _I64, C.c_int32, _P],
            "orc_slice_s1_rows": [_P, _P, _I64, _I64, _I, _I, _P],
            "orc_slice_s1_cols": [_P, _P, _I64, _I64, _I, _I, _P],
            "orc_symmetric_mod": [_I64, _I],
            "orc_slice_s2_rows": [_P, _P, _I64, _I64, _I, _P],
            "orc_slice_s2_cols": [_P, _P, _I64, _I64, _I, _P],
            "orc_igemm": [_P, _P, _I64, _I64, _I64, _P],
            "orc_s2_products": [_P, _P, _I64, _I64, _I64, _I, _P],
            "orc_s1_group_products": [_P, _P, _I64, _I64, _I64, _I, _P],
            "orc_recon_s1": [_P, _P, _P, _I64, _I64, _I, _I, _P],
            "orc_recon_s2": [_P, _P, _P, _I64, _I64, _I, _P],
            "orc_crt_digits": [_P, _I64, _I, _P],
            "orc_s1_gemm": [_P, _P, _P, _I64, _I64, _I64, _I],
            "orc_s2_gemm": [_P, _P, _P, _I64, _I64, _I64, _I],
            "orc_gemm_fp128": [_P, _P, _P, _I64, _I64, _I64],
            "orc_gemm_dd": [_P, _P, _P, _I64, _I64, _I64],
            "orc_gemm_abs": [_P, _P, _P, _I64, _I64, _I64],
            "orc_errors": [_P, _P, _I64, _P],
            "orc_num_threads": [],
        }
        restype = {"orc_to_double_256": _D, "orc_igemm": _I64, "orc_s2_products": _I64,
                   "orc_s1_group_products": _I64}
        for n, a in protos.items():
            f = getattr(L, n)
            f.argtypes = a
            f.restype = restype.get(n, _I)

    # ---- inputs and parameters
    def generate(self, rows, cols, phi, seed):
        out = np.empty((rows, cols), np.float64)
        self.lib.orc_generate_matrix(rows, cols, phi, seed, _ptr(out))
        return out

    def moduli(self):
        out = np.empty(20, np.uint8)
        self.lib.orc_moduli(_ptr(out))
        return out

    def crt_tables(self, S):
        M, Mh, part, inv = (np.zeros(4, np.uint64), np.zeros(4, np.uint64), np.zeros((20, 4), np.uint64),
                            np.zeros(20, np.uint64))
        assert self.lib.orc_crt_tables(S, _ptr(M), _ptr(Mh), _ptr(part), _ptr(inv)) == 0
        return M, Mh, part, inv

    def to_double_256(self, limbs):
        a = np.asarray(limbs, np.uint64)
        return self.lib.orc_to_double_256(_ptr(a))

    def beta(self, K):
        return self.lib.orc_scheme1_beta(K)

    def kparam(self, S, K):
        return self.lib.orc_scheme2_kparam(S, K)

    # ---- stages
    def row_shifts(self, A, offset):
        A = np.ascontiguousarray(A, np.float64)
        out = np.empty(A.shape[0], np.int32)
        self.lib.orc_row_shifts(_ptr(A), A.shape[0], A.shape[1], offset, _ptr(out))
        return out

    def col_shifts(self, B, offset):
        B = np.ascontiguousarray(B, np.float64)
        out = np.empty(B.shape[1], np.int32)
        self.lib.orc_col_shifts(_ptr(B), B.shape[0], B.shape[1], offset, _ptr(out))
        return out

    def slice_s1(self, X, neg_exp, S, beta, by_rows):
        X = np.ascontiguousarray(X, np.float64)
        e = np.ascontiguousarray(neg_exp, np.int32)
        out = np.empty((S,) + X.shape, np.int8)
        f = self.lib.orc_slice_s1_rows if by_rows else self.lib.orc_slice_s1_cols
        f(_ptr(X), _ptr(e), X.shape[0], X.shape[1], S, beta, _ptr(out))
        return out

    def slice_s2(self, X, shifts, S, by_rows):
        X = np.ascontiguousarray(X, np.float64)
        sh = np.ascontiguousarray(shifts, np.int32)
        out = np.empty((S,) + X.shape, np.int8)
        f = self.lib.orc_slice_s2_rows if by_rows else self.lib.orc_slice_s2_cols
        f(_ptr(X), _ptr(sh), X.shape[0], X.shape[1], S, _ptr(out))
        return out

    def symmetric_mod(self, a, m):
        return self.lib.orc_symmetric_mod(int(a), int(m))

    def igemm(self, A8, B8):
        A8 = np.ascontiguousarray(A8, np.int8)
        B8 = np.ascontiguousarray(B8, np.int8)
        M, K = A8.shape
        N = B8.shape[1]
        out = np.empty((M, N), np.int32)
        bad = self.lib.orc_igemm(_ptr(A8), _ptr(B8), M, N, K, _ptr(out))
        return out, bad

    def s2_products(self, As, Bs):
        As, Bs = np.ascontiguousarray(As, np.int8), np.ascontiguousarray(Bs, np.int8)
        S, M, K = As.shape
        N = Bs.shape[2]
        out = np.empty((S, M, N), np.int32)
        bad = self.lib.orc_s2_products(_ptr(As), _ptr(Bs), M, N, K, S, _ptr(out))
        assert bad == 0
        return out

    def s1_groups(self, As, Bs):
        As, Bs = np.ascontiguousarray(As, np.int8), np.ascontiguousarray(Bs, np.int8)
        S, M, K = As.shape
        N = Bs.shape[2]
        out = np.empty((S, M, N), np.int32)
        bad = self.lib.orc_s1_group_products(_ptr(As), _ptr(Bs), M, N, K, S, _ptr(out))
        assert bad == 0
        return out

    def recon_s1(self, T, shA, shB, beta):
        T = np.ascontiguousarray(T, np.int32)
        S, M, N = T.shape
        out = np.empty((M, N), np.float64)
        self.lib.orc_recon_s1(_ptr(T), _ptr(np.ascontiguousarray(shA, np.int32)),
                              _ptr(np.ascontiguousarray(shB, np.int32)), M, N, S, beta, _ptr(out))
        return out

    def recon_s2(self, Ct, shA, shB):
        Ct = np.ascontiguousarray(Ct, np.int32)
        S, M, N = Ct.shape
        out = np.empty((M, N), np.float64)
        self.lib.orc_recon_s2(_ptr(Ct), _ptr(np.ascontiguousarray(shA, np.int32)),
                              _ptr(np.ascontiguousarray(shB, np.int32)), M, N, S, _ptr(out))
        return out

    def crt_digits(self, Ct):
        Ct = np.ascontiguousarray(Ct, np.int32)
        S, M, N = Ct.shape
        out = np.empty((S, M, N), np.uint8)
        self.lib.orc_crt_digits(_ptr(Ct), M * N, S, _ptr(out))
        return out

    # ---- pipelines and end-to-end oracles
    def _mm(self, f, A, B, *extra):
        A = np.ascontiguousarray(A, np.float64)
        B = np.ascontiguousarray(B, np.float64)
        M, K = A.shape
        N = B.shape[1]
        out = np.empty((M, N), np.float64)
        f(_ptr(A), _ptr(B), _ptr(out), M, N, K, *extra)
        return out

    def s1_gemm(self, A, B, S):
        return self._mm(self.lib.orc_s1_gemm, A, B, S)

    def s2_gemm(self, A, B, S):
        return self._mm(self.lib.orc_s2_gemm, A, B, S)

    def gemm_fp128(self, A, B):
        return self._mm(self.lib.orc_gemm_fp128, A, B)

    def gemm_dd(self, A, B):
        return self._mm(self.lib.orc_gemm_dd, A, B)

    def gemm_abs(self, A, B):
        return self._mm(self.lib.orc_gemm_abs, A, B)

    def errors(self, expected, actual):
        e = np.ascontiguousarray(expected, np.float64).ravel()
        a = np.ascontiguousarray(actual, np.float64).ravel()
        out = np.empty(4, np.float64)
        self.lib.orc_errors(_ptr(e), _ptr(a), e.size, _ptr(out))
        return out


class RefBuild:
    """The reference's own sources compiled by oracle/Makefile (target `ref`)."""

    def __init__(self, path: str = REF_PATH):
        self.lib = L = C.CDLL(path)
        protos = {
            "ref_generate_matrix": ([_I, _I, _D, _I, _P], None),
            "ref_gemm_fp128": ([_P, _P, _P, _SZ, _SZ, _SZ], None),
            "ref_errors": ([_P, _P, _SZ, _P], None),
            "ref_moduli": ([_P], None),
            "ref_crt_tables": ([_I, _P, _P, _P, _P], None),
            "ref_to_double_256": ([_P], _D),
            "ref_scheme1_beta": ([_I], _I),
            "ref_scheme2_kparam": ([_I, _I], _I),
            "ref_gpu_init": ([], _I),
            "ref_gpu_row_shifts": ([_P, _I, _I, _I, _P], _I),
            "ref_gpu_col_shifts": ([_P, _I, _I, _I, _P], _I),
            "ref_gpu_slice_s1_A": ([_P, _P, _I, _I, _I, _I, _P], _I),
            "ref_gpu_slice_s1_B": ([_P, _P, _I, _I, _I, _I, _P], _I),
            "ref_gpu_slice_s2_A": ([_P, _P, _I, _I, _I, _P], _I),
            "ref_gpu_slice_s2_B": ([_P, _P, _I, _I, _I, _P], _I),
            "ref_gpu_recon_s1_group": ([_P, _P, _P, _I, _I, _I, _I, _P], _I),
            "ref_gpu_recon_s2": ([_P, _P, _P, _I, _I, _I, _P], _I),
        }
        for n, (a, r) in protos.items():
            f = getattr(L, n)
            f.argtypes = a
            f.restype = r

    def generate(self, rows, cols, phi, seed):
        out = np.empty((rows, cols), np.float64)
        self.lib.ref_generate_matrix(rows, cols, phi, seed, _ptr(out))
        return out

    def crt_tables(self, S):
        M, Mh, part, inv = (np.zeros(4, np.uint64), np.zeros(4, np.uint64), np.zeros((20, 4), np.uint64),
                            np.zeros(20, np.uint64))
        self.lib.ref_crt_tables(S, _ptr(M), _ptr(Mh), _ptr(part), _ptr(inv))
        return M, Mh, part, inv

    def moduli(self):
        out = np.empty(20, np.uint8)
        self.lib.ref_moduli(_ptr(out))
        return out

    def gemm_fp128(self, A, B):
        A = np.ascontiguousarray(A, np.float64)
        B = np.ascontiguousarray(B, np.float64)
        M, K = A.shape
        N = B.shape[1]
        out = np.empty((M, N), np.float64)
        self.lib.ref_gemm_fp128(_ptr(A), _ptr(B), _ptr(out), M, N, K)
        return out

    def errors(self, expected, actual):
        e = np.ascontiguousarray(expected, np.float64).ravel()
        a = np.ascontiguousarray(actual, np.float64).ravel()
        out = np.empty(4, np.float64)
        self.lib.ref_errors(_ptr(e), _ptr(a), e.size, _ptr(out))
        return out
