"""ctypes binding of the C ABI (include/ozablas_hip.h) and a thin Python mirror of the C++ API."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libozablas_hip.so")
MAX_SLICES = 20

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"ozablas_amd: {LIB_PATH} is missing; build it with `python -c 'import __graft_entry__ as g; g.build()'` "
        "(or `make -C ozablas_amd/csrc`). There is no fallback implementation."
    )

lib = C.CDLL(LIB_PATH)

_P = C.c_void_p
_SZ = C.c_size_t
_I = C.c_int
_D = C.c_double


class oza_timings(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("step1_ms", "step2_ms", "step3_ms", "step4_ms", "total_ms")]


# name -> argtypes (all return int status unless listed in _RESTYPE)
_PROTOS = {
    "oza_version": [],
    "oza_status_string": [_I],
    "oza_last_error": [],
    "oza_get_device_count": [C.POINTER(_I)],
    "oza_ctx_create": [_I, C.POINTER(_P)],
    "oza_ctx_destroy": [_P],
    "oza_ctx_device": [_P, C.POINTER(_I)],
    "oza_ctx_stream": [_P, C.POINTER(_P)],
    "oza_malloc": [_P, C.POINTER(_P), _SZ],
    "oza_free": [_P, _P],
    "oza_memcpy_h2d": [_P, _P, _P, _SZ],
    "oza_memcpy_d2h": [_P, _P, _P, _SZ],
    "oza_memcpy_d2d": [_P, _P, _P, _SZ],
    "oza_synchronize": [_P],
    "oza_stream_synchronize": [_P],
    "oza_blas_handle": [_P, C.POINTER(_P)],
    "oza_ctx_set_bands": [_P, _I],
    "oza_ctx_bands_for": [_P, _I, _SZ, _SZ, _SZ, _I, C.POINTER(_I)],
    "oza_ctx_event_timing": [_P, _I],
    "oza_ctx_event_times": [_P, C.POINTER(_D), C.POINTER(_I)],
    "oza_workspace_bytes": [_I, _SZ, _SZ, _SZ, _I, C.POINTER(_SZ)],
    "oza_scheme1_gemm": [_P, _SZ, _SZ, _SZ, _I, _P, _P, _P, _P, _P, _P, _P, _P, C.POINTER(oza_timings)],
    "oza_scheme2_gemm": [_P, _SZ, _SZ, _SZ, _I, _P, _P, _P, _P, _P, _P, _P, _P, C.POINTER(oza_timings)],
    "oza_stage1_shifts": [_P, _I, _SZ, _SZ, _SZ, _I, _P, _P, _P, _P],
    "oza_stage2_slices": [_P, _I, _SZ, _SZ, _SZ, _I, _P, _P, _P, _P, _P, _P],
    "oza_unpack_slices": [_P, _SZ, _SZ, _SZ, _I, _P, _P, _P, _P],
    "oza_stage3_s2_products": [_P, _SZ, _SZ, _SZ, _I, _P, _P, _P],
    "oza_stage3_s2_digits": [_P, _SZ, _SZ, _SZ, _I, _P, _P, _P],
    "oza_unpack_digits": [_P, _SZ, _SZ, _I, _P, _P],
    "oza_stage3_s1_groups": [_P, _SZ, _SZ, _SZ, _I, _P, _P, _P],
    "oza_stage4_s2_products": [_P, _SZ, _SZ, _SZ, _I, _P, _P, _P, _P],
    "oza_stage4_s2_digits": [_P, _SZ, _SZ, _SZ, _I, _P, _P, _P, _P],
    "oza_stage4_s1_groups": [_P, _SZ, _SZ, _SZ, _I, _P, _P, _P, _P],
    "oza_scheme1_beta": [_SZ, C.POINTER(_I)],
    "oza_scheme2_kparam": [_I, _SZ, C.POINTER(_I)],
    "oza_scheme2_slices_for_digits": [_SZ, _D, C.POINTER(_I)],
    "oza_scheme2_slices_for_accuracy": [_SZ, _D, _D, C.POINTER(_I)],
    "oza_generate_matrix": [_SZ, _SZ, _D, _I, _P],
    "oza_dd_gemm": [_P, _SZ, _SZ, _SZ, _P, _P, _P, _P],
    "oza_dgemm": [_P, _SZ, _SZ, _SZ, _P, _P, _P],
    "oza_error_stats": [_P, _SZ, _P, _P, _P, C.POINTER(_D)],
    "oza_mfma_ceiling": [_P, _I, C.POINTER(_D)],
    "oza_group_create": [_P, _I, C.POINTER(_SZ), C.POINTER(_SZ), C.POINTER(_SZ), C.POINTER(_I), _I, C.POINTER(_P)],
    "oza_group_create_split": [_P, _I, C.POINTER(_SZ), C.POINTER(_SZ), C.POINTER(_SZ), C.POINTER(_I), C.POINTER(_I),
                               C.POINTER(_I), C.POINTER(_P)],
    "oza_group_splits": [_P, C.POINTER(_I)],
    "oza_group_plan": [C.POINTER(_D), _I, C.POINTER(_I), _I, C.POINTER(_SZ), C.POINTER(_SZ), C.POINTER(_SZ),
                       C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_SZ)],
    "oza_pack_digits": [_P, _SZ, _SZ, _I, _P, _P],
    "oza_group_destroy": [_P],
    "oza_group_info": [_P, C.POINTER(_SZ), C.POINTER(_I)],
    "oza_group_scheme2_gemm": [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), _I, C.POINTER(oza_timings)],
    "oza_group_autotune": [_P, C.POINTER(_D)],
    "oza_group_costs": [_P, C.POINTER(_D), _I],
    "oza_gemm_workspace_bytes": [_I, _SZ, _SZ, _SZ, _I, C.POINTER(_SZ)],
    "oza_gemm": [_P, _I, _I, _I, _I, _I, _SZ, _SZ, _SZ, _D, _P, _SZ, _P, _SZ, _D, _P, _SZ, _P, _SZ],
    "oza_gemm_noalloc": [_P, _I, _I, _I, _I, _I, _SZ, _SZ, _SZ, _D, _P, _SZ, _P, _SZ, _D, _P, _SZ],
    "oza_server_create": [_P, _SZ, _I, C.POINTER(_P)],
    "oza_server_destroy": [_P],
    "oza_server_submit": [_P, _SZ, _SZ, _SZ, _I, _P, _P, _P, C.POINTER(C.c_uint64)],
    "oza_server_wait": [_P, C.c_uint64, C.POINTER(_D)],
    "oza_server_drain": [_P],
    "oza_server_stats": [_P, C.POINTER(_D)],
    "oza_server_config": [_P, _D, _D],
}
_RESTYPE = {"oza_version": C.c_char_p, "oza_status_string": C.c_char_p, "oza_last_error": C.c_char_p}

for _name, _args in _PROTOS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _RESTYPE.get(_name, _I)

EXPORTED = tuple(_PROTOS)

# status codes (include/ozablas_hip.h)
SUCCESS, INVALID_ARGUMENT, ALLOC_FAILED, HIP_ERROR, NOT_SUPPORTED = range(5)


class OzaError(RuntimeError):
    """A failed C-ABI call; ``status`` is the oza_status code."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"{lib.oza_status_string(status).decode()}: {msg}")
        self.status = status


def check(status: int) -> None:
    if status != SUCCESS:
        msg = lib.oza_last_error().decode()
        if status == INVALID_ARGUMENT:
            raise ValueError(msg)
        if status == ALLOC_FAILED:
            raise MemoryError(msg)
        raise OzaError(status, msg)


def device_count() -> int:
    n = _I(0)
    check(lib.oza_get_device_count(C.byref(n)))
    return n.value


def scheme1_beta(K: int) -> int:
    b = _I(0)
    check(lib.oza_scheme1_beta(K, C.byref(b)))
    return b.value


def scheme2_kparam(S: int, K: int) -> int:
    k = _I(0)
    check(lib.oza_scheme2_kparam(S, K, C.byref(k)))
    return k.value


def scheme2_slices_for_digits(K: int, digits: float) -> int:
    """Smallest moduli count whose truncation bound gives `digits` decimal digits relative to
    K * 2^(E_i + F_j) (the row/column scale of step 1)."""
    s = _I(0)
    check(lib.oza_scheme2_slices_for_digits(K, digits, C.byref(s)))
    return s.value


def scheme2_slices_for_accuracy(K: int, digits: float, spread: float) -> int:
    """Moduli count for `digits` componentwise digits (relative to |A||B|) on inputs whose magnitudes
    spread like exp(spread * N(0, 1)) (oza_scheme2_slices_for_accuracy)."""
    s = _I(0)
    check(lib.oza_scheme2_slices_for_accuracy(K, float(digits), float(spread), C.byref(s)))
    return s.value


def workspace_bytes(scheme: int, M: int, N: int, K: int, S: int) -> tuple:
    out = (_SZ * 5)()
    check(lib.oza_workspace_bytes(scheme, M, N, K, S, out))
    return tuple(out)


def generate_matrix(rows: int, cols: int, phi: float, seed: int) -> np.ndarray:
    """The reference generator (utils/matrix_gen.hpp:15-34), host side."""
    out = np.empty((rows, cols), dtype=np.float64)
    check(lib.oza_generate_matrix(rows, cols, phi, seed, out.ctypes.data))
    return out


class OzaTimings:
    """Mirror of struct OzaTimings (stage times in ms)."""

    __slots__ = ("step1_ms", "step2_ms", "step3_ms", "step4_ms", "total_ms")

    def __init__(self):
        for n in self.__slots__:
            setattr(self, n, 0.0)

    def _fill(self, t: oza_timings) -> None:
        for n in self.__slots__:
            setattr(self, n, float(getattr(t, n)))

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n in self.__slots__}


class DeviceBuffer:
    """A device allocation owned by an executor context."""

    def __init__(self, ex: "HipExecutor", nbytes: int):
        self.ex = ex
        self.nbytes = int(nbytes)
        p = _P()
        check(lib.oza_malloc(ex.ctx, C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    def upload(self, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes, (arr.nbytes, self.nbytes)
        check(lib.oza_memcpy_h2d(self.ex.ctx, self.ptr, arr.ctypes.data, arr.nbytes))
        return self

    def download(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes, (out.nbytes, self.nbytes)
        check(lib.oza_memcpy_d2h(self.ex.ctx, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            check(lib.oza_free(self.ex.ctx, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ex.ctx:
                lib.oza_free(self.ex.ctx, self.ptr)
        except Exception:
            pass


class HipExecutor:
    """Mirror of ozablas::HipExecutor: one library context on one device."""

    def __init__(self, device_id: int = 0):
        p = _P()
        check(lib.oza_ctx_create(device_id, C.byref(p)))
        self.ctx = p.value
        self.device_id = device_id

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr: np.ndarray) -> DeviceBuffer:
        arr = np.ascontiguousarray(arr)
        return DeviceBuffer(self, arr.nbytes).upload(arr)

    def synchronize(self) -> None:
        check(lib.oza_synchronize(self.ctx))

    def stream_synchronize(self) -> None:
        check(lib.oza_stream_synchronize(self.ctx))

    def stream(self) -> int:
        p = _P()
        check(lib.oza_ctx_stream(self.ctx, C.byref(p)))
        return p.value or 0

    def set_bands(self, bands: int) -> None:
        """Row bands of Scheme II's steps 3-4 (0 = automatic, 1 = steps in sequence, 2..8 = fixed)."""
        check(lib.oza_ctx_set_bands(self.ctx, int(bands)))

    def bands_for(self, M: int, N: int, K: int, S: int, scheme: int = 2) -> int:
        """Band count (step-3 launches) of a call of this scheme and shape on this context."""
        n = _I(0)
        check(lib.oza_ctx_bands_for(self.ctx, scheme, M, N, K, S, C.byref(n)))
        return n.value

    def event_timing(self, enable: bool) -> None:
        check(lib.oza_ctx_event_timing(self.ctx, int(enable)))

    def event_times(self):
        """(sums of {step1..4, total} ms over the recorded pipeline calls, number of calls)."""
        sums = (_D * 5)()
        n = _I(0)
        check(lib.oza_ctx_event_times(self.ctx, sums, C.byref(n)))
        return list(sums), n.value

    def close(self) -> None:
        if self.ctx:
            lib.oza_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class _Workspace:
    scheme = 0

    def __init__(self, ex: HipExecutor, M: int, N: int, K: int, slices: int):
        if ex is None:
            raise ValueError("OzaBLAS: Executor cannot be null.")
        self.ex, self.M, self.N, self.K, self.slices = ex, M, N, K, slices
        self.bytes = workspace_bytes(self.scheme, M, N, K, slices)
        self.shift_A, self.shift_B, self.A_slices, self.B_slices, self.C_tc = (ex.alloc(b) for b in self.bytes)

    def buffers(self):
        return (self.shift_A.ptr, self.shift_B.ptr, self.A_slices.ptr, self.B_slices.ptr, self.C_tc.ptr)

    def shifts(self):
        return (self.shift_A.download((self.M,), np.int32), self.shift_B.download((self.N,), np.int32))

    def planar_slices(self):
        """(A [S][M][K], B [S][K][N]) int8 in the reference's planar layout."""
        S, M, N, K = self.slices, self.M, self.N, self.K
        ap, bp = self.ex.alloc(S * M * K), self.ex.alloc(S * K * N)
        check(lib.oza_unpack_slices(self.ex.ctx, M, N, K, S, self.A_slices.ptr, self.B_slices.ptr, ap.ptr, bp.ptr))
        return ap.download((S, M, K), np.int8), bp.download((S, K, N), np.int8)

    def free(self):
        for b in (self.shift_A, self.shift_B, self.A_slices, self.B_slices, self.C_tc):
            b.free()


class WorkspaceScheme1(_Workspace):
    scheme = 1


class WorkspaceScheme2(_Workspace):
    scheme = 2


def _gemm(fn, ws, A, B, C_out, timings):
    t = oza_timings()
    check(fn(ws.ex.ctx, ws.M, ws.N, ws.K, ws.slices, A.ptr, B.ptr, C_out.ptr, *ws.buffers(),
             C.byref(t) if timings is not None else None))
    if timings is not None:
        timings._fill(t)


def ozaki_scheme1_gemm(ws: WorkspaceScheme1, A: DeviceBuffer, B: DeviceBuffer, C_out: DeviceBuffer,
                       timings: OzaTimings | None = None) -> None:
    _gemm(lib.oza_scheme1_gemm, ws, A, B, C_out, timings)


def ozaki_scheme2_gemm(ws: WorkspaceScheme2, A: DeviceBuffer, B: DeviceBuffer, C_out: DeviceBuffer,
                       timings: OzaTimings | None = None) -> None:
    _gemm(lib.oza_scheme2_gemm, ws, A, B, C_out, timings)


def _shape_arrays(shapes):
    n = len(shapes)
    return ((_SZ * n)(*[sh[0] for sh in shapes]), (_SZ * n)(*[sh[1] for sh in shapes]),
            (_SZ * n)(*[sh[2] for sh in shapes]), (_I * n)(*[sh[3] for sh in shapes]))


def group_plan(shapes, cost, cus: int, slots, tiles=None, splits=None):
    """The grouped planner on the host (oza_group_plan): for (M, N, K, S) shapes on a device with
    tile cost model ``cost`` (6 numbers), ``cus`` CUs and ``slots`` resident workgroups per tile
    shape (256, 128, 64), returns (tiles, splits, workspace bytes).  ``tiles`` / ``splits`` entries
    of 0 (or None) are the planner's choice."""
    shapes = [tuple(int(v) for v in sh) for sh in shapes]
    n = len(shapes)
    t = (_I * n)(*(tiles or [0] * n))
    s = (_I * n)(*(splits or [0] * n))
    b = _SZ(0)
    check(lib.oza_group_plan((_D * 6)(*cost), int(cus), (_I * 3)(*slots), n, *_shape_arrays(shapes), t, s,
                             C.byref(b)))
    return list(t), list(s), b.value


class GroupScheme2:
    """A planned batch of independent Scheme II GEMMs (oza_group_*), each with its own shape and
    moduli count.  ``shapes`` is a list of (M, N, K, S); ``tile`` = 0 lets the context's cost model
    choose each problem's GEMM tile edge (256, 128 or 64).  ``tiles`` / ``splits`` (one entry per
    problem, 0 = the planner's choice) fix a problem's tile edge and K split
    (oza_group_create_split)."""

    def __init__(self, ex: HipExecutor, shapes, tile: int = 0, tiles=None, splits=None):
        self.ex = ex
        self.shapes = [tuple(int(v) for v in sh) for sh in shapes]
        n = len(self.shapes)
        p = _P()
        self.handle = None
        if tiles is None and splits is None:
            check(lib.oza_group_create(ex.ctx, n, *_shape_arrays(self.shapes), tile, C.byref(p)))
        else:
            t = (_I * n)(*(tiles or [tile] * n))
            s = (_I * n)(*(splits or [0] * n))
            check(lib.oza_group_create_split(ex.ctx, n, *_shape_arrays(self.shapes), t, s, C.byref(p)))
        self.handle = p.value

    def splits(self):
        """[K split of each problem's GEMM] (1: not split)."""
        s = (_I * len(self.shapes))()
        check(lib.oza_group_splits(self.handle, s))
        return list(s)

    def __len__(self):
        return len(self.shapes)

    def info(self):
        """(workspace bytes, [tile edge per problem])."""
        b = _SZ(0)
        tiles = (_I * len(self.shapes))()
        check(lib.oza_group_info(self.handle, C.byref(b), tiles))
        return b.value, list(tiles)

    def run(self, A, B, C_out, use_graph: bool = True, timings: OzaTimings | None = None) -> None:
        """A, B, C_out: sequences of DeviceBuffer (or raw device addresses), one per problem."""
        n = len(self.shapes)

        def addrs(xs):
            assert len(xs) == n
            return (_P * n)(*[x.ptr if isinstance(x, DeviceBuffer) else x for x in xs])

        t = oza_timings()
        check(lib.oza_group_scheme2_gemm(self.handle, addrs(A), addrs(B), addrs(C_out), int(use_graph),
                                         C.byref(t) if timings is not None else None))
        if timings is not None:
            timings._fill(t)

    def free(self) -> None:
        if self.handle:
            check(lib.oza_group_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            if self.handle:
                lib.oza_group_destroy(self.handle)
        except Exception:
            pass


def group_autotune(ex: HipExecutor) -> list:
    """Measure the grouped GEMM's tile cost model on the executor's device (see oza_group_autotune)."""
    c = (_D * 6)()
    check(lib.oza_group_autotune(ex.ctx, c))
    return list(c)


def group_costs(ex: HipExecutor, costs=None) -> list:
    """Read, or with ``costs`` replace, the context's tile cost model."""
    c = (_D * 6)(*(costs if costs is not None else [0.0] * 6))
    check(lib.oza_group_costs(ex.ctx, c, int(costs is not None)))
    return list(c)


class ServerScheme2:
    """Request server (oza_server_*): Scheme II GEMMs submitted one by one, executed in continuous
    batches on the executor's device from two resident workspace arenas of workspace_bytes / 2."""

    def __init__(self, ex: HipExecutor, workspace_bytes: int, max_batch: int = 64):
        self.ex = ex
        p = _P()
        self.handle = None
        check(lib.oza_server_create(ex.ctx, int(workspace_bytes), int(max_batch), C.byref(p)))
        self.handle = p.value

    def submit(self, M: int, N: int, K: int, S: int, A, B, C_out) -> int:
        """Queue C_out = A B (device buffers or addresses); returns a ticket."""
        t = C.c_uint64(0)
        addr = [x.ptr if isinstance(x, DeviceBuffer) else x for x in (A, B, C_out)]
        check(lib.oza_server_submit(self.handle, M, N, K, S, addr[0], addr[1], addr[2], C.byref(t)))
        return t.value

    def wait(self, ticket: int) -> float:
        """Block until the request is complete; returns its latency in ms."""
        lat = _D(0.0)
        check(lib.oza_server_wait(self.handle, ticket, C.byref(lat)))
        return lat.value

    def drain(self) -> None:
        check(lib.oza_server_drain(self.handle))

    def config(self, large_ops: float, large_gemm_share: float = 1.0) -> None:
        """Requests with 2 M N K S INT8 ops above large_ops are large (own batches, one large batch
        at a time, its GEMM on large_gemm_share of the CUs); applies to later submissions."""
        check(lib.oza_server_config(self.handle, float(large_ops), float(large_gemm_share)))

    def stats(self) -> dict:
        s = (_D * 4)()
        check(lib.oza_server_stats(self.handle, s))
        return {"batches": int(s[0]), "served": int(s[1]), "mean_batch": s[2], "max_batch": int(s[3])}

    def close(self) -> None:
        if self.handle:
            check(lib.oza_server_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            if self.handle:
                lib.oza_server_destroy(self.handle)
        except Exception:
            pass


ROW_MAJOR, COL_MAJOR = 0, 1
OP_N, OP_T = 0, 1


def gemm_workspace_bytes(scheme: int, M: int, N: int, K: int, slices: int) -> int:
    b = _SZ(0)
    check(lib.oza_gemm_workspace_bytes(scheme, M, N, K, slices, C.byref(b)))
    return b.value


def gemm(ex: HipExecutor, M: int, N: int, K: int, alpha: float, A, lda: int, B, ldb: int, beta: float, C_out,
         ldc: int, transA: int = OP_N, transB: int = OP_N, layout: int = ROW_MAJOR, scheme: int = 2,
         slices: int = 16, workspace=None) -> None:
    """BLAS-style C = alpha op(A) op(B) + beta C (oza_gemm); A, B, C_out, workspace are device
    buffers or addresses.  Asynchronous on the executor's stream."""
    def addr(x):
        return x.ptr if isinstance(x, DeviceBuffer) else x

    wsz = workspace.nbytes if isinstance(workspace, DeviceBuffer) else 0
    check(lib.oza_gemm(ex.ctx, scheme, slices, layout, transA, transB, M, N, K, alpha, addr(A), lda, addr(B), ldb,
                       beta, addr(C_out), ldc, addr(workspace), wsz))
