"""The C-ABI library on the host: it loads, exports every entry point the headers declare, and its
host-only functions (generator, parameters, workspace sizing, argument checks) agree with the
oracle and the reference's rules.  No GPU kernels run here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ozablas_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oza_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ozablas_amd as oz

    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(oz.lib, n), n
        getattr(oz.lib, n)  # resolves the symbol
    # and the Python binding covers all of them
    from ozablas_amd import _lib

    assert set(names) <= set(_lib.EXPORTED), set(names) - set(_lib.EXPORTED)


def test_cpp_api_library_exports():
    lib = os.path.join(ROOT, "ozablas_amd", "lib", "libozablas.so")
    assert os.path.exists(lib)
    out = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    for sym in ["ozablas::ozaki_scheme1_gemm(ozablas::WorkspaceScheme1&, double const*, double const*, double*, ozablas::OzaTimings*)",
                "ozablas::ozaki_scheme2_gemm(ozablas::WorkspaceScheme2&, double const*, double const*, double*, ozablas::OzaTimings*)",
                "ozablas::HipExecutor::HipExecutor(int)",
                "ozablas::WorkspaceScheme1::WorkspaceScheme1(std::shared_ptr<ozablas::Executor const>, unsigned long, unsigned long, unsigned long, int)",
                "ozablas::WorkspaceScheme2::WorkspaceScheme2(std::shared_ptr<ozablas::Executor const>, unsigned long, unsigned long, unsigned long, int)",
                "ozablas::CpuExecutor::allocate(void**, unsigned long) const",
                "ozablas::GroupScheme2::GroupScheme2(std::shared_ptr<ozablas::Executor const>, "
                "std::vector<ozablas::GemmShape, std::allocator<ozablas::GemmShape> >, int)",
                "ozablas::autotune_grouped(std::shared_ptr<ozablas::Executor const> const&)"]:
        assert sym in out, sym


def test_version_and_status_strings():
    import ozablas_amd as oz

    assert b"gfx950" in oz.lib.oza_version()
    for s in range(5):
        assert oz.lib.oza_status_string(s)


def test_generator_matches_oracle(orc):
    import ozablas_amd as oz

    for rows, cols, phi, seed in [(256, 256, 0.5, 42), (31, 77, 2.0, 9)]:
        a = oz.generate_matrix(rows, cols, phi, seed)
        b = orc.generate(rows, cols, phi, seed)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_parameters_match_oracle(orc):
    import ozablas_amd as oz

    for S in range(1, 21):
        for K in [1, 3, 64, 256, 1000, 4096, 8192, 100000]:
            assert oz.scheme2_kparam(S, K) == orc.kparam(S, K)
    for K in [1, 2, 100, 4096, 8192, 2 ** 17, 2 ** 17 + 1, 2 ** 20]:
        assert oz.scheme1_beta(K) == orc.beta(K)


def test_workspace_sizing():
    import ozablas_amd as oz

    b = oz.workspace_bytes(2, 8192, 8192, 8192, 16)
    assert b == (4 * 8192, 4 * 8192, 16 * 8192 * 8192, 16 * 8192 * 8192, 16 * 8192 * 8192)
    b = oz.workspace_bytes(2, 100, 7, 33, 3)  # padded operand planes
    assert b[2] == 3 * 256 * 128 and b[3] == 3 * 256 * 128 and b[4] == 3 * 256 * 256
    b = oz.workspace_bytes(1, 100, 7, 33, 3)  # blocked INT32 group planes, Mp x Np
    assert b[4] == 4 * 3 * 256 * 256


@pytest.mark.parametrize("scheme,S,K", [(2, 0, 64), (2, 21, 64), (1, 0, 64), (2, 4, 0), (2, 4, 200000),
                                        (1, 10, 60000), (3, 4, 64)])
def test_invalid_arguments_rejected(scheme, S, K):
    import ozablas_amd as oz

    with pytest.raises(ValueError):
        oz.workspace_bytes(scheme, 64, 64, K, S)


def test_null_context_rejected():
    import ozablas_amd as oz
    from ozablas_amd import _lib

    assert oz.lib.oza_synchronize(None) == _lib.INVALID_ARGUMENT
    t = _lib.oza_timings()
    rc = oz.lib.oza_scheme2_gemm(None, 8, 8, 8, 4, None, None, None, None, None, None, None, None, C.byref(t))
    assert rc == _lib.INVALID_ARGUMENT
    assert oz.lib.oza_ctx_set_bands(None, 0) == _lib.INVALID_ARGUMENT


def test_context_creation_without_gpu_fails_cleanly():
    import ozablas_amd as oz

    if oz.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(oz.OzaError):
        oz.HipExecutor(0)


def test_slices_for_digits_is_smallest_sufficient(orc):
    import math

    import ozablas_amd as oz

    for K in [64, 1000, 2048, 8192]:
        for d in [4, 8, 10, 11, 12.5, 14, 15]:
            bits = math.ceil(d * math.log2(10))
            S = oz.scheme2_slices_for_digits(K, d)
            assert orc.kparam(S, K) >= bits
            assert S == 1 or orc.kparam(S - 1, K) < bits
    with pytest.raises(ValueError):
        oz.scheme2_slices_for_digits(8192, 40.0)


SHORTFALL = {0.0: 0.8, 0.5: 1.3, 1.0: 2.2, 1.5: 3.5, 2.0: 4.8, 3.0: 7.2, 4.0: 9.3}


def test_slices_for_accuracy_contract(orc):
    """The smallest S whose K_param covers the target plus the spread's shortfall; more digits or a
    wider spread never asks for fewer moduli; impossible targets and negative spreads fail."""
    import math

    import ozablas_amd as oz

    for K in [64, 1000, 8192]:
        for phi, short in SHORTFALL.items():
            for d in [4, 8, 11, 14]:
                need = d + short
                if orc.kparam(20, K) * math.log10(2) < need:
                    with pytest.raises(ValueError):
                        oz.scheme2_slices_for_accuracy(K, d, phi)
                    continue
                S = oz.scheme2_slices_for_accuracy(K, d, phi)
                assert orc.kparam(S, K) * math.log10(2) >= need - 1e-9
                assert S == 1 or orc.kparam(S - 1, K) * math.log10(2) < need
        prev = 0
        for phi in [0.0, 0.25, 0.75, 1.25, 2.5, 3.5]:  # between the table points
            S = oz.scheme2_slices_for_accuracy(K, 8, phi)
            assert S >= prev
            prev = S
    with pytest.raises(ValueError):
        oz.scheme2_slices_for_accuracy(1000, 8, -1.0)
    with pytest.raises(ValueError):
        oz.scheme2_slices_for_accuracy(1000, 30, 1.0)


@pytest.mark.parametrize("phi", [0.5, 1.0, 2.0, 4.0])
def test_slices_for_accuracy_delivers_componentwise_digits(orc, phi):
    """On the reference generator's inputs (fresh seeds, not those the shortfall table was measured
    on) the chosen moduli count reaches the requested digits componentwise against the FP128
    product: max |C - C_exact| / (|A||B|) <= 10^-digits."""
    import ozablas_amd as oz

    for K, d in [(64, 7), (500, 12), (2000, 9)]:
        try:
            S = oz.scheme2_slices_for_accuracy(K, d, phi)
        except ValueError:
            continue  # beyond 20 moduli
        A, B = orc.generate(128, K, phi, 1000 + K), orc.generate(K, 96, phi, 2000 + K)
        exact, absab = orc.gemm_fp128(A, B), orc.gemm_abs(A, B)
        C = orc.s2_gemm(A, B, S)
        cw = np.max(np.abs(C - exact) / np.where(absab > 0, absab, 1.0))
        assert cw <= 10.0 ** -d, (phi, K, d, S, cw)


def test_group_entry_points_reject_null_handles():
    from ozablas_amd import _lib

    L = _lib.lib
    n = (C.c_size_t * 1)(8)
    sl = (C.c_int * 1)(4)
    out = C.c_void_p()
    assert L.oza_group_create(None, 1, n, n, n, sl, 0, C.byref(out)) == _lib.INVALID_ARGUMENT
    assert L.oza_group_scheme2_gemm(None, None, None, None, 0, None) == _lib.INVALID_ARGUMENT
    assert L.oza_group_info(None, None, None) == _lib.INVALID_ARGUMENT
    assert L.oza_group_destroy(None) == _lib.SUCCESS
    assert L.oza_group_autotune(None, None) == _lib.INVALID_ARGUMENT
    assert L.oza_group_costs(None, None, 0) == _lib.INVALID_ARGUMENT


def test_server_entry_points_reject_null_handles():
    from ozablas_amd import _lib

    L = _lib.lib
    out = C.c_void_p()
    t = C.c_uint64()
    assert L.oza_server_create(None, 1 << 30, 8, C.byref(out)) == _lib.INVALID_ARGUMENT
    assert L.oza_server_submit(None, 8, 8, 8, 4, None, None, None, C.byref(t)) == _lib.INVALID_ARGUMENT
    assert L.oza_server_wait(None, 1, None) == _lib.INVALID_ARGUMENT
    assert L.oza_server_drain(None) == _lib.INVALID_ARGUMENT
    assert L.oza_server_stats(None, None) == _lib.INVALID_ARGUMENT
    assert L.oza_server_config(None, 1e12, 0.75) == _lib.INVALID_ARGUMENT
    assert L.oza_server_destroy(None) == _lib.SUCCESS


def test_gemm_front_end_host_checks():
    import ozablas_amd as oz
    from ozablas_amd import _lib

    b = oz.gemm_workspace_bytes(2, 100, 7, 33, 3)
    parts = oz.workspace_bytes(2, 100, 7, 33, 3)
    assert b == sum((x + 255) // 256 * 256 for x in parts)
    with pytest.raises(ValueError):
        oz.gemm_workspace_bytes(2, 8, 8, 8, 0)
    assert _lib.lib.oza_gemm(None, 2, 4, 0, 0, 0, 8, 8, 8, 1.0, None, 8, None, 8, 0.0, None, 8, None, 0) == \
        _lib.INVALID_ARGUMENT


COST = [6.26, 1.61, 1.55, 0.465, 0.68, 0.25]  # a measured tile cost model (CU-us per tile)
SLOTS = [256, 1024, 1024]


def test_group_planner_splits_only_what_pays():
    """The host planner (oza_group_plan): the 512-GEMM batch, whose items are short against the
    whole pass, stays (almost) unsplit; a lone long-K problem with few tiles is split; a lone large
    square problem is not; more K never gets fewer splits alone."""
    import ozablas_amd as oz
    from ozablas_amd import workloads

    shapes = [(M, N, K, S) for (M, N, K, S, _, _) in workloads.batch_shapes()]
    _, splits, _ = oz.group_plan(shapes, COST, 256, SLOTS)
    assert sum(s > 1 for s in splits) <= len(shapes) // 50
    _, s, _ = oz.group_plan([(256, 256, 65536, 16)], COST, 256, SLOTS)
    assert 8 <= s[0] <= 32
    _, s, _ = oz.group_plan([(8192, 8192, 8192, 17)], COST, 256, SLOTS)
    assert s == [1]
    prev = 1
    for K in [512, 2048, 8192, 32768]:
        _, s, _ = oz.group_plan([(128, 128, K, 12)], COST, 256, SLOTS)
        assert s[0] >= prev and (K + 127) // 128 >= 2 * s[0] or s[0] == 1
        prev = s[0]


def test_group_planner_forced_splits_and_bytes():
    """Forced splits are clamped to non-empty splits of ceil(blocks / p) K blocks; the workspace
    grows by the partial digit planes, (split - 1) x S x Mp x Np bytes (256-aligned)."""
    import ozablas_amd as oz

    def align(x):
        return (x + 255) // 256 * 256

    cases = [((7, 13, 1, 4), 4, 1), ((96, 80, 1001, 16), 3, 3), ((96, 80, 1001, 16), 5, 4),
             ((33, 17, 640, 8), 64, 5), ((300, 1, 2000, 18), 8, 8)]
    for shape, p, want in cases:
        t, s, b = oz.group_plan([shape], COST, 256, SLOTS, tiles=[128], splits=[p])
        t1, s1, b1 = oz.group_plan([shape], COST, 256, SLOTS, tiles=[128], splits=[1])
        assert t == [128] and s == [want] and s1 == [1], (shape, p, s)
        M, N, K, S = shape
        Mp, Np = -(-M // 128) * 128, -(-N // 128) * 128
        assert b - b1 == (align((want - 1) * S * Mp * Np) if want > 1 else 0), (shape, p)
    with pytest.raises(ValueError):
        oz.group_plan([(8, 8, 8, 4)], COST, 256, SLOTS, splits=[65])
    with pytest.raises(ValueError):
        oz.group_plan([(8, 8, 8, 4)], COST, 256, SLOTS, tiles=[96])
    with pytest.raises(ValueError):
        oz.group_plan([(8, 8, 8, 4)], COST, 0, SLOTS)


def test_reference_drivers_link_this_library():
    """The reference's benchmark and example programs, built unmodified against include/ by
    `make -C oracle drivers`, resolve the C++ API to this repository's libraries."""
    import glob
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    bins = sorted(glob.glob(os.path.join(root, "oracle", "_ref", "drivers", "*")))
    if not bins:
        pytest.skip("reference drivers not built (no reference checkout at build time)")
    lib = os.path.realpath(os.path.join(root, "ozablas_amd", "lib", "libozablas.so"))
    for b in bins:
        out = subprocess.run(["ldd", b], capture_output=True, text=True, check=True).stdout
        line = [l for l in out.splitlines() if "libozablas.so" in l]
        assert line, (b, out)
        assert os.path.realpath(line[0].split("=>")[1].split("(")[0].strip()) == lib, (b, line)
