"""Grouped Scheme II (oza_group_*): every problem of a batch gives the same bits as the
single-problem pipeline (oza_scheme2_gemm) and as the oracle, whatever tile shape the planner
picks and whoever else is in the batch; HIP-graph replays equal the eager run.

All calls go through the C ABI (include/ozablas_hip.h) via ozablas_amd._lib."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (M, N, K, S): ragged edges, K = 1, K past one 128-step, small and large moduli counts
MIX = [(1, 1, 1, 4), (7, 13, 5, 1), (33, 17, 64, 8), (129, 257, 300, 12), (64, 64, 128, 20),
       (200, 70, 1001, 16), (256, 256, 256, 9), (65, 300, 129, 14), (96, 80, 200, 3), (0, 5, 7, 4),
       (300, 1, 50, 18), (1, 300, 77, 7)]


def _oz():
    import ozablas_amd as oz

    return oz


def single(ex, A, B, S):
    oz = _oz()
    M, K = A.shape
    N = B.shape[1]
    ws = oz.WorkspaceScheme2(ex, M, N, K, S)
    dA, dB, dC = ex.to_device(A), ex.to_device(B), ex.alloc(8 * M * N)
    oz.ozaki_scheme2_gemm(ws, dA, dB, dC)
    ex.synchronize()
    out = dC.download((M, N), np.float64)
    ws.free()
    return out


def inputs(orc, shapes, seed0):
    out = []
    for i, (M, N, K, S) in enumerate(shapes):
        A = orc.generate(M, K, 0.5 + (i % 4) * 0.5, seed0 + 2 * i) if M and K else np.zeros((M, K))
        B = orc.generate(K, N, 0.5 + (i % 4) * 0.5, seed0 + 2 * i + 1) if N and K else np.zeros((K, N))
        out.append((A, B))
    return out


def run_group(ex, g, shapes, data, use_graph, dCs=None):
    dAs = [ex.to_device(A) for A, _ in data]
    dBs = [ex.to_device(B) for _, B in data]
    if dCs is None:
        dCs = [ex.alloc(max(8 * M * N, 8)) for (M, N, _, _) in shapes]
    g.run(dAs, dBs, dCs, use_graph=use_graph)
    ex.synchronize()
    return [c.download((M, N), np.float64) for c, (M, N, _, _) in zip(dCs, shapes)]


def same_bits(x, y):
    return np.array_equal(np.ascontiguousarray(x).view(np.uint64), np.ascontiguousarray(y).view(np.uint64))


@pytest.mark.parametrize("tile", [0, 256, 128, 64])
def test_group_matches_single_and_oracle(ex, orc, tile):
    oz = _oz()
    data = inputs(orc, MIX, 100)
    g = oz.GroupScheme2(ex, MIX, tile=tile)
    _, tiles = g.info()
    if tile:
        assert all(t == tile for t in tiles)
    outs = run_group(ex, g, MIX, data, use_graph=False)
    for (M, N, K, S), (A, B), C in zip(MIX, data, outs):
        if M == 0 or N == 0:
            continue
        assert same_bits(C, single(ex, A, B, S)), (M, N, K, S, tile)
        if M * N * K <= 300 * 300 * 300:
            assert same_bits(C, orc.s2_gemm(A, B, S)), (M, N, K, S, tile)
    g.free()


@pytest.mark.parametrize("tile", [256, 128, 64])
def test_group_chained_items(ex, orc, tile):
    """Many more GEMM items than resident workgroups (1280 items at edge 256 on 256 CUs, 5120 at
    128, 20480 at 64), so each workgroup runs a chain of items through one LDS ring, crossing from
    items of 3 K steps to 2 and to 1 (K = 300, 129, 100, 1); every problem equals its single call,
    and the oracle on a sample."""
    oz = _oz()
    ks = [300, 129, 100, 1]
    shapes = [(256, 256, ks[i % 4], 20) for i in range(64)]
    data = inputs(orc, shapes, 3000)
    g = oz.GroupScheme2(ex, shapes, tile=tile)
    outs = run_group(ex, g, shapes, data, use_graph=False)
    g.free()
    for i, ((M, N, K, S), (A, B), C) in enumerate(zip(shapes, data, outs)):
        assert same_bits(C, single(ex, A, B, S)), (i, K, tile)
        if i < 4:
            assert same_bits(C, orc.s2_gemm(A, B, S)), (i, K, tile)


def test_group_alone_equals_mixed(ex, orc):
    oz = _oz()
    data = inputs(orc, MIX, 500)
    g = oz.GroupScheme2(ex, MIX, tile=128)
    mixed = run_group(ex, g, MIX, data, use_graph=False)
    g.free()
    for i in [2, 5, 8]:
        alone = oz.GroupScheme2(ex, [MIX[i]], tile=128)
        out = run_group(ex, alone, [MIX[i]], [data[i]], use_graph=False)[0]
        assert same_bits(out, mixed[i])
        alone.free()


def test_group_graph_replays_equal_eager(ex, orc):
    """Captured once, replayed several times with eager library work and synchronizations in
    between; new pointers (a second input set) are uploaded before a replay."""
    oz = _oz()
    shapes = MIX[1:9]
    d1, d2 = inputs(orc, shapes, 900), inputs(orc, shapes, 1900)
    g = oz.GroupScheme2(ex, shapes)
    eager1 = run_group(ex, g, shapes, d1, use_graph=False)
    eager2 = run_group(ex, g, shapes, d2, use_graph=False)
    dCs = [ex.alloc(8 * M * N) for (M, N, _, _) in shapes]
    for rep in range(3):
        got = run_group(ex, g, shapes, d1, use_graph=True, dCs=dCs)
        assert all(same_bits(x, y) for x, y in zip(got, eager1)), rep
        single(ex, d1[0][0], d1[0][1], 4)  # unrelated eager work between replays
        got = run_group(ex, g, shapes, d2, use_graph=True)
        assert all(same_bits(x, y) for x, y in zip(got, eager2)), rep
    g.free()


def test_group_timings(ex, orc):
    oz = _oz()
    shapes = [(256, 256, 512, 10)] * 4
    data = inputs(orc, shapes, 7)
    g = oz.GroupScheme2(ex, shapes)
    dAs = [ex.to_device(A) for A, _ in data]
    dBs = [ex.to_device(B) for _, B in data]
    dCs = [ex.alloc(8 * 256 * 256) for _ in shapes]
    t = oz.OzaTimings()
    g.run(dAs, dBs, dCs, timings=t)
    assert t.total_ms > 0 and t.step3_ms > 0
    assert abs(t.step1_ms + t.step2_ms + t.step3_ms + t.step4_ms - t.total_ms) < 1e-2 + 1e-3 * t.total_ms
    got = dCs[0].download((256, 256), np.float64)
    assert same_bits(got, single(ex, data[0][0], data[0][1], 10))
    g.free()


def test_group_invalid_arguments(ex):
    oz = _oz()
    with pytest.raises(ValueError):
        oz.GroupScheme2(ex, [(8, 8, 8, 0)])
    with pytest.raises(ValueError):
        oz.GroupScheme2(ex, [(8, 8, 8, 21)])
    with pytest.raises(ValueError):
        oz.GroupScheme2(ex, [(8, 8, 200000, 4)])
    with pytest.raises(ValueError):
        oz.GroupScheme2(ex, [(8, 8, 8, 4)], tile=100)
    with pytest.raises(ValueError):
        oz.GroupScheme2(ex, [])
    with pytest.raises(ValueError):
        oz.GroupScheme2(ex, [(8, 8, 8, 4)], splits=[65])
    with pytest.raises(ValueError):
        oz.GroupScheme2(ex, [(8, 8, 8, 4)], tiles=[100])
    g = oz.GroupScheme2(ex, [(8, 8, 8, 4)])
    buf = ex.alloc(8 * 64)
    with pytest.raises(ValueError):
        g.run([buf], [0], [buf])
    g.free()


def test_group_autotune_picks_positive_costs(ex):
    oz = _oz()
    saved = oz.group_costs(ex)
    c = oz.group_autotune(ex)
    assert len(c) == 6 and all(v >= 0 for v in c) and all(c[2 * i + 1] > 0 for i in range(3))
    assert oz.group_costs(ex) == c
    oz.group_costs(ex, saved)


# (M, N, K, S, forced K split): the split is clamped to the 128-byte K blocks (K = 1 has one, K = 640
# five), uneven splits (8 blocks in 3 or 5), many splits, the wide residue path (S = 20), and
# unsplit problems beside split ones
SPLIT_MIX = [(96, 80, 1001, 16, 3), (33, 17, 640, 8, 64), (129, 257, 300, 12, 3), (7, 13, 1, 4, 4),
             (96, 64, 4096, 20, 16), (200, 70, 1001, 9, 5), (64, 64, 256, 11, 1), (65, 300, 129, 14, 2),
             (300, 1, 2000, 18, 8), (256, 256, 512, 7, 0)]


@pytest.mark.parametrize("tile", [256, 128, 64])
def test_group_split_k_matches_single_and_oracle(ex, orc, tile):
    """Forced K splits: every problem gives the bits of its single call and of the oracle, eager
    and graph-replayed (the partial digits are summed mod m_s before the reconstruction)."""
    oz = _oz()
    shapes = [sh[:4] for sh in SPLIT_MIX]
    data = inputs(orc, shapes, 7000)
    g = oz.GroupScheme2(ex, shapes, tiles=[tile] * len(shapes), splits=[sh[4] for sh in SPLIT_MIX])
    got_splits = g.splits()
    for (M, N, K, S, p), q in zip(SPLIT_MIX, got_splits):
        kb = (K + 127) // 128
        if p:
            assert q == -(-kb // -(-kb // min(p, kb))), (K, p, q)  # non-empty splits of ceil(kb / p) blocks
    assert max(got_splits) > 1
    eager = run_group(ex, g, shapes, data, use_graph=False)
    dCs = [ex.alloc(8 * M * N) for (M, N, _, _) in shapes]
    for rep in range(2):
        graphed = run_group(ex, g, shapes, data, use_graph=True, dCs=dCs)
        assert all(same_bits(x, y) for x, y in zip(graphed, eager)), rep
    g.free()
    for (M, N, K, S), (A, B), C in zip(shapes, data, eager):
        assert same_bits(C, single(ex, A, B, S)), (M, N, K, S, tile)
        assert same_bits(C, orc.s2_gemm(A, B, S)), (M, N, K, S, tile)


def test_group_planner_splits_a_lone_long_k_problem(ex, orc):
    """Planned alone, 256 x 256 x 32768 with 16 moduli (16 GEMM tiles for 256 CUs) is split; its
    result equals the single call, and the oracle on sampled rows and columns (step 1 depends only
    on whole rows of A and columns of B, so a block of C is the oracle's product of its rows and
    columns)."""
    oz = _oz()
    M, N, K, S = 256, 256, 32768, 16
    A, B = orc.generate(M, K, 1.0, 91), orc.generate(K, N, 1.0, 92)
    g = oz.GroupScheme2(ex, [(M, N, K, S)])
    assert g.splits()[0] > 1
    C = run_group(ex, g, [(M, N, K, S)], [(A, B)], use_graph=True)[0]
    g.free()
    assert same_bits(C, single(ex, A, B, S))
    rows, cols = np.array([0, 1, 31, 32, 100, 200, 254, 255]), np.array([0, 5, 63, 64, 127, 128, 250, 255])
    want = orc.s2_gemm(np.ascontiguousarray(A[rows]), np.ascontiguousarray(B[:, cols]), S)
    assert same_bits(C[np.ix_(rows, cols)], want)


def test_group_fused_and_two_pass_operands_special_values(ex, orc):
    """Operands with zero rows and columns, subnormals and 1e300 entries, ragged and odd K: with
    Kp <= 2048 the grouped pipeline takes A's exponent maxima and residues in one pass, above it in
    two (K = 2047 -> Kp 2048 and K = 2100 -> Kp 2176 sit either side of the switch).  Every problem
    must carry the bits of its single call and of the oracle."""
    from test_gpu_parity import special_matrix

    oz = _oz()
    rng = np.random.default_rng(77)
    shapes = [(70, 90, 129, 8), (37, 41, 2047, 18), (33, 65, 2100, 12), (8, 300, 1, 4), (129, 7, 640, 20),
              (300, 50, 1000, 16), (40, 33, 2176, 5)]
    data = [(special_matrix(rng, M, K), special_matrix(rng, K, N)) for M, N, K, _ in shapes]
    g = oz.GroupScheme2(ex, shapes)
    outs = run_group(ex, g, shapes, data, use_graph=False)
    g.free()
    for (M, N, K, S), (A, B), C in zip(shapes, data, outs):
        assert same_bits(C, single(ex, A, B, S)), (M, N, K, S)
        assert same_bits(C, orc.s2_gemm(A, B, S)), (M, N, K, S)
