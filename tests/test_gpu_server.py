"""Request server (oza_server_*): results served in a batch mix equal the single-problem pipeline
bit for bit (alone-vs-mix), admission respects the workspace budget and batch limit, and the
server drains on shutdown.  All calls go through the C ABI via ozablas_amd._lib."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REQS = [(1, 1, 1, 4), (7, 13, 5, 1), (33, 17, 64, 8), (129, 257, 300, 12), (64, 64, 128, 20),
        (200, 70, 1001, 16), (256, 256, 256, 9), (65, 300, 129, 14), (96, 80, 200, 3), (300, 1, 50, 18),
        (1, 300, 77, 7), (513, 385, 700, 11)]


def _oz():
    import ozablas_amd as oz

    return oz


def single(ex, dA, dB, M, N, K, S):
    oz = _oz()
    ws = oz.WorkspaceScheme2(ex, M, N, K, S)
    dC = ex.alloc(8 * M * N)
    oz.ozaki_scheme2_gemm(ws, dA, dB, dC)
    ex.synchronize()
    out = dC.download((M, N), np.float64)
    ws.free()
    return out


def same_bits(x, y):
    return np.array_equal(np.ascontiguousarray(x).view(np.uint64), np.ascontiguousarray(y).view(np.uint64))


def prepare(ex, orc, reqs, seed):
    out = []
    for i, (M, N, K, S) in enumerate(reqs):
        A = orc.generate(M, K, 0.5 + 0.5 * (i % 4), seed + 2 * i)
        B = orc.generate(K, N, 0.5 + 0.5 * (i % 4), seed + 2 * i + 1)
        out.append((ex.to_device(A), ex.to_device(B), ex.alloc(8 * M * N)))
    return out


@pytest.mark.parametrize("budget,max_batch", [(1 << 30, 64), (48 << 20, 3)])
def test_server_mix_equals_alone(ex, orc, budget, max_batch):
    oz = _oz()
    bufs = prepare(ex, orc, REQS, 300)
    srv = oz.ServerScheme2(ex, budget, max_batch)
    tickets = [srv.submit(M, N, K, S, a, b, c) for (M, N, K, S), (a, b, c) in zip(REQS, bufs)]
    lat = [srv.wait(t) for t in tickets]
    assert all(x >= 0 for x in lat)
    st = srv.stats()
    assert st["served"] == len(REQS) and st["max_batch"] <= max_batch
    if max_batch == 3:
        assert st["batches"] >= len(REQS) // 3
    for (M, N, K, S), (a, b, c) in zip(REQS, bufs):
        got = c.download((M, N), np.float64)
        assert same_bits(got, single(ex, a, b, M, N, K, S)), (M, N, K, S)
    srv.close()


def test_server_concurrent_submitters(ex, orc):
    """Several host threads submit and wait at once; every result is still exact."""
    oz = _oz()
    reqs = REQS * 3
    bufs = prepare(ex, orc, reqs, 700)
    srv = oz.ServerScheme2(ex, 1 << 30, 8)
    errors = []

    def client(idx):
        try:
            for i in idx:
                M, N, K, S = reqs[i]
                a, b, c = bufs[i]
                srv.wait(srv.submit(M, N, K, S, a, b, c))
        except Exception as e:  # collected and asserted below
            errors.append(e)

    threads = [threading.Thread(target=client, args=(range(j, len(reqs), 4),)) for j in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for (M, N, K, S), (a, b, c) in zip(reqs, bufs):
        assert same_bits(c.download((M, N), np.float64), single(ex, a, b, M, N, K, S))
    srv.close()


def test_server_rejects_and_drains(ex, orc):
    oz = _oz()
    srv = oz.ServerScheme2(ex, 64 << 20, 4)
    big = ex.alloc(8)
    with pytest.raises(ValueError):  # needs far more than one 32 MiB arena
        srv.submit(4096, 4096, 4096, 16, big, big, big)
    with pytest.raises(ValueError):
        srv.submit(8, 8, 8, 0, big, big, big)
    with pytest.raises(ValueError):
        srv.wait(12345)
    srv.drain()  # nothing queued: returns at once
    bufs = prepare(ex, orc, REQS[:6], 40)
    for (M, N, K, S), (a, b, c) in zip(REQS[:6], bufs):
        srv.submit(M, N, K, S, a, b, c)
    srv.drain()
    srv.close()  # tickets never waited are released with the server
    for (M, N, K, S), (a, b, c) in zip(REQS[:6], bufs):
        assert same_bits(c.download((M, N), np.float64), single(ex, a, b, M, N, K, S))


@pytest.mark.parametrize("share", [0.75, 0.25])
def test_server_large_and_small_lanes(ex, orc, share):
    """With a low threshold a third of the requests are large: they run in batches of their own on
    one arena (their GEMM on a share of the CUs) while small batches pass them on the other.  Every
    result equals its single call, and the oracle on the smaller shapes."""
    oz = _oz()
    reqs = REQS * 2
    bufs = prepare(ex, orc, reqs, 900)
    srv = oz.ServerScheme2(ex, 1 << 30, 8)
    srv.config(2.0 * 256 * 256 * 256 * 9, share)  # (256, 256, 256, 9) and larger are large
    with pytest.raises(ValueError):
        srv.config(1e9, 0.0)
    tickets = [srv.submit(M, N, K, S, a, b, c) for (M, N, K, S), (a, b, c) in zip(reqs, bufs)]
    for t in tickets:
        srv.wait(t)
    srv.close()
    for i, ((M, N, K, S), (a, b, c)) in enumerate(zip(reqs, bufs)):
        got = c.download((M, N), np.float64)
        assert same_bits(got, single(ex, a, b, M, N, K, S)), (i, M, N, K, S)
        if i < len(REQS) and M * N * K <= 300 ** 3:
            A = orc.generate(M, K, 0.5 + 0.5 * (i % 4), 900 + 2 * i)
            B = orc.generate(K, N, 0.5 + 0.5 * (i % 4), 900 + 2 * i + 1)
            assert same_bits(got, orc.s2_gemm(A, B, S)), (i, M, N, K, S)


def test_server_long_k_requests_split(ex, orc):
    """Requests with few tiles and a long K run their GEMM in K parts (the split each would get
    alone), beside unsplit requests in the same batches: every result equals its single call, and
    the oracle on sampled rows and columns with the full K."""
    oz = _oz()
    reqs = [(256, 256, 32768, 16), (100, 60, 20000, 9), (129, 257, 300, 12), (96, 300, 12000, 5)]
    splits = [oz.group_plan([r], oz.group_costs(ex), 256, [256, 1024, 1024])[1][0] for r in reqs]
    assert max(splits) > 1, splits
    data = []
    for i, (M, N, K, S) in enumerate(reqs):
        A, B = orc.generate(M, K, 1.0, 900 + 2 * i), orc.generate(K, N, 1.0, 901 + 2 * i)
        data.append((A, B, ex.to_device(A), ex.to_device(B), ex.alloc(8 * M * N)))
    srv = oz.ServerScheme2(ex, 4 << 30, 8)
    tickets = [srv.submit(M, N, K, S, a, b, c) for (M, N, K, S), (_, _, a, b, c) in zip(reqs, data)]
    for t in tickets:
        srv.wait(t)
    srv.close()
    rng = np.random.default_rng(5)
    for (M, N, K, S), (A, B, a, b, c) in zip(reqs, data):
        got = c.download((M, N), np.float64)
        assert same_bits(got, single(ex, a, b, M, N, K, S)), (M, N, K, S)
        rows = np.unique(np.r_[0, M - 1, rng.choice(M, 6, replace=False)])
        cols = np.unique(np.r_[0, N - 1, rng.choice(N, 6, replace=False)])
        want = orc.s2_gemm(np.ascontiguousarray(A[rows]), np.ascontiguousarray(B[:, cols]), S)
        assert same_bits(got[np.ix_(rows, cols)], want), (M, N, K, S)
