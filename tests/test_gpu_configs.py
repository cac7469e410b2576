"""The benchmark configurations at their full sizes, checked against the oracle.

A block C[I, J] of the product is computed exactly by the oracle from A[I, :] and B[:, J] with the
full K: step 1's row exponents depend only on whole rows of A and its column exponents only on whole
columns of B, and every later step is elementwise in (i, j) given those.  So the oracle's product of
the sampled rows and columns must equal the sampled block of the GPU result bit for bit.  Samples
take the first and last rows and columns, the rows around row-band and tile boundaries, and random
ones.

  C2        Scheme II 4096^3, phi = 1, S = 4, 12, 17, 19, 20 (19 and 20 take the wide residue path)
  headline  Scheme II 8192^3, phi = 1, S = 17 (the equal-error count), automatic row bands
  C3        Scheme I 8192^3, phi = 2 with S = 10 and phi = 4 with S = 6
  C4        the seeded batch of 512 GEMMs (M, N, K log-uniform [64, 2048]) as one grouped pass,
            eager and graph-replayed: every problem equals its single call, 16 sampled problems
            equal the oracle
  C5        32 requests of the served stream (M, N, K log-uniform [128, 8192], digits {8, 11, 14})
            through the request server: every result equals the oracle on a block

All calls go through the C ABI."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oz():
    import ozablas_amd as oz

    return oz


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def sample(n, rng, edges=(), k=12):
    idx = set(range(min(8, n))) | set(range(max(0, n - 8), n))
    for e in edges:
        idx |= {i for i in range(e - 3, e + 3) if 0 <= i < n}
    idx |= set(rng.choice(n, size=min(k, n), replace=False).tolist())
    return np.array(sorted(idx))


def gen_pair(M, N, K, phi, seeds=(42, 43)):
    oz = _oz()
    out = {}

    def work(key, r, c, seed):
        out[key] = oz.generate_matrix(r, c, phi, seed)

    th = [threading.Thread(target=work, args=("A", M, K, seeds[0])),
          threading.Thread(target=work, args=("B", K, N, seeds[1]))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    return out["A"], out["B"]


def check_block(orc, scheme, A, B, C, S, I, J, what):
    Ab = np.ascontiguousarray(A[I])
    Bb = np.ascontiguousarray(B[:, J])
    want = orc.s2_gemm(Ab, Bb, S) if scheme == 2 else orc.s1_gemm(Ab, Bb, S)
    got = C[np.ix_(I, J)]
    bad = np.count_nonzero(bits(got) != bits(want))
    assert bad == 0, f"{what}: {bad} of {got.size} sampled entries differ from the oracle"


def run_single(ex, scheme, dA, dB, M, N, K, S):
    oz = _oz()
    ws = (oz.WorkspaceScheme2 if scheme == 2 else oz.WorkspaceScheme1)(ex, M, N, K, S)
    dC = ex.alloc(8 * M * N)
    (oz.ozaki_scheme2_gemm if scheme == 2 else oz.ozaki_scheme1_gemm)(ws, dA, dB, dC)
    ex.synchronize()
    ws.free()
    C = dC.download((M, N), np.float64)
    dC.free()
    return C


@pytest.fixture(scope="module")
def pair4096():
    return gen_pair(4096, 4096, 4096, 1.0)


@pytest.mark.parametrize("S", [4, 12, 17, 19, 20])
def test_c2_scheme2_4096(ex, orc, pair4096, S):
    A, B = pair4096
    n = 4096
    ex.set_bands(0)
    dA, dB = ex.to_device(A), ex.to_device(B)
    C = run_single(ex, 2, dA, dB, n, n, n, S)
    dA.free()
    dB.free()
    rng = np.random.default_rng(S)
    edges = [256 * t for t in range(1, 16)] if ex.bands_for(n, n, n, S) > 1 else [256, 2048]
    check_block(orc, 2, A, B, C, S, sample(n, rng, edges), sample(n, rng, [256, 2048]), f"4096^3 S={S}")


@pytest.fixture(scope="module")
def pairs8192():
    return {phi: gen_pair(8192, 8192, 8192, phi) for phi in (1.0, 2.0, 4.0)}


def test_headline_scheme2_8192_s17(ex, orc, pairs8192):
    A, B = pairs8192[1.0]
    n = 8192
    ex.set_bands(0)
    nb = ex.bands_for(n, n, n, 17)
    dA, dB = ex.to_device(A), ex.to_device(B)
    C = run_single(ex, 2, dA, dB, n, n, n, 17)
    dA.free()
    dB.free()
    rng = np.random.default_rng(17)
    band_edges = [n * b // nb // 256 * 256 for b in range(1, nb)]
    check_block(orc, 2, A, B, C, 17, sample(n, rng, band_edges + [256, 4352]), sample(n, rng, [256, 4096]),
                f"8192^3 S=17, {nb} bands")


@pytest.mark.parametrize("phi,S", [(2.0, 10), (4.0, 6)])
def test_c3_scheme1_8192(ex, orc, pairs8192, phi, S):
    A, B = pairs8192[phi]
    n = 8192
    ex.set_bands(0)
    nb = ex.bands_for(n, n, n, S, scheme=1)
    dA, dB = ex.to_device(A), ex.to_device(B)
    C = run_single(ex, 1, dA, dB, n, n, n, S)
    dA.free()
    dB.free()
    rng = np.random.default_rng(100 + S)
    band_edges = [n * b // nb // 256 * 256 for b in range(1, nb)]
    check_block(orc, 1, A, B, C, S, sample(n, rng, band_edges), sample(n, rng, [256]),
                f"Scheme I 8192^3 phi={phi} S={S}")


def test_c4_batch_512(ex, orc):
    oz = _oz()
    from ozablas_amd import workloads

    ex.set_bands(0)
    shapes = workloads.batch_shapes()
    data = workloads.generate_inputs(shapes)
    probs = [(M, N, K, S) for (M, N, K, S, _, _) in shapes]
    dAs = [ex.to_device(A) for A, _ in data]
    dBs = [ex.to_device(B) for _, B in data]
    g = oz.GroupScheme2(ex, probs)
    eager = [ex.alloc(8 * M * N) for (M, N, _, _) in probs]
    g.run(dAs, dBs, eager, use_graph=False)
    graphed = [ex.alloc(8 * M * N) for (M, N, _, _) in probs]
    for _ in range(3):  # captured on the first run, replayed after
        g.run(dAs, dBs, graphed, use_graph=True)
    ex.synchronize()
    g.free()
    rng = np.random.default_rng(512)
    check = set(rng.choice(len(probs), size=16, replace=False).tolist())
    for i, ((M, N, K, S), dA, dB) in enumerate(zip(probs, dAs, dBs)):
        C = eager[i].download((M, N), np.float64)
        assert np.array_equal(bits(graphed[i].download((M, N), np.float64)), bits(C)), f"problem {i}: graph"
        assert np.array_equal(bits(run_single(ex, 2, dA, dB, M, N, K, S)), bits(C)), f"problem {i}: single call"
        if i in check:
            A, B = data[i]
            check_block(orc, 2, A, B, C, S, sample(M, rng, [64, 128, 256]), sample(N, rng, [64, 128, 256]),
                        f"problem {i} {M}x{N}x{K} S={S}")
        eager[i].free()
        graphed[i].free()


def test_c5_served_stream(ex, orc):
    oz = _oz()
    from ozablas_amd import workloads

    reqs = workloads.serve_requests(32, rate=2000.0)
    shapes = [(M, N, K, S, phi) for (_, M, N, K, S, phi, _) in reqs]
    data = workloads.generate_inputs(shapes, seed0=7000)
    bufs = [(ex.to_device(A), ex.to_device(B), ex.alloc(8 * M * N)) for (A, B), (M, N, _, _, _) in zip(data, shapes)]
    srv = oz.ServerScheme2(ex, 12 << 30, 16)
    tickets = [srv.submit(M, N, K, S, a, b, c) for (M, N, K, S, _), (a, b, c) in zip(shapes, bufs)]
    for t in tickets:
        srv.wait(t)
    srv.close()
    rng = np.random.default_rng(32)
    for i, ((M, N, K, S, _), (A, B), (a, b, c)) in enumerate(zip(shapes, data, bufs)):
        C = c.download((M, N), np.float64)
        check_block(orc, 2, A, B, C, S, sample(M, rng, [256]), sample(N, rng, [256]), f"request {i} {M}x{N}x{K} S={S}")
        a.free()
        b.free()
        c.free()
