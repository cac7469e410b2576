"""Synthetic workloads of the batched and served forms, with fixed seeds.

Batched (512 GEMMs): M, N, K each round(exp(U(ln 64, ln 2048))), phi ~ U(0.5, 2), accuracy target
digits ~ U{10..15} (componentwise, relative to |A||B|) mapped to a moduli count by
scheme2_slices_for_accuracy with the problem's phi as its spread.
Served: M, N, K log-uniform on [128, 8192], phi ~ U(0.5, 2) (optionally moved to the nearest of
the phi values the input pools were generated at), digits in {8, 11, 14}, Poisson arrivals.
Matrix values follow the reference generator (utils/matrix_gen.hpp:15-34).
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np

BATCH_SEED = 20261202
SERVE_SEED = 20261203


def _log_uniform(rng, lo, hi, size):
    return np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), size))).astype(np.int64)


def batch_shapes(n=512, seed=BATCH_SEED, lo=64, hi=2048, digits=(10, 15)):
    """[(M, N, K, S, phi, digits)] of the batched workload."""
    from . import scheme2_slices_for_accuracy

    rng = np.random.default_rng(seed)
    M, N, K = (_log_uniform(rng, lo, hi, n) for _ in range(3))
    phi = rng.uniform(0.5, 2.0, n)
    d = rng.integers(digits[0], digits[1] + 1, n)
    return [(int(M[i]), int(N[i]), int(K[i]), scheme2_slices_for_accuracy(int(K[i]), float(d[i]), float(phi[i])),
             float(phi[i]), int(d[i])) for i in range(n)]


def serve_requests(n, rate, seed=SERVE_SEED, lo=128, hi=8192, digits=(8, 11, 14), phi_levels=None):
    """[(arrival_s, M, N, K, S, phi, digits)] with Poisson arrivals at `rate` per second.
    phi_levels: when the inputs come from pools generated at a few phi values, each request's phi
    is moved to the nearest level, so that its moduli count is chosen for the data it multiplies."""
    from . import scheme2_slices_for_accuracy

    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    M, N, K = (_log_uniform(rng, lo, hi, n) for _ in range(3))
    phi = rng.uniform(0.5, 2.0, n)
    if phi_levels is not None:
        lv = np.asarray(phi_levels, dtype=np.float64)
        phi = lv[np.argmin(np.abs(phi[:, None] - lv[None, :]), axis=1)]
    d = rng.choice(np.array(digits), n)
    return [(float(t[i]), int(M[i]), int(N[i]), int(K[i]),
             scheme2_slices_for_accuracy(int(K[i]), float(d[i]), float(phi[i])), float(phi[i]), int(d[i]))
            for i in range(n)]


def generate_inputs(shapes, seed0=1000, threads=8):
    """Reference-generator A (M x K) and B (K x N) per shape (host arrays), in parallel."""
    from . import generate_matrix

    def one(i):
        M, N, K, _, phi = shapes[i][:5]
        return generate_matrix(M, K, phi, seed0 + 2 * i), generate_matrix(K, N, phi, seed0 + 2 * i + 1)

    with ThreadPoolExecutor(threads) as pool:
        return list(pool.map(one, range(len(shapes))))
