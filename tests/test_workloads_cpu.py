"""The seeded synthetic workloads of the batched and served forms (ozablas_amd.workloads): fixed
seeds give fixed shapes, sizes stay in their ranges, and accuracy targets map to moduli counts
through the library's own digits -> S rule.  Host only."""
import math

import numpy as np


def test_batch_shapes_are_seeded_and_in_range():
    from ozablas_amd import scheme2_kparam, workloads

    a, b = workloads.batch_shapes(), workloads.batch_shapes()
    assert a == b and len(a) == 512
    M = np.array([x[0] for x in a])
    assert M.min() >= 64 and M.max() <= 2048
    # log-uniform: the median sits near the geometric mean of the bounds
    assert 250 < np.median(M) < 500
    # digits -> S: the smallest S whose K_param bits cover the target plus the componentwise
    # shortfall measured at the problem's phi (piecewise linear between the calibration points)
    xs, ys = [0, 0.5, 1, 1.5, 2, 3, 4], [0.8, 1.3, 2.2, 3.5, 4.8, 7.2, 9.3]
    for Mi, Ni, Ki, S, phi, d in a:
        assert 0.5 <= phi <= 2.0 and 10 <= d <= 15
        need = d + float(np.interp(phi, xs, ys))
        assert scheme2_kparam(S, Ki) * math.log10(2) >= need - 1e-9
        assert S == 1 or scheme2_kparam(S - 1, Ki) * math.log10(2) < need
    assert workloads.batch_shapes(seed=1) != a


def test_serve_requests_poisson_and_digits():
    from ozablas_amd import workloads

    r = workloads.serve_requests(2000, 500.0)
    t = np.array([x[0] for x in r])
    assert np.all(np.diff(t) > 0)
    # mean inter-arrival 1/rate within a few percent
    assert abs(np.mean(np.diff(t)) * 500.0 - 1.0) < 0.1
    assert {x[6] for x in r} == {8, 11, 14}
    K = np.array([x[3] for x in r])
    assert K.min() >= 128 and K.max() <= 8192


def test_serve_requests_phi_levels():
    """With input pools at a few phi values, every request's phi is the nearest pool value and its
    moduli count is the rule's count at that phi (the data it multiplies), not at the drawn phi."""
    from ozablas_amd import scheme2_slices_for_accuracy, workloads

    levels = [0.5, 0.75, 1.0, 1.25, 1.5, 1.75, 2.0]
    free = workloads.serve_requests(400, 100.0)
    snapped = workloads.serve_requests(400, 100.0, phi_levels=levels)
    for (t0, M0, N0, K0, _, p0, d0), (t1, M1, N1, K1, S1, p1, d1) in zip(free, snapped):
        assert (t0, M0, N0, K0, d0) == (t1, M1, N1, K1, d1)  # same draws
        assert p1 in levels and abs(p1 - p0) <= 0.125 + 1e-12
        assert S1 == scheme2_slices_for_accuracy(K1, float(d1), p1)
