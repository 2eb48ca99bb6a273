"""VEGAS map adaptation (oracle.vegas.refine and the product's
petite_amd.vegas_train.refine_nodes), pinned by what the algorithm fixes:
a map whose increments already carry equal importance does not move; the
equal-split rule on a two-increment grid; and repeated adaptation with the
exact per-increment importance of f(x) = exp(-x) converges to the map whose
J f is constant, the quantiles of the truncated exponential."""
import numpy as np

from oracle import vegas as ov
from petite_amd.vegas_train import refine_nodes


def test_uniform_importance_is_a_fixed_point():
    nodes = np.linspace(0.0, 2.0, 41) ** 1.5
    s = np.ones(40) * 3.0
    n = np.ones(40)
    for alpha in (0.0, 0.5, 1.5):
        np.testing.assert_allclose(ov.refine(nodes, s, n, alpha), nodes, rtol=0, atol=1e-14)
        np.testing.assert_allclose(refine_nodes(nodes, s, n, alpha), nodes, rtol=0, atol=1e-14)


def test_two_increments_equal_split():
    # importance 3 : 1 on [0, 1] and [1, 2], no smoothing or damping: the new
    # node sits where 2 of the total 4 is reached, 2/3 into the first increment
    got = ov.refine([0.0, 1.0, 2.0], [3.0, 1.0], [1, 1], alpha=0.0)
    np.testing.assert_allclose(got, [0.0, 2.0 / 3.0, 2.0], rtol=1e-15)
    np.testing.assert_allclose(refine_nodes([0.0, 1.0, 2.0], [3.0, 1.0], [1, 1], alpha=0.0), got, rtol=1e-15)
    # an empty increment carries no importance: thirds of the total 4 fall
    # 2/3 into the first increment and 1/3 into the last
    got = ov.refine([0.0, 1.0, 2.0, 3.0], [2.0, 0.0, 2.0], [1, 0, 1], alpha=0.0)
    np.testing.assert_allclose(got, [0.0, 2.0 / 3.0, 7.0 / 3.0, 3.0], rtol=1e-15)


def _exact_importance(nodes):
    """Per-increment average over x of (J f)^2 for f = exp(-x), J = N dx_i."""
    N = len(nodes) - 1
    a, b = nodes[:-1], nodes[1:]
    dx = b - a
    mean_f2 = (np.exp(-2 * a) - np.exp(-2 * b)) / (2 * dx)
    return (N * dx) ** 2 * mean_f2, np.ones(N)


def test_adaptation_converges_to_quantiles():
    # the fixed point has equal importance per increment; with f varying
    # inside an increment it differs from the quantiles by O(dx^2)
    L = 5.0
    err = {}
    for N in (100, 200):
        nodes = np.linspace(0.0, L, N + 1)
        for _ in range(300):
            s, n = _exact_importance(nodes)
            nodes = ov.refine(nodes, s, n, alpha=0.5)
        s, _ = _exact_importance(nodes)
        assert s.std() / s.mean() < 1e-9
        q = -np.log(1.0 - np.arange(N + 1) / N * (1.0 - np.exp(-L)))
        err[N] = np.max(np.abs(nodes - q))
    assert err[200] < 6e-3
    assert 3.0 < err[100] / err[200] < 5.0  # second order


def test_product_refine_matches_oracle():
    rng = np.random.default_rng(7)
    for N in (1, 2, 7, 100, 1000):
        nodes = np.sort(rng.uniform(-3.0, 4.0, N + 1))
        s = rng.exponential(1.0, N) * (rng.uniform(size=N) > 0.2)
        n = rng.integers(0, 5, N)
        for alpha in (0.0, 0.5):
            np.testing.assert_allclose(refine_nodes(nodes, s, n, alpha), ov.refine(nodes, s, n, alpha),
                                       rtol=1e-12, atol=1e-12)
