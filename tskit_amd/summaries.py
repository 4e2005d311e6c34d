"""The summary-function registry of ``TreeSequence.general_stat``.

The reference's ``general_stat(W, f, output_dim, ...)`` takes any Python
callable ``f`` (python/tskit/trees.py:7917-8004) and calls it at every node
of every tree; there is no Python on the device, so here ``f`` is an entry of
a fixed registry of device summary functions, each with its parameters --
the weighted statistics of the reference's C library (c/tskit/trees.c:
3961-4152, 4801-4837).  An entry is also callable on one state vector (numpy,
on the host) so that ``general_stat``'s ``strict`` check of f(0) and f(total)
runs as in the reference; the statistics themselves run on the GPU.
"""
import numpy as np

# tsg_stat_id values of the weighted summaries (include/tsgpu.h)
TRAIT_COVARIANCE = 11
TRAIT_CORRELATION = 12
TRAIT_LINEAR_MODEL = 13
RELATEDNESS_WEIGHTED = 14
RELATEDNESS_WEIGHTED_NONCENTRED = 15


class Summary:
    """A registry summary: its id, the index tuples and FP64 parameters the
    device function takes, and the numpy form of the same function."""

    def __init__(self, stat_id, fn, index_tuples=None, dparams=None):
        self.stat_id = int(stat_id)
        self._fn = fn
        self.index_tuples = (None if index_tuples is None
                             else np.ascontiguousarray(index_tuples, dtype=np.int32).reshape(-1, 2))
        self.dparams = (None if dparams is None
                        else np.ascontiguousarray(dparams, dtype=np.float64).reshape(-1))

    def __call__(self, x):
        return self._fn(np.asarray(x, dtype=np.float64))


def trait_covariance(num_samples):
    """x^2 / (2 (n - 1)^2) per column (trees.c:3961-3974)."""
    n = float(num_samples)
    return Summary(TRAIT_COVARIANCE, lambda x: x * x / (2 * (n - 1) * (n - 1)))


def trait_correlation(num_samples):
    """x_j^2 / (2 p (1 - p) n (n - 1)), p the last column, 0 unless
    0 < p < 1 (trees.c:4026-4045)."""
    n = float(num_samples)

    def f(x):
        p = x[-1]
        if not (0.0 < p < 1.0):
            return np.zeros(len(x) - 1)
        return x[:-1] * x[:-1] / (2 * (p * (1 - p)) * n * (n - 1))

    return Summary(TRAIT_CORRELATION, f)


def trait_linear_model(num_samples, V):
    """State [traits | covariates | count]; V [traits x covariates]
    (trees.c:4108-4152)."""
    V = np.asarray(V, dtype=np.float64)
    n = float(num_samples)
    m_out, k = V.shape

    def f(x):
        m = x[-1]
        out = np.zeros(m_out)
        if not (0.0 < m < n):
            return out
        z = x[m_out:m_out + k]
        for i in range(m_out):
            a = x[i] - np.dot(z, V[i])
            denom = m - np.dot(z, z)
            out[i] = 0.0 if denom < 1e-8 else a * a / (2 * denom * denom)
        return out

    return Summary(TRAIT_LINEAR_MODEL, f, dparams=V)


def relatedness_weighted(index_tuples, total_weights):
    """(x_i - n_i p)(x_j - n_j p), p the last column, n the total weights
    (trees.c:4801-4820)."""
    idx = np.asarray(index_tuples, dtype=np.int32).reshape(-1, 2)
    tw = np.asarray(total_weights, dtype=np.float64)

    def f(x):
        p = x[-1]
        return np.array([(x[i] - tw[i] * p) * (x[j] - tw[j] * p) for i, j in idx])

    return Summary(RELATEDNESS_WEIGHTED, f, index_tuples=idx, dparams=tw)


def relatedness_weighted_noncentred(index_tuples):
    """x_i x_j (trees.c:4822-4837)."""
    idx = np.asarray(index_tuples, dtype=np.int32).reshape(-1, 2)
    return Summary(RELATEDNESS_WEIGHTED_NONCENTRED,
                   lambda x: np.array([x[i] * x[j] for i, j in idx]), index_tuples=idx)
