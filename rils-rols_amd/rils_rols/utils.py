"""Helpers of the estimator API (behaviour of rils_rols/utils.py:1-35)."""
from math import sqrt

import numpy as np
from numpy.random import RandomState
from sympy import preorder_traversal


def complexity_sympy(model):
    """Number of nodes of a sympy expression (utils.py:6-10)."""
    return sum(1 for _ in preorder_traversal(model))


def noisefy(y, noise_level, random_state):
    """y plus Gaussian noise with SD = noise_level * RMS(y) (utils.py:13-24)."""
    rms = sqrt(sum(v * v for v in y) / len(y))
    noise = RandomState(random_state).normal(0, noise_level * rms, len(y))
    return [y[i] + noise[i] for i in range(len(y))]


def logistic(yp):
    return 1.0 / (1.0 + np.exp(-yp))


def binarize(yp):
    """Class labels from predictions: 1 where yp > 0.5 (utils.py:29-31)."""
    return (yp > 0.5) * 1


def proba(yp):
    """Two-column class probabilities from the logistic of yp (utils.py:33-35)."""
    p = logistic(yp)
    return np.vstack([1 - p, p]).T
