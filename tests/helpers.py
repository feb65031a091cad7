"""Comparison helpers shared by the tests."""
import numpy as np


def relnorm(a, b):
    """||a-b||_2 / ||b||_2 over the whole array (0 when both are zero)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm((a - b).reshape(-1))
    n = np.linalg.norm(b.reshape(-1))
    return d / n if n > 0 else d


def relnorm_cols(a, b):
    """relnorm per column of (n, k) arrays (per conservative component)."""
    a = np.asarray(a).reshape(np.shape(b)[0], -1)
    b = np.asarray(b).reshape(np.shape(b)[0], -1)
    return np.array([relnorm(a[:, j], b[:, j]) for j in range(b.shape[1])])


def bitdiff(a, b):
    """number of entries whose binary64 bit patterns differ"""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return int(np.sum(a.view(np.int64) != b.view(np.int64)))
