"""Seeded synthetic gausslet coefficient sets used by the benchmark workloads.

A 1D gausslet is G(x) = sum_j c_j exp(-(3x - j)^2 / 2) over j = -(K-1)/2 ...
(K-1)/2 (gausslets/gausslets.py:8-31 of the reference).  The synthetic set is

    c_j = z_j * exp(-(4 j / 57)^2),  z ~ N(0, 1) from numpy.random.default_rng(seed)

with the literal 57 kept for every K, then symmetrised c_j <- (c_j + c_-j) / 2
and L2-normalised with the elementary overlaps <g_a|g_b> = (sqrt(pi)/3) exp(-(a-b)^2/4).
"""
import numpy as np


def overlap_matrix(K):
    j = np.arange(K) - (K - 1) // 2
    d = j[:, None] - j[None, :]
    return (np.sqrt(np.pi) / 3.0) * np.exp(-0.25 * d * d)


def synthetic_coefficients(K, seed, symmetrise=True):
    """symmetrise=False keeps the raw draws (no mirror symmetry), for the general key set."""
    if K % 2 != 1 or K < 3:
        raise ValueError("K must be odd and >= 3")
    rng = np.random.default_rng(seed)
    j = np.arange(K) - (K - 1) // 2
    c = rng.standard_normal(K) * np.exp(-(4.0 * j / 57.0) ** 2)
    if symmetrise:
        c = 0.5 * (c + c[::-1])
    c = c / np.sqrt(c @ overlap_matrix(K) @ c)
    return np.ascontiguousarray(c, dtype=np.float64)
