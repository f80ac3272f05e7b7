"""Synthetic densities: seeded sums of Gaussians.

``f(x) = sum_i a_i exp(-|x - c_i|^2 / w_i^2)`` with centres uniform in
``[-0.5, 0.5]^d``, widths log-uniform in ``[0.02, 0.1]`` and amplitudes
``N(0, 1)`` (numpy ``default_rng(seed)``), the recipe of the benchmark
configurations (cf. ``volumential/gaussian.py:51-133``).
"""

from __future__ import annotations

import numpy as np


def gaussian_mixture_params(dim, seed, n_terms=8, lo=0.02, hi=0.1, box=0.5):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-box, box, size=(n_terms, dim))
    widths = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_terms))
    amps = rng.standard_normal(n_terms)
    return centers, widths, amps


def gaussian_mixture(points, seed, n_terms=8):
    """Evaluate at ``points[dim, n]``."""
    dim = points.shape[0]
    c, w, a = gaussian_mixture_params(dim, seed, n_terms)
    out = np.zeros(points.shape[1])
    for i in range(n_terms):
        r2 = np.sum((points - c[i][:, None]) ** 2, axis=0)
        out += a[i] * np.exp(-r2 / w[i] ** 2)
    return out


def gaussian_mixture_batch(points, seeds, n_terms=8):
    return np.stack([gaussian_mixture(points, s, n_terms) for s in seeds])


def helmholtz_beam_params(dim, seed, n_beams=16, box=0.8):
    """Random-phase Gaussian beams: amplitudes a + ib with a, b ~ N(0, 1),
    centres uniform in ``[-box, box]^d``, unit propagation directions."""
    rng = np.random.default_rng(seed)
    amps = rng.standard_normal(n_beams) + 1j * rng.standard_normal(n_beams)
    centers = rng.uniform(-box, box, size=(n_beams, dim))
    dirs = rng.standard_normal((n_beams, dim))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return amps, centers, dirs


def helmholtz_beams(points, seed, k=40.0, width=0.03, n_beams=16):
    """``f(x) = sum_b A_b exp(-|x - c_b|^2 / w^2) exp(i k d_b . x)`` at
    ``points[dim, n]`` (complex128)."""
    dim = points.shape[0]
    amps, c, d = helmholtz_beam_params(dim, seed, n_beams)
    out = np.zeros(points.shape[1], dtype=np.complex128)
    for b in range(n_beams):
        r2 = np.sum((points - c[b][:, None]) ** 2, axis=0)
        out += amps[b] * np.exp(-r2 / width**2) * np.exp(1j * k * (d[b] @ points))
    return out
