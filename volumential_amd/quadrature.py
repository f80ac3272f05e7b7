"""Quadrature rules for the near-field table build.

* Tensor Gauss-Legendre nodes of each box (``numpy.polynomial.legendre``).
* The Duffy node set for a singular point inside a box: the box is split at
  the singular point into ``2^d`` sub-boxes (one per sign vector), each
  sub-box into ``d!`` simplices (one per axis permutation); each simplex is
  mapped to the unit cube by the Duffy transform
  ``u[perm[i]] = prod_{j<=i} r_j`` with Jacobian ``prod_{i<d-1} r_i^(d-1-i)``.
  ``r_0`` (the radial variable) uses a tanh-sinh rule, ``r_1..r_{d-1}`` a
  tensor Gauss-Legendre rule on [0, 1].

Follows ``volumential/singular_integral_2d.py:807-853``
(``_duffy_regular_nodes_weights``, ``_duffy_radial_nodes_weights``
"tanh-sinh-fast" branch) and
``volumential/nearfield_potential_table.py:2224-2299``
(``_get_duffy_radial_node_data``).  Node order: sign vector (slowest, in
``itertools.product([-1, 1])`` order), permutation (``itertools.permutations``
order), tail node, radial node (fastest).
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass
from functools import lru_cache

import numpy as np


def gauss_legendre_01(order: int):
    """Gauss-Legendre nodes/weights mapped to [0, 1] (ascending nodes)."""
    x, w = np.polynomial.legendre.leggauss(int(order))
    return 0.5 * (x + 1.0), 0.5 * w


def box_nodes_1d(order: int, extent: float = 1.0) -> np.ndarray:
    """1D interpolation nodes of a box ``[0, extent]``: GL nodes, ascending."""
    x, _ = np.polynomial.legendre.leggauss(int(order))
    return 0.5 * extent * (x + 1.0)


def tensor_q_points(order: int, dim: int, extent: float = 1.0) -> np.ndarray:
    """``q^d`` tensor nodes on ``[0, extent]^d``, axis 0 slowest.

    Matches the lexicographic node order of
    ``nearfield_potential_table.py:535-592`` and the per-box order of a tree
    built with ``meshgrid(indexing="ij")``.
    """
    x = box_nodes_1d(order, extent)
    grids = np.meshgrid(*([x] * dim), indexing="ij")
    return np.stack([g.ravel() for g in grids], axis=-1)


def tanh_sinh_fast(order: int):
    """Closed-form tanh-sinh rule on (0, 1).

    ``m = max(3, N)``, ``h = 1/sqrt(m)``, ``t_k = k h`` for ``|k| <= m``,
    ``rho = (1 + tanh(pi/2 sinh t)) / 2``,
    ``w = pi/4 h cosh t sech^2(pi/2 sinh t)`` with the sech^2 written as
    ``4 e^{-2|a|} / (1 + e^{-2|a|})^2``; nodes that round to 0 or 1 are
    dropped.
    """
    m = max(3, int(order))
    h = 1.0 / np.sqrt(m)
    k = np.arange(-m, m + 1, dtype=np.float64)
    t = h * k
    a = 0.5 * np.pi * np.sinh(t)
    rho = 0.5 * (1.0 + np.tanh(a))
    e = np.exp(-2.0 * np.abs(a))
    sech2 = 4.0 * e / (1.0 + e) ** 2
    w = 0.25 * np.pi * h * np.cosh(t) * sech2
    keep = (rho > 0.0) & (rho < 1.0) & np.isfinite(w)
    return rho[keep], w[keep]


def tail_rule(dim_minus_one: int, order: int):
    """Tensor GL rule on [0,1]^(d-1), first axis slowest."""
    if dim_minus_one == 0:
        return np.zeros((1, 0)), np.ones(1)
    x, w = gauss_legendre_01(order)
    gx = np.meshgrid(*([x] * dim_minus_one), indexing="ij")
    gw = np.meshgrid(*([w] * dim_minus_one), indexing="ij")
    nodes = np.stack([g.ravel() for g in gx], axis=-1)
    weights = np.prod(np.array(gw), axis=0).ravel()
    return nodes, weights


@dataclass(frozen=True)
class DuffyNodeSet:
    dim: int
    node_u: np.ndarray  # [dim, n_nodes]
    node_sign: np.ndarray  # [dim, n_nodes]  (+-1.0)
    node_jac: np.ndarray  # [n_nodes]

    @property
    def n_nodes(self) -> int:
        return int(self.node_jac.size)


@lru_cache(maxsize=16)
def duffy_node_set(dim: int, regular_order: int, radial_order: int) -> DuffyNodeSet:
    """Duffy nodes for the "tanh-sinh-fast" radial rule (vectorised)."""
    dim = int(dim)
    tail_x, tail_w = tail_rule(dim - 1, regular_order)
    rho, w_rho = tanh_sinh_fast(radial_order)
    signs = np.array(list(itertools.product([-1.0, 1.0], repeat=dim)))
    perms = list(itertools.permutations(range(dim)))

    n_tail, n_rho = tail_w.size, rho.size
    # r[i] per (tail, rho): r_0 = rho, r_1.. = tail
    r = np.empty((n_tail, n_rho, dim))
    r[:, :, 0] = rho[None, :]
    if dim > 1:
        r[:, :, 1:] = tail_x[:, None, :]
    jac = np.ones((n_tail, n_rho))
    for i in range(dim - 1):
        jac = jac * r[:, :, i] ** (dim - 1 - i)
    jac = jac * tail_w[:, None] * w_rho[None, :]
    cum = np.cumprod(r, axis=2)  # cum[..., i] = prod_{j<=i} r_j

    blocks_u, blocks_s, blocks_j = [], [], []
    for sgn in signs:
        for perm in perms:
            u = np.empty((dim, n_tail * n_rho))
            for i, ax in enumerate(perm):
                u[ax] = cum[:, :, i].ravel()
            blocks_u.append(u)
            blocks_s.append(np.repeat(sgn[:, None], n_tail * n_rho, axis=1))
            blocks_j.append(jac.ravel())
    node_u = np.ascontiguousarray(np.concatenate(blocks_u, axis=1))
    node_sign = np.ascontiguousarray(np.concatenate(blocks_s, axis=1))
    node_jac = np.ascontiguousarray(np.concatenate(blocks_j))
    for a in (node_u, node_sign, node_jac):
        a.setflags(write=False)
    return DuffyNodeSet(dim, node_u, node_sign, node_jac)


def barycentric_weights(nodes) -> np.ndarray:
    """First-kind barycentric weights, scaled by the largest magnitude.

    ``volumential/lagrange.py:45-59``.
    """
    nodes = np.asarray(nodes, dtype=np.float64)
    diffs = nodes[:, None] - nodes[None, :]
    np.fill_diagonal(diffs, 1.0)
    if np.any(diffs == 0):
        raise ValueError("nodes must be distinct")
    signs = np.prod(np.sign(diffs), axis=1)
    logw = -np.sum(np.log(np.abs(diffs)), axis=1)
    return signs * np.exp(logw - np.max(logw))
