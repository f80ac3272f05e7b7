"""List-1 interaction case gallery.

Every (target box, source box) pair of the near field of a 2:1-balanced
tree falls into one of a fixed set of *cases*: the centre offset
``c_target - c_source`` measured in units of a quarter of the source box
extent.  The gallery is the lexicographically sorted list of those integer
vectors; a case id is a position in that list.

Follows ``volumential/list1_gallery.py:118-198`` (``generate_interactions``,
``postprocess_interactions``, ``generate_list1_gallery``).  The reference
enumerates the offsets by walking a 4-level integer box tree; here the same
set is written down directly from the adjacency geometry:

* the box itself: the zero vector;
* a same-size neighbour: components in {-4, 0, 4}, not all zero;
* a half-size source: components in {-3, -1, 1, 3}, at least one +-3;
* a double-size source: components in {-6, -2, 2, 6}, at least one +-6.

Counts: 7 (1D), 33 (2D), 139 (3D).  Components span -6..6, so the encoding
base is 13 and the shift 6 (Horner order, axis 0 most significant).
"""

from __future__ import annotations

import itertools
from functools import lru_cache

import numpy as np

ENCODING_BASE = 13
ENCODING_SHIFT = 6


def _shell(values, must_have):
    """Vectors over ``values`` per axis with max |component| in ``must_have``."""

    def gen(dim):
        for vec in itertools.product(values, repeat=dim):
            if max(abs(v) for v in vec) in must_have:
                yield tuple(int(v) for v in vec)

    return gen


@lru_cache(maxsize=None)
def _gallery(dim: int):
    if dim < 1:
        raise ValueError("dimension must be positive")
    vecs = {tuple([0] * dim)}
    vecs.update(_shell((-4, 0, 4), {4})(dim))
    vecs.update(_shell((-3, -1, 1, 3), {3})(dim))
    vecs.update(_shell((-6, -2, 2, 6), {6})(dim))
    vec_list = sorted(vecs)

    case_indices = -np.ones(ENCODING_BASE**dim, dtype=np.int64)
    for case_id, vec in enumerate(vec_list):
        case_indices[case_encode(vec)] = case_id
    case_indices.setflags(write=False)
    return tuple(vec_list), case_indices


def case_encode(case_vec) -> int:
    """Horner encoding of a case vector, base 13, shift 6."""
    code = 0
    for comp in case_vec:
        code = code * ENCODING_BASE + (int(comp) + ENCODING_SHIFT)
    return int(code)


def generate_list1_gallery(dim: int):
    """Return ``(vec_list, case_encode, case_indices)``.

    Same return convention as the reference
    (``volumential/list1_gallery.py:167-198``): ``vec_list`` is the sorted
    list of case vectors (tuples), ``case_encode`` the encoder and
    ``case_indices`` an int64 lookup of length ``13**dim`` holding the case
    id of each code or -1.
    """
    vec_list, case_indices = _gallery(int(dim))
    return [tuple(v) for v in vec_list], case_encode, np.array(case_indices)


def case_vectors(dim: int) -> np.ndarray:
    """Case vectors as an int32 array of shape ``(n_cases, dim)``."""
    vec_list, _ = _gallery(int(dim))
    return np.asarray(vec_list, dtype=np.int32).reshape(len(vec_list), dim)


def target_box_size(case_vec) -> float:
    """Target box extent relative to the source box for a case.

    ``nearfield_potential_table.py:917-946``: 1 for the zero vector,
    otherwise ``(max|v| - 2) / 2`` (0.5, 1 or 2).
    """
    m = max(abs(int(v)) for v in case_vec)
    if m == 0:
        return 1.0
    return (m - 2) / 2.0
