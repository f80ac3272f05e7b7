"""Arithmetic orbit reduction of near-field tables (scalar and signed kernels).

A table entry ``(case, s, t)`` (source mode ``s``, target node ``t``) is
invariant under the hyperoctahedral group acting jointly on the case vector
and on the per-axis node indices (axis flips ``i -> q-1-i`` and axis
permutations).  Only one canonical representative per orbit is stored; an
entry is mapped to its compact row by a few integer operations:

1. per case, a signed axis permutation mapping the case vector to its
   canonical form (components ``-|v|`` ordered by decreasing ``|v|``);
2. axes whose case component is zero may be flipped freely: each such axis
   folds ``(s_k, t_k)`` to the lexicographically smaller of itself and its
   mirror image;
3. axes of equal ``|v|`` form a group and may be permuted: their
   ``(s_k, t_k)`` pairs are sorted;
4. each group's multiset of pair codes ``s_k*q + t_k`` is ranked among all
   multisets of that size, and the group ranks are combined mixed-radix;
5. the result is offset by the canonical case's first compact row.

For an axis-derivative kernel (``AxisTargetDerivative`` /
``AxisSourceDerivative``) the derivative axis is never permuted (a group of
its own) and a flip of it negates the entry: the entry is read as
``sign * compact[row]`` (the reference's signed-arithmetic-orbit kind).

Follows ``volumential/orbit_arithmetic.py``: ``_case_arithmetic_axis_descriptors``
(45-163, scalar-kernel branch), ``_canonicalize_arithmetic_entry_axes``
(372-432), ``_rank_multiset`` (529-542), ``_compact_arithmetic_case_value_count``
(545-563), ``_evaluate_compact_arithmetic_orbit_entry`` (566-634),
``build_arithmetic_case_metadata`` (665-745) and
``enumerate_scalar_arithmetic_representatives`` (748-790).  The compact row
order is the reference's, so tables written by the reference (reduced full
entry ids + values) map onto the same rows.
"""

from __future__ import annotations

import itertools
import math
from dataclasses import dataclass
from functools import lru_cache

import numpy as np

from . import gallery


@dataclass(frozen=True)
class ArithmeticCaseMetadata:
    dim: int
    quad_order: int
    canonical_case_ids: np.ndarray  # int32 [n_orbits], sorted
    case_orbit_ranks: np.ndarray  # uint16 [n_cases]
    case_axis_perm: np.ndarray  # uint8 [n_cases, dim]
    case_axis_sign: np.ndarray  # int8 [n_cases, dim]
    case_axis_group: np.ndarray  # uint8 [n_cases, dim]
    case_value_offsets: np.ndarray  # int32 [n_orbits + 1]
    axis_sign_power: np.ndarray  # uint8 [dim] (zeros for scalar kernels)

    @property
    def n_cases(self) -> int:
        return int(self.case_orbit_ranks.size)

    @property
    def n_orbits(self) -> int:
        return int(self.canonical_case_ids.size)

    @property
    def n_compact_entries(self) -> int:
        return int(self.case_value_offsets[-1])

    @property
    def metadata_bytes(self) -> int:
        """Bytes of the per-case descriptors a device evaluation needs."""
        return int(
            self.case_orbit_ranks.nbytes
            + self.case_axis_perm.nbytes
            + self.case_axis_sign.nbytes
            + self.case_axis_group.nbytes
            + self.axis_sign_power.nbytes
            + self.case_value_offsets.nbytes
        )


def axis_group_specs(dim, fixed_axes=()):
    """Axis sets the symmetry group may permute among themselves.

    ``orbit_arithmetic.py:245-251``: every derivative (fixed) axis is a group
    of its own, in increasing order, followed by the remaining free axes.
    """
    fixed = sorted(set(int(a) for a in fixed_axes))
    free = tuple(a for a in range(dim) if a not in fixed)
    return tuple((a,) for a in fixed) + ((free,) if free else ())


def case_axis_descriptors(case_vec, fixed_axes=()):
    """Signed permutation mapping ``case_vec`` to its canonical vector.

    Within each axis group (``axis_group_specs``) output axis ``k`` reads
    input axis ``perm[k]``, the group's axes ordered by ``(-|v|, axis)``.
    ``sign[k]`` is -1 for a positive component, +1 for a negative one and 0
    for a zero component, so canonical components are ``<= 0``.  Runs of
    equal ``|v|`` inside a group share a runtime group id; ids count up over
    the groups in order (``orbit_arithmetic.py:140-163``).
    """
    vec = [int(v) for v in case_vec]
    dim = len(vec)
    perm = np.empty(dim, dtype=np.uint8)
    sign = np.empty(dim, dtype=np.int8)
    group = np.empty(dim, dtype=np.uint8)
    next_gid = 0
    for out_axes in axis_group_specs(dim, fixed_axes):
        order = sorted(out_axes, key=lambda a: (-abs(vec[a]), a))
        gid = next_gid - 1
        last = None
        for out_axis, in_axis in zip(out_axes, order):
            v = vec[in_axis]
            if abs(v) != last:
                gid += 1
                last = abs(v)
            perm[out_axis] = in_axis
            sign[out_axis] = -1 if v > 0 else (1 if v < 0 else 0)
            group[out_axis] = gid
        next_gid = gid + 1
    return perm, sign, group


def canonical_case_vector(case_vec, perm, sign):
    return [0 if int(s) == 0 else int(case_vec[int(p)]) * int(s) for p, s in zip(perm, sign)]


def group_alphabet(q: int, has_zero: bool) -> int:
    pair_count = q * q
    return (pair_count + 1) // 2 if has_zero else pair_count


def case_value_count(sign_row, group_row, q: int) -> int:
    count = 1
    sign_row = np.asarray(sign_row)
    group_row = np.asarray(group_row)
    for gid in range(int(group_row.max()) + 1):
        axes = np.flatnonzero(group_row == gid)
        if axes.size == 0:
            continue
        alpha = group_alphabet(q, bool(np.any(sign_row[axes] == 0)))
        count *= math.comb(alpha + axes.size - 1, axes.size)
    return int(count)


@lru_cache(maxsize=None)
def _metadata_cached(dim: int, q: int, fixed_axes: tuple, sign_power: tuple
                     ) -> ArithmeticCaseMetadata:
    if dim not in (2, 3):
        raise NotImplementedError("arithmetic orbit reduction supports dim 2 or 3")
    vecs, _, case_indices = gallery.generate_list1_gallery(dim)
    n_cases = len(vecs)
    perm = np.empty((n_cases, dim), dtype=np.uint8)
    sign = np.empty((n_cases, dim), dtype=np.int8)
    group = np.empty((n_cases, dim), dtype=np.uint8)
    canon_ids = np.empty(n_cases, dtype=np.int32)
    for cid, vec in enumerate(vecs):
        p, s, g = case_axis_descriptors(vec, fixed_axes)
        perm[cid], sign[cid], group[cid] = p, s, g
        canon_ids[cid] = case_indices[gallery.case_encode(canonical_case_vector(vec, p, s))]
    uniq = np.unique(canon_ids)
    rank_of = {int(c): r for r, c in enumerate(uniq)}
    ranks = np.array([rank_of[int(c)] for c in canon_ids], dtype=np.uint16)
    counts = np.array(
        [case_value_count(sign[c], group[c], q) for c in uniq], dtype=np.int64
    )
    offsets = np.zeros(uniq.size + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(counts)
    if offsets[-1] > np.iinfo(np.int32).max:
        raise OverflowError("compact table exceeds int32 row range")
    return ArithmeticCaseMetadata(
        dim=dim,
        quad_order=q,
        canonical_case_ids=uniq.astype(np.int32),
        case_orbit_ranks=ranks,
        case_axis_perm=perm,
        case_axis_sign=sign,
        case_axis_group=group,
        case_value_offsets=offsets.astype(np.int32),
        axis_sign_power=np.asarray(sign_power, dtype=np.uint8),
    )


def build_case_metadata(dim: int, quad_order: int, fixed_axes=(), axis_sign_power=None
                        ) -> ArithmeticCaseMetadata:
    """Per-case descriptors (``orbit_arithmetic.py:665-745``).

    ``fixed_axes`` / ``axis_sign_power`` describe an axis-derivative kernel
    (``kernels.derivative_axes``, ``orbit_arithmetic.py:176-369``); the
    defaults give the scalar-kernel metadata.
    """
    dim = int(dim)
    power = tuple(int(v) for v in (axis_sign_power if axis_sign_power is not None
                                   else (0,) * dim))
    if len(power) != dim or any(v not in (0, 1) for v in power):
        raise ValueError("axis_sign_power must hold dim values in {0, 1}")
    fixed = tuple(sorted(set(int(a) for a in fixed_axes)))
    if any(not 0 <= a < dim for a in fixed):
        raise ValueError("fixed axis out of range")
    return _metadata_cached(dim, int(quad_order), fixed, power)


def kernel_case_metadata(kernel, quad_order: int) -> ArithmeticCaseMetadata:
    """Metadata for a kernel: its derivative axes fix and sign the group."""
    from .kernels import derivative_axes

    fixed, power = derivative_axes(kernel)
    return build_case_metadata(kernel.dim, quad_order, fixed, power)


def _combos_with_replacement(alpha: int, k: int) -> np.ndarray:
    return np.asarray(
        list(itertools.combinations_with_replacement(range(alpha), k)), dtype=np.int64
    ).reshape(-1, k)


def enumerate_representatives(meta: ArithmeticCaseMetadata) -> np.ndarray:
    """Full entry ids of the canonical representatives in compact-row order.

    ``orbit_arithmetic.py:748-790``.  Full id = ``case*n^2 + s*n + t``.
    """
    dim, q = meta.dim, meta.quad_order
    n = q**dim
    out = np.empty(meta.n_compact_entries, dtype=np.int64)
    pos = 0
    for rank, cid in enumerate(meta.canonical_case_ids.tolist()):
        groups = meta.case_axis_group[cid]
        signs = meta.case_axis_sign[cid]
        group_axes = []
        group_opts = []
        for gid in range(int(groups.max()) + 1):
            axes = np.flatnonzero(groups == gid)
            alpha = group_alphabet(q, bool(np.any(signs[axes] == 0)))
            group_axes.append(axes)
            group_opts.append(_combos_with_replacement(alpha, axes.size))
        # itertools.product order over groups: first group slowest
        sizes = [o.shape[0] for o in group_opts]
        total = int(np.prod(sizes))
        idx = np.indices(sizes).reshape(len(sizes), -1)
        s_axes = np.zeros((total, dim), dtype=np.int64)
        t_axes = np.zeros((total, dim), dtype=np.int64)
        for g, axes in enumerate(group_axes):
            codes = group_opts[g][idx[g]]  # (total, k)
            for j, ax in enumerate(axes):
                s_axes[:, ax] = codes[:, j] // q
                t_axes[:, ax] = codes[:, j] % q
        s_id = np.zeros(total, dtype=np.int64)
        t_id = np.zeros(total, dtype=np.int64)
        for ax in range(dim):
            s_id = s_id * q + s_axes[:, ax]
            t_id = t_id * q + t_axes[:, ax]
        out[pos : pos + total] = cid * n * n + s_id * n + t_id
        assert pos == int(meta.case_value_offsets[rank])
        pos += total
    assert pos == out.size
    return out


def _rank_sorted_multiset(values: np.ndarray, alpha: int) -> np.ndarray:
    """Lexicographic rank of sorted tuples ``values[:, 0] <= values[:, 1] ...``.

    Vectorised form of ``orbit_arithmetic.py:529-542``:
    ``rank = sum_i sum_{c=v_{i-1}}^{v_i - 1} C(alpha - c + rem - 1, rem)``.
    """
    m, k = values.shape
    rank = np.zeros(m, dtype=np.int64)
    if k == 0:
        return rank
    # prefix table P[rem][c] = sum_{c' < c} C(alpha - c' + rem - 1, rem)
    for i in range(k):
        rem = k - i - 1
        terms = np.array(
            [math.comb(alpha - c + rem - 1, rem) for c in range(alpha)], dtype=np.int64
        )
        prefix = np.concatenate([[0], np.cumsum(terms)])
        lo = values[:, i - 1] if i > 0 else np.zeros(m, dtype=np.int64)
        rank += prefix[values[:, i]] - prefix[lo]
    return rank


def entry_rows(meta: ArithmeticCaseMetadata, case_ids, s_ids, t_ids) -> np.ndarray:
    """Compact row of each ``(case, s, t)`` (vectorised).

    ``orbit_arithmetic.py:372-432`` followed by ``566-634``.  The entry sign
    of signed kernels is ``entry_signs``.
    """
    dim, q = meta.dim, meta.quad_order
    case_ids = np.asarray(case_ids, dtype=np.int64)
    s_ids = np.asarray(s_ids, dtype=np.int64)
    t_ids = np.asarray(t_ids, dtype=np.int64)
    case_ids, s_ids, t_ids = np.broadcast_arrays(case_ids, s_ids, t_ids)
    shape = case_ids.shape
    case_ids = case_ids.ravel()
    s_raw = np.empty((case_ids.size, dim), dtype=np.int64)
    t_raw = np.empty((case_ids.size, dim), dtype=np.int64)
    rs, rt = s_ids.ravel().copy(), t_ids.ravel().copy()
    for ax in range(dim - 1, -1, -1):
        s_raw[:, ax] = rs % q
        t_raw[:, ax] = rt % q
        rs //= q
        rt //= q

    perm = meta.case_axis_perm[case_ids].astype(np.int64)
    sign = meta.case_axis_sign[case_ids].astype(np.int64)
    grp = meta.case_axis_group[case_ids].astype(np.int64)
    rows = np.arange(case_ids.size)
    s_ax = np.empty_like(s_raw)
    t_ax = np.empty_like(t_raw)
    for out in range(dim):
        s_in = s_raw[rows, perm[:, out]]
        t_in = t_raw[rows, perm[:, out]]
        s_f = q - 1 - s_in
        t_f = q - 1 - t_in
        neg = sign[:, out] < 0
        zero = sign[:, out] == 0
        fold = zero & ((s_f < s_in) | ((s_f == s_in) & (t_f < t_in)))
        flip = neg | fold
        s_ax[:, out] = np.where(flip, s_f, s_in)
        t_ax[:, out] = np.where(flip, t_f, t_in)

    def cswap(i, j):
        same = grp[:, i] == grp[:, j]
        less = (s_ax[:, j] < s_ax[:, i]) | (
            (s_ax[:, j] == s_ax[:, i]) & (t_ax[:, j] < t_ax[:, i])
        )
        sw = same & less
        for arr in (s_ax, t_ax):
            a_i = arr[:, i].copy()
            arr[sw, i] = arr[sw, j]
            arr[sw, j] = a_i[sw]

    if dim == 2:
        cswap(0, 1)
    else:
        cswap(0, 1)
        cswap(1, 2)
        cswap(0, 1)
        cswap(0, 2)

    pair = s_ax * q + t_ax
    compact = np.zeros(case_ids.size, dtype=np.int64)
    for gid in range(dim):
        in_g = grp == gid
        size = in_g.sum(axis=1)
        has_zero = np.any(in_g & (sign == 0), axis=1)
        g_rank = np.zeros(case_ids.size, dtype=np.int64)
        g_count = np.ones(case_ids.size, dtype=np.int64)
        for k in range(1, dim + 1):
            for hz in (False, True):
                sel = (size == k) & (has_zero == hz)
                if not np.any(sel):
                    continue
                alpha = group_alphabet(q, hz)
                vals = np.where(in_g[sel], pair[sel], np.iinfo(np.int64).max)
                vals = np.sort(vals, axis=1)[:, :k]
                g_rank[sel] = _rank_sorted_multiset(vals, alpha)
                g_count[sel] = math.comb(alpha + k - 1, k)
        compact = compact * g_count + g_rank
    offsets = meta.case_value_offsets.astype(np.int64)
    row = offsets[meta.case_orbit_ranks[case_ids].astype(np.int64)] + compact
    return row.reshape(shape)


def entry_signs(meta: ArithmeticCaseMetadata, case_ids, s_ids, t_ids) -> np.ndarray:
    """Entry sign (+1/-1, int8) of each ``(case, s, t)``: the reconstructed
    value is ``sign * compact[row]``.  ``orbit_arithmetic.py:372-432``: every
    axis the canonicalisation flips (negative case component, or the zero-
    component fold) negates the entry when its input axis has sign power 1.
    All +1 for scalar kernels."""
    dim, q = meta.dim, meta.quad_order
    case_ids, s_ids, t_ids = np.broadcast_arrays(
        np.asarray(case_ids, dtype=np.int64), np.asarray(s_ids, dtype=np.int64),
        np.asarray(t_ids, dtype=np.int64))
    shape = case_ids.shape
    power = np.asarray(meta.axis_sign_power, dtype=np.int64)
    if not np.any(power):
        return np.ones(shape, dtype=np.int8)
    case_ids = case_ids.ravel()
    rs, rt = s_ids.ravel().copy(), t_ids.ravel().copy()
    s_raw = np.empty((case_ids.size, dim), dtype=np.int64)
    t_raw = np.empty((case_ids.size, dim), dtype=np.int64)
    for ax in range(dim - 1, -1, -1):
        s_raw[:, ax] = rs % q
        t_raw[:, ax] = rt % q
        rs //= q
        rt //= q
    perm = meta.case_axis_perm[case_ids].astype(np.int64)
    sign = meta.case_axis_sign[case_ids].astype(np.int64)
    rows = np.arange(case_ids.size)
    neg = np.zeros(case_ids.size, dtype=bool)
    for out in range(dim):
        s_in = s_raw[rows, perm[:, out]]
        t_in = t_raw[rows, perm[:, out]]
        s_f, t_f = q - 1 - s_in, q - 1 - t_in
        fold = (sign[:, out] == 0) & ((s_f < s_in) | ((s_f == s_in) & (t_f < t_in)))
        flip = (sign[:, out] < 0) | fold
        neg ^= flip & (power[perm[:, out]] == 1)
    return np.where(neg, -1, 1).astype(np.int8).reshape(shape)


def case_row_map(meta: ArithmeticCaseMetadata, case_id: int) -> np.ndarray:
    """Dense ``[t, s]`` map of compact rows for one case (int64)."""
    n = meta.quad_order**meta.dim
    t, s = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return entry_rows(meta, np.full(t.shape, case_id), s, t)
