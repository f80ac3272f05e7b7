"""Oracle: list-1 gallery and arithmetic orbit reduction (scalar restatement).

Follows, statement by statement:
* ``list1_gallery.py:29-198`` -- the 4-level integer tree walk;
* ``orbit_arithmetic.py:29-42`` (axis encode/decode), ``45-163``
  (``_case_arithmetic_axis_descriptors``, scalar-kernel branch),
  ``166-173``, ``372-432`` (``_canonicalize_arithmetic_entry_axes``),
  ``435-506`` (``_evaluate_arithmetic_orbit_entry``), ``529-542``
  (``_rank_multiset``), ``545-563``, ``566-634``
  (``_evaluate_compact_arithmetic_orbit_entry``), ``665-745``
  (``build_arithmetic_case_metadata``), ``748-790``
  (``enumerate_scalar_arithmetic_representatives``);
* ``list1.py:604-902`` -- the closed-form group ranks of the device code
  (``rows_for_case``, vectorised over (s, t)).

``orbit_min_ids`` is an independent brute-force oracle: it applies all
``2^d d!`` signed axis permutations to (case vector, s, t) and returns the
smallest full entry id of each orbit -- the representative the reference's
union-find picks (``nearfield_potential_table.py:1646-1751``).
"""

from __future__ import annotations

import itertools
import math

import numpy as np


# ---------------------------------------------------------------- gallery
class _Box:
    def __init__(self, center, radius, child_nlevels):
        self.center = center
        self.radius = radius
        self.children = []
        if child_nlevels:
            cr = radius // 2
            d = len(center)
            for i in range(2**d):
                cc = center.copy()
                for k in range(d):
                    cc[k] += (2 * int(bool(i & (1 << k))) - 1) * cr
                self.children.append(_Box(cc, cr, child_nlevels - 1))


def _boxes_on_level(box, lev):
    if lev:
        for ch in box.children:
            yield from _boxes_on_level(ch, lev - 1)
    else:
        yield box


def _all_boxes(box):
    yield box
    for ch in box.children:
        yield from _all_boxes(ch)


def gallery(dim):
    """``(vec_list, case_indices)`` by the reference's tree walk."""
    root = _Box(np.array([8] * dim, np.int64), 8, 3)
    lo, hi = 8 >> 2, 2 * 8 - (8 >> 2)
    targets = [
        b for b in _boxes_on_level(root, 2) if np.min(b.center) > lo and np.max(b.center) < hi
    ]
    vecs = set()
    for tb in targets:
        for sb in _all_boxes(root):
            if np.max(np.abs(tb.center - sb.center) - (tb.radius + sb.radius)) == 0:
                vecs.add(tuple(int(v) for v in tb.center - sb.center))
    vecs.add(tuple([0] * dim))
    vec_list = sorted(vecs)
    base = 13
    case_indices = -np.ones(base**dim, dtype=np.int64)
    for cid, v in enumerate(vec_list):
        case_indices[encode(v)] = cid
    return vec_list, case_indices


def encode(vec):
    code = 0
    for c in vec:
        code = code * 13 + (int(c) + 6)
    return code


# ------------------------------------------------------- axis descriptors
def decode_axes(mode_id, q, dim):
    axes = [0] * dim
    r = int(mode_id)
    for ax in range(dim - 1, -1, -1):
        axes[ax] = r % q
        r //= q
    return axes


def encode_axes(axes, q):
    r = 0
    for a in axes:
        r = r * q + int(a)
    return r


def case_axis_descriptors(case_vec, fixed_axes=()):
    """``orbit_arithmetic.py:45-163`` with ``axis_groups`` = one group per
    fixed (derivative) axis, then the free axes (``245-251``); the
    directional-derivative branch is not restated."""
    vec = [int(v) for v in case_vec]
    dim = len(vec)
    fixed = sorted(set(fixed_axes))
    specs = [(a,) for a in fixed]
    free = tuple(a for a in range(dim) if a not in fixed)
    if free:
        specs.append(free)
    perm = [0] * dim
    sign = [0] * dim
    group = [0] * dim
    nxt = 0
    for out_axes in specs:
        free_order = sorted(out_axes, key=lambda i: (-abs(vec[i]), i))
        last = None
        cur = nxt - 1
        for out_axis, in_axis in zip(out_axes, free_order):
            v = vec[in_axis]
            if last != abs(v):
                cur += 1
                last = abs(v)
            perm[out_axis] = in_axis
            sign[out_axis] = -1 if v > 0 else (1 if v < 0 else 0)
            group[out_axis] = cur
        nxt = cur + 1
    return perm, sign, group


def case_metadata(dim, q, fixed_axes=(), sign_power=None):
    vecs, case_indices = gallery(dim)
    n = len(vecs)
    perm = np.empty((n, dim), np.uint8)
    sign = np.empty((n, dim), np.int8)
    group = np.empty((n, dim), np.uint8)
    canon = np.empty(n, np.int32)
    for cid, v in enumerate(vecs):
        p, s, g = case_axis_descriptors(v, fixed_axes)
        perm[cid], sign[cid], group[cid] = p, s, g
        cv = [0 if sg == 0 else v[pi] * sg for pi, sg in zip(p, s)]
        canon[cid] = case_indices[encode(cv)]
    uniq = np.unique(canon)
    rank_of = {int(c): r for r, c in enumerate(uniq)}
    ranks = np.array([rank_of[int(c)] for c in canon], np.uint16)
    counts = [value_count(sign[c], group[c], q) for c in uniq]
    offsets = np.zeros(len(uniq) + 1, np.int64)
    offsets[1:] = np.cumsum(counts)
    return {
        "vecs": vecs,
        "case_indices": case_indices,
        "canonical_case_ids": uniq.astype(np.int32),
        "case_orbit_ranks": ranks,
        "case_axis_perm": perm,
        "case_axis_sign": sign,
        "case_axis_group": group,
        "case_value_offsets": offsets.astype(np.int32),
        "axis_sign_power": np.asarray(sign_power if sign_power is not None else [0] * dim,
                                      np.uint8),
        "n_compact_entries": int(offsets[-1]),
    }


def value_count(sign, group, q):
    pc = q * q
    fpc = (pc + 1) // 2
    cnt = 1
    group = np.asarray(group, np.int64)
    for gid in range(int(group.max()) + 1):
        axes = np.flatnonzero(group == gid)
        if len(axes) == 0:
            continue
        a = fpc if np.any(np.asarray(sign)[axes] == 0) else pc
        cnt *= math.comb(int(a) + len(axes) - 1, len(axes))
    return int(cnt)


# --------------------------------------------------------- entry mapping
def canonicalize(case_id, s, t, meta, q, dim):
    sa, ta, groups, signs, _ = canonicalize_signed(case_id, s, t, meta, q, dim)
    return sa, ta, groups, signs


def canonicalize_signed(case_id, s, t, meta, q, dim):
    """``orbit_arithmetic.py:372-432``, with the entry sign of axis
    derivatives (directional derivatives not restated)."""
    s_raw = decode_axes(s, q, dim)
    t_raw = decode_axes(t, q, dim)
    power = meta.get("axis_sign_power", [0] * dim)
    sa, ta, groups, signs = [], [], [], []
    esign = 1
    for out in range(dim):
        ia = int(meta["case_axis_perm"][case_id, out])
        sg = int(meta["case_axis_sign"][case_id, out])
        sv, tv = s_raw[ia], t_raw[ia]
        applied = 1
        if sg < 0:
            sv, tv = q - 1 - sv, q - 1 - tv
            applied = -1
        elif sg == 0:
            fs, ft = q - 1 - sv, q - 1 - tv
            if (fs, ft) < (sv, tv):
                sv, tv = fs, ft
                applied = -1
        if int(power[ia]):
            esign *= applied
        sa.append(sv)
        ta.append(tv)
        groups.append(int(meta["case_axis_group"][case_id, out]))
        signs.append(sg)
    return sa, ta, groups, signs, esign


def arithmetic_entry(case_id, s, t, meta, q, dim):
    """Non-compact form ``orbit_rank*n^2 + canon_s*n + canon_t`` (+ sign 1)."""
    sa, ta, groups, _ = canonicalize(case_id, s, t, meta, q, dim)

    def cswap(i, j):
        if groups[i] != groups[j]:
            return
        if (sa[j], ta[j]) < (sa[i], ta[i]):
            sa[i], sa[j] = sa[j], sa[i]
            ta[i], ta[j] = ta[j], ta[i]

    if dim == 2:
        cswap(0, 1)
    else:
        cswap(0, 1)
        cswap(1, 2)
        cswap(0, 1)
        cswap(0, 2)
    n = q**dim
    return int(meta["case_orbit_ranks"][case_id]) * n * n + encode_axes(sa, q) * n + encode_axes(
        ta, q
    ), 1


def rank_multiset(values, alpha):
    rank = 0
    lo = 0
    k = len(values)
    for i, v in enumerate(values):
        rem = k - i - 1
        for c in range(lo, v):
            rank += math.comb(alpha - c + rem - 1, rem)
        lo = v
    return rank


def compact_entry(case_id, s, t, meta, q, dim):
    """``(row, entry sign)``, ``orbit_arithmetic.py:566-634``."""
    sa, ta, groups, signs, esign = canonicalize_signed(case_id, s, t, meta, q, dim)
    codes = [a * q + b for a, b in zip(sa, ta)]
    pc = q * q
    fpc = (pc + 1) // 2
    comp = 0
    for gid in range(max(groups) + 1):
        axes = [a for a, g in enumerate(groups) if g == gid]
        if not axes:
            continue
        alpha = fpc if any(signs[a] == 0 for a in axes) else pc
        vals = sorted(codes[a] for a in axes)
        comp = comp * math.comb(alpha + len(axes) - 1, len(axes)) + rank_multiset(vals, alpha)
    rank = int(meta["case_orbit_ranks"][case_id])
    return int(meta["case_value_offsets"][rank]) + comp, esign


def representatives(meta, q, dim):
    pc = q * q
    fpc = (pc + 1) // 2
    n = q**dim
    out = []
    for cid in meta["canonical_case_ids"].tolist():
        groups = meta["case_axis_group"][cid]
        signs = meta["case_axis_sign"][cid]
        opts = []
        for gid in range(int(groups.max()) + 1):
            axes = np.flatnonzero(groups == gid).tolist()
            a = fpc if np.any(signs[axes] == 0) else pc
            opts.append((axes, tuple(itertools.combinations_with_replacement(range(a), len(axes)))))
        for combo in itertools.product(*(o for _, o in opts)):
            sa = [0] * dim
            ta = [0] * dim
            for (axes, _), vals in zip(opts, combo):
                for ax, code in zip(axes, vals):
                    sa[ax] = code // q
                    ta[ax] = code % q
            out.append(cid * n * n + encode_axes(sa, q) * n + encode_axes(ta, q))
    return np.asarray(out, np.int64)


def rows_for_case(meta, case_id, q, dim):
    """Dense ``rows[s, t]`` for one case using the device-code closed forms
    of ``list1.py:604-902`` (min/max/sum of each group's pair codes)."""
    n = q**dim
    s = np.arange(n)[:, None] * np.ones((1, n), np.int64)
    t = np.ones((n, 1), np.int64) * np.arange(n)[None, :]
    if dim == 2:
        sr = [s // q, s % q]
        tr = [t // q, t % q]
    else:
        sr = [s // (q * q), (s // q) % q, s % q]
        tr = [t // (q * q), (t // q) % q, t % q]
    perm = meta["case_axis_perm"][case_id].astype(int)
    sign = meta["case_axis_sign"][case_id].astype(int)
    grp = meta["case_axis_group"][case_id].astype(int)
    src, tgt = [], []
    for ax in range(dim):
        sp, tp = sr[perm[ax]], tr[perm[ax]]
        sfix = np.where(sign[ax] < 0, q - 1 - sp, sp)
        tfix = np.where(sign[ax] < 0, q - 1 - tp, tp)
        sfl, tfl = q - 1 - sp, q - 1 - tp
        zf = (sign[ax] == 0) & ((sfl < sp) | ((sfl == sp) & (tfl < tp)))
        src.append(np.where(zf, sfl, sfix))
        tgt.append(np.where(zf, tfl, tfix))
    pair = [src[a] * q + tgt[a] for a in range(dim)]
    pc = q * q
    fpc = (pc + 1) // 2
    ranks, counts = [], []
    for g in range(dim):
        members = [a for a in range(dim) if grp[a] == g]
        size = len(members)
        if size == 0:
            ranks.append(0)
            counts.append(1)
            continue
        A = fpc if any(sign[a] == 0 for a in members) else pc
        vals = [pair[a] for a in members]
        lo = np.minimum.reduce(vals)
        hi = np.maximum.reduce(vals)
        mid = sum(vals) - lo - hi
        if size == 1:
            r, c = lo, A
        elif size == 2:
            r = A * (A + 1) // 2 - (A - lo) * (A - lo + 1) // 2 + (hi - lo)
            c = A * (A + 1) // 2
        else:
            r = (
                A * (A + 1) * (A + 2) // 6
                - (A - lo) * (A - lo + 1) * (A - lo + 2) // 6
                + (A - lo) * (A - lo + 1) // 2
                - (A - mid) * (A - mid + 1) // 2
                + (hi - mid)
            )
            c = A * (A + 1) * (A + 2) // 6
        ranks.append(r)
        counts.append(c)
    comp = ranks[0]
    for g in range(1, dim):
        comp = comp * counts[g] + ranks[g]
    rank = int(meta["case_orbit_ranks"][case_id])
    return int(meta["case_value_offsets"][rank]) + comp


def signs_for_case(meta, case_id, q, dim):
    """Dense ``signs[s, t]`` of one case in the device-code form of
    ``list1.py:799-832`` (``entry_sign_factor``: the transform sign of every
    axis whose permuted input axis has sign power 1)."""
    n = q**dim
    s = np.arange(n)[:, None] * np.ones((1, n), np.int64)
    t = np.ones((n, 1), np.int64) * np.arange(n)[None, :]
    sr = [(s // q ** (dim - 1 - a)) % q for a in range(dim)]
    tr = [(t // q ** (dim - 1 - a)) % q for a in range(dim)]
    perm = meta["case_axis_perm"][case_id].astype(int)
    sign = meta["case_axis_sign"][case_id].astype(int)
    power = meta.get("axis_sign_power", [0] * dim)
    out = np.ones((n, n), np.int64)
    for ax in range(dim):
        sp, tp = sr[perm[ax]], tr[perm[ax]]
        sfl, tfl = q - 1 - sp, q - 1 - tp
        zf = (sign[ax] == 0) & ((sfl < sp) | ((sfl == sp) & (tfl < tp)))
        tsign = np.where((sign[ax] < 0) | zf, -1, 1)
        if int(power[perm[ax]]):
            out = out * tsign
    return out


# ----------------------------------------------------- brute-force orbits
def orbit_min_ids(dim, q):
    """Smallest full entry id of the orbit of every entry (dense, int64)."""
    vecs, case_indices = gallery(dim)
    n = q**dim
    nc = len(vecs)
    axes = np.array([decode_axes(m, q, dim) for m in range(n)], np.int64)  # [n, dim]
    vec_arr = np.asarray(vecs, np.int64)
    best = None
    for sgn in itertools.product((-1, 1), repeat=dim):
        sgn = np.asarray(sgn)
        for perm in itertools.permutations(range(dim)):
            perm = list(perm)
            m_ax = axes.copy()
            m_ax[:, sgn < 0] = q - 1 - m_ax[:, sgn < 0]
            m_ax = m_ax[:, perm]
            mode_map = np.zeros(n, np.int64)
            for ax in range(dim):
                mode_map = mode_map * q + m_ax[:, ax]
            cv = (vec_arr * sgn)[:, perm]
            case_map = np.array([case_indices[encode(v)] for v in cv], np.int64)
            assert np.all(case_map >= 0)
            ids = (
                case_map[:, None, None] * n * n
                + mode_map[None, :, None] * n
                + mode_map[None, None, :]
            )
            best = ids if best is None else np.minimum(best, ids)
    return best.reshape(nc * n * n)


def group_transforms(dim, fixed_axes=(), sign_power=None):
    """The symmetry group of a kernel as (axis signs, axis permutation,
    value sign): ``nearfield_potential_table.py:55-140``
    (``_kernel_symmetry_meta``) for axis derivatives -- a permutation must fix
    every derivative axis, and a flip of a derivative axis negates values."""
    power = list(sign_power) if sign_power is not None else [0] * dim
    out = []
    for sgn in itertools.product((-1, 1), repeat=dim):
        for perm in itertools.permutations(range(dim)):
            if any(perm[a] != a for a in fixed_axes):
                continue
            vs = 1
            for a in range(dim):
                if power[a]:
                    vs *= sgn[a]
            out.append((sgn, perm, vs))
    return out


def transform_entries(dim, q, sgn, perm):
    """Full-id map of one group element: ``new_id[old_id]`` (int64, dense)."""
    vecs, case_indices = gallery(dim)
    n = q**dim
    axes = np.array([decode_axes(m, q, dim) for m in range(n)], np.int64)
    sgn = np.asarray(sgn)
    perm = list(perm)
    m_ax = axes.copy()
    m_ax[:, sgn < 0] = q - 1 - m_ax[:, sgn < 0]
    m_ax = m_ax[:, perm]
    mode_map = np.zeros(n, np.int64)
    for ax in range(dim):
        mode_map = mode_map * q + m_ax[:, ax]
    cv = (np.asarray(vecs, np.int64) * sgn)[:, perm]
    case_map = np.array([case_indices[encode(v)] for v in cv], np.int64)
    assert np.all(case_map >= 0)
    ids = case_map[:, None, None] * n * n + mode_map[None, :, None] * n + mode_map[None, None, :]
    return ids.reshape(-1)
