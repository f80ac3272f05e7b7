"""Box trees with per-leaf tensor Gauss-Legendre nodes, and their list-1 CSR.

Builds the inputs of the list-1 path the reference gets from ``meshgen`` and
boxtree's traversal (``meshgen.py:391-444``): leaf boxes of a (possibly
adaptive, 2:1-balanced) ``2^d``-tree on a cube, ``q^d`` nodes per leaf in
lexicographic order (axis 0 slowest, ``tree_interactive_build.py:979-998``),
and for each leaf the list of leaves touching it, itself included (list 1).

Only leaves carry nodes, so the box arrays below index leaves directly:
``box_target_starts[b] = b * q^d``, ``box_target_counts_nonchild[b] = q^d``.
Leaves are ordered by level, then lexicographic integer index.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

from . import quadrature


@dataclass
class BoxTree:
    dim: int
    quad_order: int
    root_vertex: np.ndarray  # [dim]
    root_extent: float
    leaf_levels: np.ndarray  # uint8 [nleaves]
    leaf_index: np.ndarray  # int64 [nleaves, dim] integer index at its level
    # derived
    box_centers: np.ndarray = field(init=False)  # [dim, nboxes]
    box_levels: np.ndarray = field(init=False)
    box_target_starts: np.ndarray = field(init=False)
    box_target_counts_nonchild: np.ndarray = field(init=False)
    target_boxes: np.ndarray = field(init=False)
    neighbor_source_boxes_starts: np.ndarray = field(init=False)
    neighbor_source_boxes_lists: np.ndarray = field(init=False)
    points: np.ndarray = field(init=False)  # [dim, n_points]
    weights: np.ndarray = field(init=False)  # [n_points]
    local_ids: np.ndarray = field(init=False)  # [n_points] box-local node id

    def __post_init__(self):
        self.leaf_levels = np.asarray(self.leaf_levels, dtype=np.uint8)
        self.leaf_index = np.asarray(self.leaf_index, dtype=np.int64)
        n = self.n_q_points
        h = self.root_extent / (2.0 ** self.leaf_levels.astype(np.float64))
        lower = self.root_vertex[None, :] + self.leaf_index * h[:, None]
        self.box_centers = np.ascontiguousarray((lower + 0.5 * h[:, None]).T)
        self.box_levels = self.leaf_levels.copy()
        nb = self.nboxes
        self.box_target_starts = (np.arange(nb, dtype=np.int64) * n).astype(np.int32)
        self.box_target_counts_nonchild = np.full(nb, n, dtype=np.int32)
        self.target_boxes = np.arange(nb, dtype=np.int32)
        starts, lists = _list1_csr(self.dim, self.leaf_levels, self.leaf_index)
        self.neighbor_source_boxes_starts = starts
        self.neighbor_source_boxes_lists = lists
        x1, w1 = quadrature.gauss_legendre_01(self.quad_order)
        ref = quadrature.tensor_q_points(self.quad_order, self.dim, 1.0)  # [n, dim]
        wref = np.prod(
            np.stack(np.meshgrid(*([w1] * self.dim), indexing="ij")).reshape(self.dim, -1), axis=0
        )
        pts = lower[:, None, :] + ref[None, :, :] * h[:, None, None]
        self.points = np.ascontiguousarray(pts.reshape(-1, self.dim).T)
        self.weights = (wref[None, :] * (h**self.dim)[:, None]).reshape(-1)
        self.local_ids = np.tile(np.arange(n, dtype=np.int32), nb)

    @property
    def n_q_points(self) -> int:
        return self.quad_order**self.dim

    @property
    def nboxes(self) -> int:
        return int(self.leaf_levels.size)

    @property
    def n_points(self) -> int:
        return self.nboxes * self.n_q_points

    @property
    def nlevels(self) -> int:
        return int(self.leaf_levels.max()) + 1

    @property
    def n_list1_pairs(self) -> int:
        return int(self.neighbor_source_boxes_starts[-1])

    ntargets = n_points
    sources_are_targets = True


def _list1_csr(dim, levels, index):
    """Leaves touching each leaf (itself included), sorted by box id."""
    levels = np.asarray(levels, dtype=np.int64)
    index = np.asarray(index, dtype=np.int64)
    nb = levels.size
    if nb == 0:
        return np.zeros(1, np.int32), np.zeros(0, np.int32)
    lmax = int(levels.max())
    lookup = {}
    for b in range(nb):
        lookup[(int(levels[b]), *map(int, index[b]))] = b
    if np.all(levels == levels[0]):
        return _list1_uniform(dim, int(levels[0]), index, lookup)
    offs = list(itertools.product((-1, 0, 1), repeat=dim))
    starts = np.zeros(nb + 1, dtype=np.int64)
    out = []
    for b in range(nb):
        lv = int(levels[b])
        ib = index[b]
        nbrs = set()
        span = 1 << lv
        for o in offs:
            nbi = ib + np.asarray(o)
            if np.any(nbi < 0) or np.any(nbi >= span):
                continue
            key = (lv, *map(int, nbi))
            if key in lookup:
                nbrs.add(lookup[key])
                continue
            if lv > 0:
                pkey = (lv - 1, *map(int, nbi // 2))
                if pkey in lookup:
                    nbrs.add(lookup[pkey])
                    continue
            if lv < lmax:
                # finer neighbours: children of nbi touching box ib
                for ch in itertools.product((0, 1), repeat=dim):
                    ci = 2 * nbi + np.asarray(ch)
                    # closed boxes [ci, ci+1] (level lv+1) and [2ib, 2ib+2]
                    if np.all((ci + 1 >= 2 * ib) & (ci <= 2 * ib + 2)):
                        ckey = (lv + 1, *map(int, ci))
                        if ckey in lookup:
                            nbrs.add(lookup[ckey])
                        else:
                            raise ValueError("tree is not 2:1 balanced")
                continue
            raise ValueError("tree leaves do not tile the domain")
        lst = sorted(nbrs)
        out.extend(lst)
        starts[b + 1] = starts[b] + len(lst)
    return starts.astype(np.int32), np.asarray(out, dtype=np.int32)


def _list1_uniform(dim, level, index, lookup):
    span = 1 << level
    nb = index.shape[0]
    lin = np.zeros(nb, dtype=np.int64)
    for ax in range(dim):
        lin = lin * span + index[:, ax]
    inv = np.full(span**dim, -1, dtype=np.int64)
    inv[lin] = np.arange(nb)
    offs = np.array(list(itertools.product((-1, 0, 1), repeat=dim)), dtype=np.int64)
    cand = index[:, None, :] + offs[None, :, :]  # [nb, 3^d, dim]
    valid = np.all((cand >= 0) & (cand < span), axis=2)
    clin = np.zeros(cand.shape[:2], dtype=np.int64)
    for ax in range(dim):
        clin = clin * span + np.clip(cand[:, :, ax], 0, span - 1)
    nbr = np.where(valid, inv[clin], -1)
    nbr = np.sort(np.where(nbr >= 0, nbr, np.iinfo(np.int64).max), axis=1)
    counts = valid.sum(axis=1)
    starts = np.zeros(nb + 1, dtype=np.int64)
    starts[1:] = np.cumsum(counts)
    lists = nbr[nbr != np.iinfo(np.int64).max]
    return starts.astype(np.int32), lists.astype(np.int32)


def uniform_tree(dim, quad_order, level, a=-1.0, b=1.0) -> BoxTree:
    """All leaves at ``level`` on ``[a, b]^dim``."""
    span = 1 << int(level)
    idx = np.stack(
        np.meshgrid(*([np.arange(span)] * dim), indexing="ij"), axis=-1
    ).reshape(-1, dim)
    return BoxTree(
        dim=dim,
        quad_order=quad_order,
        root_vertex=np.full(dim, float(a)),
        root_extent=float(b - a),
        leaf_levels=np.full(idx.shape[0], level, dtype=np.uint8),
        leaf_index=idx,
    )


def adaptive_tree(dim, quad_order, refine, min_level=1, max_level=6, a=-1.0, b=1.0) -> BoxTree:
    """Refine leaves while ``refine(level, lower, h) -> bool``; then 2:1-balance.

    Balance includes edge and corner neighbours (any two touching leaves
    differ by at most one level), which list 1 relies on.
    """
    root_vertex = np.full(dim, float(a))
    extent = float(b - a)
    leaves = {(0, (0,) * dim)}
    changed = True
    while changed:
        changed = False
        new = set()
        for lv, ix in leaves:
            h = extent / 2**lv
            lower = root_vertex + np.asarray(ix) * h
            if lv < max_level and (lv < min_level or refine(lv, lower, h)):
                for ch in itertools.product((0, 1), repeat=dim):
                    new.add((lv + 1, tuple(2 * i + c for i, c in zip(ix, ch))))
                changed = True
            else:
                new.add((lv, ix))
        leaves = new
    leaves = _balance(dim, leaves)
    items = sorted(leaves)
    return BoxTree(
        dim=dim,
        quad_order=quad_order,
        root_vertex=root_vertex,
        root_extent=extent,
        leaf_levels=np.array([lv for lv, _ in items], dtype=np.uint8),
        leaf_index=np.array([ix for _, ix in items], dtype=np.int64).reshape(-1, dim),
    )


def adaptive_tree_batched(dim, quad_order, refine_batch, min_level=1, max_level=6, a=-1.0,
                          b=1.0) -> BoxTree:
    """Level-by-level refinement with a vectorised criterion
    ``refine_batch(level, lowers[N, dim], h) -> bool[N]``; then 2:1-balance."""
    root_vertex = np.full(dim, float(a))
    extent = float(b - a)
    done = []  # (level, index) leaves that stay
    cur = np.zeros((1, dim), dtype=np.int64)
    for lv in range(0, max_level + 1):
        if cur.shape[0] == 0:
            break
        h = extent / 2**lv
        if lv >= max_level:
            ref = np.zeros(cur.shape[0], dtype=bool)
        elif lv < min_level:
            ref = np.ones(cur.shape[0], dtype=bool)
        else:
            ref = np.asarray(refine_batch(lv, root_vertex[None, :] + cur * h, h), dtype=bool)
        done.extend((lv, tuple(map(int, ix))) for ix in cur[~ref])
        kids = cur[ref]
        ch = np.array(list(itertools.product((0, 1), repeat=dim)), dtype=np.int64)
        cur = (2 * kids[:, None, :] + ch[None, :, :]).reshape(-1, dim)
    leaves = _balance(dim, set(done))
    items = sorted(leaves)
    return BoxTree(
        dim=dim,
        quad_order=quad_order,
        root_vertex=root_vertex,
        root_extent=extent,
        leaf_levels=np.array([lv for lv, _ in items], dtype=np.uint8),
        leaf_index=np.array([ix for _, ix in items], dtype=np.int64).reshape(-1, dim),
    )


def interpolation_refiner(func, dim, quad_order, tol, relative_to=None):
    """Batched criterion: refine a box while ``max |f - I_q f| > tol`` on the
    ``(q+1)^d`` Gauss-Legendre check grid, ``I_q`` the tensor Lagrange
    interpolant on the box's ``q^d`` nodes; ``func(points[dim, m])``.
    With ``relative_to`` the tolerance is ``tol * relative_to``."""
    x_q, _ = quadrature.gauss_legendre_01(quad_order)
    x_c, _ = quadrature.gauss_legendre_01(quad_order + 1)
    w = quadrature.barycentric_weights(x_q)
    # P[c, j] = l_j(x_c): 1D interpolation matrix from q nodes to q+1 checks
    diff = x_c[:, None] - x_q[None, :]
    P = (w[None, :] / diff) / np.sum(w[None, :] / diff, axis=1, keepdims=True)
    ref_q = quadrature.tensor_q_points(quad_order, dim, 1.0)
    ref_c = quadrature.tensor_q_points(quad_order + 1, dim, 1.0)
    thr = tol * (1.0 if relative_to is None else float(relative_to))

    def refine_batch(level, lowers, h):
        N = lowers.shape[0]
        out = np.zeros(N, dtype=bool)
        for b0 in range(0, N, 4096):
            lw = lowers[b0:b0 + 4096]
            pq = lw[:, None, :] + h * ref_q[None, :, :]
            pc = lw[:, None, :] + h * ref_c[None, :, :]
            fq = func(pq.reshape(-1, dim).T).reshape(lw.shape[0], *([quad_order] * dim))
            fc = func(pc.reshape(-1, dim).T).reshape(lw.shape[0], *([quad_order + 1] * dim))
            interp = fq
            for ax in range(dim):
                interp = np.moveaxis(np.tensordot(interp, P, axes=([1 + ax], [1])), -1, 1 + ax)
            err = np.abs(fc - interp).reshape(lw.shape[0], -1).max(axis=1)
            out[b0:b0 + 4096] = err > thr
        return out

    return refine_batch


class Forest:
    """Several trees on the same root box as one set of box arrays (box and
    point ids offset tree by tree), so one plan -- one launch per eval --
    covers all of them.  List 1 never crosses trees."""

    sources_are_targets = True

    def __init__(self, trees):
        t0 = trees[0]
        for t in trees:
            if (t.dim, t.quad_order, t.root_extent) != (t0.dim, t0.quad_order, t0.root_extent) or \
                    not np.array_equal(t.root_vertex, t0.root_vertex):
                raise ValueError("forest trees must share dimension, order and root box")
        self.trees = list(trees)
        self.dim, self.quad_order = t0.dim, t0.quad_order
        self.root_vertex, self.root_extent = t0.root_vertex, t0.root_extent
        boff = np.cumsum([0] + [t.nboxes for t in trees])
        poff = np.cumsum([0] + [t.n_points for t in trees])
        self.tree_box_offsets, self.tree_point_offsets = boff, poff
        cat = np.concatenate
        self.box_centers = np.ascontiguousarray(cat([t.box_centers for t in trees], axis=1))
        self.box_levels = cat([t.box_levels for t in trees])
        self.box_target_starts = cat([t.box_target_starts.astype(np.int64) + poff[i]
                                      for i, t in enumerate(trees)]).astype(np.int32)
        self.box_target_counts_nonchild = cat([t.box_target_counts_nonchild for t in trees])
        self.target_boxes = cat([t.target_boxes.astype(np.int64) + boff[i]
                                 for i, t in enumerate(trees)]).astype(np.int32)
        starts, lists, acc = [np.zeros(1, np.int64)], [], 0
        for i, t in enumerate(trees):
            starts.append(t.neighbor_source_boxes_starts[1:].astype(np.int64) + acc)
            lists.append(t.neighbor_source_boxes_lists.astype(np.int64) + boff[i])
            acc += int(t.neighbor_source_boxes_starts[-1])
        self.neighbor_source_boxes_starts = cat(starts).astype(np.int32)
        self.neighbor_source_boxes_lists = cat(lists).astype(np.int32)
        self.points = np.ascontiguousarray(cat([t.points for t in trees], axis=1))
        self.weights = cat([t.weights for t in trees])
        self.local_ids = cat([t.local_ids for t in trees])

    n_q_points = property(lambda self: self.quad_order**self.dim)
    nboxes = property(lambda self: int(self.box_levels.size))
    n_points = property(lambda self: int(self.local_ids.size))
    nlevels = property(lambda self: int(self.box_levels.max()) + 1)
    n_list1_pairs = property(lambda self: int(self.neighbor_source_boxes_starts[-1]))
    ntargets = n_points


def _balance(dim, leaves):
    leaves = set(leaves)
    offs = [o for o in itertools.product((-1, 0, 1), repeat=dim) if any(o)]
    while True:
        leaf_set = leaves
        lmax = max(lv for lv, _ in leaf_set)
        by_level = {}
        for lv, ix in leaf_set:
            by_level.setdefault(lv, set()).add(ix)

        def covering_leaf_level(lv, ix):
            # level of the leaf containing cell (lv, ix), or None if finer
            for l2 in range(lv, -1, -1):
                anc = tuple(i >> (lv - l2) for i in ix)
                if anc in by_level.get(l2, ()):
                    return l2
            return None

        to_split = set()
        for lv in range(lmax, 1, -1):
            for ix in by_level.get(lv, ()):
                span = 1 << lv
                for o in offs:
                    nix = tuple(i + d for i, d in zip(ix, o))
                    if any(v < 0 or v >= span for v in nix):
                        continue
                    cl = covering_leaf_level(lv, nix)
                    if cl is not None and cl < lv - 1:
                        anc = tuple(i >> (lv - cl) for i in nix)
                        to_split.add((cl, anc))
        if not to_split:
            return leaves
        for lv, ix in to_split:
            leaves.discard((lv, ix))
            for ch in itertools.product((0, 1), repeat=dim):
                leaves.add((lv + 1, tuple(2 * i + c for i, c in zip(ix, ch))))


def list1_csr_gpu(dim, leaf_levels, leaf_index, stream=None):
    """List-1 CSR (leaves touching each leaf, itself included, sorted by id)
    built on the GPU by ``vl1_csr_builder_*``; returns CUDA int32 tensors
    ``(starts[n + 1], lists[n_pairs])``.  Same result as the host builder of
    :class:`BoxTree` (``_list1_csr``)."""
    import torch

    from . import _capi

    L = _capi.lib()
    lv = np.ascontiguousarray(leaf_levels, dtype=np.uint8)
    ix = np.ascontiguousarray(leaf_index, dtype=np.int32).reshape(lv.size, dim)
    handle = _capi.vp()
    npairs = _capi.i64(0)
    sh = _capi.stream_handle(stream)
    _capi.check(L.vl1_csr_builder_create(_capi.C.byref(handle), int(dim), int(lv.size),
                                         lv.ctypes.data, ix.ctypes.data, _capi.C.byref(npairs), sh),
                "vl1_csr_builder_create")
    try:
        starts = torch.empty(lv.size + 1, dtype=torch.int32, device="cuda")
        lists = torch.empty(max(int(npairs.value), 1), dtype=torch.int32, device="cuda")
        _capi.check(L.vl1_csr_builder_fill(handle, starts.data_ptr(), lists.data_ptr(), sh),
                    "vl1_csr_builder_fill")
    finally:
        L.vl1_csr_builder_destroy(handle)
    return starts, lists[: int(npairs.value)]
