"""Pins of the list-1 oracles (numpy and C/OpenMP) on the CPU.

* completeness: constant kernel, f = 1 on a 2^d-box mesh gives 2^d at every
  target (reference: test_nearfield_interaction_completeness.py:175-176);
* multi-RHS superposition and linearity (test_volume_fmm.py:6892-7011);
* the numpy oracle against a brute-force dense sum over every
  (target, source) pair built from the full table definition;
* the C oracle against the numpy oracle (uniform and adaptive trees, all
  three scaling policies);
* the float case-id encoding agrees with the exact integer case vector.
"""

import numpy as np
import pytest

from oracle import cpu as ocpu
from oracle import list1_ref, orbit_ref
from volumential_amd import densities, orbit, tree as vtree


def synthetic_table(dim, q, seed=0, n_tables=1):
    """Random compact table values (the oracles do not care where they
    came from); mode normalisers = GL weight outer product."""
    m = orbit.build_case_metadata(dim, q)
    rng = np.random.default_rng(seed)
    data = rng.standard_normal((n_tables, m.n_compact_entries))
    w = np.polynomial.legendre.leggauss(q)[1] * 0.5
    nm = np.multiply.outer(w, w).reshape(-1) if dim == 2 else np.zeros(q**dim)
    return data, np.tile(nm, (n_tables, 1))


def test_completeness_constant_kernel():
    for dim in (2, 3):
        q = 2
        tr = vtree.uniform_tree(dim, q, 1, a=-1, b=1)  # 2^d boxes of extent 1
        meta = orbit_ref.case_metadata(dim, q)
        w = np.polynomial.legendre.leggauss(q)[1] * 0.5
        wt = np.prod(np.stack(np.meshgrid(*([w] * dim), indexing="ij")).reshape(dim, -1), axis=0)
        # constant-kernel table on [0,1]: T(c, s, t) = prod GL weights of s
        m = orbit.build_case_metadata(dim, q)
        reps = orbit.enumerate_representatives(m)
        n = q**dim
        data = wt[(reps % (n * n)) // n][None, :]
        f = np.ones(tr.n_points)
        out = list1_ref.near_field_from_tree(
            tr, data, np.zeros((1, n)), meta, f, mode="canonical", scaling_power=dim,
            table_root_extent=1.0,
        )
        assert np.allclose(out, 2**dim, rtol=0, atol=1e-13)


def brute_force(tr, full_table, coefs, scal_fn):
    """Dense definition: y[t] = sum_{s in list1} f[s] T(case, s, t) * scal."""
    n = tr.n_q_points
    out = np.zeros((coefs.shape[0], tr.n_points))
    _, ci = orbit_ref.gallery(tr.dim)
    for i, tb in enumerate(tr.target_boxes):
        for sb in tr.neighbor_source_boxes_lists[
            tr.neighbor_source_boxes_starts[i] : tr.neighbor_source_boxes_starts[i + 1]
        ]:
            c = list1_ref.case_id_exact(
                tr.box_centers[:, tb], tr.box_centers[:, sb], tr.box_levels[sb], tr.root_extent, ci
            )
            T = full_table[c]  # [s, t]
            sl = slice(sb * n, (sb + 1) * n)
            out[:, tb * n : (tb + 1) * n] += coefs[:, sl] @ T * scal_fn(tr.box_levels[sb])
    return out


@pytest.mark.parametrize("dim,q,level", [(2, 3, 2), (3, 2, 2)])
def test_numpy_oracle_matches_dense_definition(dim, q, level):
    tr = vtree.uniform_tree(dim, q, level)
    data, nm = synthetic_table(dim, q)
    m = orbit.build_case_metadata(dim, q)
    n = q**dim
    c, s, t = np.meshgrid(np.arange(m.n_cases), np.arange(n), np.arange(n), indexing="ij")
    full = data[0][orbit.entry_rows(m, c, s, t)]  # [case, s, t]
    coefs = densities.gaussian_mixture_batch(tr.points, [1, 2])
    H = 2.0
    ref = brute_force(tr, full, coefs, lambda lv: (tr.root_extent / 2**lv / H) ** 2)
    meta = orbit_ref.case_metadata(dim, q)
    out = list1_ref.near_field_from_tree(
        tr, data, nm * 0, meta, coefs, mode="canonical", scaling_power=2, table_root_extent=H
    )
    assert np.allclose(out, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_superposition_and_linearity():
    tr = vtree.uniform_tree(2, 3, 2)
    data, nm = synthetic_table(2, 3, seed=4)
    meta = orbit_ref.case_metadata(2, 3)
    f = densities.gaussian_mixture_batch(tr.points, [7, 8, 9])
    kw = dict(mode="canonical", scaling_power=2, log_displacement=True, table_root_extent=2.0)
    single = [list1_ref.near_field_from_tree(tr, data, nm, meta, f[i], **kw)[0] for i in range(3)]
    batched = list1_ref.near_field_from_tree(tr, data, nm, meta, f, **kw)
    assert np.allclose(batched, np.stack(single), rtol=1e-14, atol=1e-14)
    comb = list1_ref.near_field_from_tree(tr, data, nm, meta, 2 * f[0] - 3 * f[1], **kw)[0]
    assert np.allclose(comb, 2 * single[0] - 3 * single[1], rtol=1e-12, atol=1e-12)


def _adaptive(dim, q):
    def refine(lv, lower, h):
        c = lower + 0.5 * h
        return np.linalg.norm(c - 0.23) < 1.2 * h

    return vtree.adaptive_tree(dim, q, refine, min_level=1, max_level=4)


@pytest.mark.parametrize("dim,q", [(2, 3), (3, 2)])
@pytest.mark.parametrize("mode", ["canonical", "per_level"])
def test_c_oracle_matches_numpy(dim, q, mode):
    tr = _adaptive(dim, q)
    assert len(set(tr.box_levels.tolist())) > 1
    n_tables = 1 if mode == "canonical" else tr.nlevels
    data, nm = synthetic_table(dim, q, seed=2, n_tables=n_tables)
    meta = orbit_ref.case_metadata(dim, q)
    f = densities.gaussian_mixture_batch(tr.points, [3, 4])
    kw = dict(mode=mode, scaling_power=2, log_displacement=(dim == 2 and mode == "canonical"),
              table_root_extent=2.0)
    a = list1_ref.near_field_from_tree(tr, data, nm, meta, f, **kw)
    b = ocpu.near_field(tr, data, nm, meta, f, nthreads=2, **kw)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_float_case_encoding_matches_integer():
    for dim in (2, 3):
        tr = _adaptive(dim, 2)
        _, ci = orbit_ref.gallery(dim)
        for i, tb in enumerate(tr.target_boxes):
            for sb in tr.neighbor_source_boxes_lists[
                tr.neighbor_source_boxes_starts[i] : tr.neighbor_source_boxes_starts[i + 1]
            ]:
                a, _ = list1_ref.case_id_float(
                    tr.box_centers[:, tb], tr.box_centers[:, sb], tr.box_levels[sb],
                    tr.root_extent, ci,
                )
                b = list1_ref.case_id_exact(
                    tr.box_centers[:, tb], tr.box_centers[:, sb], tr.box_levels[sb],
                    tr.root_extent, ci,
                )
                assert a == b >= 0


def test_list1_csr_adjacency_bruteforce():
    tr = _adaptive(2, 2)
    lo = tr.box_centers.T - 0.5 * (tr.root_extent / 2.0 ** tr.box_levels.astype(float))[:, None]
    hi = lo + (tr.root_extent / 2.0 ** tr.box_levels.astype(float))[:, None]
    for i in range(tr.nboxes):
        touch = np.all((lo <= hi[i] + 1e-12) & (hi >= lo[i] - 1e-12), axis=1)
        expect = np.flatnonzero(touch)
        got = tr.neighbor_source_boxes_lists[
            tr.neighbor_source_boxes_starts[i] : tr.neighbor_source_boxes_starts[i + 1]
        ]
        assert np.array_equal(np.sort(got), expect)


def _integer_touch_lists(levels, index):
    """List 1 from integer cell indices alone: leaves a (level la, index ia)
    and b touch iff on every axis the closed intervals
    [ia 2^(L-la), (ia+1) 2^(L-la)] and [ib 2^(L-lb), (ib+1) 2^(L-lb)] on the
    finest grid L overlap (a second formulation of the geometric test of
    ``list1_ref.list1_csr_bruteforce``, exact in integers)."""
    L = int(levels.max())
    scale = (1 << (L - levels.astype(np.int64)))[:, None]
    lo = index.astype(np.int64) * scale
    hi = lo + scale
    out = []
    for a in range(levels.size):
        touch = np.all((lo <= hi[a]) & (hi >= lo[a]), axis=1)
        out.append(np.flatnonzero(touch))
    return out


def _c2_style_tree(max_level=4):
    fden = (lambda P: densities.gaussian_mixture(P, 0))
    return vtree.adaptive_tree_batched(3, 2, vtree.interpolation_refiner(fden, 3, 2, 1e-3),
                                       min_level=2, max_level=max_level)


@pytest.mark.parametrize("kind", ["uniform3", "adaptive3", "c2style", "adaptive2", "forest"])
def test_list1_csr_matches_oracle_bruteforce(kind):
    """The tree builder's list-1 CSR (the one every parity test hands to the
    GPU) equals the oracle's own geometric list 1 and the integer-index
    formulation, in 3D (uniform level 3, a point-refined adaptive octree and
    a C2-style density-refined 2:1 octree), 2D, and on a forest of two
    beam-refined quadtrees (list 1 never crosses trees)."""
    if kind == "uniform3":
        tr = vtree.uniform_tree(3, 2, 3)
    elif kind == "adaptive3":
        tr = _adaptive(3, 2)
    elif kind == "c2style":
        tr = _c2_style_tree()
        assert tr.nlevels >= 4
    elif kind == "adaptive2":
        tr = _adaptive(2, 2)
    else:
        trees = [vtree.adaptive_tree_batched(
            2, 4, vtree.interpolation_refiner(
                (lambda P, s=s: densities.helmholtz_beams(P, s, k=40.0)), 2, 4, 1e-5,
                relative_to=4.0), min_level=3, max_level=6) for s in (0, 1)]
        tr = vtree.Forest(trees)
    starts, lists = list1_ref.list1_csr_bruteforce(tr.box_centers, tr.box_levels, tr.root_extent,
                                                   tr.target_boxes,
                                                   box_group=list1_ref.tree_groups(tr))
    if kind == "forest":
        offs = tr.tree_box_offsets
        ints = [v + offs[i] for i, t in enumerate(tr.trees)
                for v in _integer_touch_lists(t.leaf_levels, t.leaf_index)]
    else:
        ints = _integer_touch_lists(tr.leaf_levels, tr.leaf_index)
    for i, tb in enumerate(tr.target_boxes):
        got = np.sort(tr.neighbor_source_boxes_lists[
            tr.neighbor_source_boxes_starts[i]:tr.neighbor_source_boxes_starts[i + 1]])
        want = lists[starts[i]:starts[i + 1]]
        assert np.array_equal(got, want), (kind, i)
        assert np.array_equal(want, ints[tb]), (kind, i)


@pytest.mark.parametrize("dim,level", [(3, 3), (3, 4), (2, 5)])
def test_oracle_csr_uniform_counts(dim, level):
    """Closed form on a uniform tree: a box at integer index i has
    prod_axes (3 - [i_a = 0] - [i_a = 2^L - 1]) list-1 boxes, P = (3 2^L -
    2)^d pairs in all (94^3 = 830,584 at level 5), and the relation is
    symmetric."""
    tr = vtree.uniform_tree(dim, 1, level)
    starts, lists = list1_ref.list1_csr_bruteforce(tr.box_centers, tr.box_levels, tr.root_extent,
                                                   tr.target_boxes)
    m = 2**level
    edge = (tr.leaf_index == 0) | (tr.leaf_index == m - 1)
    want = np.prod(3 - edge.astype(int) - (m == 1), axis=1)
    assert np.array_equal(np.diff(starts), want)
    assert starts[-1] == (3 * m - 2) ** dim
    pairs = {(int(t), int(s)) for i, t in enumerate(tr.target_boxes)
             for s in lists[starts[i]:starts[i + 1]]}
    assert all((s, t) in pairs for t, s in pairs)


@pytest.mark.parametrize("mode", ["canonical", "per_level"])
def test_oracle_exterior_term(mode):
    """Potential kind 2 (list1.py:1097-1143): with exterior normalisers 0
    the oracle is the kind-1 sum; over a zero table it is each target's
    density times exterior_nmlz[its box's table level, its mode] times its
    box's scaling, on an adaptive tree with targets on several levels."""
    from oracle import list1_ref
    from tests.gpu_helpers import adaptive, synthetic_table_levels

    tr = adaptive(3, 2)
    n_tables = tr.nlevels if mode == "per_level" else 1
    tables = synthetic_table_levels(3, 2, n_tables=n_tables, seed=4)
    meta = orbit_ref.case_metadata(3, 2)
    data = np.stack([t.compact_values() for t in tables])
    nm = np.stack([t.mode_normalizers for t in tables])
    rng = np.random.default_rng(2)
    ext = rng.standard_normal((n_tables, tr.n_q_points))
    f = rng.standard_normal((2, tr.n_points))
    kw = dict(mode=mode, scaling_power=2, table_root_extent=2.0, own_csr=False)
    one = list1_ref.near_field_from_tree(tr, data, nm, meta, f, **kw)
    z = list1_ref.near_field_from_tree(tr, data, nm, meta, f, exterior_nmlz=np.zeros_like(ext),
                                       **kw)
    assert np.array_equal(one, z)
    got = list1_ref.near_field_from_tree(tr, np.zeros_like(data), nm, meta, f, exterior_nmlz=ext,
                                         **kw)
    exp = np.zeros_like(f)
    lv = np.asarray(tr.box_levels)
    assert len(np.unique(lv[tr.target_boxes])) >= 2
    for tb in tr.target_boxes:
        s0, c = int(tr.box_target_starts[tb]), int(tr.box_target_counts_nonchild[tb])
        ids = np.asarray(tr.local_ids[s0:s0 + c])
        h = tr.root_extent / 2.0 ** lv[tb]
        if mode == "canonical":
            e = ext[0, ids] * (h / 2.0) ** 2
        else:
            e = ext[min(int(lv[tb]), n_tables - 1), ids]
        exp[:, s0:s0 + c] = f[:, s0:s0 + c] * e[None, :]
    np.testing.assert_allclose(got, exp, rtol=1e-15, atol=0)
