"""The oracle's ``dense`` reconstruction kind (list1.py:1018-1045,
1111-1116), pinned on the CPU:

- identity maps over the full table expanded from the compact one give the
  arithmetic-orbit oracle's result (that path is pinned by the reference's
  known-answer cases and brute-force orbits);
- random signed-permutation maps per source mode give the same result
  exactly (a scalar kernel's table is invariant under the group; a map
  applied to the wrong index, s and t swapped, or the case map skipped
  breaks this);
- entries without data contribute nothing, the 2D log displacement
  included: dropping every entry of one case equals the arithmetic oracle
  on the list-1 CSR without that case's pairs, and dropping all entries
  gives zero.
"""

import numpy as np
import pytest

from oracle import list1_ref, orbit_ref
from tests.dense_kind import dense_arrays, oracle_meta
from tests.gpu_helpers import synthetic_table_levels
from volumential_amd import densities, tree as vtree


def _setup(dim, q, level, seed=0):
    tr = vtree.uniform_tree(dim, q, level)
    tables = synthetic_table_levels(dim, q, seed=seed)
    meta = orbit_ref.case_metadata(dim, q)
    data = tables[0].compact_values()[None, :]
    nm = tables[0].mode_normalizers[None, :]
    f = densities.gaussian_mixture_batch(tr.points, range(2))
    return tr, tables, meta, data, nm, f


def _run(tr, table, nm, meta, f, logd, starts=None, lists=None):
    return list1_ref.near_field(
        dim=tr.dim, q=tr.quad_order, box_centers=tr.box_centers, box_levels=tr.box_levels,
        box_starts=tr.box_target_starts, box_counts=tr.box_target_counts_nonchild,
        target_boxes=tr.target_boxes,
        nbr_starts=tr.neighbor_source_boxes_starts if starts is None else starts,
        nbr_lists=tr.neighbor_source_boxes_lists if lists is None else lists,
        source_mode_ids=tr.local_ids, target_point_ids=tr.local_ids, table_data=table,
        mode_nmlz=nm, meta=meta, mode="canonical", log_displacement=logd, root_extent=tr.root_extent,
        table_root_extent=2.0, mode_coefs=f)


@pytest.mark.parametrize("dim,q,level", [(2, 3, 3), (3, 2, 2)])
def test_dense_identity_equals_orbit(dim, q, level):
    tr, _, meta, data, nm, f = _setup(dim, q, level)
    logd = dim == 2
    ref = _run(tr, data, nm, meta, f, logd)
    arr = dense_arrays(data, meta, q, dim)
    got = _run(tr, arr["table_data_combined"], nm, oracle_meta(meta, arr), f, logd)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14 * np.abs(ref).max())


@pytest.mark.parametrize("dim,q,level", [(2, 4, 3), (3, 3, 2)])
def test_dense_group_maps_invariant(dim, q, level):
    tr, _, meta, data, nm, f = _setup(dim, q, level, seed=2)
    arr0 = dense_arrays(data, meta, q, dim)
    arr1 = dense_arrays(data, meta, q, dim, maps="group", seed=5)
    assert not np.array_equal(arr1["mode_qpoint_map"], arr0["mode_qpoint_map"])
    a = _run(tr, arr0["table_data_combined"], nm, oracle_meta(meta, arr0), f, dim == 2)
    b = _run(tr, arr1["table_data_combined"], nm, oracle_meta(meta, arr1), f, dim == 2)
    np.testing.assert_array_equal(a, b)
    # a plausible slip: the target node mapped with its own transform
    bad = dict(arr1)
    n = q**dim
    bad["mode_qpoint_map"] = np.tile(np.diag(arr1["mode_qpoint_map"]), (n, 1))  # [s][t] = [t][t]
    c = _run(tr, arr1["table_data_combined"], nm, oracle_meta(meta, bad), f, dim == 2)
    assert np.abs(c - a).max() > 1e-6 * np.abs(a).max()


def test_dense_missing_entries_drop_their_pairs():
    tr, _, meta, data, nm, f = _setup(2, 3, 3, seed=4)
    st = np.asarray(tr.neighbor_source_boxes_starts)
    ls = np.asarray(tr.neighbor_source_boxes_lists)
    tb = np.asarray(tr.target_boxes)
    cases = np.array([
        list1_ref.case_id_float(tr.box_centers[:, tb[i]], tr.box_centers[:, s],
                                tr.box_levels[s], tr.root_extent, meta["case_indices"])[0]
        for i in range(len(tb)) for s in ls[st[i]:st[i + 1]]])
    drop = int(np.bincount(cases).argmax())
    keep = cases != drop
    pair_tgt = np.repeat(np.arange(len(tb)), np.diff(st))
    st2 = np.concatenate([[0], np.cumsum(np.bincount(pair_tgt[keep], minlength=len(tb)))])
    ref = _run(tr, data, nm, meta, f, True, starts=st2, lists=ls[keep])
    # identity maps: case c reads only its own entries
    arr = dense_arrays(data, meta, 3, 2, drop_cases=[drop])
    got = _run(tr, arr["table_data_combined"], nm, oracle_meta(meta, arr), f, True)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14 * np.abs(ref).max())
    allm = dense_arrays(data, meta, 3, 2, drop_cases=range(len(meta["case_orbit_ranks"])))
    z = _run(tr, allm["table_data_combined"], nm, oracle_meta(meta, allm), f, True)
    assert np.all(z == 0.0)
