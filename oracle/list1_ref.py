"""Oracle: the ``near_field`` list-1 kernel semantics in numpy.

Restates volumential/list1.py:1057-1148 (kernel body), 439-459 and
1083-1091 (float case-id encoding with bias and +0.5 truncation), 278-437
(scaling policies), 604-902 (arithmetic orbit reconstruction and the
entry sign of axis-derivative kernels, via ``orbit_ref.rows_for_case`` /
``signs_for_case``) and 1018-1045 (the ``dense`` kind, ``dense_entries``;
``meta["kind"] == "dense"``).  Loops over (target box, source box) pairs in
Python and contracts each pair's q^d x q^d block with numpy, so it is meant
for small trees.  ``mode_coefs`` may hold several densities ``[R, n]``.
"""

from __future__ import annotations

import numpy as np

from . import orbit_ref


def case_id_float(ct, cs, level, root_extent, case_indices, bias=1e-10, shift=6.0, base=13):
    sext = root_extent * (1.0 / (2**int(level)))
    code = 0.0
    for d in range(len(ct)):
        code = code * base + ((ct[d] - cs[d]) / sext * 4.0 + shift + bias)
    vid = int(code + 0.5)
    if vid < 0 or vid >= case_indices.size:
        return -1, sext
    return int(case_indices[vid]), sext


def case_id_exact(ct, cs, level, root_extent, case_indices):
    """Integer case vector from rounded offsets (exact for dyadic trees)."""
    sext = root_extent / 2**int(level)
    vec = np.rint((np.asarray(ct) - np.asarray(cs)) / sext * 4.0).astype(int)
    return int(case_indices[orbit_ref.encode(vec)])


def near_field(
    *,
    dim,
    q,
    box_centers,
    box_levels,
    box_starts,
    box_counts,
    target_boxes,
    nbr_starts,
    nbr_lists,
    source_mode_ids,
    target_point_ids,
    table_data,
    mode_nmlz,
    meta,
    mode,
    scaling_power=2,
    log_displacement=False,
    root_extent,
    table_root_extent,
    table_starting_level=0,
    mode_coefs,
    result=None,
    bias=1e-10,
    level_scaling=None,
    level_disp=None,
    exterior_nmlz=None,
):
    """``mode``: "canonical", "fixed", "per_level" or "custom"
    (list1.py:361-437).  "custom" is the caller's code strings evaluated per
    source level: table level 0, scaling ``level_scaling[level]``,
    displacement ``level_disp[level] * mode_nmlz[0, sid]`` (``sid`` the
    source's position in its box, as the kernel's loop index,
    list1.py:1094-1112)."""
    coefs = np.atleast_2d(np.asarray(mode_coefs))
    R, npts = coefs.shape
    table = np.atleast_2d(np.asarray(table_data))
    nmlz = np.atleast_2d(np.asarray(mode_nmlz))
    n_tables = table.shape[0]
    out = np.zeros((R, npts), dtype=np.result_type(coefs, table)) if result is None else (
        np.atleast_2d(np.array(result, copy=True))
    )
    case_indices = np.asarray(meta["case_indices"])
    rows_cache = {}
    for i, tb in enumerate(np.asarray(target_boxes)):
        tbeg, nt = int(box_starts[tb]), int(box_counts[tb])
        tmodes = np.asarray(target_point_ids[tbeg : tbeg + nt])
        ct = box_centers[:, tb]
        acc = np.zeros((R, nt), dtype=out.dtype)
        for sb in np.asarray(nbr_lists[nbr_starts[i] : nbr_starts[i + 1]]):
            lvl = int(box_levels[sb])
            cid, sext = case_id_float(ct, box_centers[:, sb], lvl, root_extent, case_indices, bias)
            if cid < 0:
                raise RuntimeError("invalid list-1 case")
            if mode == "canonical":
                num = sext**scaling_power
                den = table_root_extent**scaling_power
                scal = num / den
                lev = 0
            elif mode == "fixed":
                scal, lev = 1.0, 0
            elif mode == "custom":
                scal, lev = float(level_scaling[lvl]), 0
            else:
                scal = 1.0
                rel = lvl - table_starting_level
                lev = 0 if lvl < table_starting_level else (rel if rel < n_tables else n_tables - 1)
            sbeg, ns = int(box_starts[sb]), int(box_counts[sb])
            smodes = np.asarray(source_mode_ids[sbeg : sbeg + ns])
            if meta.get("kind") == "dense":
                val, has = dense_entries(meta, table[lev], cid, smodes, tmodes, q**dim)
                integ = val * scal
                if mode == "canonical" and log_displacement:
                    disp = -0.5 / np.pi * scal * np.log(sext / table_root_extent) * nmlz[lev, :ns]
                    integ = integ + disp[:, None]
                elif mode == "custom":
                    integ = integ + (float(level_disp[lvl]) * nmlz[0, :ns])[:, None]
                acc += coefs[:, sbeg : sbeg + ns] @ np.where(has, integ, 0.0)
                continue
            if cid not in rows_cache:
                rows_cache[cid] = (orbit_ref.rows_for_case(meta, cid, q, dim),
                                   orbit_ref.signs_for_case(meta, cid, q, dim))
            rmap, smap = rows_cache[cid]
            rows = rmap[np.ix_(smodes, tmodes)]  # [ns, nt]
            integ = table[lev][rows] * smap[np.ix_(smodes, tmodes)] * scal
            if mode == "canonical" and log_displacement:
                disp = -0.5 / np.pi * scal * np.log(sext / table_root_extent) * nmlz[lev, :ns]
                integ = integ + disp[:, None]
            elif mode == "custom":
                integ = integ + (float(level_disp[lvl]) * nmlz[0, :ns])[:, None]
            acc += coefs[:, sbeg : sbeg + ns] @ integ
        if exterior_nmlz is not None:
            # potential kind 2 (list1.py:1097-1105, 1141-1143): the target's
            # density times its exterior normaliser, scaled as its own box
            lvl_t = int(box_levels[tb])
            text = root_extent / 2.0**lvl_t
            if mode == "canonical":
                scal_t, lev_t = (text / table_root_extent) ** scaling_power, 0
            elif mode == "fixed":
                scal_t, lev_t = 1.0, 0
            elif mode == "custom":
                scal_t, lev_t = float(level_scaling[lvl_t]), 0
            else:
                rel = lvl_t - table_starting_level
                scal_t = 1.0
                lev_t = 0 if lvl_t < table_starting_level else (rel if rel < n_tables else
                                                                 n_tables - 1)
            ext = np.atleast_2d(np.asarray(exterior_nmlz))[lev_t, tmodes] * scal_t
            acc += coefs[:, tbeg : tbeg + nt] * ext[None, :]
        out[:, tbeg : tbeg + nt] = acc
    return out


def dense_entries(meta, table_row, cid, smodes, tmodes, n):
    """The ``dense`` reconstruction kind (list1.py:1018-1045) for one case:
    value[s, t] = T[ids[f]] * table_entry_scales[f] * mode_case_scale[s, c]
    with f = (mode_case_map[s, c] n + mode_qpoint_map[s, s]) n
    + mode_qpoint_map[s, t], and has[s, t] = ids[f] >= 0 (list1.py:1111-1116:
    no entry, no contribution).  Returns (value, has), [len(smodes),
    len(tmodes)]."""
    mqm = np.asarray(meta["mode_qpoint_map"]).reshape(n, n)
    mcm = np.asarray(meta["mode_case_map"]).reshape(n, -1)
    mcs = np.asarray(meta["mode_case_scale"]).reshape(n, -1)
    ids = np.asarray(meta["table_entry_ids"]).reshape(-1)
    tes = np.asarray(meta["table_entry_scales"]).reshape(-1)
    s = np.asarray(smodes)[:, None]
    t = np.asarray(tmodes)[None, :]
    f = (mcm[s, cid].astype(np.int64) * n + mqm[s, s]) * n + mqm[s, t]
    eid = ids[f]
    has = eid >= 0
    val = np.where(has, np.asarray(table_row)[np.where(has, eid, 0)], 0.0) * tes[f] * mcs[s, cid]
    return val, has


def list1_csr_bruteforce(box_centers, box_levels, root_extent, target_boxes, chunk=256,
                         box_group=None):
    """List 1 of each target leaf by geometry alone: every leaf box whose
    closed cell touches the target's (a shared face, edge or corner, or the
    box itself), sorted by box id.  For the 2:1-balanced leaf trees used
    here this is the list boxtree's traversal produces (meshgen.py:391-444,
    tree_interactive_build.py:202-470); it is computed from the box centres
    and levels only, independent of the tree builder's own lists.
    ``box_group`` (optional, one label per box) keeps boxes of different
    trees of a forest apart.  Returns (starts int32 [n_targets + 1], lists
    int32)."""
    centers = np.asarray(box_centers, dtype=np.float64)
    h = root_extent / 2.0 ** np.asarray(box_levels, dtype=np.float64)
    lo = centers.T - 0.5 * h[:, None]
    hi = centers.T + 0.5 * h[:, None]
    eps = 1e-9 * root_extent
    tb = np.asarray(target_boxes, dtype=np.int64)
    lists = []
    for c0 in range(0, tb.size, chunk):
        t = tb[c0:c0 + chunk]
        touch = np.all((lo[None, :, :] <= hi[t][:, None, :] + eps)
                       & (hi[None, :, :] >= lo[t][:, None, :] - eps), axis=2)
        if box_group is not None:
            g = np.asarray(box_group)
            touch &= g[None, :] == g[t][:, None]
        for row in touch:
            lists.append(np.flatnonzero(row))
    starts = np.zeros(tb.size + 1, dtype=np.int64)
    starts[1:] = np.cumsum([len(v) for v in lists])
    flat = np.concatenate(lists) if lists else np.zeros(0, np.int64)
    return starts.astype(np.int32), flat.astype(np.int32)


def tree_groups(tree):
    """Tree label of every box of a forest (trees merged side by side, with
    ``tree_box_offsets``), or None for a single tree."""
    offs = getattr(tree, "tree_box_offsets", None)
    if offs is None:
        return None
    return np.searchsorted(np.asarray(offs), np.arange(len(tree.box_levels)), side="right") - 1


def near_field_from_tree(tree, table_data, mode_nmlz, meta, mode_coefs, *, mode, scaling_power=2,
                         log_displacement=False, table_root_extent, table_starting_level=0,
                         own_csr=True, level_scaling=None, level_disp=None, exterior_nmlz=None):
    """The kernel semantics on `tree`.  With `own_csr` (default) the list-1
    CSR of the tree's target boxes is recomputed by ``list1_csr_bruteforce``
    instead of taken from the tree builder."""
    if own_csr:
        nbr_starts, nbr_lists = list1_csr_bruteforce(tree.box_centers, tree.box_levels,
                                                     tree.root_extent, tree.target_boxes,
                                                     box_group=tree_groups(tree))
    else:
        nbr_starts = tree.neighbor_source_boxes_starts
        nbr_lists = tree.neighbor_source_boxes_lists
    return near_field(
        dim=tree.dim,
        q=tree.quad_order,
        box_centers=tree.box_centers,
        box_levels=tree.box_levels,
        box_starts=tree.box_target_starts,
        box_counts=tree.box_target_counts_nonchild,
        target_boxes=tree.target_boxes,
        nbr_starts=nbr_starts,
        nbr_lists=nbr_lists,
        source_mode_ids=tree.local_ids,
        target_point_ids=tree.local_ids,
        table_data=table_data,
        mode_nmlz=mode_nmlz,
        meta=meta,
        mode=mode,
        scaling_power=scaling_power,
        log_displacement=log_displacement,
        root_extent=tree.root_extent,
        table_root_extent=table_root_extent,
        table_starting_level=table_starting_level,
        mode_coefs=mode_coefs,
        level_scaling=level_scaling,
        level_disp=level_disp,
        exterior_nmlz=exterior_nmlz,
    )
