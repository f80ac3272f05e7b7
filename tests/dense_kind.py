"""Inputs of the reference's ``dense`` reconstruction kind for tests (not
collected by pytest).

``dense_arrays`` turns a compact (arithmetic-orbit) table into the arrays
``NearFieldFromCSR`` takes for ``reconstruction_kind="dense"``
(expansion_wrangler_fpnd.py:1887-2066, 2980-2995): the full table
T[lev][c][s][t] expanded through the oracle's orbit rows, entry ids over the
full entries, and mode maps that are either the identity or, per source
mode, a random element of the signed-permutation group acting on the mode
axes and the case vectors as nearfield_potential_table.py:1536-1571 does.
A scalar kernel's table is invariant under that group, so every choice of
maps reconstructs the same blocks.
"""

from itertools import permutations, product

import numpy as np

from oracle import orbit_ref


def full_table(data, meta, q, dim):
    """[n_tables][n_cases * n * n] values of every (case, s, t)."""
    data = np.atleast_2d(np.asarray(data))
    n = q**dim
    n_cases = len(meta["case_orbit_ranks"])
    out = np.empty((data.shape[0], n_cases, n, n), dtype=data.dtype)
    for c in range(n_cases):
        rows = orbit_ref.rows_for_case(meta, c, q, dim)
        sg = orbit_ref.signs_for_case(meta, c, q, dim)
        out[:, c] = data[:, rows] * sg
    return out.reshape(data.shape[0], -1)


def group_maps(meta, q, dim, seed):
    """mode_qpoint_map [n][n] and mode_case_map [n][n_cases]: source mode s
    takes a random signed axis permutation g_s (mode m -> g_s(m), case c ->
    the case of g_s(vec_c))."""
    rng = np.random.default_rng(seed)
    n = q**dim
    vecs = np.asarray(meta["vecs"], dtype=np.int64)
    ci = np.asarray(meta["case_indices"])
    axes = np.stack(np.unravel_index(np.arange(n), (q,) * dim), axis=1)  # axis 0 slowest
    elems = [(np.array(sg), list(pm)) for sg in product([-1, 1], repeat=dim)
             for pm in permutations(range(dim))]
    mqm = np.empty((n, n), np.int32)
    mcm = np.empty((n, len(vecs)), np.int32)
    for s in range(n):
        sg, pm = elems[rng.integers(len(elems))]
        m = axes.copy()
        m[:, sg < 0] = q - 1 - m[:, sg < 0]
        mqm[s] = np.ravel_multi_index(m[:, pm].T, (q,) * dim)
        cv = (vecs * sg)[:, pm]
        mcm[s] = ci[[orbit_ref.encode(v) for v in cv]]
    return mqm, mcm


def dense_arrays(data, meta, q, dim, maps="identity", seed=0, drop_cases=()):
    """Keyword arrays of the dense kind; ``drop_cases``: cases whose entries
    get id -1 (no entry)."""
    data = np.atleast_2d(np.asarray(data))
    n = q**dim
    n_cases = len(meta["case_orbit_ranks"])
    full = full_table(data, meta, q, dim)
    ids = np.arange(n_cases * n * n, dtype=np.int32)
    for c in drop_cases:
        ids[c * n * n:(c + 1) * n * n] = -1
    if maps == "identity":
        mqm = np.tile(np.arange(n, dtype=np.int32), (n, 1))
        mcm = np.tile(np.arange(n_cases, dtype=np.int32), (n, 1))
    else:
        mqm, mcm = group_maps(meta, q, dim, seed)
    return {
        "table_data_combined": full,
        "table_entry_ids": ids,
        "table_entry_scales": np.ones(ids.size, dtype=data.dtype),
        "mode_qpoint_map": mqm,
        "mode_case_map": mcm,
        "mode_case_scale": np.ones((n, n_cases), dtype=data.dtype),
    }


def oracle_meta(meta, arrays):
    """The oracle's ``meta`` for the dense kind."""
    return {"kind": "dense", "case_indices": meta["case_indices"], **{
        k: arrays[k] for k in ("table_entry_ids", "table_entry_scales", "mode_qpoint_map",
                               "mode_case_map", "mode_case_scale")}}
