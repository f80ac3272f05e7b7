"""Shared inputs for the GPU parity tests (not collected by pytest)."""

import numpy as np

from volumential_amd import orbit


def synthetic_table_levels(dim, q, n_tables=1, seed=0, extent=2.0):
    """Built-looking tables with random canonical values (parity of the
    evaluation does not depend on where the values come from)."""
    from volumential_amd.kernels import LaplaceKernel
    from volumential_amd.table import NearFieldInteractionTable

    m = orbit.build_case_metadata(dim, q)
    reps = orbit.enumerate_representatives(m)
    rng = np.random.default_rng(seed)
    levels = []
    for lev in range(n_tables):
        t = NearFieldInteractionTable(q, dim, LaplaceKernel(dim), source_box_extent=extent / 2**lev)
        t.build_normalizer_table()
        t.set_reduced_table_data(reps, rng.standard_normal(reps.size))
        levels.append(t)
    return levels


def rel_errors(a, b):
    a = np.atleast_2d(np.asarray(a))
    b = np.atleast_2d(np.asarray(b))
    l2 = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    linf = np.max(np.abs(a - b), axis=1) / np.max(np.abs(b), axis=1)
    return float(l2.max()), float(linf.max())


def adaptive(dim, q, max_level=4, point=0.23):
    from volumential_amd import tree as vtree

    def refine(lv, lower, h):
        c = lower + 0.5 * h
        return np.linalg.norm(c - point) < 1.2 * h

    return vtree.adaptive_tree(dim, q, refine, min_level=1, max_level=max_level)
