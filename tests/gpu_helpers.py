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


def beam_forest(seeds, q=6, tol=1e-5, min_level=4, max_level=6, k=40.0):
    """Adaptive quadtrees refined around Gaussian-beam densities (one per
    seed), merged into one forest."""
    from volumential_amd import densities
    from volumential_amd import tree as vtree

    trees = []
    for s in seeds:
        f = (lambda P, s=s: densities.helmholtz_beams(P, s, k=k))
        trees.append(vtree.adaptive_tree_batched(
            2, q, vtree.interpolation_refiner(f, 2, q, tol, relative_to=4.0),
            min_level=min_level, max_level=max_level))
    return vtree.Forest(trees)


def helmholtz_tables(dim, q, k, first_level, n_levels, root_extent=2.0, rule=(10, 31)):
    """Per-level Helmholtz tables built on the GPU, source extents
    root_extent * 2^-(first_level + lev)."""
    from volumential_amd.kernels import HelmholtzKernel
    from volumential_amd.table import DuffyBuildConfig, NearFieldInteractionTable

    out = []
    for lev in range(n_levels):
        t = NearFieldInteractionTable(q, dim, HelmholtzKernel(dim, k),
                                      source_box_extent=root_extent / 2 ** (first_level + lev))
        t.build_table(DuffyBuildConfig(regular_quad_order=rule[0], radial_quad_order=rule[1]))
        out.append(t)
    return out
