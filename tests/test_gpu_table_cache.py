"""GPU evaluation of tables loaded from the table cache.

Tables built on the GPU (Duffy build) are stored in the reference's SQLite
cache format (schema 2.1.0, ``table_manager.py:665-740``), read back from
the file, and evaluated by the plan and plain routes; the results match the
oracle run on the tables as built (FP64 <= 1e-12, split forms <= 1e-10 per
RHS).  One canonical 2D table with the log term, and a per-level set of 3D
tables (one record per source-box level) over an adaptive tree.
"""

import numpy as np
import pytest
import torch

from oracle import list1_ref, orbit_ref
from tests.gpu_helpers import adaptive, rel_errors
from volumential_amd import densities, tree as vtree
from volumential_amd.kernels import LaplaceKernel
from volumential_amd.table import DuffyBuildConfig, NearFieldInteractionTable
from volumential_amd.table_cache import TableCache
from volumential_amd.wrangler import NearFieldWrangler

pytestmark = pytest.mark.gpu


def _built(dim, q, level):
    t = NearFieldInteractionTable(q, dim, LaplaceKernel(dim), source_box_extent=2.0 / 2**level,
                                  source_box_level=level)
    return t.build_table(DuffyBuildConfig(regular_quad_order=10, radial_quad_order=31))


def _evaluate(tr, tables, f, form):
    w = NearFieldWrangler(tr, tables if len(tables) > 1 else tables[0],
                          direct_method="plain" if form == "plain" else "plan")
    fd = torch.as_tensor(f, device="cuda").contiguous()
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    if form == "plain":
        pot, _ = w.eval_direct_batched(*args, fd)
    else:
        pot, _ = w.eval_direct_plan(*args, fd, form=form, n_slices=5)
    return pot[0].cpu().numpy()


def _round_trip(tmp_path, built, dim, q):
    path = tmp_path / "tables.sqlite"
    cache = TableCache(path, root_extent=2.0)
    for t in built:
        cache.save(t, "Laplace", t.source_box_level)
    cache.close()
    cache = TableCache(path, root_extent=2.0)
    loaded = [cache.load(dim, "Laplace", q, t.source_box_level, LaplaceKernel(dim)) for t in built]
    cache.close()
    for a, b in zip(built, loaded):
        assert np.array_equal(a.compact_values(), b.compact_values())
        assert np.array_equal(a.mode_normalizers, b.mode_normalizers)
    return loaded


@pytest.mark.parametrize("form,tol", [("f64", 1e-12), ("split_i8", 1e-10), ("plain", 1e-12)])
def test_cached_2d_table_evaluates(cuda, tmp_path, form, tol):
    print(f"case: cached 2D q=4 table, uniform level 4, form={form}", flush=True)
    built = [_built(2, 4, 0)]
    loaded = _round_trip(tmp_path, built, 2, 4)
    tr = vtree.uniform_tree(2, 4, 4)
    f = densities.gaussian_mixture_batch(tr.points, range(200, 206))
    got = _evaluate(tr, loaded, f, form)
    meta = orbit_ref.case_metadata(2, 4)
    ref = list1_ref.near_field_from_tree(
        tr, built[0].compact_values()[None], built[0].mode_normalizers[None], meta, f,
        mode="canonical", scaling_power=2, log_displacement=True, table_root_extent=2.0)
    l2, linf = rel_errors(got, ref)
    print(f"rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= tol and linf <= tol, (form, l2, linf)


@pytest.mark.parametrize("form,tol", [("f64", 1e-12), ("split", 1e-10)])
def test_cached_3d_per_level_tables_evaluate(cuda, tmp_path, form, tol):
    print(f"case: cached 3D q=3 per-level tables, adaptive tree, form={form}", flush=True)
    tr = adaptive(3, 3)
    built = [_built(3, 3, lev) for lev in range(tr.nlevels)]
    loaded = _round_trip(tmp_path, built, 3, 3)
    f = densities.gaussian_mixture_batch(tr.points, [3, 4, 5])
    got = _evaluate(tr, loaded, f, form)
    meta = orbit_ref.case_metadata(3, 3)
    ref = list1_ref.near_field_from_tree(
        tr, np.stack([t.compact_values() for t in built]),
        np.stack([t.mode_normalizers for t in built]), meta, f, mode="per_level",
        scaling_power=2, table_root_extent=2.0, table_starting_level=0)
    l2, linf = rel_errors(got, ref)
    print(f"rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= tol and linf <= tol, (form, l2, linf)
