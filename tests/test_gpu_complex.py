"""GPU parity of complex (Helmholtz) list-1 evaluation.

Complex case blocks run as the real block [[Re T, -Im T], [Im T, Re T]] on
[Re f; Im f], in both plan forms.  Checked against the numpy oracle (complex
arithmetic, the reference's per-level table semantics, list1.py:361-437):

* 2D Helmholtz k=40, a forest of adaptive quadtrees refined around random
  Gaussian beams, per-level tables (levels 4..), several complex densities:
  FP64 form rel-L2 / rel-Linf <= 1e-12, split form <= 1e-10 per density;
* eval_direct on a complex kernel (FP64 plan route) == the oracle;
* 3D Helmholtz on a uniform tree with one fixed-level table.
"""

import numpy as np
import pytest
import torch

from oracle import list1_ref, orbit_ref
from tests.gpu_helpers import beam_forest, helmholtz_tables, rel_errors
from volumential_amd import densities, tree as vtree
from volumential_amd.wrangler import NearFieldWrangler

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOLSPLIT = 1e-10


def _oracle(tr, tables, f, mode):
    meta = orbit_ref.case_metadata(tr.dim, tr.quad_order)
    data = np.stack([t.compact_values() for t in tables])
    nm = np.stack([t.mode_normalizers for t in tables]).astype(np.complex128)
    tsl = int(round(np.log2(tr.root_extent / tables[0].source_box_extent)))
    return list1_ref.near_field_from_tree(
        tr, data, nm, meta, f, mode=mode, scaling_power=0, log_displacement=False,
        table_root_extent=tables[0].source_box_extent, table_starting_level=tsl)


def _forest_case():
    fo = beam_forest([0, 1], q=6, tol=1e-5, min_level=4, max_level=6)
    tables = helmholtz_tables(2, 6, 40.0, 4, fo.nlevels - 4)
    f = np.stack([densities.helmholtz_beams(fo.points, s) for s in (10, 11, 12)])
    return fo, tables, f


@pytest.fixture(scope="module")
def forest_case(cuda):
    return _forest_case()


@pytest.mark.parametrize("form,tol", [("f64", TOL64), ("split", TOLSPLIT)])
def test_helmholtz2d_forest(forest_case, form, tol):
    print(f"case: Helmholtz 2D forest form={form}", flush=True)
    fo, tables, f = forest_case
    w = NearFieldWrangler(fo, tables)
    pot, _ = w.eval_direct_plan(fo.target_boxes, fo.neighbor_source_boxes_starts,
                                fo.neighbor_source_boxes_lists,
                                torch.as_tensor(f, device="cuda"), form=form, n_slices=5)
    got = pot[0].cpu().numpy()
    assert got.dtype == np.complex128
    ref = _oracle(fo, tables, f, "per_level")
    l2, linf = rel_errors(got, ref)
    print(f"{form}: rel L2 {l2:.2e} Linf {linf:.2e}")
    assert l2 <= tol and linf <= tol, (l2, linf)


def test_helmholtz2d_eval_direct(forest_case):
    print("case: Helmholtz 2D eval_direct", flush=True)
    fo, tables, f = forest_case
    w = NearFieldWrangler(fo, tables)
    pot, _ = w.eval_direct(fo.target_boxes, fo.neighbor_source_boxes_starts,
                           fo.neighbor_source_boxes_lists, torch.as_tensor(f[0], device="cuda"))
    ref = _oracle(fo, tables, f[:1], "per_level")
    l2, linf = rel_errors(pot[0].cpu().numpy(), ref)
    assert l2 <= TOL64 and linf <= TOL64, (l2, linf)


@pytest.mark.parametrize("form,tol", [("f64", TOL64), ("split", TOLSPLIT)])
def test_helmholtz3d_uniform_fixed_table(cuda, form, tol):
    print(f"case: Helmholtz 3D uniform level 2 form={form}", flush=True)
    tr = vtree.uniform_tree(3, 3, 2)
    tables = helmholtz_tables(3, 3, 5.0, 2, 1)
    rng = np.random.default_rng(5)
    f = rng.standard_normal((2, tr.n_points)) + 1j * rng.standard_normal((2, tr.n_points))
    w = NearFieldWrangler(tr, tables[0])
    pot, _ = w.eval_direct_plan(tr.target_boxes, tr.neighbor_source_boxes_starts,
                                tr.neighbor_source_boxes_lists, torch.as_tensor(f, device="cuda"),
                                form=form, n_slices=5)
    ref = _oracle(tr, tables, f, "fixed")
    l2, linf = rel_errors(pot[0].cpu().numpy(), ref)
    print(f"{form}: rel L2 {l2:.2e} Linf {linf:.2e}")
    assert l2 <= tol and linf <= tol, (l2, linf)


@pytest.mark.parametrize("form,tol", [("f64", TOL64), ("split", TOLSPLIT)])
def test_helmholtz3d_q6_row_groups(cuda, form, tol):
    """3D Helmholtz q = 6: 432 rows per complex block, beyond one tile, so the
    plan cuts each block into two row groups of 216 rows (one tile each,
    the full 432 columns); both forms vs the oracle.  Smooth densities (the
    split form's error bar is relative to the largest |T| |f| products, so
    white-noise densities with heavy cancellation are not its workload)."""
    print(f"case: Helmholtz 3D q=6 uniform level 2 (row groups) form={form}", flush=True)
    tr = vtree.uniform_tree(3, 6, 2)
    tables = helmholtz_tables(3, 6, 4.0, 2, 1)
    g = densities.gaussian_mixture_batch(tr.points, [71, 72, 73, 74])
    f = g[0::2] + 1j * g[1::2]
    w = NearFieldWrangler(tr, tables[0])
    pot, _ = w.eval_direct_plan(tr.target_boxes, tr.neighbor_source_boxes_starts,
                                tr.neighbor_source_boxes_lists, torch.as_tensor(f, device="cuda"),
                                form=form, n_slices=5)
    plan = next(iter(w._plans.values()))
    boxes_per_tile = 64 // f.shape[0]  # two row-group tiles per group of target boxes
    assert plan.stats()["tiles"] == 2 * -(-len(tr.target_boxes) // boxes_per_tile)
    ref = _oracle(tr, tables, f, "fixed")
    l2, linf = rel_errors(pot[0].cpu().numpy(), ref)
    print(f"{form}: rel L2 {l2:.2e} Linf {linf:.2e}")
    assert l2 <= tol and linf <= tol, (l2, linf)
