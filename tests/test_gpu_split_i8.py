"""GPU parity of the int8 split form (VL1_FORM_SPLIT_I8) and its error
estimate.

* vs the CPU oracle: rel-L2 and rel-Linf <= 1e-10 per RHS at 6 slices, on the
  headline shape (3D Laplace q=6, 64 RHS, 1 box x 64 RHS tiles) at uniform
  levels 2 and 3, adaptive canonical and per-level trees, the 2D log term,
  several boxes per tile, and complex (Helmholtz) data;
* the error falls by more than 50x per added slice;
* the per-RHS estimate (vl1_split_opts.error_est) is at least the measured
  error against the FP64 form for smooth and white-noise densities at 3 to
  5 slices (at 2 slices, 15 bits, the size of the first dropped diagonal no
  longer follows the last kept one closely enough on smooth data), and at 6
  slices below 1e-10 on the headline shape and at least the measured error.
"""

import numpy as np
import pytest
import torch

from oracle import cpu as ocpu
from oracle import list1_ref, orbit_ref
from tests.gpu_helpers import adaptive, beam_forest, helmholtz_tables, rel_errors, synthetic_table_levels
from volumential_amd import densities, tree as vtree
from volumential_amd.kernels import LaplaceKernel
from volumential_amd.table import DuffyBuildConfig, NearFieldInteractionTable
from volumential_amd.wrangler import NearFieldWrangler

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _laplace_table(dim, q, extent=2.0, rule=(10, 31)):
    t = NearFieldInteractionTable(q, dim, LaplaceKernel(dim), source_box_extent=extent)
    return t.build_table(DuffyBuildConfig(regular_quad_order=rule[0], radial_quad_order=rule[1]))


def _plan(tr, tables):
    w = NearFieldWrangler(tr, tables if len(tables) > 1 else tables[0])
    return w.get_plan(w.target_kernels[0], tr.target_boxes, tr.neighbor_source_boxes_starts,
                      tr.neighbor_source_boxes_lists)


def _eval(plan, f, form="split_i8", n_slices=None, est=False):
    fd = torch.as_tensor(np.atleast_2d(f), device="cuda").contiguous()
    e = torch.zeros((fd.shape[0], 2), dtype=torch.float64, device="cuda") if est else None
    out = plan.eval(fd, form=form, n_slices=n_slices, error_est=e)
    torch.cuda.synchronize()
    return out.cpu().numpy(), (e.cpu().numpy() if est else None)


def _oracle(tr, tables, f, mode, power=2, logd=False):
    meta = orbit_ref.case_metadata(tr.dim, tr.quad_order)
    data = np.stack([t.compact_values() for t in tables])
    nm = np.stack([t.mode_normalizers for t in tables])
    tsl = int(round(np.log2(tr.root_extent / tables[0].source_box_extent)))
    return ocpu.near_field(tr, data, nm, meta, f, mode=mode, scaling_power=power,
                           log_displacement=logd, table_root_extent=tables[0].source_box_extent,
                           table_starting_level=tsl)


def _per_rhs(a, b):
    a, b = np.atleast_2d(a), np.atleast_2d(b)
    return (np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1),
            np.abs(a - b).max(axis=1) / np.abs(b).max(axis=1))


def _sub_tree(tr, sel):
    import copy

    sub = copy.copy(tr)
    starts = np.zeros(sel.size + 1, np.int32)
    lists = []
    for i, b in enumerate(sel):
        seg = tr.neighbor_source_boxes_lists[tr.neighbor_source_boxes_starts[b]:
                                             tr.neighbor_source_boxes_starts[b + 1]]
        lists.append(seg)
        starts[i + 1] = starts[i] + seg.size
    sub.target_boxes = np.asarray(tr.target_boxes, np.int32)[sel]
    sub.neighbor_source_boxes_starts = starts
    sub.neighbor_source_boxes_lists = np.concatenate(lists).astype(np.int32)
    return sub


@pytest.mark.parametrize("level", [2, 3])
def test_headline_shape_vs_oracle(cuda, level):
    """3D Laplace q=6, GPU-built table, 64 seeded 8-Gaussian densities
    (seeds 1000..1063), one evaluation, 6 int8 slices, every RHS against the
    C oracle (level 3: a seeded sample of 64 of 512 target boxes); the
    estimate is below 1e-10 and at least the error over all targets."""
    print(f"case: int8 split headline shape 3D q=6 level {level} R=64 S=6", flush=True)
    tr = vtree.uniform_tree(3, 6, level)
    table = _laplace_table(3, 6)
    f = densities.gaussian_mixture_batch(tr.points, range(1000, 1064))
    plan = _plan(tr, [table])
    got, est = _eval(plan, f, est=True)
    dense, _ = _eval(plan, f, form="f64")
    sel = np.arange(tr.target_boxes.size, dtype=np.int32)
    if sel.size > 64:
        sel = np.sort(np.random.default_rng(11).choice(sel, 64, replace=False)).astype(np.int32)
    meta = orbit_ref.case_metadata(3, 6)
    ref = ocpu.near_field(_sub_tree(tr, sel), table.compact_values(), table.mode_normalizers,
                          meta, f, mode="canonical", table_root_extent=2.0)
    idx = np.concatenate([np.arange(tr.box_target_starts[b], tr.box_target_starts[b] + 216)
                          for b in np.asarray(tr.target_boxes)[sel]])
    l2, linf = rel_errors(got[:, idx], ref[:, idx])
    print(f"int8 S=6: rel L2 {l2:.2e} Linf {linf:.2e}; estimate max Linf {est[:, 0].max():.2e} "
          f"L2 {est[:, 1].max():.2e}")
    assert l2 <= TOL and linf <= TOL, (l2, linf)
    assert est[:, 0].max() <= TOL and est[:, 1].max() <= TOL, est.max(axis=0)
    # the estimate is relative to all of the plan's targets: check it
    # against the error over all points (vs the FP64 form, itself at 1e-14
    # of the oracle)
    assert rel_errors(dense[:, idx], ref[:, idx])[0] <= 1e-12
    e_l2, e_li = _per_rhs(got, dense)
    print(f"vs FP64 form, all points: Linf {e_li.max():.2e} L2 {e_l2.max():.2e}")
    assert np.all(est[:, 0] >= e_li) and np.all(est[:, 1] >= e_l2), (est, e_li, e_l2)


@pytest.mark.parametrize("case", ["adaptive3_canonical", "adaptive2_per_level", "log2d",
                                  "boxes_per_tile", "q5_adaptive"])
def test_vs_oracle(cuda, case):
    print(f"case: int8 split {case}", flush=True)
    mode, logd = "canonical", False
    if case == "adaptive3_canonical":
        tr = adaptive(3, 4)
        tables = [_laplace_table(3, 4)]
        f = densities.gaussian_mixture_batch(tr.points, range(6))
    elif case == "adaptive2_per_level":
        tr = adaptive(2, 6)
        tables = synthetic_table_levels(2, 6, n_tables=tr.nlevels, seed=8)
        f = densities.gaussian_mixture_batch(tr.points, [21, 22, 23, 24])
        mode = "per_level"
    elif case == "log2d":
        tr = vtree.uniform_tree(2, 4, 5)
        tables = [_laplace_table(2, 4, rule=(20, 41))]
        f = np.stack([np.exp(-np.sum(tr.points**2, axis=0) / 0.01),
                      densities.gaussian_mixture(tr.points, 5)])
        logd = True
    elif case == "boxes_per_tile":  # 8 RHS: tiles of 8 boxes x 8 RHS
        tr = vtree.uniform_tree(3, 6, 2)
        tables = [_laplace_table(3, 6)]
        f = densities.gaussian_mixture_batch(tr.points, range(1000, 1008))
    else:  # C2-style: q=5 adaptive, density-refined, 2:1 balanced
        fden = (lambda P: densities.gaussian_mixture(P, 0))
        tr = vtree.adaptive_tree_batched(3, 5, vtree.interpolation_refiner(fden, 3, 5, 1e-4),
                                         min_level=2, max_level=4)
        tables = [_laplace_table(3, 5)]
        f = densities.gaussian_mixture_batch(tr.points, [0, 1])
    got, est = _eval(_plan(tr, tables), f, est=True)
    ref = _oracle(tr, tables, f, mode, 2, logd)
    l2, linf = rel_errors(got, ref)
    print(f"int8 S=6: rel L2 {l2:.2e} Linf {linf:.2e}; estimate {est.max(axis=0)}")
    assert l2 <= TOL and linf <= TOL, (l2, linf)


def test_complex_forest_vs_oracle(cuda):
    """Per-level Helmholtz tables (complex), two beam-refined quadtrees in one
    plan, 4 complex densities: int8 split (real and imaginary parts cut
    alike) against the numpy oracle (c128)."""
    print("case: int8 split complex forest", flush=True)
    fo = beam_forest([0, 1], q=6, tol=1e-5, min_level=4, max_level=6)
    tables = helmholtz_tables(2, 6, 40.0, 4, fo.nlevels - 4)
    f = np.stack([densities.helmholtz_beams(fo.points, 60 + r) for r in range(4)])
    got, est = _eval(_plan(fo, tables), f, est=True)
    meta = orbit_ref.case_metadata(2, 6)
    ref = list1_ref.near_field_from_tree(
        fo, np.stack([t.compact_values() for t in tables]),
        np.stack([t.mode_normalizers for t in tables]), meta, f, mode="per_level",
        table_root_extent=tables[0].source_box_extent, table_starting_level=4)
    l2, linf = rel_errors(got, ref)
    print(f"int8 S=6 complex: rel L2 {l2:.2e} Linf {linf:.2e}; estimate {est.max(axis=0)}")
    assert l2 <= TOL and linf <= TOL, (l2, linf)


@pytest.mark.parametrize("kind", ["gaussian", "white"])
def test_estimate_covers_error(cuda, kind):
    """At 3..5 slices (errors well above FP64 round-off) the per-RHS estimate
    is at least the measured error against the FP64 form, in Linf and L2,
    for smooth and white-noise densities; the error falls by more than 50x
    per slice."""
    print(f"case: int8 estimate vs error, {kind} densities", flush=True)
    tr = vtree.uniform_tree(3, 4, 3)
    tables = [_laplace_table(3, 4)]
    if kind == "gaussian":
        f = densities.gaussian_mixture_batch(tr.points, range(200, 208))
    else:
        f = np.random.default_rng(5).standard_normal((8, tr.n_points))
    plan = _plan(tr, tables)
    ref, _ = _eval(plan, f, form="f64")
    prev = None
    for S in (3, 4, 5):
        got, est = _eval(plan, f, n_slices=S, est=True)
        a, b = np.atleast_2d(got), np.atleast_2d(ref)
        l2 = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
        li = np.abs(a - b).max(axis=1) / np.abs(b).max(axis=1)
        print(f"S={S}: error Linf {li.max():.2e} L2 {l2.max():.2e}; estimate Linf "
              f"{est[:, 0].min():.2e}..{est[:, 0].max():.2e} L2 {est[:, 1].min():.2e}..{est[:, 1].max():.2e}",
              flush=True)
        assert np.all(est[:, 0] >= li), (S, est[:, 0], li)
        assert np.all(est[:, 1] >= l2), (S, est[:, 1], l2)
        if prev is not None:
            assert li.max() < prev / 50, (S, li.max(), prev)
        prev = li.max()


def test_wrangler_guard(cuda):
    """eval_direct_plan(form="split_i8") raises the slice count of a density
    whose estimate exceeds tol and falls back to the FP64 form past 6
    slices: at tol = 1e-10 every RHS meets 1e-10 against the oracle (3D q=6,
    uniform level 2, 64 RHS); at tol = 1e-16 every RHS takes the FP64 form
    (<= 1e-12)."""
    print("case: int8 guarded wrangler evaluation, 3D q=6 level 2, 64 RHS", flush=True)
    tr = vtree.uniform_tree(3, 6, 2)
    table = _laplace_table(3, 6)
    f = densities.gaussian_mixture_batch(tr.points, range(1000, 1064))
    w = NearFieldWrangler(tr, table)
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    fd = torch.as_tensor(f, device="cuda").contiguous()
    ref = _oracle(tr, [table], f, "canonical")
    got = w.eval_direct_plan(*args, fd, form="split_i8", tol=1e-10)[0][0].cpu().numpy()
    rep = w.last_split_report[0]
    l2, linf = _per_rhs(got, ref)
    print(f"slices {np.bincount(rep['slices'])}, fallback {int(rep['fallback_f64'].sum())}, "
          f"error max Linf {linf.max():.2e} L2 {l2.max():.2e}, estimate max {rep['estimate'].max():.2e}")
    assert linf.max() <= TOL and l2.max() <= TOL
    ok = ~rep["fallback_f64"]
    assert np.all(rep["estimate"][ok].max(axis=1) <= TOL)
    got2 = w.eval_direct_plan(*args, fd, form="split_i8", tol=1e-16)[0][0].cpu().numpy()
    rep2 = w.last_split_report[0]
    assert rep2["fallback_f64"].all() and np.all(rep2["slices"] == 6)
    l2, linf = _per_rhs(got2, ref)
    assert linf.max() <= 1e-12 and l2.max() <= 1e-12, (linf.max(), l2.max())
