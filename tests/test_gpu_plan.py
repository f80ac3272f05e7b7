"""GPU parity of the batched plan forms (vl1_plan_eval).

* FP64 dense MFMA form vs the CPU oracles: rel-L2 and rel-Linf <= 1e-12 per
  RHS (uniform, adaptive canonical / per-level, 2D log term, permuted ids);
* Ozaki-split BF16 form vs the CPU oracle: rel-L2 and rel-Linf <= 1e-10 per
  RHS with 5 slices (the split-form bar), error decreasing with the slice count;
* the split form at a larger size vs the FP64 form (itself pinned above).
"""

import numpy as np
import pytest
import torch

from oracle import cpu as ocpu
from oracle import list1_ref, orbit_ref
from tests.gpu_helpers import adaptive, rel_errors, synthetic_table_levels
from volumential_amd import densities, tree as vtree
from volumential_amd.kernels import LaplaceKernel
from volumential_amd.table import DuffyBuildConfig, NearFieldInteractionTable
from volumential_amd.wrangler import NearFieldWrangler

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOLSPLIT = 1e-10


def _plan_eval(tr, tables, f, form, n_slices=5):
    w = NearFieldWrangler(tr, tables if len(tables) > 1 else tables[0])
    pot, _ = w.eval_direct_plan(tr.target_boxes, tr.neighbor_source_boxes_starts,
                                tr.neighbor_source_boxes_lists,
                                torch.as_tensor(np.atleast_2d(f), device="cuda"),
                                form=form, n_slices=n_slices)
    return pot[0].cpu().numpy()


def _oracle(tr, tables, f, mode, power=2, logd=False):
    meta = orbit_ref.case_metadata(tr.dim, tr.quad_order)
    data = np.stack([t.compact_values() for t in tables])
    nm = np.stack([t.mode_normalizers for t in tables])
    tsl = int(round(np.log2(tr.root_extent / tables[0].source_box_extent)))
    return list1_ref.near_field_from_tree(
        tr, data, nm, meta, f, mode=mode, scaling_power=power, log_displacement=logd,
        table_root_extent=tables[0].source_box_extent, table_starting_level=tsl,
    )


def _laplace_table(dim, q, extent=2.0, rule=(10, 31)):
    t = NearFieldInteractionTable(q, dim, LaplaceKernel(dim), source_box_extent=extent)
    return t.build_table(DuffyBuildConfig(regular_quad_order=rule[0], radial_quad_order=rule[1]))


@pytest.mark.parametrize("n_rhs", [1, 5, 64])
def test_f64_uniform_3d_q6(cuda, n_rhs):
    print(f"case: plan f64 3D q=6 level 2 R={n_rhs}", flush=True)
    tr = vtree.uniform_tree(3, 6, 2)
    tables = synthetic_table_levels(3, 6, seed=1)
    f = densities.gaussian_mixture_batch(tr.points, range(1000, 1000 + n_rhs))
    got = _plan_eval(tr, tables, f, "f64")
    meta = orbit_ref.case_metadata(3, 6)
    ref = ocpu.near_field(tr, tables[0].compact_values(), tables[0].mode_normalizers, meta, f,
                          mode="canonical", table_root_extent=2.0)
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL64 and linf <= TOL64, (l2, linf)


def test_f64_rhs_multiple_of_64(cuda):
    print("case: plan f64 3D q=3 level 2 R=128", flush=True)
    tr = vtree.uniform_tree(3, 3, 2)
    tables = synthetic_table_levels(3, 3, seed=2)
    f = densities.gaussian_mixture_batch(tr.points, range(128))
    got = _plan_eval(tr, tables, f, "f64")
    ref = _oracle(tr, tables, f, "canonical")
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL64 and linf <= TOL64, (l2, linf)


def test_f64_2d_log_term(cuda):
    print("case: plan f64 2D q=4 level 5 log", flush=True)
    tr = vtree.uniform_tree(2, 4, 5)
    tables = synthetic_table_levels(2, 4, seed=3)
    f = np.stack([np.exp(-np.sum(tr.points**2, axis=0) / 0.01),
                  densities.gaussian_mixture(tr.points, 5)])
    got = _plan_eval(tr, tables, f, "f64")
    ref = _oracle(tr, tables, f, "canonical", 2, True)
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL64 and linf <= TOL64, (l2, linf)


@pytest.mark.parametrize("dim,q,mode", [(2, 6, "per_level"), (3, 5, "per_level"),
                                        (3, 4, "canonical"), (2, 3, "canonical")])
def test_f64_adaptive(cuda, dim, q, mode):
    print(f"case: plan f64 adaptive {dim}D q={q} {mode}", flush=True)
    tr = adaptive(dim, q)
    n_tables = tr.nlevels if mode == "per_level" else 1
    tables = synthetic_table_levels(dim, q, n_tables=n_tables, seed=4)
    f = densities.gaussian_mixture_batch(tr.points, [7, 8, 9])
    got = _plan_eval(tr, tables, f, "f64")
    ref = _oracle(tr, tables, f, mode, 2, dim == 2 and mode == "canonical")
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL64 and linf <= TOL64, (l2, linf)


def test_f64_permuted_ids(cuda):
    print("case: plan f64 permuted ids", flush=True)
    tr = vtree.uniform_tree(2, 3, 3)
    rng = np.random.default_rng(9)
    n = tr.n_q_points
    tr.local_ids = np.concatenate([rng.permutation(n) for _ in range(tr.nboxes)]).astype(np.int32)
    tables = synthetic_table_levels(2, 3, seed=2)
    f = densities.gaussian_mixture_batch(tr.points, [31, 32])
    got = _plan_eval(tr, tables, f, "f64")
    ref = _oracle(tr, tables, f, "canonical", 2, True)
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL64 and linf <= TOL64, (l2, linf)


@pytest.mark.parametrize("dim,q,level", [(3, 6, 2), (3, 5, 2), (2, 6, 4)])
def test_split_vs_oracle(cuda, dim, q, level):
    print(f"case: plan split {dim}D q={q} level {level}", flush=True)
    tr = vtree.uniform_tree(dim, q, level)
    tables = [_laplace_table(dim, q)]
    f = densities.gaussian_mixture_batch(tr.points, range(1000, 1008))
    got = _plan_eval(tr, tables, f, "split", 5)
    ref = _oracle(tr, tables, f, "canonical", 2, dim == 2)
    l2, linf = rel_errors(got, ref)
    print(f"split S=5 rel L2 {l2:.2e} Linf {linf:.2e}")
    assert l2 <= TOLSPLIT and linf <= TOLSPLIT, (l2, linf)


def test_split_error_decreases_with_slices(cuda):
    print("case: plan split slices 1..5", flush=True)
    tr = vtree.uniform_tree(3, 4, 3)
    tables = [_laplace_table(3, 4)]
    f = densities.gaussian_mixture_batch(tr.points, range(4))
    ref = _plan_eval(tr, tables, f, "f64")
    errs = []
    for s in range(1, 6):
        got = _plan_eval(tr, tables, f, "split", s)
        errs.append(rel_errors(got, ref)[0])
    print("rel L2 by slices:", errs)
    assert all(errs[i + 1] < errs[i] * 1e-1 for i in range(3)), errs
    assert errs[4] <= TOLSPLIT


def test_split_adaptive_per_level(cuda):
    print("case: plan split adaptive per-level", flush=True)
    tr = adaptive(3, 4)
    tables = [_laplace_table(3, 4, extent=2.0 / 2**lev) for lev in range(tr.nlevels)]
    f = densities.gaussian_mixture_batch(tr.points, range(6))
    got = _plan_eval(tr, tables, f, "split", 5)
    ref = _plan_eval(tr, tables, f, "f64")
    ref_o = _oracle(tr, tables, f, "per_level")
    assert rel_errors(ref, ref_o)[0] <= TOL64
    l2, linf = rel_errors(got, ref_o)
    assert l2 <= TOLSPLIT and linf <= TOLSPLIT, (l2, linf)


@pytest.mark.parametrize("form", ["f64", "split"])
def test_autotune_keeps_results(cuda, form):
    """vl1_plan_tune times every valid variant (several boxes per tile here,
    so both tile orders are candidates) and later evaluations use the pick;
    results stay at the oracle's bar."""
    print(f"case: autotune adaptive 2D q=6 form={form}", flush=True)
    tr = adaptive(2, 6)
    tables = synthetic_table_levels(2, 6, n_tables=tr.nlevels, seed=8)
    f = densities.gaussian_mixture_batch(tr.points, [21, 22, 23, 24])
    w = NearFieldWrangler(tr, tables)
    plan = w.get_plan(w.target_kernels[0], tr.target_boxes, tr.neighbor_source_boxes_starts,
                      tr.neighbor_source_boxes_lists)
    fd = torch.as_tensor(f, device="cuda").contiguous()
    info = plan.tune(fd, form=form, n_slices=5, reps=2)
    print("tuned:", info)
    # 36 nodes -> 64-row blocks: one row tile per wave is a candidate in both
    # forms, and several boxes per tile make both tile orders candidates
    assert set(info["times_ms"]) == set(range(8)) and info["variant"] in info["times_ms"]
    got = plan.eval(fd, form=form, n_slices=5).cpu().numpy()
    ref = _oracle(tr, tables, f, "per_level")
    l2, linf = rel_errors(got, ref)
    tol = TOL64 if form == "f64" else TOLSPLIT
    assert l2 <= tol and linf <= tol, (l2, linf)


def test_every_variant_matches_oracle(cuda):
    """Each tile-kernel variant, pinned with vl1_plan_set_variant, meets the
    oracle's bar in both forms (the tuner itself only checks its pick); the
    one-row-tile split kernel (its own product order) agrees with the
    two-row one at every slice count."""
    print("case: all variants, adaptive 2D q=6, 4 RHS", flush=True)
    tr = adaptive(2, 6)
    tables = synthetic_table_levels(2, 6, n_tables=tr.nlevels, seed=8)
    f = densities.gaussian_mixture_batch(tr.points, [31, 32, 33, 34])
    w = NearFieldWrangler(tr, tables)
    plan = w.get_plan(w.target_kernels[0], tr.target_boxes, tr.neighbor_source_boxes_starts,
                      tr.neighbor_source_boxes_lists)
    fd = torch.as_tensor(f, device="cuda").contiguous()
    ref = _oracle(tr, tables, f, "per_level")
    for form, tol in (("f64", TOL64), ("split", TOLSPLIT)):
        for v in range(8):
            plan.set_variant(v, form=form, n_slices=5, n_rhs=4)
            got = plan.eval(fd, form=form, n_slices=5).cpu().numpy()
            l2, linf = rel_errors(got, ref)
            print(f"form={form} variant={v}: {l2:.2e} {linf:.2e}", flush=True)
            assert l2 <= tol and linf <= tol, (form, v, l2, linf)
    for S in range(1, 7):
        outs = []
        for v in (0, 1):
            plan.set_variant(v, form="split", n_slices=S, n_rhs=4)
            outs.append(plan.eval(fd, form="split", n_slices=S).cpu().numpy())
        l2, linf = rel_errors(outs[1], outs[0])
        print(f"slices={S}: one-row vs two-row {l2:.2e} {linf:.2e}", flush=True)
        assert l2 <= 1e-10 and linf <= 1e-10, (S, l2, linf)
    plan.set_variant(None, form="split", n_slices=5, n_rhs=4)


def test_phase_aligned_variants_repeatable(cuda):
    """Phase alignment (tiles start at the table block their XCD is on, read
    from a word other workgroups rewrite while they run) only reorders a
    tile's steps: repeated evaluations of a 16,384-box plan with it on agree
    with the fixed-order variant to rounding, every time.  A workgroup whose
    waves took different starting steps (each reading the word on its own)
    mixed one step's densities with another's table block, now and then."""
    print("case: phase-aligned variants, repeated, uniform 2D q=6 level 7, 16 RHS", flush=True)
    tr = vtree.uniform_tree(2, 6, 7)
    tables = synthetic_table_levels(2, 6, seed=3)
    f = densities.gaussian_mixture_batch(tr.points, range(50, 66))
    w = NearFieldWrangler(tr, tables[0])
    plan = w.get_plan(w.target_kernels[0], tr.target_boxes, tr.neighbor_source_boxes_starts,
                      tr.neighbor_source_boxes_lists)
    fd = torch.as_tensor(f, device="cuda").contiguous()
    for form in ("f64", "split"):
        worst = 0.0
        for v in (0, 1, 2, 3):
            # fixed step order, same row tiling (the split kernel with one
            # row tile per wave has its own product order)
            plan.set_variant(4 | (v & 1), form=form, n_slices=5, n_rhs=16)
            base = plan.eval(fd, form=form, n_slices=5)
            plan.set_variant(v, form=form, n_slices=5, n_rhs=16)
            for _ in range(12):
                got = plan.eval(fd, form=form, n_slices=5)
                d = ((got - base).abs().amax(dim=1) / base.abs().amax(dim=1)).max()
                worst = max(worst, float(d))
        print(f"form={form}: worst rel Linf vs fixed order over 48 evaluations {worst:.2e}",
              flush=True)
        assert worst <= 1e-13, (form, worst)
        plan.set_variant(None, form=form, n_slices=5, n_rhs=16)


@pytest.mark.parametrize("form,tol", [("f64", TOL64), ("split", TOLSPLIT)])
def test_c2_style_adaptive_canonical(cuda, form, tol):
    """A small instance of the adaptive 3D workload: q=5, refined while the
    interpolation error of the seed-0 8-Gaussian density exceeds 1e-4 (depth
    2..4), 2:1 balanced, one canonical Laplace table; GPU-built table."""
    print(f"case: C2-style adaptive 3D q=5 form={form}", flush=True)
    fden = (lambda P: densities.gaussian_mixture(P, 0))
    tr = vtree.adaptive_tree_batched(3, 5, vtree.interpolation_refiner(fden, 3, 5, 1e-4),
                                     min_level=2, max_level=4)
    assert tr.nlevels >= 4
    tables = [_laplace_table(3, 5)]
    f = densities.gaussian_mixture_batch(tr.points, [0, 1])
    got = _plan_eval(tr, tables, f, form, 5)
    ref = _oracle(tr, tables, f, "canonical")
    l2, linf = rel_errors(got, ref)
    print(f"{form}: rel L2 {l2:.2e} Linf {linf:.2e}")
    assert l2 <= tol and linf <= tol, (l2, linf)


def _sub_tree(tr, sel):
    """Copy of `tr` whose target boxes are the entries `sel` of its target-box
    list (with their list-1 sources)."""
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


@pytest.mark.parametrize("level,slices", [(3, 5)])
def test_headline_shape_vs_oracle(cuda, level, slices):
    """The benchmark's evaluation shape: 3D Laplace q=6 (216 nodes, blocks of
    224 rows, two row tiles per wave), a GPU-built table, 64 seeded
    8-Gaussian densities (seeds 1000..1063) in one plan evaluation (tiles of
    1 box x 64 RHS), the BF16 split form (5 slices) and dense FP64 against
    the C oracle, every RHS: rel-L2 and rel-Linf per RHS <= 1e-10 (split)
    and <= 1e-12 (FP64), on a seeded sample of 64 of the 512 target boxes
    (interior and boundary) of uniform level 3.  At level 2 (boxes of extent
    0.5 against Gaussian widths 0.02..0.1) five BF16 slices measured 1.1e-10
    rel Linf on the worst RHS, and six do not fit the BF16 kernel's LDS at
    224 rows."""
    print(f"case: headline shape 3D q=6 level {level} R=64 S={slices}", flush=True)
    tr = vtree.uniform_tree(3, 6, level)
    table = _laplace_table(3, 6)
    f = densities.gaussian_mixture_batch(tr.points, range(1000, 1064))
    w = NearFieldWrangler(tr, table)
    fd = torch.as_tensor(f, device="cuda").contiguous()
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    got = {form: w.eval_direct_plan(*args, fd, form=form, n_slices=slices)[0][0].cpu().numpy()
           for form in ("f64", "split")}
    sel = np.arange(tr.target_boxes.size, dtype=np.int32)
    if sel.size > 64:
        sel = np.sort(np.random.default_rng(11).choice(sel, 64, replace=False)).astype(np.int32)
    meta = orbit_ref.case_metadata(3, 6)
    ref = ocpu.near_field(_sub_tree(tr, sel), table.compact_values(), table.mode_normalizers,
                          meta, f, mode="canonical", table_root_extent=2.0)
    idx = np.concatenate([np.arange(tr.box_target_starts[b], tr.box_target_starts[b] + 216)
                          for b in np.asarray(tr.target_boxes)[sel]])
    for form, tol in (("f64", TOL64), ("split", TOLSPLIT)):
        l2, linf = rel_errors(got[form][:, idx], ref[:, idx])
        print(f"{form}: rel L2 {l2:.2e} Linf {linf:.2e} (64 RHS, {idx.size} points)")
        assert l2 <= tol and linf <= tol, (form, l2, linf)


def test_staging_alloc_failure_recovers(cuda):
    """A staging request beyond the device's memory fails with VL1_E_HIP
    before any kernel runs, leaves the plan's buffers as they were (no
    recorded capacity without memory), and the next evaluation allocates
    and matches the oracle."""
    from volumential_amd import _capi

    print("case: staging allocation failure, 3D q=3 level 1", flush=True)
    tr = vtree.uniform_tree(3, 3, 1)  # 8 boxes, 27 nodes -> 32 staged rows
    tables = synthetic_table_levels(3, 3, seed=12)
    w = NearFieldWrangler(tr, tables[0])
    plan = w.get_plan(w.target_kernels[0], tr.target_boxes, tr.neighbor_source_boxes_starts,
                      tr.neighbor_source_boxes_lists)
    f = densities.gaussian_mixture_batch(tr.points, [1, 2])
    fd = torch.as_tensor(f, device="cuda").contiguous()
    plan.eval(fd, form="split", n_slices=5)
    torch.cuda.synchronize()
    before = plan.stats()["device_bytes"]
    R = (1 << 28) - 64  # 8 boxes x R columns stays below 2^31
    need = 8 * 5 * R * 32 * 2
    assert need > 2 * torch.cuda.get_device_properties(0).total_memory
    L = _capi.lib()
    opts = _capi.SplitOpts(n_slices=5, n_products=0)
    st = L.vl1_plan_stage(plan.handle, _capi.FORM_SPLIT_BF16, fd.data_ptr(), fd.shape[1], R, opts,
                          _capi.stream_handle(None))
    assert st == _capi.VL1_E_HIP, st
    assert plan.stats()["device_bytes"] == before
    got = plan.eval(fd, form="split", n_slices=5).cpu().numpy()
    ref = _oracle(tr, tables, f, "canonical")
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOLSPLIT and linf <= TOLSPLIT, (l2, linf)


@pytest.mark.parametrize("bad", ["negative", "too_large", "starts"])
def test_plan_rejects_bad_csr(cuda, bad):
    """Plan creation range-checks the host CSR (ValueError, nothing indexed)."""
    print(f"case: plan bad CSR {bad}", flush=True)
    from volumential_amd.payload import DevicePayload, prepare_table_data_and_entry_map
    from volumential_amd.plan import List1Plan

    tr = vtree.uniform_tree(2, 3, 2)
    t = synthetic_table_levels(2, 3, seed=1)[0]
    data, nmlz, _, _, _, _ = prepare_table_data_and_entry_map([t])
    pay = DevicePayload(2, 3, data, nmlz, t.case_indices, t.arithmetic_metadata())
    tb = np.array(tr.target_boxes, np.int32)
    st = np.array(tr.neighbor_source_boxes_starts, np.int32)
    if bad == "negative":
        tb[3] = -1
    elif bad == "too_large":
        tb[0] = tr.nboxes
    else:
        st[2], st[3] = st[3], st[2]
    with pytest.raises(ValueError):
        List1Plan(tr, pay, scaling_mode=0, scaling_power=2, log_displacement=True,
                  table_root_extent=2.0, target_boxes=tb, starts=st)
