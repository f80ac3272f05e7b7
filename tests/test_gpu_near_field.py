"""GPU parity of the wrangler's eval_direct drop-in against the CPU oracles,
through both list-1 routes: the cached FP64 plan (default, vl1_plan_* with
FORM_F64) and the plain FP64 kernel (vl1_near_field_f64).

Tolerance (FP64 form): rel-L2 and rel-Linf <= 1e-12 per RHS.
"""

import numpy as np
import pytest
import torch

from oracle import cpu as ocpu
from oracle import list1_ref, orbit_ref
from tests.gpu_helpers import adaptive, rel_errors, synthetic_table_levels
from volumential_amd import densities, tree as vtree
from volumential_amd.wrangler import NearFieldWrangler

TOL = 1e-12

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["plan", "plain"])
def method(request):
    return request.param


def _run(tr, tables, coefs, method, **kw):
    w = NearFieldWrangler(tr, tables if len(tables) > 1 else tables[0], direct_method=method,
                          **kw)
    c = torch.as_tensor(coefs, device="cuda")
    if c.dim() == 1:
        pot, _ = w.eval_direct(tr.target_boxes, tr.neighbor_source_boxes_starts,
                               tr.neighbor_source_boxes_lists, c)
    else:
        pot, _ = w.eval_direct_batched(tr.target_boxes, tr.neighbor_source_boxes_starts,
                                       tr.neighbor_source_boxes_lists, c)
    return pot[0].cpu().numpy()


def _oracle(tr, tables, coefs, mode, power=2, logd=False):
    meta = orbit_ref.case_metadata(tr.dim, tr.quad_order)
    data = np.stack([t.compact_values() for t in tables])
    nm = np.stack([t.mode_normalizers for t in tables])
    tsl = int(round(np.log2(tr.root_extent / tables[0].source_box_extent)))
    return list1_ref.near_field_from_tree(
        tr, data, nm, meta, coefs, mode=mode, scaling_power=power, log_displacement=logd,
        table_root_extent=tables[0].source_box_extent, table_starting_level=tsl,
    )


def test_laplace2d_q4_level5_log_term(cuda, method):
    """C1 geometry: 2D, q=4, uniform level 5 (1024 boxes), canonical scaling
    with the FP64 log displacement, single RHS."""
    print(f"case: 2D q=4 level 5, canonical + log, {method}", flush=True)
    tr = vtree.uniform_tree(2, 4, 5)
    tables = synthetic_table_levels(2, 4)
    f = np.exp(-np.sum(tr.points**2, axis=0) / 0.01)
    got = _run(tr, tables, f, method)
    ref = _oracle(tr, tables, f, "canonical", 2, True)[0]
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL and linf <= TOL, (l2, linf)


@pytest.mark.parametrize("n_rhs", [1, 3, 6, 20])
def test_laplace3d_q6_level2_batched(cuda, n_rhs, method):
    print(f"case: 3D q=6 level 2, R={n_rhs}, {method}", flush=True)
    tr = vtree.uniform_tree(3, 6, 2)
    tables = synthetic_table_levels(3, 6, seed=1)
    f = densities.gaussian_mixture_batch(tr.points, range(1000, 1000 + n_rhs))
    got = _run(tr, tables, f if n_rhs > 1 else f[0], method)
    meta = orbit_ref.case_metadata(3, 6)
    ref = ocpu.near_field(tr, tables[0].compact_values(), tables[0].mode_normalizers, meta, f,
                          mode="canonical", table_root_extent=2.0)
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL and linf <= TOL, (l2, linf)


@pytest.mark.parametrize("dim,q", [(2, 3), (2, 6), (3, 2), (3, 5)])
def test_adaptive_per_level_tables(cuda, dim, q, method):
    print(f"case: adaptive {dim}D q={q} per-level tables, {method}", flush=True)
    tr = adaptive(dim, q)
    tables = synthetic_table_levels(dim, q, n_tables=tr.nlevels, seed=5)
    f = densities.gaussian_mixture_batch(tr.points, [11, 12])
    got = _run(tr, tables, f, method)
    ref = _oracle(tr, tables, f, "per_level")
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL and linf <= TOL, (l2, linf)


@pytest.mark.parametrize("dim,q", [(2, 4), (3, 3)])
def test_adaptive_canonical_scaling(cuda, dim, q, method):
    print(f"case: adaptive {dim}D q={q} canonical, {method}", flush=True)
    tr = adaptive(dim, q, max_level=5)
    tables = synthetic_table_levels(dim, q, seed=6)
    f = densities.gaussian_mixture_batch(tr.points, [21, 22, 23])
    got = _run(tr, tables, f, method)
    ref = _oracle(tr, tables, f, "canonical", 2, dim == 2)
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL and linf <= TOL, (l2, linf)


def test_permuted_box_local_ids(cuda, method):
    """Particles inside a box in any order (source_mode_ids/target_point_ids)."""
    print(f"case: permuted local ids, {method}", flush=True)
    tr = vtree.uniform_tree(2, 3, 3)
    rng = np.random.default_rng(9)
    n = tr.n_q_points
    perm = np.concatenate([rng.permutation(n) for _ in range(tr.nboxes)]).astype(np.int32)
    tr.local_ids = perm
    tables = synthetic_table_levels(2, 3, seed=2)
    f = densities.gaussian_mixture_batch(tr.points, [31, 32])
    got = _run(tr, tables, f, method)
    ref = _oracle(tr, tables, f, "canonical", 2, True)
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL and linf <= TOL, (l2, linf)


def test_only_listed_targets_overwritten(cuda, method):
    """Targets of unlisted boxes keep the caller's zeros."""
    from volumential_amd.list1 import NearFieldFromCSR  # noqa: F401

    print(f"case: partial target list, {method}", flush=True)
    tr = vtree.uniform_tree(2, 3, 3)
    tables = synthetic_table_levels(2, 3, seed=3)
    f = densities.gaussian_mixture(tr.points, 41)
    w = NearFieldWrangler(tr, tables[0], direct_method=method)
    sel = np.arange(0, tr.nboxes, 3)
    starts = np.zeros(sel.size + 1, np.int32)
    lists = []
    for k, b in enumerate(sel):
        seg = tr.neighbor_source_boxes_lists[
            tr.neighbor_source_boxes_starts[b] : tr.neighbor_source_boxes_starts[b + 1]
        ]
        lists.extend(seg)
        starts[k + 1] = starts[k] + len(seg)
    pot, _ = w.eval_direct(sel.astype(np.int32), starts, np.asarray(lists, np.int32),
                           torch.as_tensor(f, device="cuda"))
    got = pot[0].cpu().numpy()
    full = _oracle(tr, tables, f, "canonical", 2, True)[0]
    n = tr.n_q_points
    mask = np.zeros(tr.n_points, bool)
    for b in sel:
        mask[b * n : (b + 1) * n] = True
    assert np.all(got[~mask] == 0.0)
    l2, linf = rel_errors(got[mask], full[mask])
    assert l2 <= TOL and linf <= TOL


def test_error_paths(cuda, method):
    print(f"case: error paths, {method}", flush=True)
    tr = vtree.uniform_tree(2, 3, 2)
    tables = synthetic_table_levels(2, 3)
    f = torch.ones(tr.n_points, dtype=torch.float64, device="cuda")
    # a source box far away in the negative direction -> invalid case code
    w = NearFieldWrangler(tr, tables[0], direct_method=method)
    bad_lists = np.array([tr.nboxes - 1], np.int32)
    with pytest.raises(RuntimeError, match="invalid case id"):
        w.eval_direct(np.array([0], np.int32), np.array([0, 1], np.int32), bad_lists, f)
    # fixed single table (no inferred scaling) at a level that does not match
    w2 = NearFieldWrangler(tr, tables[0], list1_extra_kwargs={"infer_kernel_scaling": False},
                           direct_method=method)
    with pytest.raises(RuntimeError, match="level mismatch"):
        w2.eval_direct(tr.target_boxes, tr.neighbor_source_boxes_starts,
                       tr.neighbor_source_boxes_lists, f)
    # layout: a box with the wrong node count
    tr.box_target_counts_nonchild = tr.box_target_counts_nonchild.copy()
    tr.box_target_counts_nonchild[3] = 5
    with pytest.raises(ValueError, match="quadrature points per active"):
        w.eval_direct(tr.target_boxes, tr.neighbor_source_boxes_starts,
                      tr.neighbor_source_boxes_lists, f)


def test_plan_route_rhs_chunks_and_repeat_calls(cuda):
    """The plan route splits R = 70 densities into tile-kernel chunks (64 + 6)
    and reuses its cached plan on a second call; both agree with the plain
    kernel to 1e-13 and the second call with the oracle to the FP64 bar."""
    print("case: plan route, R=70 then R=3", flush=True)
    tr = vtree.uniform_tree(2, 4, 4)
    tables = synthetic_table_levels(2, 4, seed=8)
    wp = NearFieldWrangler(tr, tables[0])
    wq = NearFieldWrangler(tr, tables[0], direct_method="plain")
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    for seeds in (range(70), range(500, 503)):
        f = torch.as_tensor(densities.gaussian_mixture_batch(tr.points, seeds), device="cuda")
        a = wp.eval_direct_batched(*args, f)[0][0].cpu().numpy()
        b = wq.eval_direct_batched(*args, f)[0][0].cpu().numpy()
        l2, linf = rel_errors(a, b)
        assert l2 <= 1e-13 and linf <= 1e-13, (len(seeds), l2, linf)
    assert len(wp._plans) == 1
    ref = _oracle(tr, tables, f.cpu().numpy(), "canonical", 2, True)
    l2, linf = rel_errors(a, ref)
    assert l2 <= TOL and linf <= TOL, (l2, linf)


def test_plan_cache_evicts_least_recent(cuda):
    """With room for one plan, a second CSR evicts (and frees) the first;
    evaluating the first again rebuilds it with the same result."""
    print("case: plan cache eviction, 2D q=4 L4", flush=True)
    tr = vtree.uniform_tree(2, 4, 4)
    tables = synthetic_table_levels(2, 4, seed=9)
    w = NearFieldWrangler(tr, tables[0])
    w.max_cached_plans = 1
    f = torch.as_tensor(densities.gaussian_mixture_batch(tr.points, [3])[0], device="cuda")
    full = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    n_half = len(tr.target_boxes) // 2
    starts = np.asarray(tr.neighbor_source_boxes_starts)
    half = (np.asarray(tr.target_boxes)[:n_half], starts[:n_half + 1],
            np.asarray(tr.neighbor_source_boxes_lists)[:starts[n_half]])
    a = w.eval_direct(*full, f)[0][0].clone()
    first = next(iter(w._plans.values()))
    w.eval_direct(*half, f)
    assert len(w._plans) == 1 and next(iter(w._plans.values())) is not first
    b = w.eval_direct(*full, f)[0][0]
    assert len(w._plans) == 1
    assert torch.equal(a, b)


def test_plan_route_box_size_beyond_tiles(cuda):
    """3D q=7 (343 nodes per box) is beyond one tile: the plan cuts each block
    into two row groups (172 rows each, all 343 columns), one tile each; the
    FP64 form meets the FP64 bar and the split form the split bar."""
    print("case: plan route, 3D q=7 row groups", flush=True)
    tr = vtree.uniform_tree(3, 7, 1)
    tables = synthetic_table_levels(3, 7, seed=4)
    f = densities.gaussian_mixture_batch(tr.points, [61, 62])
    w = NearFieldWrangler(tr, tables[0])
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    got = w.eval_direct(*args, torch.as_tensor(f[0], device="cuda"))[0]
    assert not w._plain_only
    plan = next(iter(w._plans.values()))
    assert plan.stats()["tiles"] == 2 * -(-len(tr.target_boxes) // 64)  # 1 RHS: 64 boxes a tile
    meta = orbit_ref.case_metadata(3, 7)
    ref = ocpu.near_field(tr, tables[0].compact_values(), tables[0].mode_normalizers, meta, f,
                          mode="canonical", table_root_extent=2.0)
    l2, linf = rel_errors(got[0].cpu().numpy(), ref[0])
    assert l2 <= TOL and linf <= TOL, (l2, linf)
    sp, _ = w.eval_direct_plan(*args, torch.as_tensor(f, device="cuda"), form="split")
    l2, linf = rel_errors(sp[0].cpu().numpy(), ref)
    assert l2 <= 1e-10 and linf <= 1e-10, (l2, linf)


def test_single_precision_wrangler_rounds_fp64_result(cuda):
    """A float32 wrangler returns float32 (complex64 for complex kernels)
    potentials: the FP64 evaluation rounded (expansion_wrangler_fpnd.py:
    3283-3297 picks f32 / c64 for such wranglers)."""
    print("case: float32 wrangler", flush=True)
    tr = vtree.uniform_tree(2, 3, 3)
    tables = synthetic_table_levels(2, 3, seed=12)
    f = torch.as_tensor(densities.gaussian_mixture(tr.points, 3), device="cuda")
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    p64, _ = NearFieldWrangler(tr, tables[0]).eval_direct(*args, f)
    p32, _ = NearFieldWrangler(tr, tables[0], dtype=np.float32).eval_direct(*args, f)
    assert p32[0].dtype == torch.float32
    assert torch.equal(p32[0], p64[0].to(torch.float32))


@pytest.mark.parametrize("gather", [False, True])
def test_plain_table_paths(cuda, gather, monkeypatch):
    """The plain kernel's two table paths -- the expanded per-(level, case)
    blocks (default) and the staged orbit block with row-map gathers
    (VL1_PLAIN_TABLE=gather, used beyond 8 GiB) -- on real, complex and
    signed tables, against the oracle."""
    from volumential_amd.kernels import AxisTargetDerivative, HelmholtzKernel, LaplaceKernel
    from volumential_amd.table import DuffyBuildConfig, NearFieldInteractionTable

    print(f"case: plain table path gather={gather}", flush=True)
    if gather:
        monkeypatch.setenv("VL1_PLAIN_TABLE", "gather")
    tr = vtree.uniform_tree(2, 4, 3)
    f = densities.gaussian_mixture_batch(tr.points, [90, 91, 92])
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    meta = orbit_ref.case_metadata(2, 4)
    # real canonical (2D log term)
    tables = synthetic_table_levels(2, 4, seed=21)
    got = NearFieldWrangler(tr, tables[0], direct_method="plain").eval_direct_batched(
        *args, torch.as_tensor(f, device="cuda"))[0][0].cpu().numpy()
    ref = _oracle(tr, tables, f, "canonical", 2, True)
    assert max(rel_errors(got, ref)) <= TOL
    # signed (gradient) table
    g = NearFieldInteractionTable(4, 2, AxisTargetDerivative(1, LaplaceKernel(2)),
                                  source_box_extent=2.0)
    g.build_table(DuffyBuildConfig(regular_quad_order=8, radial_quad_order=21))
    got = NearFieldWrangler(tr, g, direct_method="plain").eval_direct_batched(
        *args, torch.as_tensor(f, device="cuda"))[0][0].cpu().numpy()
    ref = ocpu.near_field(tr, g.compact_values(), g.mode_normalizers,
                          orbit_ref.case_metadata(2, 4, (1,), (0, 1)), f, mode="canonical",
                          scaling_power=1, table_root_extent=2.0)
    assert max(rel_errors(got, ref)) <= TOL
    # complex per-level table at the tree's level (fixed single table)
    h = NearFieldInteractionTable(4, 2, HelmholtzKernel(2, 9.0), source_box_extent=0.25,
                                  source_box_level=3)
    h.build_table(DuffyBuildConfig(regular_quad_order=8, radial_quad_order=21))
    fc = f[0] + 1j * f[1]
    got = NearFieldWrangler(tr, h, direct_method="plain").eval_direct(
        *args, torch.as_tensor(fc, device="cuda"))[0][0].cpu().numpy()
    ref = list1_ref.near_field_from_tree(
        tr, h.compact_values()[None], h.mode_normalizers[None], meta, fc, mode="fixed",
        table_root_extent=0.25, table_starting_level=3)[0]
    assert max(rel_errors(got, ref)) <= TOL
