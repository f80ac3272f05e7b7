"""The plain route's FP64-MFMA kernel (vl1_near_field_f64 with 16 or more
densities on real tables) against the oracle and against the route's
scalar kernel (VL1_PLAIN_KERNEL=scalar), on the cases where its tiling
differs: one or two RHS groups of 64 (the second partial), boxes of 9 to
216 nodes (1 to 8 waves, one or two row tiles per wave), permuted, ragged,
repeated and invalid particle mode ids, per-level tables, the 2D log
displacement, and dense-kind payloads with missing entries.

Tolerance: rel-L2 and rel-Linf <= 1e-12 per RHS against the oracle, and
<= 1e-13 against the scalar kernel (same FP64 products, another order).
"""

import numpy as np
import pytest
import torch

from oracle import cpu as ocpu
from oracle import list1_ref, orbit_ref
from tests.dense_kind import dense_arrays, oracle_meta
from tests.gpu_helpers import adaptive, rel_errors, synthetic_table_levels
from volumential_amd import densities, gallery, tree as vtree
from volumential_amd.kernels import LaplaceKernel
from volumential_amd.list1 import NearFieldFromCSR
from volumential_amd.payload import prepare_table_data_and_entry_map
from volumential_amd.wrangler import NearFieldWrangler

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _both(monkeypatch, run):
    """(MFMA result, scalar-kernel result) of `run()`."""
    monkeypatch.delenv("VL1_PLAIN_KERNEL", raising=False)
    a = run()
    monkeypatch.setenv("VL1_PLAIN_KERNEL", "scalar")
    b = run()
    monkeypatch.delenv("VL1_PLAIN_KERNEL")
    return a, b


def _wrangler_plain(tr, tables, f):
    w = NearFieldWrangler(tr, tables if len(tables) > 1 else tables[0], direct_method="plain")
    pot, _ = w.eval_direct_batched(tr.target_boxes, tr.neighbor_source_boxes_starts,
                                   tr.neighbor_source_boxes_lists,
                                   torch.as_tensor(f, device="cuda"))
    return pot[0].cpu().numpy()


def _oracle(tr, tables, f, mode, logd=False, **kw):
    meta = orbit_ref.case_metadata(tr.dim, tr.quad_order)
    data = np.stack([t.compact_values() for t in tables])
    nm = np.stack([t.mode_normalizers for t in tables])
    tsl = int(round(np.log2(tr.root_extent / tables[0].source_box_extent)))
    return list1_ref.near_field_from_tree(
        tr, data, nm, meta, f, mode=mode, log_displacement=logd,
        table_root_extent=tables[0].source_box_extent, table_starting_level=tsl, **kw)


def _check(got, ref, tol=TOL):
    l2, linf = rel_errors(got, ref)
    assert l2 <= tol and linf <= tol, (l2, linf)


@pytest.mark.parametrize("n_rhs", [16, 70])
def test_2d_log_term_rhs_groups(cuda, monkeypatch, n_rhs):
    print(f"case: plain MFMA 2D q=4 level 4 canonical+log R={n_rhs}", flush=True)
    tr = vtree.uniform_tree(2, 4, 4)
    tables = synthetic_table_levels(2, 4, seed=21)
    f = densities.gaussian_mixture_batch(tr.points, range(n_rhs))
    a, b = _both(monkeypatch, lambda: _wrangler_plain(tr, tables, f))
    _check(a, b, 1e-13)
    _check(a, _oracle(tr, tables, f, "canonical", True))


def test_3d_q6_r64_two_row_tiles(cuda, monkeypatch):
    print("case: plain MFMA 3D q=6 level 2 R=64", flush=True)
    tr = vtree.uniform_tree(3, 6, 2)
    tables = synthetic_table_levels(3, 6, seed=22)
    f = densities.gaussian_mixture_batch(tr.points, range(100, 164))
    a, b = _both(monkeypatch, lambda: _wrangler_plain(tr, tables, f))
    _check(a, b, 1e-13)
    meta = orbit_ref.case_metadata(3, 6)
    _check(a, ocpu.near_field(tr, tables[0].compact_values(), tables[0].mode_normalizers, meta,
                              f, mode="canonical", table_root_extent=2.0))


@pytest.mark.parametrize("dim,q", [(2, 3), (3, 3)])
def test_adaptive_per_level(cuda, monkeypatch, dim, q):
    print(f"case: plain MFMA adaptive {dim}D q={q} per-level tables R=20", flush=True)
    tr = adaptive(dim, q)
    tables = synthetic_table_levels(dim, q, n_tables=tr.nlevels, seed=23)
    f = densities.gaussian_mixture_batch(tr.points, range(200, 220))
    a, b = _both(monkeypatch, lambda: _wrangler_plain(tr, tables, f))
    _check(a, b, 1e-13)
    _check(a, _oracle(tr, tables, f, "per_level"))


def _nf_direct(tr, tables, f, counts, src_ids, tgt_ids):
    """NearFieldFromCSR with the arithmetic kind and the given per-box
    particle counts and mode ids (the wrangler's layout checks bypassed)."""
    data, nmlz, _, _, _, info = prepare_table_data_and_entry_map(tables)
    t0 = tables[0]
    nf = NearFieldFromCSR(LaplaceKernel(tr.dim), {
        "n_tables": 1, "n_q_points": t0.n_q_points, "n_table_entries": data.shape[1],
        "n_cases": t0.n_cases, "quad_order": t0.quad_order,
        "reconstruction_kind": info["kind"]})
    res = torch.zeros(f.shape, dtype=torch.float64, device="cuda")
    out, _ = nf(
        None, result=res, box_centers=tr.box_centers, box_levels=tr.box_levels,
        box_source_counts_nonchild=counts, box_source_starts=tr.box_target_starts,
        box_target_counts_nonchild=counts, box_target_starts=tr.box_target_starts,
        case_indices=t0.case_indices, encoding_base=gallery.ENCODING_BASE,
        encoding_shift=gallery.ENCODING_SHIFT, mode_nmlz_combined=nmlz,
        neighbor_source_boxes_starts=tr.neighbor_source_boxes_starts,
        root_extent=tr.root_extent, neighbor_source_boxes_lists=tr.neighbor_source_boxes_lists,
        mode_coefs=torch.as_tensor(f, device="cuda"), source_mode_ids=src_ids,
        target_boxes=tr.target_boxes, target_point_ids=tgt_ids, table_data_combined=data,
        table_root_extent=t0.source_box_extent, table_starting_level=0,
        arithmetic_case_orbit_ranks=info["case_orbit_ranks"],
        arithmetic_case_axis_perm=info["case_axis_perm"],
        arithmetic_case_axis_sign=info["case_axis_sign"],
        arithmetic_case_axis_group=info["case_axis_group"],
        arithmetic_case_value_offsets=info["case_value_offsets"],
        arithmetic_axis_sign_power=info["axis_sign_power"])
    return out.cpu().numpy()


def test_ragged_permuted_ids(cuda, monkeypatch):
    """Boxes with fewer particles than nodes and particles in any mode
    order: the MFMA rows and K run over particles, so the oracle (which
    sums over a box's particles) agrees."""
    print("case: plain MFMA ragged boxes, permuted ids, 2D q=3 R=32", flush=True)
    tr = vtree.uniform_tree(2, 3, 3)
    n = tr.n_q_points
    rng = np.random.default_rng(5)
    counts = np.asarray(tr.box_target_counts_nonchild).copy()
    counts[rng.choice(tr.nboxes, tr.nboxes // 3, replace=False)] = rng.integers(1, n, tr.nboxes // 3)
    ids = np.concatenate([rng.permutation(n) for _ in range(tr.nboxes)]).astype(np.int32)
    tables = synthetic_table_levels(2, 3, seed=24)
    f = densities.gaussian_mixture_batch(tr.points, range(300, 332))
    a, b = _both(monkeypatch, lambda: _nf_direct(tr, tables, f, counts, ids, ids))
    _check(a, b, 1e-13)
    meta = orbit_ref.case_metadata(2, 3)
    ref = list1_ref.near_field(
        dim=2, q=3, box_centers=tr.box_centers, box_levels=tr.box_levels,
        box_starts=tr.box_target_starts, box_counts=counts, target_boxes=tr.target_boxes,
        nbr_starts=tr.neighbor_source_boxes_starts, nbr_lists=tr.neighbor_source_boxes_lists,
        source_mode_ids=ids, target_point_ids=ids, table_data=tables[0].compact_values()[None],
        mode_nmlz=tables[0].mode_normalizers[None], meta=meta, mode="canonical",
        log_displacement=True, root_extent=tr.root_extent, table_root_extent=2.0,
        mode_coefs=f, result=np.zeros_like(f))
    mask = np.zeros(tr.n_points, bool)
    for bx in range(tr.nboxes):
        s0 = tr.box_target_starts[bx]
        mask[s0:s0 + counts[bx]] = True
    _check(a[:, mask], ref[:, mask])
    assert np.all(a[:, ~mask] == 0.0)


def test_repeated_and_invalid_ids_match_scalar_kernel(cuda, monkeypatch):
    """Repeated mode ids sum, invalid ones (-1, >= n) contribute nothing and
    get no output: the same as the scalar kernel."""
    print("case: plain MFMA repeated/invalid ids, 3D q=3 level 2 R=20", flush=True)
    tr = vtree.uniform_tree(3, 3, 2)
    n = tr.n_q_points
    rng = np.random.default_rng(6)
    ids = np.asarray(tr.local_ids).copy()
    k = rng.choice(ids.size, ids.size // 10, replace=False)
    ids[k[: k.size // 3]] = -1
    ids[k[k.size // 3: 2 * k.size // 3]] = n + 3
    ids[k[2 * k.size // 3:]] = rng.integers(0, n, k.size - 2 * (k.size // 3))
    tables = synthetic_table_levels(3, 3, seed=25)
    f = densities.gaussian_mixture_batch(tr.points, range(400, 420))
    counts = np.asarray(tr.box_target_counts_nonchild)
    a, b = _both(monkeypatch, lambda: _nf_direct(tr, tables, f, counts, ids, ids))
    _check(a, b, 1e-13)
    assert np.all(a[:, (ids < 0) | (ids >= n)] == 0.0)


def test_dense_kind_missing_entries(cuda, monkeypatch):
    """Missing entries (NaN in the expanded blocks) read as zero: 3D, no
    displacement, R=24, against the oracle's dense kind."""
    print("case: plain MFMA dense kind, missing entries, 3D q=3 level 2 R=24", flush=True)
    tr = vtree.uniform_tree(3, 3, 2)
    t = synthetic_table_levels(3, 3, seed=26)[0]
    meta = orbit_ref.case_metadata(3, 3)
    arr = dense_arrays(t.compact_values()[None, :], meta, 3, 3, drop_cases=[3, 40, 70])
    nm = t.mode_normalizers[None, :]
    f = densities.gaussian_mixture_batch(tr.points, range(600, 624))
    shapes = {"n_tables": 1, "n_q_points": 27, "n_table_entries": arr["table_data_combined"].shape[1],
              "n_cases": len(meta["case_orbit_ranks"]), "quad_order": 3}

    def run():
        nf = NearFieldFromCSR(LaplaceKernel(3), shapes)
        res = torch.zeros(f.shape, dtype=torch.float64, device="cuda")
        out, _ = nf(
            None, result=res, box_centers=tr.box_centers, box_levels=tr.box_levels,
            box_source_counts_nonchild=tr.box_target_counts_nonchild,
            box_source_starts=tr.box_target_starts,
            box_target_counts_nonchild=tr.box_target_counts_nonchild,
            box_target_starts=tr.box_target_starts, case_indices=meta["case_indices"],
            encoding_base=gallery.ENCODING_BASE, encoding_shift=gallery.ENCODING_SHIFT,
            mode_nmlz_combined=nm, neighbor_source_boxes_starts=tr.neighbor_source_boxes_starts,
            root_extent=tr.root_extent, neighbor_source_boxes_lists=tr.neighbor_source_boxes_lists,
            mode_coefs=torch.as_tensor(f, device="cuda"), source_mode_ids=tr.local_ids,
            target_boxes=tr.target_boxes, target_point_ids=tr.local_ids, table_root_extent=2.0,
            table_starting_level=0, **arr)
        return out.cpu().numpy()

    a, b = _both(monkeypatch, run)
    _check(a, b, 1e-13)
    ref = list1_ref.near_field(
        dim=3, q=3, box_centers=tr.box_centers, box_levels=tr.box_levels,
        box_starts=tr.box_target_starts, box_counts=tr.box_target_counts_nonchild,
        target_boxes=tr.target_boxes, nbr_starts=tr.neighbor_source_boxes_starts,
        nbr_lists=tr.neighbor_source_boxes_lists, source_mode_ids=tr.local_ids,
        target_point_ids=tr.local_ids, table_data=arr["table_data_combined"], mode_nmlz=nm,
        meta=oracle_meta(meta, arr), mode="canonical", root_extent=tr.root_extent,
        table_root_extent=2.0, mode_coefs=f)
    _check(a, ref)
