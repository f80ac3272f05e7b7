"""GPU parity of the reference's ``dense`` reconstruction kind
(list1.py:1018-1045, the default of NearFieldFromCSR, list1.py:120) against
the oracle, through both routes: ``NearFieldFromCSR`` (plain FP64 kernel)
and a ``List1Plan`` built on a dense payload (FP64, BF16 and int8 split
forms).

Inputs: compact synthetic tables expanded to the full (case, s, t) layout
(``tests/dense_kind.py``), read through identity maps or random per-mode
signed-permutation maps, with missing entries, and complex values with
complex entry scales.  Tolerance: FP64 routes <= 1e-12 relative per RHS
(L2 and Linf); int8 split <= 1e-10; BF16 split (5 slices) <= 3e-10.
"""

import numpy as np
import pytest
import torch

from oracle import cpu as ocpu
from oracle import list1_ref, orbit_ref
from tests.dense_kind import dense_arrays, oracle_meta
from tests.gpu_helpers import rel_errors, synthetic_table_levels
from volumential_amd import _capi, densities, gallery, tree as vtree
from volumential_amd.kernels import LaplaceKernel
from volumential_amd.list1 import NearFieldFromCSR
from volumential_amd.payload import DenseDevicePayload
from volumential_amd.plan import List1Plan

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _plain(tr, arrays, nm, meta, f, kernel):
    n = tr.quad_order**tr.dim
    shapes = {"n_tables": 1, "n_q_points": n,
              "n_table_entries": int(np.asarray(arrays["table_data_combined"]).shape[-1]),
              "n_cases": len(meta["case_orbit_ranks"]), "quad_order": tr.quad_order}
    nf = NearFieldFromCSR(kernel, shapes)  # no reconstruction_kind: the reference's default
    assert nf.reconstruction_kind == "dense"
    cplx = np.iscomplexobj(arrays["table_data_combined"])
    fd = torch.as_tensor(f, device="cuda", dtype=torch.complex128 if cplx else torch.float64)
    res = torch.zeros_like(fd)
    out, _ = nf(
        None, result=res, box_centers=tr.box_centers, box_levels=tr.box_levels,
        box_source_counts_nonchild=tr.box_target_counts_nonchild,
        box_source_starts=tr.box_target_starts,
        box_target_counts_nonchild=tr.box_target_counts_nonchild,
        box_target_starts=tr.box_target_starts, case_indices=meta["case_indices"],
        encoding_base=gallery.ENCODING_BASE, encoding_shift=gallery.ENCODING_SHIFT,
        mode_nmlz_combined=nm, neighbor_source_boxes_starts=tr.neighbor_source_boxes_starts,
        root_extent=tr.root_extent, neighbor_source_boxes_lists=tr.neighbor_source_boxes_lists,
        mode_coefs=fd, source_mode_ids=tr.local_ids, target_boxes=tr.target_boxes,
        target_point_ids=tr.local_ids, table_root_extent=2.0, table_starting_level=0,
        **arrays)
    return out.cpu().numpy()


def _oracle(tr, arrays, nm, meta, f, logd):
    return list1_ref.near_field(
        dim=tr.dim, q=tr.quad_order, box_centers=tr.box_centers, box_levels=tr.box_levels,
        box_starts=tr.box_target_starts, box_counts=tr.box_target_counts_nonchild,
        target_boxes=tr.target_boxes, nbr_starts=tr.neighbor_source_boxes_starts,
        nbr_lists=tr.neighbor_source_boxes_lists, source_mode_ids=tr.local_ids,
        target_point_ids=tr.local_ids, table_data=arrays["table_data_combined"], mode_nmlz=nm,
        meta=oracle_meta(meta, arrays), mode="canonical", log_displacement=logd,
        root_extent=tr.root_extent, table_root_extent=2.0, mode_coefs=f)


@pytest.mark.parametrize("maps", ["identity", "group"])
def test_plain_dense_2d_log(cuda, maps):
    print(f"case: dense kind plain 2D q=4 level 4 canonical+log maps={maps}", flush=True)
    tr = vtree.uniform_tree(2, 4, 4)
    t = synthetic_table_levels(2, 4, seed=11)[0]
    meta = orbit_ref.case_metadata(2, 4)
    arr = dense_arrays(t.compact_values()[None, :], meta, 4, 2, maps=maps, seed=3)
    nm = t.mode_normalizers[None, :]
    f = densities.gaussian_mixture_batch(tr.points, range(3))
    got = _plain(tr, arr, nm, meta, f, LaplaceKernel(2))
    ref = _oracle(tr, arr, nm, meta, f, True)
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL and linf <= TOL, (l2, linf)


def test_plain_dense_missing_entries_skip_displacement(cuda):
    """Two cases without entries: their pairs add nothing, the 2D log
    displacement included (list1.py:1111-1116)."""
    print("case: dense kind plain 2D q=3 level 4, two cases missing, log term", flush=True)
    tr = vtree.uniform_tree(2, 3, 4)
    t = synthetic_table_levels(2, 3, seed=12)[0]
    meta = orbit_ref.case_metadata(2, 3)
    ci = np.asarray(meta["case_indices"])
    drop = [int(ci[orbit_ref.encode([0, 4])]), int(ci[orbit_ref.encode([-4, -4])])]
    arr = dense_arrays(t.compact_values()[None, :], meta, 3, 2, drop_cases=drop)
    nm = t.mode_normalizers[None, :]
    f = densities.gaussian_mixture_batch(tr.points, range(20, 24))
    got = _plain(tr, arr, nm, meta, f, LaplaceKernel(2))
    ref = _oracle(tr, arr, nm, meta, f, True)
    full = _oracle(tr, dense_arrays(t.compact_values()[None, :], meta, 3, 2), nm, meta, f, True)
    assert np.abs(full - ref).max() > 1e-3 * np.abs(full).max()  # the drop matters
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL and linf <= TOL, (l2, linf)


def test_plain_dense_complex_scales(cuda):
    """Complex values with complex entry scales and +-1 mode-case scales:
    layout values are the full table divided by both, so the product the
    reconstruction forms is the full table again."""
    print("case: dense kind plain 3D q=3 level 2 complex scales", flush=True)
    tr = vtree.uniform_tree(3, 3, 2)
    meta = orbit_ref.case_metadata(3, 3)
    rng = np.random.default_rng(4)
    t = synthetic_table_levels(3, 3, seed=13)[0]
    comp = t.compact_values() * (1.0 + 0.5j) + 0.3j * rng.standard_normal(t.compact_values().size)
    arr = dense_arrays(comp[None, :], meta, 3, 3)
    n, nc = 27, len(meta["case_orbit_ranks"])
    phase = np.exp(2j * np.pi * rng.random(nc * n * n))
    mcs = rng.choice([-1.0, 1.0], size=(n, nc)).astype(np.complex128)
    full = arr["table_data_combined"][0].reshape(nc, n, n)
    lay = full / phase.reshape(nc, n, n) / mcs.T[:, :, None]
    arr.update(table_data_combined=lay.reshape(1, -1), table_entry_scales=phase,
               mode_case_scale=mcs)
    nm = t.mode_normalizers[None, :].astype(np.complex128)
    f = (densities.gaussian_mixture_batch(tr.points, range(2))
         * (1.0 - 0.7j)).astype(np.complex128)
    got = _plain(tr, arr, nm, meta, f, LaplaceKernel(3))
    ref = _oracle(tr, arr, nm, meta, f, False)
    l2, linf = rel_errors(got, ref)
    assert l2 <= TOL and linf <= TOL, (l2, linf)


# BF16 split, 5 slices, on this random-valued table at level 2: 1.7e-10
# Linf (the blocks are right: the FP64 form meets 1e-12 on them)
@pytest.mark.parametrize("form,tol", [("f64", 1e-12), ("split", 3e-10), ("split_i8", 1e-10)])
def test_plan_dense_payload_3d(cuda, form, tol):
    print(f"case: dense kind plan 3D q=6 level 2 R=64 group maps form={form}", flush=True)
    tr = vtree.uniform_tree(3, 6, 2)
    t = synthetic_table_levels(3, 6, seed=14)[0]
    meta = orbit_ref.case_metadata(3, 6)
    arr = dense_arrays(t.compact_values()[None, :], meta, 6, 3, maps="group", seed=9)
    pay = DenseDevicePayload(3, 6, arr["table_data_combined"], t.mode_normalizers[None, :],
                             meta["case_indices"], arr["table_entry_ids"],
                             arr["table_entry_scales"], arr["mode_qpoint_map"],
                             arr["mode_case_map"], arr["mode_case_scale"],
                             n_cases=len(meta["case_orbit_ranks"]))
    assert not pay.has_missing
    plan = List1Plan(tr, pay, scaling_mode=_capi.SCALING_CANONICAL, scaling_power=2,
                     table_root_extent=2.0)
    f = densities.gaussian_mixture_batch(tr.points, range(500, 564))
    got = plan.eval(torch.as_tensor(f, device="cuda"), form=form).cpu().numpy()
    ref = ocpu.near_field(tr, t.compact_values(), t.mode_normalizers, meta, f,
                          mode="canonical", table_root_extent=2.0)
    l2, linf = rel_errors(got, ref)
    print(f"{form}: rel L2 {l2:.2e} Linf {linf:.2e}")
    assert l2 <= tol and linf <= tol, (l2, linf)


def test_plan_dense_missing_with_log_term_refused(cuda):
    print("case: dense kind plan 2D q=3 level 3, missing entries + log term refused", flush=True)
    tr = vtree.uniform_tree(2, 3, 3)
    t = synthetic_table_levels(2, 3, seed=15)[0]
    meta = orbit_ref.case_metadata(2, 3)
    arr = dense_arrays(t.compact_values()[None, :], meta, 3, 2, drop_cases=[0])
    pay = DenseDevicePayload(2, 3, arr["table_data_combined"], t.mode_normalizers[None, :],
                             meta["case_indices"], arr["table_entry_ids"],
                             arr["table_entry_scales"], arr["mode_qpoint_map"],
                             arr["mode_case_map"], arr["mode_case_scale"],
                             n_cases=len(meta["case_orbit_ranks"]))
    assert pay.has_missing
    with pytest.raises(Exception, match="missing entries"):
        List1Plan(tr, pay, scaling_mode=_capi.SCALING_CANONICAL, scaling_power=2,
                  log_displacement=True, table_root_extent=2.0)
