"""GPU parity of the custom single-table scaling policy (caller code strings,
list1.py:389-404), plain route (vl1_near_field_f64) and plan forms.

* the reference's own inferred 2D Laplace rules given as strings reproduce
  the inferred policy (<= 1e-14) and the oracle's canonical result
  (<= 1e-12);
* level-dependent rules (scaling and displacement factors that change with
  the source level) match the oracle's custom branch: FP64 routes
  <= 1e-12, int8 split <= 1e-10 per RHS (rel L2 and Linf), on an adaptive
  tree with source boxes on several levels; also in 3D without a
  displacement.
"""

import math

import numpy as np
import pytest
import torch

from oracle import list1_ref, orbit_ref
from tests.gpu_helpers import adaptive, rel_errors, synthetic_table_levels
from volumential_amd import densities
from volumential_amd.scaling_code import level_arrays
from volumential_amd.wrangler import NearFieldWrangler

pytestmark = pytest.mark.gpu

LAP = "BOX_extent * BOX_extent / (table_root_extent * table_root_extent)"
LAP2D_DISP = ("-0.5 / PI * scaling * log(BOX_extent / table_root_extent) * "
              "mode_nmlz[table_lev, sid]").replace("PI", str(math.pi))
VAR_SCAL = "(1.0 + 0.25 * sbox_level) * " + LAP
VAR_DISP = "(0.5 - 0.1 * sbox_level) * scaling * mode_nmlz[table_lev, sid]"


def _eval(tr, table, f, codes, route, form="f64"):
    kw = {} if codes is None else {"kernel_scaling_code": codes[0],
                                   "kernel_displacement_code": codes[1]}
    w = NearFieldWrangler(tr, table, list1_extra_kwargs=kw,
                          direct_method="plain" if route == "plain" else "plan")
    fd = torch.as_tensor(f, device="cuda").contiguous()
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    if route == "plan_form":
        pot, _ = w.eval_direct_plan(*args, fd, form=form, n_slices=5)
    else:
        pot, _ = w.eval_direct_batched(*args, fd)
    return pot[0].cpu().numpy()


def _oracle(tr, table, f, codes=None):
    meta = orbit_ref.case_metadata(tr.dim, tr.quad_order)
    kw = dict(table_root_extent=2.0, own_csr=False)
    if codes is None:
        return list1_ref.near_field_from_tree(
            tr, table.compact_values()[None], table.mode_normalizers[None], meta, f,
            mode="canonical", scaling_power=2, log_displacement=tr.dim == 2, **kw)
    s, d = level_arrays(codes[0], codes[1], tr.root_extent, 2.0)
    return list1_ref.near_field_from_tree(
        tr, table.compact_values()[None], table.mode_normalizers[None], meta, f,
        mode="custom", level_scaling=s, level_disp=d, **kw)


@pytest.mark.parametrize("route", ["plain", "plan"])
def test_reference_rules_as_strings(cuda, route):
    print(f"case: custom strings = inferred 2D Laplace rules, adaptive 2D q=4, {route}",
          flush=True)
    tr = adaptive(2, 4)
    table = synthetic_table_levels(2, 4, seed=21)[0]
    f = densities.gaussian_mixture_batch(tr.points, [3, 4, 5])
    got = _eval(tr, table, f, (LAP, LAP2D_DISP), route)
    inferred = _eval(tr, table, f, None, route)
    l2, linf = rel_errors(got, inferred)
    print(f"vs inferred policy: {l2:.2e} {linf:.2e}", flush=True)
    assert l2 <= 1e-14 and linf <= 1e-14
    l2, linf = rel_errors(got, _oracle(tr, table, f))
    print(f"vs oracle canonical: {l2:.2e} {linf:.2e}", flush=True)
    assert l2 <= 1e-12 and linf <= 1e-12


@pytest.mark.parametrize("route,form,tol", [("plain", "f64", 1e-12), ("plan", "f64", 1e-12),
                                            ("plan_form", "split_i8", 1e-10),
                                            ("plan_form", "split", 1e-10)])
def test_level_dependent_rules_2d(cuda, route, form, tol):
    print(f"case: level-dependent custom rules, adaptive 2D q=4, {route} {form}", flush=True)
    tr = adaptive(2, 4)
    table = synthetic_table_levels(2, 4, seed=22)[0]
    f = densities.gaussian_mixture_batch(tr.points, [6, 7, 8, 9])
    got = _eval(tr, table, f, (VAR_SCAL, VAR_DISP), route, form)
    l2, linf = rel_errors(got, _oracle(tr, table, f, (VAR_SCAL, VAR_DISP)))
    print(f"vs oracle custom: {l2:.2e} {linf:.2e}", flush=True)
    assert l2 <= tol and linf <= tol


@pytest.mark.parametrize("route", ["plain", "plan"])
def test_level_dependent_rules_3d(cuda, route):
    print(f"case: level-dependent custom scaling, adaptive 3D q=3, {route}", flush=True)
    tr = adaptive(3, 3)
    table = synthetic_table_levels(3, 3, seed=23)[0]
    f = densities.gaussian_mixture_batch(tr.points, [1, 2])
    codes = ("pow(BOX_extent / table_root_extent, 2) * (2.0 - 0.125 * sbox_level)", "0.0")
    got = _eval(tr, table, f, codes, route)
    l2, linf = rel_errors(got, _oracle(tr, table, f, codes))
    print(f"vs oracle custom: {l2:.2e} {linf:.2e}", flush=True)
    assert l2 <= 1e-12 and linf <= 1e-12
