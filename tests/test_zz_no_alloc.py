"""vl1_plan_eval_noalloc (the allocation-free plan evaluation) against
vl1_plan_eval, through the C ABI: two calls in a row on one plan, the second
with fewer densities, in both forms; results agree to 1e-13 relative per
RHS (the same kernels on the same data; only the order of a tile's steps,
which follows a racy L2 hint, can change the last bits), and the split
result meets the oracle's 1e-10 bar.  Oversized and unprepared requests
return errors without launching anything."""

import numpy as np
import pytest
import torch

from oracle import cpu as ocpu
from oracle import orbit_ref
from tests.gpu_helpers import rel_errors, synthetic_table_levels
from volumential_amd import _capi, densities, tree as vtree
from volumential_amd.plan import FORMS
from volumential_amd.wrangler import NearFieldWrangler

pytestmark = pytest.mark.gpu


def _call(fn, plan, form, f, y, S):
    opts = _capi.SplitOpts(n_slices=S, n_products=0)
    return fn(plan.handle, FORMS[form], f.data_ptr(), f.shape[1], f.shape[0], y.data_ptr(),
              y.shape[1], opts, _capi.stream_handle(None))


@pytest.mark.parametrize("form", ["f64", "split"])
def test_noalloc_matches_eval(cuda, form):
    print(f"case: noalloc vs eval 3D q=4 level 3 form={form} (R=12, then R=5)", flush=True)
    tr = vtree.uniform_tree(3, 4, 3)
    tables = synthetic_table_levels(3, 4, seed=21)
    w = NearFieldWrangler(tr, tables[0])
    plan = w.get_plan(w.target_kernels[0], tr.target_boxes, tr.neighbor_source_boxes_starts,
                      tr.neighbor_source_boxes_lists)
    L = _capi.lib()
    S = 5
    meta = orbit_ref.case_metadata(3, 4)
    for R in (12, 5):
        f = densities.gaussian_mixture_batch(tr.points, range(40 + R, 40 + 2 * R))
        fd = torch.as_tensor(f, device="cuda").contiguous()
        y_ref = torch.zeros_like(fd)
        y_na = torch.full_like(fd, float("nan"))  # every target must be written
        _capi.check(_call(L.vl1_plan_eval, plan, form, fd, y_ref, S), "vl1_plan_eval")
        _capi.check(_call(L.vl1_plan_eval_noalloc, plan, form, fd, y_na, S),
                    "vl1_plan_eval_noalloc")
        torch.cuda.synchronize()
        a, b = y_na.cpu().numpy(), y_ref.cpu().numpy()
        l2, linf = rel_errors(a, b)
        print(f"R={R}: noalloc vs eval rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
        assert l2 <= 1e-13 and linf <= 1e-13, (R, l2, linf)
        ref = ocpu.near_field(tr, tables[0].compact_values(), tables[0].mode_normalizers, meta, f,
                              mode="canonical", table_root_extent=2.0)
        l2, linf = rel_errors(a, ref)
        tol = 1e-12 if form == "f64" else 1e-10
        assert l2 <= tol and linf <= tol, (R, l2, linf)


def test_noalloc_rejects(cuda):
    print("case: noalloc errors (unprepared, oversized)", flush=True)
    tr = vtree.uniform_tree(3, 6, 3)  # 512 boxes, 224 staged rows
    tables = synthetic_table_levels(3, 6, seed=22)
    w = NearFieldWrangler(tr, tables[0])
    plan = w.get_plan(w.target_kernels[0], tr.target_boxes, tr.neighbor_source_boxes_starts,
                      tr.neighbor_source_boxes_lists)
    L = _capi.lib()
    f = torch.as_tensor(densities.gaussian_mixture_batch(tr.points, range(64)),
                        device="cuda").contiguous()
    y = torch.zeros_like(f)
    # no slice set for 3 slices yet: refused, nothing built
    st = _call(L.vl1_plan_eval_noalloc, plan, "split", f[:4].contiguous(), y[:4].contiguous(), 3)
    assert st == _capi.VL1_E_UNSUPPORTED and "not prepared" in _capi.last_error()
    # 512 boxes x 64 RHS at 5 slices: 73 MB of staging, above the 64 MiB array
    _capi.check(_call(L.vl1_plan_eval, plan, "split", f, y, 5))
    st = _call(L.vl1_plan_eval_noalloc, plan, "split", f, y, 5)
    assert st == _capi.VL1_E_UNSUPPORTED and "scratch" in _capi.last_error()
    torch.cuda.synchronize()
