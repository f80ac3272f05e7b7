"""HIP-graph capture of a plan evaluation (List1Plan.capture).

The captured graph (density staging + tile kernel, no allocation, no
memset) is replayed several times with new densities written between the
replays by ordinary eager work; every replay must match an eager evaluation
of the same densities to 1e-13 relative (same kernels and inputs; the
order of a tile's steps follows a racy L2 hint, so FP64 sums may differ in
the last bits), for the FP64, BF16 and int8 forms and for real and complex data.
"""

import numpy as np
import pytest
import torch

from tests.gpu_helpers import beam_forest, helmholtz_tables, synthetic_table_levels
from volumential_amd import densities, tree as vtree
from volumential_amd.wrangler import NearFieldWrangler

pytestmark = pytest.mark.gpu


def _replays(plan, make_f, dtype, form):
    f = make_f(0)
    out = torch.zeros_like(f)
    graph = plan.capture(f, out, form=form, n_slices=5)
    for it in range(1, 4):
        f.copy_(make_f(it))          # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        ref = plan.eval(f.clone(), form=form, n_slices=5)
        torch.cuda.synchronize()
        err = (out - ref).abs().max() / ref.abs().max()
        assert err <= 1e-13, f"replay {it} differs from the eager evaluation ({err})"
        assert out.abs().max() > 0


@pytest.mark.parametrize("form", ["f64", "split", "split_i8"])
def test_graph_replay_real(cuda, form):
    print(f"case: graph replay real 3D q=6 level 2 form={form}", flush=True)
    tr = vtree.uniform_tree(3, 6, 2)
    tables = synthetic_table_levels(3, 6, seed=3)
    w = NearFieldWrangler(tr, tables[0])
    plan = w.get_plan(w.target_kernels[0], tr.target_boxes, tr.neighbor_source_boxes_starts,
                      tr.neighbor_source_boxes_lists)

    def make_f(it):
        return torch.as_tensor(densities.gaussian_mixture_batch(tr.points, range(10 * it, 10 * it + 8)),
                               device="cuda").contiguous()

    _replays(plan, make_f, torch.float64, form)


@pytest.mark.parametrize("form", ["f64", "split", "split_i8"])
def test_graph_replay_complex(cuda, form):
    print(f"case: graph replay complex 2D forest form={form}", flush=True)
    fo = beam_forest([3], q=6, tol=1e-4, min_level=4, max_level=5)
    tables = helmholtz_tables(2, 6, 40.0, 4, fo.nlevels - 4)
    w = NearFieldWrangler(fo, tables)
    plan = w.get_plan(w.target_kernels[0], fo.target_boxes, fo.neighbor_source_boxes_starts,
                      fo.neighbor_source_boxes_lists)

    def make_f(it):
        return torch.as_tensor(np.stack([densities.helmholtz_beams(fo.points, 50 + 4 * it + r)
                                         for r in range(4)]), device="cuda").contiguous()

    _replays(plan, make_f, torch.complex128, form)


def test_graph_outlives_regrowth_and_eviction(cuda):
    """A graph stays valid after the plan's buffers are outgrown (more
    densities, another slice count, the other form) and after the wrangler
    drops the plan from its cache: the graph holds the plan, and outgrown
    buffers are retired rather than freed while the plan lives."""
    print("case: graph lifetime 3D q=4 level 2 (R=4 S=5 captured; R=128, S=4, f64 eager)",
          flush=True)
    tr = vtree.uniform_tree(3, 4, 2)
    tables = synthetic_table_levels(3, 4, seed=5)
    w = NearFieldWrangler(tr, tables[0])
    w.max_cached_plans = 1
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    plan = w.get_plan(w.target_kernels[0], *args)

    def dens(seeds):
        return torch.as_tensor(densities.gaussian_mixture_batch(tr.points, seeds),
                               device="cuda").contiguous()

    f = dens(range(4))
    out = torch.zeros_like(f)
    graph = plan.capture(f, out, form="split", n_slices=5)
    plan.eval(dens(range(100, 228)), form="split", n_slices=5)  # grows the staging buffers
    plan.eval(f.clone(), form="split", n_slices=4)             # a second slice set
    plan.eval(dens(range(300, 428)), form="f64")                # FP64 staging
    half = len(tr.target_boxes) // 2
    st = np.asarray(tr.neighbor_source_boxes_starts)
    w.get_plan(w.target_kernels[0], np.asarray(tr.target_boxes)[:half], st[:half + 1],
               np.asarray(tr.neighbor_source_boxes_lists)[:st[half]])  # evicts `plan`
    assert len(w._plans) == 1
    del plan
    for it in range(2):
        f.copy_(dens(range(10 + 4 * it, 14 + 4 * it)))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        ref = graph.plan.eval(f.clone(), form="split", n_slices=5)
        torch.cuda.synchronize()
        err = float((out - ref).abs().max() / ref.abs().max())
        print(f"replay {it}: {err:.2e}")
        assert err <= 1e-13
