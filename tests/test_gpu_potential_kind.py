"""Potential kind 2 (list1.py:101-109, 586-595, 1097-1143) on the GPU: the
plain route's list-1 sums plus each target's density times its exterior
mode normaliser, scaled as its own box, vs the oracle's term
(``oracle.list1_ref`` with ``exterior_nmlz``, pinned in
``tests/test_oracle_list1.py``): rel <= 1e-12 per RHS, canonical 3D and
per-level 2D scaling on adaptive trees, one density (the VALU kernel) and
20 densities (the FP64-MFMA kernel).  The 2D log displacement is refused.
"""

import numpy as np
import pytest
import torch

from oracle import list1_ref, orbit_ref
from tests.gpu_helpers import adaptive, rel_errors, synthetic_table_levels
from volumential_amd import densities
from volumential_amd.kernels import LaplaceKernel
from volumential_amd.wrangler import NearFieldWrangler

pytestmark = pytest.mark.gpu


def _tables(dim, q, n_tables, seed):
    tables = synthetic_table_levels(dim, q, n_tables=n_tables, seed=seed)
    rng = np.random.default_rng(seed + 100)
    for t in tables:
        t.kernel_exterior_normalizers = rng.standard_normal(t.n_q_points)
    return tables


@pytest.mark.parametrize("dim,q,mode,R", [(3, 3, "canonical", 1), (3, 3, "canonical", 20),
                                          (2, 4, "per_level", 1), (2, 4, "per_level", 20)])
def test_potential_kind_2_matches_oracle(cuda, dim, q, mode, R):
    print(f"case: potential kind 2, adaptive {dim}D q={q} {mode}, R={R}", flush=True)
    tr = adaptive(dim, q)
    n_tables = tr.nlevels if mode == "per_level" else 1
    tables = _tables(dim, q, n_tables, seed=31 + dim)
    w = NearFieldWrangler(tr, tables if n_tables > 1 else tables[0], potential_kind=2)
    f = densities.gaussian_mixture_batch(tr.points, range(40, 40 + R))
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    got = w.eval_direct_batched(*args, torch.as_tensor(f, device="cuda"))[0][0].cpu().numpy()
    assert not w._plans, "potential kind 2 evaluates by the plain route"
    meta = orbit_ref.case_metadata(dim, q)
    data = np.stack([t.compact_values() for t in tables])
    nm = np.stack([t.mode_normalizers for t in tables])
    ext = np.stack([t.kernel_exterior_normalizers for t in tables])
    ref = list1_ref.near_field_from_tree(tr, data, nm, meta, f, mode=mode, scaling_power=2,
                                         table_root_extent=tables[0].source_box_extent,
                                         own_csr=False, exterior_nmlz=ext)
    kind1 = list1_ref.near_field_from_tree(tr, data, nm, meta, f, mode=mode, scaling_power=2,
                                           table_root_extent=tables[0].source_box_extent,
                                           own_csr=False)
    assert rel_errors(ref, kind1)[0] > 1e-3  # the exterior term is not negligible here
    l2, linf = rel_errors(got, ref)
    print(f"vs oracle: rel L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= 1e-12 and linf <= 1e-12


def test_potential_kind_2_refuses_the_log_displacement(cuda):
    tr = adaptive(2, 4)
    tables = _tables(2, 4, 1, seed=7)
    w = NearFieldWrangler(tr, tables[0], potential_kind=2)
    f = densities.gaussian_mixture_batch(tr.points, [1])
    args = (tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists)
    with pytest.raises(NotImplementedError):
        w.eval_direct_batched(*args, torch.as_tensor(f, device="cuda"))
