"""GPU list-1 CSR producer (vl1_csr_builder_*) vs the host builder.

The host builder (tree._list1_csr) is checked against brute-force box
adjacency in tests/test_oracle_list1.py; here the GPU lists must equal it
exactly (same neighbours, same order) on uniform and adaptive 2:1 trees in
2D and 3D, and an unbalanced leaf set must be rejected (ValueError)."""

import numpy as np
import pytest

from tests.gpu_helpers import adaptive, beam_forest
from volumential_amd import tree as vtree

pytestmark = pytest.mark.gpu


def _check(tr):
    st, ls = vtree.list1_csr_gpu(tr.dim, tr.leaf_levels, tr.leaf_index)
    assert np.array_equal(st.cpu().numpy(), tr.neighbor_source_boxes_starts)
    assert np.array_equal(ls.cpu().numpy(), tr.neighbor_source_boxes_lists)


@pytest.mark.parametrize("dim,level", [(2, 4), (3, 3)])
def test_uniform(cuda, dim, level):
    print(f"case: csr uniform {dim}D level {level}", flush=True)
    _check(vtree.uniform_tree(dim, 2, level))


@pytest.mark.parametrize("dim,max_level", [(2, 6), (3, 4)])
def test_adaptive(cuda, dim, max_level):
    print(f"case: csr adaptive {dim}D max level {max_level}", flush=True)
    _check(adaptive(dim, 2, max_level=max_level))


def test_beam_trees(cuda):
    print("case: csr beam trees", flush=True)
    fo = beam_forest([0, 1], q=4, tol=1e-5, min_level=3, max_level=7)
    for t in fo.trees:
        _check(t)


def test_unbalanced_rejected(cuda):
    print("case: csr unbalanced leaf set", flush=True)
    # level-1 quadrant next to level-3 leaves: not 2:1 balanced
    leaves = [(1, (0, 0)), (1, (1, 0)), (1, (0, 1))]
    leaves += [(3, (i, j)) for i in range(4, 8) for j in range(4, 8)]
    leaves += [(2, (2, 3)), (2, (3, 2))]
    leaves = [lv for lv in leaves if not (lv[0] == 2 and lv[1] == (3, 2))]
    lv = np.array([l for l, _ in leaves], np.uint8)
    ix = np.array([i for _, i in leaves], np.int32)
    with pytest.raises(ValueError):
        vtree.list1_csr_gpu(2, lv, ix)
