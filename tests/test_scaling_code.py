"""Custom single-table scaling policy (caller code strings), CPU side.

* the evaluator reproduces, per source level, the reference's own inferred
  rules written as code strings (list1.py:278-355: Laplace 2D/3D
  (h/H)^2 with the 2D log displacement, constant 3D (h/H)^3, derivatives
  h/H) -- the closed forms of the canonical policy;
* unsafe or unsupported expressions are refused, and so are displacements
  that are not a multiple of mode_nmlz[table_lev, sid];
* the oracle's custom branch, given the canonical policy's per-level values,
  equals its canonical branch on an adaptive 2D tree (several source
  levels, log term) -- a wrong level index or displacement factor fails it;
* NearFieldFromCSR takes the strings as the reference does (both needed,
  inference switched off, "custom_single_table" policy).
"""

import math

import numpy as np
import pytest

from oracle import list1_ref, orbit_ref
from tests.gpu_helpers import adaptive, synthetic_table_levels
from volumential_amd import densities
from volumential_amd.kernels import HelmholtzKernel, LaplaceKernel
from volumential_amd.list1 import NearFieldFromCSR
from volumential_amd.scaling_code import MAX_LEVELS, level_arrays

# the reference's inferred rules as strings (list1.py:278-355), BOX = source box
LAP = "BOX_extent * BOX_extent / (table_root_extent * table_root_extent)"
LAP2D_DISP = ("-0.5 / PI * scaling * log(BOX_extent / table_root_extent) * "
              "mode_nmlz[table_lev, sid]").replace("PI", str(math.pi))
CST3 = ("BOX_extent * BOX_extent * BOX_extent / "
        "(table_root_extent * table_root_extent * table_root_extent)")
GRAD = "BOX_extent / table_root_extent"


@pytest.mark.parametrize("root,H", [(2.0, 2.0), (1.0, 0.25), (3.0, 1.5)])
def test_reference_rules_per_level(root, H):
    lev = np.arange(MAX_LEVELS)
    h = root / 2.0**lev
    s, d = level_arrays(LAP, LAP2D_DISP, root, H)
    assert np.array_equal(s, h * h / (H * H))
    np.testing.assert_allclose(d, -0.5 / np.pi * (h / H) ** 2 * np.log(h / H), rtol=1e-15,
                               atol=1e-300)
    s3, d3 = level_arrays(CST3, "0.0", root, H)
    np.testing.assert_allclose(s3, (h / H) ** 3, rtol=1e-15)
    assert not d3.any()
    sg, _ = level_arrays(GRAD, "0.0", root, H)
    np.testing.assert_allclose(sg, h / H, rtol=1e-15)


def test_level_and_conditional_names():
    s, d = level_arrays("1.0 + sbox_level if sbox_level < 3 else 0.5",
                        "(2.0 - sbox_level) * mode_nmlz[table_lev, sid]", 2.0, 2.0)
    assert list(s[:5]) == [1.0, 2.0, 3.0, 0.5, 0.5]
    assert list(d[:3]) == [2.0, 1.0, 0.0]


@pytest.mark.parametrize("code", ["__import__('os').getcwd()", "sbox_extent.real", "foo * 2",
                                  "(lambda: 1)()", "'a'", "open('x')", "[1][0]"])
def test_refuses_unsupported_expressions(code):
    with pytest.raises(ValueError):
        level_arrays(code, "0.0", 2.0, 2.0)


def test_refuses_unsupported_displacements():
    with pytest.raises(NotImplementedError):
        level_arrays("1.0", "1.0 + mode_nmlz[table_lev, sid]", 2.0, 2.0)
    with pytest.raises(NotImplementedError):
        level_arrays("1.0", "mode_nmlz[table_lev, sid] ** 2", 2.0, 2.0)
    with pytest.raises(ValueError):  # the normaliser belongs to the displacement
        level_arrays("mode_nmlz[table_lev, sid]", "0.0", 2.0, 2.0)
    with pytest.raises(ValueError):  # log(0) at every level
        level_arrays("log(0.0)", "0.0", 2.0, 2.0)


def test_oracle_custom_branch_equals_canonical():
    tr = adaptive(2, 3)
    assert len(np.unique(tr.box_levels[tr.neighbor_source_boxes_lists])) >= 3
    tables = synthetic_table_levels(2, 3, seed=11)
    meta = orbit_ref.case_metadata(2, 3)
    f = densities.gaussian_mixture_batch(tr.points, [1, 2])
    kw = dict(table_root_extent=2.0, own_csr=False)
    canon = list1_ref.near_field_from_tree(
        tr, tables[0].compact_values()[None], tables[0].mode_normalizers[None], meta, f,
        mode="canonical", scaling_power=2, log_displacement=True, **kw)
    s, d = level_arrays(LAP, LAP2D_DISP, tr.root_extent, 2.0)
    cust = list1_ref.near_field_from_tree(
        tr, tables[0].compact_values()[None], tables[0].mode_normalizers[None], meta, f,
        mode="custom", level_scaling=s, level_disp=d, **kw)
    np.testing.assert_allclose(cust, canon, rtol=1e-13, atol=1e-15 * np.abs(canon).max())


def test_operator_takes_code_strings():
    shapes = {"n_tables": 1, "n_q_points": 9, "n_table_entries": 10, "n_cases": 33}
    nf = NearFieldFromCSR(LaplaceKernel(2), shapes, kernel_scaling_code=LAP,
                          kernel_displacement_code=LAP2D_DISP)
    pol = nf.get_kernel_scaling_policy()
    assert pol.mode == "custom_single_table" and not nf.extra_kwargs["infer_kernel_scaling"]
    s, d = nf.custom_level_arrays(2.0, 2.0)
    assert s[1] == 0.25 and d[0] == 0.0
    # a Helmholtz kernel with its own rules is accepted (no inference needed)
    NearFieldFromCSR(HelmholtzKernel(2, 1.0), shapes, kernel_scaling_code="1.0",
                     kernel_displacement_code="0.0")
    with pytest.raises(ValueError):
        NearFieldFromCSR(LaplaceKernel(2), shapes, kernel_scaling_code=LAP)
    with pytest.raises(ValueError):
        NearFieldFromCSR(LaplaceKernel(2), shapes, kernel_scaling_code="foo",
                         kernel_displacement_code="0.0")
