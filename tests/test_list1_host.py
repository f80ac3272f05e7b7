"""Host-side logic of NearFieldFromCSR (CPU, no GPU calls).

* the case-encoding drift warning (list1.py:468-530): silent on an exact
  dyadic tree, and on a perturbed box centre it reports the largest drift
  with the pair and dimension that produced it;
* the kernel scaling policy (list1.py:361-437): canonical for Laplace,
  constant and axis derivatives of Laplace (powers 2, d and 1; the 2D log
  term only for Laplace), fixed single table / per-level tables otherwise,
  and Helmholtz with inferred scaling rejected (list1.py:183-195).
"""

import logging

import numpy as np
import pytest

from volumential_amd import tree as vtree
from volumential_amd.kernels import (AxisSourceDerivative, AxisTargetDerivative,
                                     ConstantKernel, HelmholtzKernel, LaplaceKernel)
from volumential_amd.list1 import NearFieldFromCSR, warn_case_encoding_drift


def _drift(tr, tol):
    return warn_case_encoding_drift(tr.dim, tr.box_centers, tr.box_levels, tr.root_extent,
                                    tr.target_boxes, tr.neighbor_source_boxes_starts,
                                    tr.neighbor_source_boxes_lists, 6.0, tol)


def test_drift_silent_on_exact_tree(caplog):
    tr = vtree.uniform_tree(3, 2, 3)
    with caplog.at_level(logging.WARNING):
        d = _drift(tr, 1e-12)
    assert d <= 1e-12
    assert not caplog.records
    assert _drift(tr, None) is None


def test_drift_reports_worst_pair(caplog):
    tr = vtree.uniform_tree(2, 2, 3)
    h = tr.root_extent / 8
    tr.box_centers = tr.box_centers.copy()
    tr.box_centers[1, 10] += 3e-7 * h  # box 10, dim 1: drift 4 * 3e-7
    with caplog.at_level(logging.WARNING):
        d = _drift(tr, 1e-8)
    assert abs(d - 1.2e-6) < 1e-9
    msg = caplog.records[-1].getMessage()
    assert "dim 1" in msg and ("target box 10" in msg or "source box 10" in msg)


@pytest.mark.parametrize("kernel,power,logd", [
    (LaplaceKernel(2), 2, True), (LaplaceKernel(3), 2, False), (ConstantKernel(3), 3, False),
    (AxisTargetDerivative(0, LaplaceKernel(2)), 1, False),
    (AxisSourceDerivative(2, LaplaceKernel(3)), 1, False)])
def test_inferred_scaling_policy(kernel, power, logd):
    nf = NearFieldFromCSR(kernel, {"n_tables": 1, "n_q_points": 9 if kernel.dim == 2 else 27,
                                   "n_table_entries": 10})
    pol = nf.get_kernel_scaling_policy()
    assert pol.mode == "canonical_single_table"
    assert (pol.scaling_power, pol.log_displacement) == (power, logd)


def test_policy_fallbacks_and_helmholtz_rejection():
    shapes = {"n_tables": 1, "n_q_points": 36, "n_table_entries": 10}
    assert NearFieldFromCSR(HelmholtzKernel(2, 5.0), shapes).get_kernel_scaling_policy().mode == \
        "fixed_single_table"
    shapes3 = dict(shapes, n_tables=3)
    assert NearFieldFromCSR(HelmholtzKernel(2, 5.0), shapes3).get_kernel_scaling_policy().mode == \
        "per_level_tables"
    with pytest.raises(RuntimeError, match="infer_kernel_scaling is unsupported"):
        NearFieldFromCSR(HelmholtzKernel(2, 5.0), shapes, infer_kernel_scaling=True)


def test_split_support_with_derivative_wrappers():
    """``_split_target_kernel_support_status`` / ``_split_base_table_support_status``
    (expansion_wrangler_fpnd.py:3964-4073) with axis derivative wrappers:
    the tables must hold the same wrapper around the Laplace kernel; deeper
    chains are rejected here (no nested derivative tables)."""
    from volumential_amd import _capi, tree as vtree
    from volumential_amd.kernels import (AxisSourceDerivative, AxisTargetDerivative,
                                         HelmholtzKernel, LaplaceKernel)
    from volumential_amd.table import NearFieldInteractionTable
    from volumential_amd.wrangler import NearFieldWrangler, split_wrapper_chain

    tr = vtree.uniform_tree(2, 3, 2)
    helm = HelmholtzKernel(2, 3.0)
    cases = [
        (AxisTargetDerivative(1, helm), AxisTargetDerivative(1, LaplaceKernel(2)), True),
        (AxisSourceDerivative(0, helm), AxisSourceDerivative(0, LaplaceKernel(2)), True),
        (AxisSourceDerivative(0, helm), AxisTargetDerivative(0, LaplaceKernel(2)), False),
        (AxisTargetDerivative(0, helm), LaplaceKernel(2), False),
        (AxisTargetDerivative(0, AxisTargetDerivative(1, helm)), LaplaceKernel(2), False),
        (AxisTargetDerivative(0, LaplaceKernel(2)), AxisTargetDerivative(0, LaplaceKernel(2)), False),
    ]
    for knl, tk, ok in cases:
        t = NearFieldInteractionTable(3, 2, tk, source_box_extent=2.0)
        w = NearFieldWrangler(tr, {repr(knl): [t]}, target_kernels=[knl])
        assert w.helmholtz_split == ok, (knl, tk, w._split_support_status())
        if ok:
            assert repr(w._policy_kernel(knl)) == repr(tk)
            with pytest.raises(RuntimeError, match="not supported"):
                NearFieldWrangler(tr, {repr(knl): [NearFieldInteractionTable(
                    3, 2, LaplaceKernel(2), source_box_extent=2.0)]}, target_kernels=[knl],
                    helmholtz_split=True)
    assert split_wrapper_chain(helm) == (helm, _capi.SPLIT_DERIV_NONE, 0)
    assert split_wrapper_chain(AxisTargetDerivative(1, helm)) == (helm, _capi.SPLIT_DERIV_TARGET, 1)
    assert split_wrapper_chain(AxisSourceDerivative(0, helm)) == (helm, _capi.SPLIT_DERIV_SOURCE, 0)


def test_auto_split_imaginary_coverage_warning(caplog):
    """"auto" with |Im k| h above rho_imag_split_max (default 8) warns and
    keeps the split (expansion_wrangler_fpnd.py:2550-2568)."""
    import logging

    from volumential_amd import tree as vtree
    from volumential_amd.kernels import HelmholtzKernel, LaplaceKernel
    from volumential_amd.table import NearFieldInteractionTable
    from volumential_amd.wrangler import NearFieldWrangler

    tr = vtree.uniform_tree(2, 3, 2)  # coarsest active box h = 0.5
    lap = NearFieldInteractionTable(3, 2, LaplaceKernel(2), source_box_extent=2.0)
    yuk = HelmholtzKernel(2, 20.0j)  # rho_imag = 10
    with caplog.at_level(logging.WARNING, logger="volumential_amd.wrangler"):
        w = NearFieldWrangler(tr, {repr(yuk): [lap]}, target_kernels=[yuk],
                              helmholtz_split_order="auto")
    assert w.helmholtz_split and "exceeds configured split coverage" in caplog.text
    caplog.clear()
    with caplog.at_level(logging.WARNING, logger="volumential_amd.wrangler"):
        NearFieldWrangler(tr, {repr(yuk): [lap]}, target_kernels=[yuk],
                          helmholtz_split_order="auto",
                          helmholtz_split_auto_config={"rho_imag_split_max": 12.0})
    assert "split coverage" not in caplog.text
    with caplog.at_level(logging.WARNING, logger="volumential_amd.wrangler"):
        NearFieldWrangler(tr, {repr(yuk): [lap]}, target_kernels=[yuk],
                          helmholtz_split_order="auto",
                          helmholtz_split_auto_config={"disable_split_if_outside_coverage": True})
    assert "Keeping split mode enabled" in caplog.text


def test_split_order_options():
    """helmholtz_split_order / _smooth_quad_order resolution
    (expansion_wrangler_fpnd.py:2520-2568, 2649-2747): integers >= 1,
    "auto" through the rho ladder (and the smooth order rule), clamped to 1
    with several output kernels, kept above 1 under a derivative wrapper;
    the smooth order defaults to q outside "auto"; order > 1 asks for the
    density values (src_func) in the correction."""
    from volumential_amd import tree as vtree
    from volumential_amd.kernels import AxisTargetDerivative, HelmholtzKernel, LaplaceKernel
    from volumential_amd.table import NearFieldInteractionTable
    from volumential_amd.wrangler import NearFieldWrangler

    tr = vtree.uniform_tree(2, 3, 2)  # coarsest active box h = 0.5
    lap = NearFieldInteractionTable(3, 2, LaplaceKernel(2), source_box_extent=2.0)
    helm = HelmholtzKernel(2, 7.0)

    def make(knls=(helm,), **kw):
        return NearFieldWrangler(tr, {repr(k): [lap] for k in knls}, target_kernels=list(knls),
                                 **kw)

    w = make(helmholtz_split_order=3)
    assert (w.helmholtz_split_order, w.helmholtz_split_smooth_quad_order) == (3, 3)
    w = make(helmholtz_split_order="auto")
    assert w.helmholtz_split_order == 6                 # rho = 3.5 in (2, 4]
    assert w.helmholtz_split_smooth_quad_order == 3 + 5 + 1  # + (m-1), + ceil(0.5 (3.5 - 3))
    w = make(helmholtz_split_order=2, helmholtz_split_smooth_quad_order=5)
    assert w.helmholtz_split_smooth_quad_order == 5
    with pytest.raises(ValueError):
        make(helmholtz_split_order=0)
    two = (helm, HelmholtzKernel(2, 3.0))
    assert make(two, helmholtz_split_order=4).helmholtz_split_order == 1
    dk = AxisTargetDerivative(0, helm)
    dt = NearFieldInteractionTable(3, 2, AxisTargetDerivative(0, LaplaceKernel(2)),
                                   source_box_extent=2.0)
    assert NearFieldWrangler(tr, {repr(dk): [dt]}, target_kernels=[dk]).helmholtz_split
    # order > 1 under a derivative wrapper (the wrapped terms, 5018-5051)
    assert NearFieldWrangler(tr, {repr(dk): [dt]}, target_kernels=[dk],
                             helmholtz_split_order=2).helmholtz_split_order == 2
    w = make(helmholtz_split_order=2)
    with pytest.raises(RuntimeError, match="src_func"):
        w.eval_direct_helmholtz_split_correction(
            tr.target_boxes, tr.neighbor_source_boxes_starts, tr.neighbor_source_boxes_lists,
            np.zeros(tr.n_points, complex))
    # no split: the order is irrelevant (1)
    assert make(helmholtz_split=False, helmholtz_split_order=5).helmholtz_split_order == 1
    # the 2D log term's beta mode (3709-3786): "p2p" by default, "table", else ValueError
    assert make(helmholtz_split_order=2).helmholtz_split_beta_mode == "p2p"
    cfg = {"power_log_single_table_beta_mode": " Table "}
    assert make(helmholtz_split_order=2,
                helmholtz_split_auto_config=cfg).helmholtz_split_beta_mode == "table"
    with pytest.raises(ValueError, match="beta_mode"):
        make(helmholtz_split_order=2,
             helmholtz_split_auto_config={"power_log_single_table_beta_mode": "quad"})
    # the legacy subtraction (3573-3578, 3806-3822): order 1 only
    w = make(helmholtz_split_order1_legacy_subtraction=True)
    assert w._legacy()
    assert not make(helmholtz_split_order=2,
                    helmholtz_split_order1_legacy_subtraction=True)._legacy()
    # q = 3 and 5 Gauss nodes share the midpoint (5125-5138); 4 does not
    with pytest.raises(ValueError, match="shares nodes"):
        make(helmholtz_split_order1_legacy_subtraction=True, helmholtz_split_smooth_quad_order=5)
    assert make(helmholtz_split_order1_legacy_subtraction=True,
                helmholtz_split_smooth_quad_order=4)._legacy()
    assert NearFieldWrangler(tr, {repr(dk): [dt]}, target_kernels=[dk],
                             helmholtz_split_order1_legacy_subtraction=True)._legacy()
