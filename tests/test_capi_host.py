"""The C-ABI library loads, exports every declared symbol, and its host twin
of the device compact-row function agrees with the orbit enumeration."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from volumential_amd import _capi, orbit

HEADER = Path(__file__).resolve().parents[1] / "include" / "vl1.h"


def test_header_symbols_exported():
    L = _capi.lib()
    declared = set(re.findall(r"^\s*(?:const\s+)?\w+[\s\*]+(vl1_\w+)\s*\(", HEADER.read_text(), re.M))
    assert declared, "no declarations parsed"
    assert declared == set(_capi.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(L, name), name
    assert L.vl1_abi_version() == 11


def test_status_mapping():
    L = _capi.lib()
    with pytest.raises(ValueError):
        _capi.check(L.vl1_payload_create(C.byref(_capi.vp()), None, None))
    with pytest.raises(ValueError):
        _capi.check(L.vl1_duffy_build(None, None))
    assert "null" in _capi.last_error()
    with pytest.raises(ValueError):
        _capi.check(L.vl1_plan_set_variant(None, 1, 5, 4, 0))
    assert "null plan" in _capi.last_error()


@pytest.mark.parametrize("dim,q", [(2, 4), (2, 6), (3, 3), (3, 6)])
def test_host_rows_match_enumeration(dim, q):
    L = _capi.lib()
    m = orbit.build_case_metadata(dim, q)
    n = q**dim
    rng = np.random.default_rng(dim * 10 + q)
    for c in rng.choice(m.n_cases, size=min(12, m.n_cases), replace=False):
        s = rng.integers(0, n, 2000).astype(np.int32)
        t = rng.integers(0, n, 2000).astype(np.int32)
        out = np.empty(s.size, np.int64)
        perm = np.ascontiguousarray(m.case_axis_perm[c])
        sign = np.ascontiguousarray(m.case_axis_sign[c])
        grp = np.ascontiguousarray(m.case_axis_group[c])
        _capi.check(
            L.vl1_orbit_rows_host(dim, q, perm.ctypes.data, sign.ctypes.data, grp.ctypes.data,
                                  s.ctypes.data, t.ctypes.data, s.size, out.ctypes.data)
        )
        off = m.case_value_offsets[m.case_orbit_ranks[c]]
        assert np.array_equal(out + off, orbit.entry_rows(m, np.full(s.size, c), s, t))


@pytest.mark.parametrize("dim,q,fixed,power", [(2, 4, (0,), (1, 0)), (2, 5, (1,), (0, 1)),
                                               (3, 3, (2,), (0, 0, 1)),
                                               (3, 4, (0, 1), (1, 1, 0))])
def test_host_rows_and_signs_signed_kernels(dim, q, fixed, power):
    """The device compact-row and entry-sign functions (host twins) agree
    with the vectorised orbit arithmetic for axis-derivative metadata."""
    L = _capi.lib()
    m = orbit.build_case_metadata(dim, q, fixed, power)
    n = q**dim
    sp = np.ascontiguousarray(m.axis_sign_power, dtype=np.uint8)
    rng = np.random.default_rng(dim * 100 + q)
    for c in rng.choice(m.n_cases, size=min(10, m.n_cases), replace=False):
        s = rng.integers(0, n, 3000).astype(np.int32)
        t = rng.integers(0, n, 3000).astype(np.int32)
        perm = np.ascontiguousarray(m.case_axis_perm[c])
        sign = np.ascontiguousarray(m.case_axis_sign[c])
        grp = np.ascontiguousarray(m.case_axis_group[c])
        rows = np.empty(s.size, np.int64)
        sg = np.empty(s.size, np.int8)
        _capi.check(L.vl1_orbit_rows_host(dim, q, perm.ctypes.data, sign.ctypes.data,
                                          grp.ctypes.data, s.ctypes.data, t.ctypes.data,
                                          s.size, rows.ctypes.data))
        _capi.check(L.vl1_orbit_signs_host(dim, q, perm.ctypes.data, sign.ctypes.data,
                                           sp.ctypes.data, s.ctypes.data, t.ctypes.data,
                                           s.size, sg.ctypes.data))
        cc = np.full(s.size, c)
        off = m.case_value_offsets[m.case_orbit_ranks[c]]
        assert np.array_equal(rows + off, orbit.entry_rows(m, cc, s, t))
        assert np.array_equal(sg, orbit.entry_signs(m, cc, s, t))


def test_split_correction_rejects_bad_derivative_wrapper():
    """vl1_split_correction validates the derivative wrapper before any
    device work (the dummy device pointers are never touched)."""
    L = _capi.lib()
    dummy = 256
    base = dict(dim=2, split_order=1, series_nmax=5, n_tgt_boxes=1, n_rhs=1, max_box_points=4,
                k_re=3.0, k_im=0.0, n_points=16, target_boxes=dummy, nbr_starts=dummy,
                nbr_lists=dummy, box_starts=dummy, box_counts=dummy, points=dummy,
                strengths=dummy, ld_strengths=32, result=dummy, ld_result=32)
    for kind, axis, what in ((3, 0, "deriv_kind"), (-1, 0, "deriv_kind"),
                             (_capi.SPLIT_DERIV_TARGET, 2, "deriv_axis"),
                             (_capi.SPLIT_DERIV_SOURCE, -1, "deriv_axis")):
        args = _capi.SplitArgs(**base, deriv_kind=kind, deriv_axis=axis)
        with pytest.raises(ValueError):
            _capi.check(L.vl1_split_correction(C.byref(args), None), "vl1_split_correction")
        assert what in _capi.last_error()


def test_split_correction_host_validates_derivative_options():
    """HelmholtzSplitCorrection rejects a bad wrapper before any device use."""
    from volumential_amd import tree as vtree
    from volumential_amd.split import HelmholtzSplitCorrection

    tr = vtree.uniform_tree(2, 3, 2)
    with pytest.raises(ValueError, match="deriv_kind"):
        HelmholtzSplitCorrection(tr, 3.0, deriv_kind=7)
    with pytest.raises(ValueError, match="deriv_axis"):
        HelmholtzSplitCorrection(tr, 3.0, deriv_kind=_capi.SPLIT_DERIV_TARGET, deriv_axis=2)
