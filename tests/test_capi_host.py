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
    assert L.vl1_abi_version() == 4


def test_status_mapping():
    L = _capi.lib()
    with pytest.raises(ValueError):
        _capi.check(L.vl1_payload_create(C.byref(_capi.vp()), None, None))
    with pytest.raises(ValueError):
        _capi.check(L.vl1_duffy_build(None, None))
    assert "null" in _capi.last_error()


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
