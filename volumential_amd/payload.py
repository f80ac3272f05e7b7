"""List-1 payload preparation: pack table levels for the device.

Follows ``_prepare_table_data_and_entry_map`` and
``_build_fast_scalar_arithmetic_orbit_reconstruction``
(volumential/expansion_wrangler_fpnd.py:1451-1660, 1667-1726, 1805-2066)
for the scalar- and signed-arithmetic-orbit kinds, and the device upload/caching of
``_get_cached_nearfield_payload`` (2939-3158).
"""

from __future__ import annotations

import hashlib
from collections import OrderedDict

import numpy as np

from . import _capi, orbit


def _as_levels(tables):
    if not isinstance(tables, (list, tuple)):
        tables = [tables]
    if not tables:
        raise ValueError("table_levels cannot be empty")
    return list(tables)


def prepare_table_data_and_entry_map(table_levels):
    """Pack ``table_data_combined[lev, row]`` in compact-row order.

    Returns ``(table_data_combined, mode_nmlz_combined,
    exterior_mode_nmlz_combined, table_entry_ids, table_entry_scales,
    reconstruction_info)`` like the reference; the dense-map outputs are
    empty for the arithmetic kind.
    """
    levels = _as_levels(table_levels)
    t0 = levels[0]
    flags = [bool(getattr(t, "table_data_is_symmetry_reduced", False)) for t in levels]
    if any(f != flags[0] for f in flags[1:]):
        raise RuntimeError("mixed full/reduced near-field table storage across levels")
    if not flags[0]:
        raise RuntimeError("near-field table has not been built (no reduced storage)")
    for t in levels[1:]:
        if (t.dim, t.quad_order) != (t0.dim, t0.quad_order):
            raise RuntimeError("near-field levels disagree on dimension or order")
    meta = t0.arithmetic_metadata()
    for t in levels[1:]:
        if not np.array_equal(t.arithmetic_metadata().axis_sign_power, meta.axis_sign_power):
            raise RuntimeError("near-field levels disagree on the kernel's derivative axes")
    layout_ids = orbit.enumerate_representatives(meta)
    signed = bool(np.any(meta.axis_sign_power))
    dtype = np.dtype(t0.dtype)
    data = np.zeros((len(levels), layout_ids.size), dtype=dtype)
    nmlz = np.zeros((len(levels), t0.n_q_points), dtype=t0.mode_normalizers.dtype)
    ext = np.zeros((len(levels), t0.n_q_points), dtype=t0.kernel_exterior_normalizers.dtype)
    for lev, t in enumerate(levels):
        try:
            vals = t.compact_values()
        except RuntimeError as exc:
            raise RuntimeError(
                "near-field reconstruction layout references missing data"
            ) from exc
        if not np.all(np.isfinite(vals)):
            raise RuntimeError("near-field table contains non-finite entries")
        data[lev] = vals
        nmlz[lev] = t.mode_normalizers
        ext[lev] = t.kernel_exterior_normalizers
    info = {
        "kind": "signed-arithmetic-orbit" if signed else "scalar-arithmetic-orbit",
        "case_orbit_ranks": meta.case_orbit_ranks,
        "case_axis_perm": meta.case_axis_perm,
        "case_axis_sign": meta.case_axis_sign,
        "case_axis_group": meta.case_axis_group,
        "case_value_offsets": meta.case_value_offsets,
        "axis_sign_power": meta.axis_sign_power,
        "axis_direction_signs": np.zeros(t0.dim, dtype=np.int8),
        "direction_sign_axis": -1,
        "layout_entry_ids": layout_ids,
        "layout_entry_scales": np.ones(layout_ids.size, dtype=np.int8),
        "n_case_orbits": meta.n_orbits,
        "metadata_bytes": meta.metadata_bytes,
        "dense_arithmetic_layout_entry_count": meta.n_orbits * t0.n_pairs,
        "compact_value_entry_count": int(layout_ids.size),
        "representative_entry_count": int(layout_ids.size),
        "distinct_layout_entry_count": int(np.unique(layout_ids).size),
        "duplicate_layout_entry_count": 0,
        "unused_layout_entry_count": 0,
        "sign_convention_conflict_count": 0,
    }
    return (
        data,
        nmlz,
        ext,
        np.empty(0, dtype=np.int32),
        np.empty(0, dtype=dtype),
        info,
    )


class DevicePayload:
    """Owns a ``vl1_payload`` (device table data, descriptors, row maps)."""

    def __init__(self, dim, quad_order, table_data, mode_nmlz, case_indices, meta, stream=None):
        L = _capi.lib()
        table_data = np.ascontiguousarray(table_data)
        self.is_complex = np.iscomplexobj(table_data)
        if self.is_complex:
            table_data = np.ascontiguousarray(table_data.astype(np.complex128)).view(np.float64)
            mode_nmlz = np.ascontiguousarray(np.asarray(mode_nmlz, np.complex128)).view(np.float64)
        else:
            table_data = np.ascontiguousarray(table_data, dtype=np.float64)
            mode_nmlz = np.ascontiguousarray(np.asarray(mode_nmlz).real, dtype=np.float64)
        n_tables = int(np.asarray(self._rows(table_data, self.is_complex)).shape[0])
        n_entries = int(meta.n_compact_entries)
        case_indices = np.ascontiguousarray(case_indices, dtype=np.int64)
        self._keep = [
            table_data,
            mode_nmlz,
            case_indices,
            np.ascontiguousarray(meta.case_orbit_ranks, np.uint16),
            np.ascontiguousarray(meta.case_axis_perm, np.uint8),
            np.ascontiguousarray(meta.case_axis_sign, np.int8),
            np.ascontiguousarray(meta.case_axis_group, np.uint8),
            np.ascontiguousarray(meta.case_value_offsets, np.int32),
            np.ascontiguousarray(meta.axis_sign_power, np.uint8),
        ]
        desc = _capi.PayloadDesc(
            dim=int(dim),
            quad_order=int(quad_order),
            n_cases=int(meta.n_cases),
            n_case_orbits=int(meta.n_orbits),
            n_tables=n_tables,
            is_complex=int(self.is_complex),
            n_table_entries=n_entries,
            table_data=self._keep[0].ctypes.data,
            mode_nmlz=self._keep[1].ctypes.data,
            case_indices=self._keep[2].ctypes.data,
            case_orbit_ranks=self._keep[3].ctypes.data,
            case_axis_perm=self._keep[4].ctypes.data,
            case_axis_sign=self._keep[5].ctypes.data,
            case_axis_group=self._keep[6].ctypes.data,
            case_value_offsets=self._keep[7].ctypes.data,
            axis_sign_power=self._keep[8].ctypes.data,
        )
        handle = _capi.vp()
        _capi.check(
            L.vl1_payload_create(_capi.C.byref(handle), desc, _capi.stream_handle(stream)),
            "vl1_payload_create",
        )
        self.handle = handle
        self.dim, self.quad_order = int(dim), int(quad_order)
        self.n_tables, self.n_entries = n_tables, n_entries
        self.meta = meta

    @staticmethod
    def _rows(table_data, is_complex):
        t = np.asarray(table_data)
        return t.reshape(t.shape[0], -1) if t.ndim > 1 else t.reshape(1, -1)

    @property
    def device_bytes(self) -> int:
        return int(_capi.lib().vl1_payload_device_bytes(self.handle))

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            _capi.lib().vl1_payload_destroy(self.handle)
            self.handle = _capi.vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fingerprint(arr, sample_bytes=4096):
    """Content key of a table array (``expansion_wrangler_fpnd.py:99-112``)."""
    narr = np.ascontiguousarray(arr)
    raw = narr.view(np.uint8).reshape(-1)
    sample = bytes(raw) if raw.size <= 2 * sample_bytes else bytes(raw[:sample_bytes]) + bytes(
        raw[-sample_bytes:]
    )
    return (str(narr.dtype), narr.shape, hashlib.sha256(sample).hexdigest())


class PayloadCache:
    """LRU of device payloads (the reference keeps 16)."""

    def __init__(self, capacity=16):
        self.capacity = int(capacity)
        self._items: OrderedDict = OrderedDict()

    def get(self, key, factory):
        if key in self._items:
            self._items.move_to_end(key)
            return self._items[key]
        val = factory()
        self._items[key] = val
        while len(self._items) > self.capacity:
            _, old = self._items.popitem(last=False)
            old.close()
        return val
