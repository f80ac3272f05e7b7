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


class DenseDevicePayload:
    """Owns a ``vl1_payload`` of the reference's ``dense`` reconstruction
    kind (``list1.py:1018-1045``): the arrays of
    ``expansion_wrangler_fpnd.py:1887-2066, 2980-2995`` are copied to the
    device and every (level, case) block is expanded once there."""

    def __init__(self, dim, quad_order, table_data, mode_nmlz, case_indices, table_entry_ids,
                 table_entry_scales, mode_qpoint_map, mode_case_map, mode_case_scale,
                 n_cases, stream=None):
        L = _capi.lib()
        table_data = np.asarray(table_data)
        n = int(quad_order) ** int(dim)
        self.is_complex = bool(np.iscomplexobj(table_data) or np.iscomplexobj(table_entry_scales)
                               or np.iscomplexobj(mode_case_scale))
        vdt = np.complex128 if self.is_complex else np.float64

        def vals(a, shape):
            a = np.asarray(a)
            if not self.is_complex:
                a = a.real
            return np.ascontiguousarray(a.astype(vdt).reshape(shape))

        data = np.atleast_2d(table_data)
        data = vals(data.reshape(data.shape[0], -1), (data.shape[0], -1))
        n_tables, n_entries = data.shape
        nmlz = vals(np.asarray(mode_nmlz).reshape(n_tables, -1)[:, :n], (n_tables, n))
        ids = np.ascontiguousarray(np.asarray(table_entry_ids).reshape(-1), dtype=np.int32)
        if ids.size != int(n_cases) * n * n:
            raise ValueError(f"table_entry_ids must hold n_cases * n_q_points^2 = "
                             f"{int(n_cases) * n * n} entries, got {ids.size}")
        self._keep = [
            data.view(np.float64) if self.is_complex else data,
            nmlz.view(np.float64) if self.is_complex else nmlz,
            np.ascontiguousarray(case_indices, dtype=np.int64),
            ids,
            vals(table_entry_scales, (-1,)),
            np.ascontiguousarray(np.asarray(mode_qpoint_map).reshape(n, n), dtype=np.int32),
            np.ascontiguousarray(np.asarray(mode_case_map).reshape(n, int(n_cases)),
                                 dtype=np.int32),
            vals(mode_case_scale, (n, int(n_cases))),
        ]
        for k in (4, 7):
            if self.is_complex:
                self._keep[k] = self._keep[k].view(np.float64)
        if self._keep[4].size != ids.size * (2 if self.is_complex else 1):
            raise ValueError("table_entry_scales must match table_entry_ids")
        desc = _capi.DensePayloadDesc(
            dim=int(dim), quad_order=int(quad_order), n_cases=int(n_cases),
            n_tables=int(n_tables), is_complex=int(self.is_complex),
            n_table_entries=int(n_entries),
            table_data=self._keep[0].ctypes.data, mode_nmlz=self._keep[1].ctypes.data,
            case_indices=self._keep[2].ctypes.data, table_entry_ids=self._keep[3].ctypes.data,
            table_entry_scales=self._keep[4].ctypes.data,
            mode_qpoint_map=self._keep[5].ctypes.data, mode_case_map=self._keep[6].ctypes.data,
            mode_case_scale=self._keep[7].ctypes.data,
        )
        handle = _capi.vp()
        _capi.check(L.vl1_payload_create_dense(_capi.C.byref(handle), desc,
                                               _capi.stream_handle(stream)),
                    "vl1_payload_create_dense")
        self._keep = None  # copied by the library
        self.handle = handle
        self.dim, self.quad_order = int(dim), int(quad_order)
        self.n_tables, self.n_entries = int(n_tables), int(n_entries)
        self.has_missing = bool(L.vl1_payload_has_missing(handle))

    device_bytes = DevicePayload.device_bytes
    close = DevicePayload.close
    __del__ = DevicePayload.__del__


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
