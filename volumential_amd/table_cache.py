"""Near-field table cache in the reference's on-disk format.

Tables built here can be stored in, and tables built by the reference read
from, the reference's SQLite table cache (schema 2.1.0,
``table_manager.py:62-73, 665-740``).  One row per (dim, kernel_type,
q_order, source_box_level) holds the geometry columns and a ``payload``
blob, an ``np.savez`` archive with the arrays
``q_points``, ``mode_normalizers``, ``kernel_exterior_normalizers``,
``interaction_case_vecs``, ``case_indices``,
``table_data_is_symmetry_reduced`` and either ``reduced_entry_ids`` +
``reduced_data`` or a dense ``data`` array (``table_manager.py:182-228``).
Payloads are read with ``allow_pickle=False``.
"""

from __future__ import annotations

import sqlite3
from io import BytesIO

import numpy as np

from .table import NearFieldInteractionTable

SCHEMA_VERSION = "2.1.0"
BUILD_METHOD = "DuffyRadial"

_CACHE_COLUMNS = (
    "dim", "kernel_type", "q_order", "source_box_level", "n_q_points", "quad_order", "n_pairs",
    "source_box_extent", "source_box_level_stored", "case_encoding_base", "case_encoding_shift",
    "build_method", "kernel_type_cached", "payload",
)


def table_payload(table: NearFieldInteractionTable) -> bytes:
    """The npz payload of a built table (orbit-reduced storage)."""
    ids, vals = table.get_reduced_table_data()
    payload = {
        "q_points": np.asarray(table.q_points),
        "mode_normalizers": np.asarray(table.mode_normalizers),
        "kernel_exterior_normalizers": np.asarray(table.kernel_exterior_normalizers),
        "interaction_case_vecs": np.asarray(table.interaction_case_vecs),
        "case_indices": np.asarray(table.case_indices),
        "table_data_is_symmetry_reduced": np.array([1], dtype=np.int8),
        "reduced_entry_ids": np.asarray(ids, dtype=np.int64),
        "reduced_data": np.asarray(vals),
    }
    with BytesIO() as fh:
        np.savez(fh, **payload)
        return fh.getvalue()


def table_from_payload(blob: bytes, quad_order, dim, kernel, source_box_extent,
                       source_box_level=None) -> NearFieldInteractionTable:
    """A table from an npz payload; reduced or dense storage.  The gallery
    stored with the payload must be this package's (same case order)."""
    with BytesIO(blob) as fh:
        with np.load(fh, allow_pickle=False) as z:
            arrs = {k: z[k] for k in z.files}
    t = NearFieldInteractionTable(quad_order, dim, kernel, source_box_extent=source_box_extent,
                                  source_box_level=source_box_level)
    if not np.allclose(arrs["q_points"], t.q_points, rtol=0, atol=1e-14 * source_box_extent):
        raise ValueError("payload q_points do not match this table's nodes")
    if not np.array_equal(np.asarray(arrs["interaction_case_vecs"]),
                          np.asarray(t.interaction_case_vecs)) or \
            not np.array_equal(arrs["case_indices"], t.case_indices):
        raise ValueError("payload gallery (case vectors / indices) differs from this package's")
    if "reduced_entry_ids" in arrs and "reduced_data" in arrs:
        t.set_reduced_table_data(arrs["reduced_entry_ids"], arrs["reduced_data"])
    elif "data" in arrs:
        data = np.asarray(arrs["data"])
        ids = np.flatnonzero(np.isfinite(data))
        t.set_reduced_table_data(ids, data[ids])
    else:
        raise KeyError("payload is missing table data arrays")
    t.mode_normalizers[:] = arrs["mode_normalizers"]
    t.kernel_exterior_normalizers[:] = arrs["kernel_exterior_normalizers"]
    t.is_built = True
    return t


def _case_encoding(table):
    comps = {c for v in table.interaction_case_vecs for c in v}
    lo, hi = min(comps), max(comps)
    return int(hi - lo + 1), int(-lo)


class TableCache:
    """SQLite table cache, reference schema 2.1.0 (readable by either side)."""

    def __init__(self, filename, root_extent=None):
        self.filename = str(filename)
        self.db = sqlite3.connect(self.filename)
        self.db.row_factory = sqlite3.Row
        self._init_schema()
        stored = self._meta("root_extent")
        if root_extent is not None:
            if stored is None:
                self._set_meta("root_extent", float(root_extent))
            elif abs(float(stored) - float(root_extent)) > 1e-15:
                raise RuntimeError(f"The table cache file {self.filename} was built with "
                                   f"root_extent = {stored}, which is different from the "
                                   f"requested value {root_extent}")
        self.root_extent = float(root_extent if root_extent is not None else (stored or 2.0))

    def _init_schema(self):
        self.db.execute("""
            CREATE TABLE IF NOT EXISTS nearfield_cache (
                dim INTEGER NOT NULL, kernel_type TEXT NOT NULL, q_order INTEGER NOT NULL,
                source_box_level INTEGER NOT NULL,
                n_q_points INTEGER NOT NULL, quad_order INTEGER NOT NULL, n_pairs INTEGER NOT NULL,
                source_box_extent REAL NOT NULL, source_box_level_stored INTEGER NOT NULL,
                case_encoding_base INTEGER, case_encoding_shift INTEGER,
                build_method TEXT, kernel_type_cached TEXT, payload BLOB,
                PRIMARY KEY (dim, kernel_type, q_order, source_box_level))""")
        self.db.execute("""
            CREATE TABLE IF NOT EXISTS nearfield_cache_kwargs (
                dim INTEGER NOT NULL, kernel_type TEXT NOT NULL, q_order INTEGER NOT NULL,
                source_box_level INTEGER NOT NULL, key TEXT NOT NULL,
                value_type TEXT NOT NULL, value_text TEXT NOT NULL,
                PRIMARY KEY (dim, kernel_type, q_order, source_box_level, key),
                FOREIGN KEY (dim, kernel_type, q_order, source_box_level)
                    REFERENCES nearfield_cache (dim, kernel_type, q_order, source_box_level)
                    ON DELETE CASCADE)""")
        self.db.execute("""
            CREATE TABLE IF NOT EXISTS nearfield_cache_meta (
                key TEXT PRIMARY KEY, value_type TEXT NOT NULL, value_text TEXT NOT NULL)""")
        ver = self._meta("schema_version")
        if ver is None:
            self._set_meta("schema_version", SCHEMA_VERSION)
        elif ver != SCHEMA_VERSION:
            raise RuntimeError(f"table cache schema {ver} is not {SCHEMA_VERSION}")
        self.db.commit()

    # scalar meta values as (value_type, value_text), table_manager.py:268-305
    def _meta(self, key):
        row = self.db.execute("SELECT value_type, value_text FROM nearfield_cache_meta WHERE key=?",
                              (key,)).fetchone()
        if row is None:
            return None
        ty, txt = row["value_type"], row["value_text"]
        return {"int": int, "float": float, "str": str, "complex": complex,
                "bool": lambda v: bool(int(v))}[ty](txt)

    def _set_meta(self, key, value):
        if isinstance(value, bool):
            ty, txt = "bool", int(value)
        elif isinstance(value, int):
            ty, txt = "int", value
        elif isinstance(value, float):
            ty, txt = "float", repr(value)
        else:
            ty, txt = "str", str(value)
        self.db.execute("INSERT OR REPLACE INTO nearfield_cache_meta (key, value_type, value_text) "
                        "VALUES (?, ?, ?)", (key, ty, str(txt)))

    def source_box_extent(self, source_box_level):
        return self.root_extent * 2.0 ** (-int(source_box_level))

    def save(self, table, kernel_type, source_box_level):
        lvl = int(source_box_level)
        if abs(table.source_box_extent - self.source_box_extent(lvl)) > 1e-15:
            raise ValueError("table extent does not match root_extent * 2^-level")
        base, shift = _case_encoding(table)
        self.db.execute(
            f"INSERT INTO nearfield_cache ({', '.join(_CACHE_COLUMNS)}) VALUES "
            f"({', '.join('?' * len(_CACHE_COLUMNS))}) "
            "ON CONFLICT(dim, kernel_type, q_order, source_box_level) DO UPDATE SET "
            + ", ".join(f"{c}=excluded.{c}" for c in _CACHE_COLUMNS[4:]),
            (table.dim, str(kernel_type), table.quad_order, lvl, table.n_q_points, table.quad_order,
             table.n_pairs, table.source_box_extent, lvl, base, shift, BUILD_METHOD,
             repr(table.integral_knl), table_payload(table)))
        self.db.commit()

    def load(self, dim, kernel_type, q_order, source_box_level, kernel):
        row = self.db.execute(
            "SELECT * FROM nearfield_cache WHERE dim=? AND kernel_type=? AND q_order=? "
            "AND source_box_level=?", (int(dim), str(kernel_type), int(q_order),
                                       int(source_box_level))).fetchone()
        if row is None:
            raise KeyError("missing table record")
        if row["build_method"] is not None and row["build_method"] != BUILD_METHOD:
            raise KeyError("cached build_method is unsupported; expected " + BUILD_METHOD)
        if not row["payload"]:
            raise KeyError("table cache payload is missing")
        t = table_from_payload(bytes(row["payload"]), int(q_order), int(dim), kernel,
                               float(row["source_box_extent"]), int(source_box_level))
        if (t.n_q_points, t.n_pairs) != (row["n_q_points"], row["n_pairs"]):
            raise ValueError("cache record sizes do not match the table")
        return t

    def close(self):
        self.db.close()
