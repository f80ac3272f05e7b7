"""Table cache interchange in the reference's format (CPU).

* the SQLite columns are those of the reference's schema 2.1.0
  (table_manager.py:665-740) and the schema_version meta value is "2.1.0";
* a table round-trips through the cache (reduced storage, complex too),
  value for value, including normalisers;
* the npz payload holds the reference's array names
  (table_manager.py:182-228) and loads with allow_pickle=False;
* a payload with dense `data` (NaN where not stored), as the reference writes
  for unreduced tables, reads back to the same values.
"""

import sqlite3
from io import BytesIO

import numpy as np
import pytest

from volumential_amd import orbit
from volumential_amd.kernels import HelmholtzKernel, LaplaceKernel
from volumential_amd.table import NearFieldInteractionTable
from volumential_amd.table_cache import TableCache, table_from_payload, table_payload

REFERENCE_COLUMNS = [
    "dim", "kernel_type", "q_order", "source_box_level", "n_q_points", "quad_order", "n_pairs",
    "source_box_extent", "source_box_level_stored", "case_encoding_base", "case_encoding_shift",
    "build_method", "kernel_type_cached", "payload",
]


def _table(kernel, q, dim, extent, seed):
    t = NearFieldInteractionTable(q, dim, kernel, source_box_extent=extent)
    t.build_normalizer_table()
    reps = orbit.enumerate_representatives(t.arithmetic_metadata())
    rng = np.random.default_rng(seed)
    vals = rng.standard_normal(reps.size)
    if kernel.is_complex_valued:
        vals = vals + 1j * rng.standard_normal(reps.size)
    t.set_reduced_table_data(reps, vals)
    t.is_built = True
    return t


@pytest.mark.parametrize("kernel,q,dim,level", [(LaplaceKernel(3), 3, 3, 0),
                                                (HelmholtzKernel(2, 40.0), 6, 2, 4)])
def test_round_trip(tmp_path, kernel, q, dim, level):
    cache = TableCache(tmp_path / "tables.sqlite", root_extent=2.0)
    t = _table(kernel, q, dim, 2.0 / 2**level, seed=q)
    cache.save(t, "Laplace" if dim == 3 else "Helmholtz", level)
    cache.close()
    cache = TableCache(tmp_path / "tables.sqlite", root_extent=2.0)
    u = cache.load(dim, "Laplace" if dim == 3 else "Helmholtz", q, level, kernel)
    i0, v0 = t.get_reduced_table_data()
    i1, v1 = u.get_reduced_table_data()
    assert np.array_equal(i0, i1) and np.array_equal(v0, v1)
    assert np.array_equal(u.mode_normalizers, t.mode_normalizers)
    assert np.array_equal(u.compact_values(), t.compact_values())
    cols = [r[1] for r in sqlite3.connect(tmp_path / "tables.sqlite").execute(
        "PRAGMA table_info(nearfield_cache)")]
    assert cols == REFERENCE_COLUMNS
    assert cache._meta("schema_version") == "2.1.0"
    with pytest.raises(RuntimeError):
        TableCache(tmp_path / "tables.sqlite", root_extent=1.0)


def test_payload_arrays_and_dense_form():
    t = _table(LaplaceKernel(2), 4, 2, 2.0, seed=1)
    blob = table_payload(t)
    with np.load(BytesIO(blob), allow_pickle=False) as z:
        names = set(z.files)
    assert names == {"q_points", "mode_normalizers", "kernel_exterior_normalizers",
                     "interaction_case_vecs", "case_indices", "table_data_is_symmetry_reduced",
                     "reduced_entry_ids", "reduced_data"}
    # dense payload, NaN outside the stored entries
    ids, vals = t.get_reduced_table_data()
    data = np.full(t.n_cases * t.n_pairs, np.nan)
    data[ids] = vals
    with np.load(BytesIO(blob), allow_pickle=False) as z:
        arrs = {k: z[k] for k in z.files}
    del arrs["reduced_entry_ids"], arrs["reduced_data"]
    arrs["data"] = data
    arrs["table_data_is_symmetry_reduced"] = np.array([0], np.int8)
    fh = BytesIO()
    np.savez(fh, **arrs)
    u = table_from_payload(fh.getvalue(), 4, 2, LaplaceKernel(2), 2.0)
    assert np.array_equal(u.compact_values(), t.compact_values())


def test_signed_table_any_representative_convention(tmp_path):
    """A gradient table stored at other orbit members (as another builder
    may pick them) reads back to the same compact values: each stored entry
    e lands on its row as sign(e) * value(e).  Round trip through the cache."""
    from oracle import orbit_ref
    from volumential_amd.kernels import AxisTargetDerivative

    kernel = AxisTargetDerivative(0, LaplaceKernel(2))
    t = _table(kernel, 3, 2, 2.0, seed=5)
    meta = t.arithmetic_metadata()
    n = t.n_q_points
    compact = t.compact_values()
    # largest member of every orbit under the kernel's group, with its value
    ids = np.arange(t.n_cases * n * n)
    best = ids.copy()
    for sgn, perm, _ in orbit_ref.group_transforms(2, (0,), (1, 0)):
        best = np.maximum(best, orbit_ref.transform_entries(2, 3, sgn, perm))
    members = np.unique(best)
    c, s, tt = members // (n * n), (members % (n * n)) // n, members % n
    vals = orbit.entry_signs(meta, c, s, tt) * compact[orbit.entry_rows(meta, c, s, tt)]
    assert not np.array_equal(np.sort(members), np.sort(t.get_reduced_table_data()[0]))
    u = NearFieldInteractionTable(3, 2, kernel, source_box_extent=2.0)
    u.build_normalizer_table()
    u.set_reduced_table_data(members, vals)
    assert np.array_equal(u.compact_values(), compact)
    cache = TableCache(tmp_path / "g.sqlite", root_extent=2.0)
    cache.save(u, "LaplaceGradX", 0)
    w = cache.load(2, "LaplaceGradX", 3, 0, kernel)
    assert np.array_equal(w.compact_values(), compact)
    for e in members[:50]:
        assert w.get_entry_data(int(e)) == u.get_entry_data(int(e))
