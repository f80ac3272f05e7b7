"""Near-field interaction tables (orbit-reduced storage, device Duffy build).

Mirrors ``NearFieldInteractionTable`` (volumential/nearfield_potential_table.py:
436-3619) for the parts the list-1 path needs:

* geometry: ``q_points`` (GL tensor nodes on ``[0, H]^d``, axis 0 slowest,
  535-592), case vectors and lookup (``list1_gallery``), case target points
  (``_get_case_target_points``, 1352-1382);
* orbit-reduced storage: ``reduced_entry_ids`` (sorted full ids) + values
  (``set_reduced_table_data`` / ``get_entry_data_for_full_indices``,
  1032-1083);
* ``build_table``: every canonical entry of every case integrated in one
  device launch (``vl1_duffy_build``), the batched Duffy builder of
  2621-2702 with the node rule of 2224-2299;
* 2D mode normalisers: the GL weight outer product (2744-2770).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import gallery, orbit, quadrature
from .kernels import ConstantKernel, HelmholtzKernel, LaplaceKernel


@dataclass(frozen=True)
class DuffyBuildConfig:
    """``nearfield_potential_table.py:197-206`` (fixed-order rules only)."""

    radial_rule: str = "tanh-sinh-fast"
    regular_quad_order: int = 20
    radial_quad_order: int = 61


class NearFieldInteractionTable:
    def __init__(
        self,
        quad_order,
        dim=2,
        kernel=None,
        source_box_extent=1.0,
        dtype=np.float64,
        source_box_level=None,
    ):
        self.quad_order = int(quad_order)
        self.dim = int(dim)
        if kernel is None:
            kernel = ConstantKernel(self.dim)
        if kernel.dim != self.dim:
            raise ValueError("kernel dimension does not match table dimension")
        self.integral_knl = kernel
        if kernel.is_complex_valued:
            dtype = np.complex128
        self.dtype = np.dtype(dtype).type
        if not source_box_extent > 0:
            raise ValueError("source_box_extent must be positive")
        self.source_box_extent = float(source_box_extent)
        self.source_box_level = source_box_level
        self.center = np.full(self.dim, 0.5 * self.source_box_extent)
        self.n_q_points = self.quad_order**self.dim
        self.n_pairs = self.n_q_points**2
        self.interaction_case_vecs, self.case_encode, self.case_indices = (
            gallery.generate_list1_gallery(self.dim)
        )
        self.n_cases = len(self.interaction_case_vecs)
        self.q_points = quadrature.tensor_q_points(
            self.quad_order, self.dim, self.source_box_extent
        )
        self.mode_normalizers = np.zeros(self.n_q_points, dtype=self.dtype)
        self.kernel_exterior_normalizers = np.zeros(self.n_q_points, dtype=self.dtype)
        self.reduced_entry_ids = None
        self._data = None
        self._compact = None
        self.table_data_is_symmetry_reduced = False
        self.is_built = False
        self.last_build_seconds = None

    # ------------------------------------------------------------ geometry
    @property
    def interp_nodes(self) -> np.ndarray:
        """1D interpolation nodes of the source box (last-axis nodes)."""
        return np.ascontiguousarray(self.q_points[: self.quad_order, self.dim - 1])

    def case_target_points(self) -> np.ndarray:
        """``[dim, n_cases, n_q_points]`` target node coordinates per case."""
        vecs = np.asarray(self.interaction_case_vecs, dtype=np.float64)
        offsets = (self.q_points - self.center).T  # [dim, n]
        out = np.empty((self.dim, self.n_cases, self.n_q_points))
        for c, vec in enumerate(vecs):
            cntr = 0.5 * self.source_box_extent + vec / 4.0 * self.source_box_extent
            size = gallery.target_box_size(vec)
            out[:, c, :] = cntr[:, None] + size * offsets
        return out

    def arithmetic_metadata(self) -> orbit.ArithmeticCaseMetadata:
        """Orbit metadata of this table's kernel (signed for axis derivatives)."""
        return orbit.kernel_case_metadata(self.integral_knl, self.quad_order)

    def representative_entry_ids(self) -> np.ndarray:
        """Canonical representatives in compact-row order."""
        return orbit.enumerate_representatives(self.arithmetic_metadata())

    def decode_index(self, entry_id):
        case_id = entry_id // self.n_pairs
        pair = entry_id % self.n_pairs
        return {
            "case_index": case_id,
            "source_mode_index": pair // self.n_q_points,
            "target_point_index": pair % self.n_q_points,
        }

    # ------------------------------------------------------------- storage
    @property
    def data(self):
        return self._data

    def set_reduced_table_data(self, entry_ids, values):
        """Store values by full entry id, sorted (reference 1032-1056)."""
        entry_ids = np.asarray(entry_ids, dtype=np.int64)
        values = np.asarray(values, dtype=self.dtype)
        if entry_ids.ndim != 1 or entry_ids.shape != values.shape:
            raise ValueError("entry_ids and values must be matching 1D arrays")
        order = np.argsort(entry_ids, kind="stable")
        entry_ids, values = entry_ids[order], values[order]
        if entry_ids.size:
            if entry_ids[0] < 0 or entry_ids[-1] >= self.n_cases * self.n_pairs:
                raise ValueError("entry_ids must be valid full table entry IDs")
            if np.any(entry_ids[:-1] == entry_ids[1:]):
                raise ValueError("entry_ids must be unique")
        self.reduced_entry_ids = np.ascontiguousarray(entry_ids)
        self._data = np.ascontiguousarray(values)
        self._compact = None
        self.table_data_is_symmetry_reduced = True

    def get_reduced_table_data(self):
        return self.reduced_entry_ids, self._data

    def get_entry_data_for_full_indices(self, entry_ids):
        entry_ids = np.asarray(entry_ids, dtype=np.int64)
        if self._data is None:
            raise RuntimeError("table data has not been built")
        pos = np.searchsorted(self.reduced_entry_ids, entry_ids)
        pos = np.minimum(pos, len(self.reduced_entry_ids) - 1)
        if not np.array_equal(self.reduced_entry_ids[pos], entry_ids):
            raise RuntimeError(
                "missing canonical table entry data for compact orbit-reduced table"
            )
        return self._data[pos]

    def compact_values(self) -> np.ndarray:
        """Values in compact-row order (the device layout).

        Every stored entry ``e`` lands on its row as ``sign(e) * value(e)``
        (``sign`` is +1 for scalar kernels), so any choice of stored orbit
        members works -- this package's representatives or the reference's.
        """
        if self._data is None:
            raise RuntimeError("table data has not been built")
        if getattr(self, "_compact", None) is not None:
            return self._compact
        meta = self.arithmetic_metadata()
        ids = self.reduced_entry_ids
        case = ids // self.n_pairs
        pair = ids % self.n_pairs
        s, t = pair // self.n_q_points, pair % self.n_q_points
        rows = orbit.entry_rows(meta, case, s, t)
        sg = orbit.entry_signs(meta, case, s, t)
        out = np.zeros(meta.n_compact_entries, dtype=self._data.dtype)
        hit = np.zeros(meta.n_compact_entries, dtype=bool)
        out[rows] = sg * self._data
        hit[rows] = True
        if not hit.all():
            raise RuntimeError(
                "missing canonical table entry data for compact orbit-reduced table"
            )
        self._compact = out
        return out

    def get_entry_data(self, entry_id):
        """Value of any full entry, reconstructed through its orbit."""
        info = self.decode_index(int(entry_id))
        args = (self.arithmetic_metadata(), info["case_index"], info["source_mode_index"],
                info["target_point_index"])
        row = orbit.entry_rows(*args)
        return int(orbit.entry_signs(*args)) * self.compact_values()[int(row)]

    # --------------------------------------------------------------- build
    def build_normalizer_table(self):
        """2D mode normalisers: integrals of the modes over the box."""
        if self.dim != 2:
            return
        _, w = np.polynomial.legendre.leggauss(self.quad_order)
        w = 0.5 * self.source_box_extent * w
        self.mode_normalizers[:] = np.multiply.outer(w, w).reshape(-1)

    def duffy_inputs(self, entry_ids=None):
        """Per-entry target points, decomposition vertices and mode axes."""
        if entry_ids is None:
            entry_ids = self.representative_entry_ids()
        entry_ids = np.asarray(entry_ids, dtype=np.int64)
        case = entry_ids // self.n_pairs
        pair = entry_ids % self.n_pairs
        s = pair // self.n_q_points
        t = pair % self.n_q_points
        tp = self.case_target_points()[:, case, t]
        dv = np.clip(tp, 0.0, self.source_box_extent)
        axes = np.empty((self.dim, entry_ids.size), dtype=np.int32)
        rem = s.copy()
        for ax in range(self.dim - 1, -1, -1):
            axes[ax] = rem % self.quad_order
            rem //= self.quad_order
        return (
            np.ascontiguousarray(tp),
            np.ascontiguousarray(dv),
            np.ascontiguousarray(axes),
        )

    def build_table(self, build_config=None, stream=None):
        """Build every canonical entry on the GPU in one launch."""
        from .duffy import duffy_values

        cfg = build_config or DuffyBuildConfig()
        if cfg.radial_rule != "tanh-sinh-fast":
            raise NotImplementedError("only the tanh-sinh-fast radial rule is implemented")
        import time

        t0 = time.perf_counter()
        self.build_normalizer_table()
        entry_ids = self.representative_entry_ids()
        values = duffy_values(
            self,
            entry_ids,
            int(cfg.regular_quad_order),
            int(cfg.radial_quad_order),
            stream=stream,
        )
        self.set_reduced_table_data(entry_ids, values)
        self.is_built = True
        self.last_build_seconds = time.perf_counter() - t0
        return self


def kernel_from_name(name: str, dim: int, **kw):
    key = name.lower()
    if key in ("laplace", "lapknl"):
        return LaplaceKernel(dim)
    if key in ("constant", "const", "cstknl"):
        return ConstantKernel(dim)
    if key in ("helmholtz", "helmknl"):
        return HelmholtzKernel(dim, float(kw.get("k", 1.0)))
    raise ValueError(f"unknown kernel type {name!r}")
