"""``NearFieldFromCSR``: the list-1 evaluator operator.

Same constructor and call convention as volumential's
``NearFieldFromCSR`` (volumential/list1.py:91-215, 223-437, 1260-1487):
``NearFieldFromCSR(integral_kernel, table_data_shapes, potential_kind=1,
**list1_extra_kwargs)`` and ``__call__(queue, result=..., box_centers=...,
...)`` returning ``(result, event)``.  Arrays may be torch tensors (device or
host) or numpy arrays; the evaluation always runs on the GPU through
``vl1_near_field_f64``.  All four scaling policies of the reference are
supported: inferred, fixed single table, per-level tables, and the caller's
``kernel_scaling_code`` / ``kernel_displacement_code`` strings (this library
has no run-time code generation: the strings are evaluated per source-box
level on the host, see ``scaling_code``).
"""

from __future__ import annotations

import logging
from dataclasses import dataclass

import numpy as np
import torch

from . import _capi, gallery, orbit
from .kernels import inferred_scaling, supports_inferred_scaling
from .payload import DenseDevicePayload, DevicePayload, PayloadCache, fingerprint
from .scaling_code import level_arrays

CASE_ENCODING_BIAS_DEFAULT = 1.0e-10

logger = logging.getLogger(__name__)


def warn_case_encoding_drift(dim, box_centers, box_levels, root_extent, target_boxes, starts,
                             lists, encoding_shift, warn_tol):
    """Log a warning when a list-1 pair's unrounded case-vector component
    ``(c_t - c_s) / h_s * 4 + shift`` is farther than ``warn_tol`` from an
    integer (``list1.py:468-530``, enabled by ``case_encoding_warn_tol``).
    Vectorised over all pairs; returns the largest drift."""
    if warn_tol is None:
        return None
    tb = np.asarray(_to_host(target_boxes), dtype=np.int64)
    st = np.asarray(_to_host(starts), dtype=np.int64)
    ls = np.asarray(_to_host(lists), dtype=np.int64)
    if tb.size == 0 or ls.size == 0:
        return 0.0
    centers = np.asarray(_to_host(box_centers), dtype=np.float64)
    levels = np.asarray(_to_host(box_levels), dtype=np.int64)
    tgt = np.repeat(tb, np.diff(st))
    ext = float(root_extent) * 0.5 ** levels[ls]
    raw = (centers[:dim, tgt] - centers[:dim, ls]) / ext[None, :] * 4.0 + float(encoding_shift)
    drift = np.abs(raw - np.rint(raw))
    worst = np.unravel_index(int(np.argmax(drift)), drift.shape)
    max_drift = float(drift[worst])
    if max_drift > float(warn_tol):
        d, k = int(worst[0]), int(worst[1])
        logger.warning(
            "List1 case encoding drift %.3e exceeds %.3e for target box %d, source box %d, "
            "dim %d (raw %.17g)", max_drift, float(warn_tol), int(tgt[k]), int(ls[k]), d,
            float(raw[d, k]))
    return max_drift


@dataclass(frozen=True)
class KernelScalingPolicy:
    """``list1.py:49-61``."""

    kernel_name: str
    mode: str
    infer_kernel_scaling: bool
    single_table_scaling_supported: bool
    reference_table_level: int | None
    scaling_power: int
    log_displacement: bool
    notes: str


_PAYLOADS = PayloadCache(16)


def _to_device(a, dtype):
    if isinstance(a, torch.Tensor):
        t = a
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(a)))
    t = t.to(dtype=dtype)
    if t.device.type != "cuda":
        t = t.to("cuda")
    return t.contiguous()


def _to_host(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    if hasattr(a, "get"):
        return a.get()
    return np.asarray(a)


class NearFieldFromCSR:
    default_name = "near_field_from_csr"

    def __init__(self, integral_kernel, table_data_shapes, potential_kind=1, **kwargs):
        """``potential_kind`` (list1.py:101-109): 1, the volume potentials; 2,
        the inverse-type (hypersingular) potentials, which add each
        target's density times its exterior mode normaliser scaled as its
        box (``exterior_mode_nmlz_combined`` at call time; real tables
        without a displacement term)."""
        if potential_kind not in (1, 2):
            raise ValueError("Unsupported potential kind %d" % potential_kind)
        self.integral_kernel = integral_kernel
        self.potential_kind = potential_kind
        self.n_tables = int(table_data_shapes["n_tables"])
        self.n_q_points = int(table_data_shapes["n_q_points"])
        self.n_table_entries = int(table_data_shapes["n_table_entries"])
        self.n_cases = int(table_data_shapes.get("n_cases", 0))
        # the reference's default kind (list1.py:120)
        self.reconstruction_kind = table_data_shapes.get("reconstruction_kind", "dense")
        if self.reconstruction_kind == "generated-orbit":
            raise NotImplementedError(
                "reconstruction kind 'generated-orbit' (transform scans and lookup tables, "
                "list1.py:903-1017) is not part of this library: pass the same table as "
                "'dense' (table_entry_ids / table_entry_scales with the mode maps) or as an "
                "arithmetic orbit kind"
            )
        if self.reconstruction_kind not in ("dense", "scalar-arithmetic-orbit",
                                            "signed-arithmetic-orbit"):
            raise ValueError(f"unknown reconstruction kind {self.reconstruction_kind!r}")
        self.dim = integral_kernel.dim
        self.quad_order = int(
            table_data_shapes.get("quad_order", round(self.n_q_points ** (1.0 / self.dim)))
        )
        self.extra_kwargs = dict(kwargs)
        self.extra_kwargs.setdefault("case_encoding_bias", CASE_ENCODING_BIAS_DEFAULT)
        self.extra_kwargs["case_encoding_bias"] = float(self.extra_kwargs["case_encoding_bias"])
        if "infer_kernel_scaling" not in self.extra_kwargs:
            self.extra_kwargs["infer_kernel_scaling"] = (
                self.n_tables == 1 and supports_inferred_scaling(integral_kernel)
            )
        if (self.extra_kwargs["infer_kernel_scaling"]
                and not supports_inferred_scaling(integral_kernel)
                and "kernel_scaling_code" not in self.extra_kwargs):
            raise RuntimeError(
                "infer_kernel_scaling is unsupported for kernel "
                f"{integral_kernel!r}. For Helmholtz, K_k(h r) = h^{{-1}} K_{{k h}}(r): "
                "exact table scaling requires changing the effective wave number with box "
                "size, which a single fixed-k table cannot represent. Use multi-level "
                "tables."
            )
        # user rules replace the inferred ones; both strings are needed
        # (list1.py:198-205)
        if "kernel_scaling_code" in kwargs or "kernel_displacement_code" in kwargs:
            self.extra_kwargs["infer_kernel_scaling"] = False
            if not ("kernel_scaling_code" in kwargs and "kernel_displacement_code" in kwargs):
                raise ValueError("kernel_scaling_code and kernel_displacement_code must be "
                                 "given together")
            # parse now: a malformed string fails at construction
            level_arrays(kwargs["kernel_scaling_code"], kwargs["kernel_displacement_code"],
                         2.0, 2.0)
        self._custom_levels = {}
        if self.extra_kwargs.get("case_encoding_warn_tol") is not None:
            self.extra_kwargs["case_encoding_warn_tol"] = float(
                self.extra_kwargs["case_encoding_warn_tol"])
        bs = int(self.extra_kwargs.get("list1_target_batch_size", 1))
        if bs < 1:
            raise ValueError("list1_target_batch_size must be positive")
        self._level_check_key = None

    # -------------------------------------------------------------- policy
    def get_kernel_scaling_policy(self) -> KernelScalingPolicy:
        """``list1.py:361-437``."""
        kname = repr(self.integral_kernel)
        if self.extra_kwargs.get("infer_kernel_scaling", False):
            p, logd = inferred_scaling(self.integral_kernel)
            return KernelScalingPolicy(
                kname, "canonical_single_table", True, True, 0, p, logd,
                "Uses table level 0 for all source-box levels with inferred exact scaling.",
            )
        if "kernel_scaling_code" in self.extra_kwargs:
            return KernelScalingPolicy(
                kname, "custom_single_table", False, True, 0, 0, False,
                "Uses user-provided scaling and displacement code with table level 0.",
            )
        if self.n_tables == 1:
            return KernelScalingPolicy(
                kname, "fixed_single_table", False, False, 0, 0, False,
                "Uses the only cached table level without kernel scaling; runtime checks "
                "reject mixed source levels or a table level mismatch.",
            )
        return KernelScalingPolicy(
            kname, "per_level_tables", False, False, None, 0, False,
            "Selects one cached table level per source-box level without kernel scaling.",
        )

    def _mode_code(self, policy):
        return {
            "canonical_single_table": _capi.SCALING_CANONICAL,
            "fixed_single_table": _capi.SCALING_FIXED,
            "per_level_tables": _capi.SCALING_PER_LEVEL,
            "custom_single_table": _capi.SCALING_CUSTOM,
        }[policy.mode]

    def custom_level_arrays(self, root_extent, table_root_extent):
        """``(level_scaling, level_disp)`` of the caller's code strings for
        source boxes of each level (custom policy), or ``(None, None)``."""
        if "kernel_scaling_code" not in self.extra_kwargs:
            return None, None
        key = (float(root_extent), float(table_root_extent))
        if key not in self._custom_levels:
            self._custom_levels[key] = level_arrays(
                self.extra_kwargs["kernel_scaling_code"],
                self.extra_kwargs["kernel_displacement_code"], *key)
        return self._custom_levels[key]

    # ------------------------------------------------------------ payload
    def _dense_payload(self, kw):
        """Device payload of the ``dense`` kind (list1.py:1018-1045)."""
        names = ("table_data_combined", "mode_nmlz_combined", "case_indices", "table_entry_ids",
                 "table_entry_scales", "mode_qpoint_map", "mode_case_map", "mode_case_scale")
        missing = [k for k in names if kw.get(k) is None]
        if missing:
            raise ValueError("the 'dense' reconstruction kind needs " + ", ".join(missing))
        host = {k: _to_host(kw[k]) for k in names}
        key = ("dense", repr(self.integral_kernel), self.quad_order) + tuple(
            fingerprint(np.asarray(host[k])) for k in names)

        def make():
            return DenseDevicePayload(
                self.dim, self.quad_order,
                np.asarray(host["table_data_combined"]).reshape(self.n_tables, -1),
                np.asarray(host["mode_nmlz_combined"]).reshape(self.n_tables, -1),
                host["case_indices"], host["table_entry_ids"], host["table_entry_scales"],
                host["mode_qpoint_map"], host["mode_case_map"], host["mode_case_scale"],
                n_cases=self.n_cases or np.asarray(host["mode_case_map"]).size // self.n_q_points)

        return _PAYLOADS.get(key, make)

    def _payload(self, kw):
        if self.reconstruction_kind == "dense":
            return self._dense_payload(kw)
        table_data = _to_host(kw["table_data_combined"])
        sp = kw.get("arithmetic_axis_sign_power")
        sign_power = (np.zeros(self.dim, np.uint8) if sp is None
                      else np.ascontiguousarray(_to_host(sp), dtype=np.uint8).reshape(self.dim))
        if self.reconstruction_kind == "scalar-arithmetic-orbit" and np.any(sign_power):
            raise ValueError("scalar-arithmetic-orbit payloads have no axis sign powers")
        key = (
            repr(self.integral_kernel),
            self.quad_order,
            fingerprint(table_data),
            fingerprint(_to_host(kw["mode_nmlz_combined"])),
            tuple(int(v) for v in sign_power),
        )

        def make():
            meta = orbit.ArithmeticCaseMetadata(
                dim=self.dim,
                quad_order=self.quad_order,
                canonical_case_ids=np.zeros(
                    int(np.asarray(_to_host(kw["arithmetic_case_value_offsets"])).size) - 1,
                    dtype=np.int32,
                ),
                case_orbit_ranks=np.asarray(_to_host(kw["arithmetic_case_orbit_ranks"]), np.uint16),
                case_axis_perm=np.asarray(_to_host(kw["arithmetic_case_axis_perm"]), np.uint8),
                case_axis_sign=np.asarray(_to_host(kw["arithmetic_case_axis_sign"]), np.int8),
                case_axis_group=np.asarray(_to_host(kw["arithmetic_case_axis_group"]), np.uint8),
                case_value_offsets=np.asarray(
                    _to_host(kw["arithmetic_case_value_offsets"]), np.int32
                ),
                axis_sign_power=sign_power,
            )
            if int(kw.get("arithmetic_direction_sign_axis", -1)) >= 0:
                raise NotImplementedError(
                    "directional source derivatives (arithmetic_direction_sign_axis) are "
                    "not part of this library; axis derivatives are")
            return DevicePayload(
                self.dim,
                self.quad_order,
                table_data.reshape(self.n_tables, -1),
                _to_host(kw["mode_nmlz_combined"]).reshape(self.n_tables, -1),
                _to_host(kw["case_indices"]),
                meta,
            )

        return _PAYLOADS.get(key, make)

    # ---------------------------------------------------------------- call
    def __call__(self, queue=None, **kw):
        policy = self.get_kernel_scaling_policy()
        result = kw.pop("result")
        mode_coefs = kw.pop("mode_coefs")
        payload = self._payload(kw)
        if int(kw.get("encoding_base", gallery.ENCODING_BASE)) != gallery.ENCODING_BASE:
            raise ValueError("encoding_base must be 13")

        target_boxes = kw["target_boxes"]
        lists = kw["neighbor_source_boxes_lists"]
        box_levels = kw["box_levels"]
        tsl = int(kw["table_starting_level"])
        if policy.mode == "fixed_single_table":
            key = (id(lists), id(box_levels), tsl)
            if key != self._level_check_key:
                src = _to_host(lists).astype(np.int64)
                if src.size:
                    lv = _to_host(box_levels)[src]
                    lo, hi = int(lv.min()), int(lv.max())
                    if lo != hi:
                        raise RuntimeError(
                            "Single near-field table without scaling cannot be used with "
                            f"source boxes on multiple levels ({lo}..{hi}); build per-level "
                            "tables or provide explicit scaling rules."
                        )
                    if lo != tsl:
                        raise RuntimeError(
                            "Single near-field table level mismatch: source level "
                            f"{lo} but table_starting_level is {tsl}. Build the table at "
                            "the active source level or provide scaling rules."
                        )
                self._level_check_key = key

        warn_case_encoding_drift(
            self.dim, kw["box_centers"], box_levels, kw["root_extent"], target_boxes,
            kw["neighbor_source_boxes_starts"], lists,
            kw.get("encoding_shift", gallery.ENCODING_SHIFT),
            self.extra_kwargs.get("case_encoding_warn_tol"))
        coefs_2d = mode_coefs if isinstance(mode_coefs, torch.Tensor) else torch.as_tensor(
            np.asarray(mode_coefs)
        )
        host_result = not (isinstance(result, torch.Tensor) and result.device.type == "cuda")
        # complex tables evaluate in complex128 (expansion_wrangler_fpnd.py:3283-3297);
        # the C ABI takes interleaved (re, im) doubles
        vdt = torch.complex128 if payload.is_complex else torch.float64
        d_res = _to_device(result, vdt)
        d_coef = _to_device(coefs_2d, vdt)
        n_rhs = 1 if d_coef.dim() == 1 else int(d_coef.shape[0])
        vw = 2 if payload.is_complex else 1
        p_res = torch.view_as_real(d_res) if payload.is_complex else d_res
        p_coef = torch.view_as_real(d_coef) if payload.is_complex else d_coef
        centers = _to_device(kw["box_centers"], torch.float64)
        arrs = {
            k: _to_device(kw[k], torch.int32)
            for k in (
                "box_source_counts_nonchild",
                "box_source_starts",
                "box_target_counts_nonchild",
                "box_target_starts",
                "neighbor_source_boxes_starts",
                "neighbor_source_boxes_lists",
                "target_boxes",
                "source_mode_ids",
                "target_point_ids",
            )
        }
        d_levels = _to_device(box_levels, torch.uint8)
        err = torch.zeros(1, dtype=torch.int32, device="cuda")
        n_tgt = int(target_boxes.shape[0] if hasattr(target_boxes, "shape") else len(target_boxes))
        args = _capi.NearFieldArgs(
            scaling_mode=self._mode_code(policy),
            scaling_power=int(policy.scaling_power),
            log_displacement=int(policy.log_displacement),
            root_extent=float(kw["root_extent"]),
            table_root_extent=float(kw["table_root_extent"]),
            table_starting_level=tsl,
            encoding_shift=float(kw.get("encoding_shift", gallery.ENCODING_SHIFT)),
            encoding_base=gallery.ENCODING_BASE,
            case_encoding_bias=self.extra_kwargs["case_encoding_bias"],
            n_tgt_boxes=n_tgt,
            aligned_nboxes=int(centers.shape[1]),
            n_rhs=n_rhs,
            ld_coefs=int(d_coef.shape[-1]) * vw,
            ld_result=int(d_res.shape[-1]) * vw,
            result=p_res.data_ptr(),
            mode_coefs=p_coef.data_ptr(),
            box_centers=centers.data_ptr(),
            box_levels=d_levels.data_ptr(),
            box_source_counts_nonchild=arrs["box_source_counts_nonchild"].data_ptr(),
            box_source_starts=arrs["box_source_starts"].data_ptr(),
            box_target_counts_nonchild=arrs["box_target_counts_nonchild"].data_ptr(),
            box_target_starts=arrs["box_target_starts"].data_ptr(),
            neighbor_source_boxes_starts=arrs["neighbor_source_boxes_starts"].data_ptr(),
            neighbor_source_boxes_lists=arrs["neighbor_source_boxes_lists"].data_ptr(),
            target_boxes=arrs["target_boxes"].data_ptr(),
            source_mode_ids=arrs["source_mode_ids"].data_ptr(),
            target_point_ids=arrs["target_point_ids"].data_ptr(),
            error_flags=err.data_ptr(),
        )
        if self.potential_kind == 2:
            ext = kw.get("exterior_mode_nmlz_combined")
            if ext is None:
                raise ValueError("potential_kind=2 needs exterior_mode_nmlz_combined")
            d_ext = _to_device(ext, torch.float64).reshape(-1).contiguous()
            if payload.is_complex or d_ext.numel() != self.n_tables * self.n_q_points:
                raise ValueError("exterior_mode_nmlz_combined must hold n_tables x n_q_points "
                                 "real values (potential_kind=2 takes real tables)")
            args.exterior_nmlz = d_ext.data_ptr()
            self._keep_ext = d_ext
        lsc, ldisp = self.custom_level_arrays(kw["root_extent"], kw["table_root_extent"])
        if lsc is not None:
            key = (float(kw["root_extent"]), float(kw["table_root_extent"]), "device")
            if key not in self._custom_levels:
                self._custom_levels[key] = (torch.as_tensor(lsc, device="cuda"),
                                            torch.as_tensor(ldisp, device="cuda"))
            d_lsc, d_ldisp = self._custom_levels[key]
            args.level_scaling = d_lsc.data_ptr()
            args.level_disp = d_ldisp.data_ptr()
            args.log_displacement = int(bool(np.any(ldisp != 0.0)))
        L = _capi.lib()
        _capi.check(
            L.vl1_near_field_f64(payload.handle, args, _capi.stream_handle()),
            "vl1_near_field_f64",
        )
        torch.cuda.current_stream().synchronize()
        if int(err.item()) != 0:
            raise RuntimeError(
                "list-1 case encoding produced an invalid case id (box geometry is not a "
                "2:1-balanced list-1 neighbourhood)"
            )
        if host_result:
            out = d_res.cpu().numpy()
            if isinstance(result, np.ndarray):
                result[...] = out.reshape(result.shape)
                return result, None
            return out, None
        if d_res.data_ptr() != result.data_ptr():
            result.copy_(d_res)
        return result, None
