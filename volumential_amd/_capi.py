"""ctypes binding of ``libvl1.so`` (the C ABI declared in ``include/vl1.h``).

The library is loaded after ``torch`` so that both share torch's HIP
runtime (same SONAME).  There is no fallback: a missing or stale library
raises immediately.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch  # noqa: F401  (must be loaded before libvl1: shared HIP runtime)

LIB_PATH = Path(__file__).resolve().parent / "libvl1.so"

VL1_OK = 0
VL1_E_ARG = -1
VL1_E_LAYOUT = -2
VL1_E_LEVEL = -3
VL1_E_COVERAGE = -4
VL1_E_TABLE = -5
VL1_E_NONFINITE = -6
VL1_E_HIP = -7
VL1_E_UNSUPPORTED = -8

KERNEL_CONSTANT = 0
KERNEL_LAPLACE = 1
KERNEL_HELMHOLTZ = 2
KERNEL_LAPLACE_TARGET_DERIV = 3
KERNEL_LAPLACE_SOURCE_DERIV = 4
KERNEL_HELMHOLTZ_TARGET_DERIV = 5
KERNEL_HELMHOLTZ_SOURCE_DERIV = 6
KERNEL_POWER = 7
KERNEL_POWER_LOG = 8
KERNEL_POWER_TARGET_DERIV = 9
KERNEL_POWER_SOURCE_DERIV = 10
KERNEL_POWER_LOG_TARGET_DERIV = 11
KERNEL_POWER_LOG_SOURCE_DERIV = 12

SPLIT_DERIV_NONE = 0
SPLIT_DERIV_TARGET = 1
SPLIT_DERIV_SOURCE = 2

SCALING_CANONICAL = 0
SCALING_FIXED = 1
SCALING_PER_LEVEL = 2
SCALING_CUSTOM = 3
MAX_LEVELS = 32

FORM_F64 = 0
FORM_SPLIT_BF16 = 1
FORM_SPLIT_I8 = 2

EXPORTED_SYMBOLS = (
    "vl1_last_error",
    "vl1_abi_version",
    "vl1_payload_create",
    "vl1_payload_create_dense",
    "vl1_payload_has_missing",
    "vl1_payload_destroy",
    "vl1_payload_device_bytes",
    "vl1_near_field_f64",
    "vl1_plan_create",
    "vl1_plan_create_helmholtz_split",
    "vl1_plan_destroy",
    "vl1_plan_eval",
    "vl1_plan_stage",
    "vl1_plan_contract",
    "vl1_plan_stats",
    "vl1_plan_eval_noalloc",
    "vl1_plan_tune",
    "vl1_plan_set_variant",
    "vl1_csr_builder_create",
    "vl1_csr_builder_fill",
    "vl1_csr_builder_destroy",
    "vl1_duffy_build",
    "vl1_split_correction",
    "vl1_orbit_rows_host",
    "vl1_orbit_signs_host",
)

vp = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
f64 = C.c_double


class PayloadDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("quad_order", C.c_int),
        ("n_cases", C.c_int),
        ("n_case_orbits", C.c_int),
        ("n_tables", C.c_int),
        ("is_complex", C.c_int),
        ("n_table_entries", i64),
        ("table_data", vp),
        ("mode_nmlz", vp),
        ("case_indices", vp),
        ("case_orbit_ranks", vp),
        ("case_axis_perm", vp),
        ("case_axis_sign", vp),
        ("case_axis_group", vp),
        ("case_value_offsets", vp),
        ("axis_sign_power", vp),
    ]


class DensePayloadDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("quad_order", C.c_int),
        ("n_cases", C.c_int),
        ("n_tables", C.c_int),
        ("is_complex", C.c_int),
        ("n_table_entries", i64),
        ("table_data", vp),
        ("mode_nmlz", vp),
        ("case_indices", vp),
        ("table_entry_ids", vp),
        ("table_entry_scales", vp),
        ("mode_qpoint_map", vp),
        ("mode_case_map", vp),
        ("mode_case_scale", vp),
    ]


class NearFieldArgs(C.Structure):
    _fields_ = [
        ("scaling_mode", C.c_int),
        ("scaling_power", C.c_int),
        ("log_displacement", C.c_int),
        ("root_extent", f64),
        ("table_root_extent", f64),
        ("table_starting_level", C.c_int),
        ("encoding_shift", f64),
        ("encoding_base", C.c_int),
        ("case_encoding_bias", f64),
        ("n_tgt_boxes", C.c_int),
        ("aligned_nboxes", C.c_int),
        ("n_rhs", C.c_int),
        ("ld_coefs", i64),
        ("ld_result", i64),
        ("result", vp),
        ("mode_coefs", vp),
        ("box_centers", vp),
        ("box_levels", vp),
        ("box_source_counts_nonchild", vp),
        ("box_source_starts", vp),
        ("box_target_counts_nonchild", vp),
        ("box_target_starts", vp),
        ("neighbor_source_boxes_starts", vp),
        ("neighbor_source_boxes_lists", vp),
        ("target_boxes", vp),
        ("source_mode_ids", vp),
        ("target_point_ids", vp),
        ("error_flags", vp),
        ("level_scaling", vp),
        ("level_disp", vp),
        ("exterior_nmlz", vp),
    ]


class PlanDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("n_boxes", C.c_int),
        ("n_tgt_boxes", C.c_int),
        ("n_points", i64),
        ("box_centers", vp),
        ("box_levels", vp),
        ("box_starts", vp),
        ("box_counts", vp),
        ("target_boxes", vp),
        ("nbr_starts", vp),
        ("nbr_lists", vp),
        ("local_ids", vp),
        ("root_extent", f64),
        ("table_root_extent", f64),
        ("table_starting_level", C.c_int),
        ("scaling_mode", C.c_int),
        ("scaling_power", C.c_int),
        ("log_displacement", C.c_int),
        ("case_encoding_bias", f64),
        ("boxes_per_tile", C.c_int),
        ("level_scaling", vp),
        ("level_disp", vp),
    ]


class HelmholtzSplitBlocks(C.Structure):
    _fields_ = [
        ("k_re", f64),
        ("k_im", f64),
        ("series_nmax", C.c_int),
        ("unit_nodes", vp),
        ("unit_weights", vp),
        ("case_vecs", vp),
        ("deriv_kind", C.c_int),
        ("deriv_axis", C.c_int),
        ("split_order", C.c_int),
        ("n_terms", C.c_int),
        ("n_term_entries", C.c_int64),
        ("term_tables", vp),
        ("term_powers", vp),
        ("term_logs", vp),
        ("term_coefs", vp),
        ("smooth_order", C.c_int),
        ("smooth_nodes", vp),
        ("smooth_weights", vp),
        ("smooth_interp", vp),
    ]


class SplitOpts(C.Structure):
    _fields_ = [("n_slices", C.c_int), ("n_products", C.c_int), ("error_est", C.c_void_p)]


class DuffyArgs(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("quad_order", C.c_int),
        ("kernel_kind", C.c_int),
        ("helmholtz_k", f64),
        ("n_entries", i64),
        ("n_nodes", i64),
        ("nodes_per_sign", i64),
        ("source_box_extent", f64),
        ("target_points", vp),
        ("decomposition", vp),
        ("mode_axes", vp),
        ("node_u", vp),
        ("node_sign", vp),
        ("node_jac", vp),
        ("interp_nodes", vp),
        ("bary_w", vp),
        ("result", vp),
        ("deriv_axis", C.c_int),
        ("chunk_start", vp),
        ("n_chunks", i64),
        ("power", C.c_int),
    ]


class SplitArgs(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("split_order", C.c_int),
        ("series_nmax", C.c_int),
        ("n_tgt_boxes", C.c_int),
        ("n_rhs", C.c_int),
        ("max_box_points", C.c_int),
        ("k_re", f64),
        ("k_im", f64),
        ("n_points", i64),
        ("target_boxes", vp),
        ("nbr_starts", vp),
        ("nbr_lists", vp),
        ("box_starts", vp),
        ("box_counts", vp),
        ("points", vp),
        ("strengths", vp),
        ("ld_strengths", i64),
        ("result", vp),
        ("ld_result", i64),
        ("deriv_kind", C.c_int),
        ("deriv_axis", C.c_int),
        ("coefficients", vp),
        ("src_points", vp),
        ("src_box_starts", vp),
        ("src_box_counts", vp),
        ("n_src_points", i64),
        ("max_src_box_points", C.c_int),
        ("legacy_subtraction", C.c_int),
    ]


_lib = None


def lib():
    """Load ``libvl1.so`` once; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m volumential_amd.build`"
        )
    L = C.CDLL(str(LIB_PATH))
    L.vl1_last_error.restype = C.c_char_p
    L.vl1_abi_version.restype = C.c_int
    L.vl1_payload_create.argtypes = [C.POINTER(vp), C.POINTER(PayloadDesc), vp]
    L.vl1_payload_create_dense.argtypes = [C.POINTER(vp), C.POINTER(DensePayloadDesc), vp]
    L.vl1_payload_has_missing.argtypes = [vp]
    L.vl1_payload_destroy.argtypes = [vp]
    L.vl1_payload_device_bytes.argtypes = [vp]
    L.vl1_payload_device_bytes.restype = i64
    L.vl1_near_field_f64.argtypes = [vp, C.POINTER(NearFieldArgs), vp]
    L.vl1_plan_create.argtypes = [C.POINTER(vp), vp, C.POINTER(PlanDesc), vp]
    L.vl1_plan_create_helmholtz_split.argtypes = [C.POINTER(vp), vp, C.POINTER(PlanDesc),
                                                  C.POINTER(HelmholtzSplitBlocks), vp]
    L.vl1_plan_destroy.argtypes = [vp]
    L.vl1_plan_eval.argtypes = [vp, C.c_int, vp, i64, C.c_int, vp, i64, C.POINTER(SplitOpts), vp]
    L.vl1_plan_stage.argtypes = [vp, C.c_int, vp, i64, C.c_int, C.POINTER(SplitOpts), vp]
    L.vl1_plan_contract.argtypes = [vp, C.c_int, C.c_int, vp, i64, C.POINTER(SplitOpts), vp]
    L.vl1_plan_stats.argtypes = [vp, C.POINTER(i64)]
    L.vl1_plan_eval_noalloc.argtypes = [vp, C.c_int, vp, i64, C.c_int, vp, i64,
                                        C.POINTER(SplitOpts), vp]
    L.vl1_plan_tune.argtypes = [vp, C.c_int, vp, i64, C.c_int, vp, i64, C.POINTER(SplitOpts),
                                C.c_int, C.POINTER(i32), C.POINTER(f64), vp]
    L.vl1_plan_set_variant.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.vl1_duffy_build.argtypes = [C.POINTER(DuffyArgs), vp]
    L.vl1_split_correction.argtypes = [C.POINTER(SplitArgs), vp]
    L.vl1_csr_builder_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, vp, vp, C.POINTER(i64), vp]
    L.vl1_csr_builder_fill.argtypes = [vp, vp, vp, vp]
    L.vl1_csr_builder_destroy.argtypes = [vp]
    L.vl1_orbit_rows_host.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, i64, vp]
    L.vl1_orbit_signs_host.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, i64, vp]
    if L.vl1_abi_version() != 11:
        raise RuntimeError("libvl1.so ABI version mismatch; rebuild the library")
    _lib = L
    return L


def last_error() -> str:
    return lib().vl1_last_error().decode("utf-8", "replace")


def check(status: int, what: str = "") -> None:
    """Map a vl1_status to the exception type the reference raises."""
    if status == VL1_OK:
        return
    msg = last_error() or f"{what} failed with status {status}"
    if status in (VL1_E_ARG, VL1_E_LAYOUT):
        raise ValueError(msg)
    if status == VL1_E_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def ptr(t) -> int:
    """Raw address of a torch tensor or numpy array (0 for None)."""
    if t is None:
        return 0
    if isinstance(t, torch.Tensor):
        return t.data_ptr()
    return t.ctypes.data


def stream_handle(stream=None) -> int:
    if stream is None:
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream)
