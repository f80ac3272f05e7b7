"""Build and call the C/OpenMP oracle (``list1_cpu.c``) -- test/baseline only."""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "list1_cpu.c"
LIB = HERE / "liboracle_list1.so"

# x86-64-v3 (AVX2 + FMA) so the library also runs on the GPU hosts' CPUs.
CFLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=gnu11"]


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", *CFLAGS, str(SRC), "-o", str(tmp), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.oracle_near_field_f64.restype = C.c_int
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


def near_field(tree, table_data, mode_nmlz, meta, coefs, *, mode, scaling_power=2,
               log_displacement=False, table_root_extent, table_starting_level=0,
               nthreads=0, bias=1e-10, own_csr=True):
    """Evaluate all densities ``coefs[R, n]``; returns ``[R, n]``.  With
    `own_csr` (default) the list-1 CSR of the tree's target boxes is
    recomputed by geometry (``list1_ref.list1_csr_bruteforce``) rather than
    taken from the tree builder."""
    from .list1_ref import list1_csr_bruteforce, tree_groups

    L = lib()
    if own_csr:
        nstarts, nlists = list1_csr_bruteforce(tree.box_centers, tree.box_levels,
                                               tree.root_extent, tree.target_boxes,
                                               box_group=tree_groups(tree))
    else:
        nstarts, nlists = tree.neighbor_source_boxes_starts, tree.neighbor_source_boxes_lists
    coefs = np.ascontiguousarray(np.atleast_2d(coefs), dtype=np.float64)
    R, npts = coefs.shape
    table = np.ascontiguousarray(np.atleast_2d(table_data), dtype=np.float64)
    nmlz = np.ascontiguousarray(np.atleast_2d(mode_nmlz).real, dtype=np.float64)
    out = np.zeros((R, npts))
    arrs = dict(
        centers=np.ascontiguousarray(tree.box_centers, np.float64),
        levels=np.ascontiguousarray(tree.box_levels, np.uint8),
        starts=np.ascontiguousarray(tree.box_target_starts, np.int32),
        counts=np.ascontiguousarray(tree.box_target_counts_nonchild, np.int32),
        tboxes=np.ascontiguousarray(tree.target_boxes, np.int32),
        nstarts=np.ascontiguousarray(nstarts, np.int32),
        nlists=np.ascontiguousarray(nlists, np.int32),
        ids=np.ascontiguousarray(tree.local_ids, np.int32),
        cidx=np.ascontiguousarray(meta["case_indices"], np.int64),
        ranks=np.ascontiguousarray(meta["case_orbit_ranks"], np.uint16),
        perm=np.ascontiguousarray(meta["case_axis_perm"], np.uint8),
        sign=np.ascontiguousarray(meta["case_axis_sign"], np.int8),
        group=np.ascontiguousarray(meta["case_axis_group"], np.uint8),
        offs=np.ascontiguousarray(meta["case_value_offsets"], np.int32),
    )
    mode_code = {"canonical": 0, "fixed": 1, "per_level": 2}[mode]
    power = np.asarray(meta.get("axis_sign_power", np.zeros(tree.dim)), dtype=np.int64)
    sp_mask = int(sum(int(v & 1) << a for a, v in enumerate(power)))
    st = L.oracle_near_field_f64(
        C.c_int(tree.dim), C.c_int(tree.quad_order), C.c_int(len(tree.target_boxes)),
        C.c_int(tree.nboxes), _p(arrs["centers"]), _p(arrs["levels"]), _p(arrs["starts"]),
        _p(arrs["counts"]), _p(arrs["tboxes"]), _p(arrs["nstarts"]), _p(arrs["nlists"]),
        _p(arrs["ids"]), _p(arrs["ids"]), _p(arrs["cidx"]), C.c_int(arrs["cidx"].size),
        _p(arrs["ranks"]), _p(arrs["perm"]), _p(arrs["sign"]), _p(arrs["group"]),
        _p(arrs["offs"]), _p(table), C.c_int64(table.shape[1]), C.c_int(table.shape[0]),
        _p(nmlz), C.c_int(mode_code), C.c_int(scaling_power), C.c_int(int(log_displacement)),
        C.c_double(tree.root_extent), C.c_double(table_root_extent),
        C.c_int(table_starting_level), C.c_double(bias), C.c_int(R), _p(coefs),
        C.c_int64(npts), _p(out), C.c_int64(npts), C.c_int(nthreads), C.c_int(sp_mask),
    )
    if st != 0:
        raise RuntimeError("oracle: invalid list-1 case id")
    return out
