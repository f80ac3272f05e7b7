"""Batched list-1 evaluation plans (``vl1_plan_*``).

A plan binds one tree + list-1 CSR to one device payload: it computes the
case / level / scaling of every (target box, source box) pair once, expands
the per-case dense blocks, and then evaluates many densities per call
(``mode_coefs[R, n_points]``) with

* ``form="f64"``   dense FP64 blocks on FP64 MFMA (rel. error ~1e-15);
* ``form="split"`` Ozaki-split BF16 MFMA with FP32 accumulation and FP64
  cross-box sums (``n_slices`` BF16 pieces per operand);
* ``form="split_i8"`` Ozaki split on int8 MFMA: ``n_slices`` int8 digits per
  operand (default 6), exact int32 sums over all of a tile's list-1 boxes,
  one FP64 recombination per output, and an error estimate per density.

This is the batched counterpart of the per-RHS ``eval_direct`` loop of
``volume_fmm.py:983-1031``.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _capi

FORMS = {"f64": _capi.FORM_F64, "split": _capi.FORM_SPLIT_BF16, "split_i8": _capi.FORM_SPLIT_I8}
DEFAULT_SLICES = {"f64": 0, "split": 5, "split_i8": 6}


class PlanGraph:
    """A captured plan evaluation: the HIP graph plus references to the plan
    and the tensors it reads and writes, which must outlive it."""

    def __init__(self, graph, plan, tensors):
        self.graph = graph
        self.plan = plan
        self.tensors = tensors

    def replay(self):
        self.graph.replay()


class List1Plan:
    def __init__(self, tree, payload, *, scaling_mode, scaling_power=2, log_displacement=False,
                 table_root_extent, table_starting_level=0, case_encoding_bias=1e-10,
                 target_boxes=None, starts=None, lists=None, stream=None,
                 helmholtz_split=None, level_scaling=None, level_disp=None):
        """``helmholtz_split``: ``{"k": k, "series_nmax": N}`` folds the
        Helmholtz split correction into the blocks of a real Laplace payload
        (``vl1_plan_create_helmholtz_split``); the plan then evaluates complex
        densities to the split's full list-1 value.  Optional keys
        ``deriv_kind`` (``_capi.SPLIT_DERIV_*``) and ``deriv_axis`` describe
        an axis derivative wrapper of the output kernel (the payload then
        holds the wrapped Laplace table).  ``split_order`` (m > 1) with
        ``terms``, a list of ``(compact_table, power, is_log, coefficient)``
        -- the m - 1 tabulated terms, tables in the payload's compact layout
        -- and ``smooth`` = ``(nodes, weights, interp)`` (the smooth
        quadrature's per-axis rule on [0, 1] and the Lagrange basis of the
        box nodes there, ``interp[j][i] = l_i(node_j)``) fold the order-m
        split and the smooth quadrature into the blocks.  ``level_scaling`` /
        ``level_disp`` (``_capi.MAX_LEVELS`` values each): the per-level
        scaling and displacement factor of the custom policy
        (``scaling_mode=_capi.SCALING_CUSTOM``)."""
        L = _capi.lib()
        tb = np.ascontiguousarray(tree.target_boxes if target_boxes is None else target_boxes,
                                  dtype=np.int32)
        st = np.ascontiguousarray(
            tree.neighbor_source_boxes_starts if starts is None else starts, dtype=np.int32)
        ls = np.ascontiguousarray(
            tree.neighbor_source_boxes_lists if lists is None else lists, dtype=np.int32)
        self._keep = dict(
            centers=np.ascontiguousarray(tree.box_centers, np.float64),
            levels=np.ascontiguousarray(tree.box_levels, np.uint8),
            starts=np.ascontiguousarray(tree.box_target_starts, np.int32),
            counts=np.ascontiguousarray(tree.box_target_counts_nonchild, np.int32),
            tb=tb,
            st=st,
            ls=ls,
            ids=np.ascontiguousarray(tree.local_ids, np.int32),
        )
        if int(scaling_mode) == _capi.SCALING_CUSTOM:
            if level_scaling is None or level_disp is None:
                raise ValueError("the custom scaling mode needs level_scaling and level_disp")
            for name, arr in (("lsc", level_scaling), ("ldisp", level_disp)):
                a = np.ascontiguousarray(arr, np.float64)
                if a.shape != (_capi.MAX_LEVELS,):
                    raise ValueError(f"level arrays must hold {_capi.MAX_LEVELS} values")
                self._keep[name] = a
        k = self._keep
        desc = _capi.PlanDesc(
            dim=int(tree.dim),
            n_boxes=int(k["centers"].shape[1]),
            n_tgt_boxes=int(tb.size),
            n_points=int(k["ids"].size),
            box_centers=k["centers"].ctypes.data,
            box_levels=k["levels"].ctypes.data,
            box_starts=k["starts"].ctypes.data,
            box_counts=k["counts"].ctypes.data,
            target_boxes=tb.ctypes.data,
            nbr_starts=st.ctypes.data,
            nbr_lists=ls.ctypes.data,
            local_ids=k["ids"].ctypes.data,
            root_extent=float(tree.root_extent),
            table_root_extent=float(table_root_extent),
            table_starting_level=int(table_starting_level),
            scaling_mode=int(scaling_mode),
            scaling_power=int(scaling_power),
            log_displacement=int(bool(log_displacement)),
            case_encoding_bias=float(case_encoding_bias),
            boxes_per_tile=0,
            level_scaling=k["lsc"].ctypes.data if "lsc" in k else None,
            level_disp=k["ldisp"].ctypes.data if "ldisp" in k else None,
        )
        handle = _capi.vp()
        if helmholtz_split is None:
            _capi.check(
                L.vl1_plan_create(_capi.C.byref(handle), payload.handle, desc,
                                  _capi.stream_handle(stream)),
                "vl1_plan_create",
            )
        else:
            from . import gallery, quadrature

            xi, wi = quadrature.gauss_legendre_01(payload.quad_order)
            vecs, _, _ = gallery.generate_list1_gallery(int(tree.dim))
            sk = complex(helmholtz_split["k"])
            k["split_nodes"] = np.ascontiguousarray(xi, np.float64)
            k["split_weights"] = np.ascontiguousarray(wi, np.float64)
            k["case_vecs"] = np.ascontiguousarray(np.asarray(vecs, np.int32))
            terms = list(helmholtz_split.get("terms", ()))
            smooth = helmholtz_split.get("smooth")
            sp = _capi.HelmholtzSplitBlocks(
                k_re=sk.real, k_im=sk.imag, series_nmax=int(helmholtz_split["series_nmax"]),
                unit_nodes=k["split_nodes"].ctypes.data,
                unit_weights=k["split_weights"].ctypes.data,
                case_vecs=k["case_vecs"].ctypes.data,
                deriv_kind=int(helmholtz_split.get("deriv_kind", _capi.SPLIT_DERIV_NONE)),
                deriv_axis=int(helmholtz_split.get("deriv_axis", 0)),
                split_order=int(helmholtz_split.get("split_order", 1)),
                n_terms=len(terms),
            )
            if terms:
                k["term_tables"] = np.ascontiguousarray(
                    np.stack([np.asarray(t[0], np.float64).ravel() for t in terms]))
                k["term_powers"] = np.ascontiguousarray([int(t[1]) for t in terms], np.int32)
                k["term_logs"] = np.ascontiguousarray([int(bool(t[2])) for t in terms], np.int32)
                k["term_coefs"] = np.ascontiguousarray(
                    [[complex(t[3]).real, complex(t[3]).imag] for t in terms], np.float64)
                sp.n_term_entries = int(k["term_tables"].shape[1])
                sp.term_tables = k["term_tables"].ctypes.data
                sp.term_powers = k["term_powers"].ctypes.data
                sp.term_logs = k["term_logs"].ctypes.data
                sp.term_coefs = k["term_coefs"].ctypes.data
            if smooth is not None:
                nodes, weights, interp = smooth
                k["smooth_nodes"] = np.ascontiguousarray(nodes, np.float64)
                k["smooth_weights"] = np.ascontiguousarray(weights, np.float64)
                k["smooth_interp"] = np.ascontiguousarray(interp, np.float64)
                ms = k["smooth_nodes"].size
                if (k["smooth_weights"].shape != (ms,)
                        or k["smooth_interp"].shape != (ms, payload.quad_order)):
                    raise ValueError("smooth quadrature: nodes, weights and interp disagree")
                sp.smooth_order = ms
                sp.smooth_nodes = k["smooth_nodes"].ctypes.data
                sp.smooth_weights = k["smooth_weights"].ctypes.data
                sp.smooth_interp = k["smooth_interp"].ctypes.data
            _capi.check(
                L.vl1_plan_create_helmholtz_split(_capi.C.byref(handle), payload.handle, desc, sp,
                                                  _capi.stream_handle(stream)),
                "vl1_plan_create_helmholtz_split",
            )
        self.handle = handle
        self.payload = payload  # keep alive
        self.n_points = int(k["ids"].size)
        self.is_complex = bool(payload.is_complex) or helmholtz_split is not None

    def eval(self, mode_coefs: torch.Tensor, out: torch.Tensor | None = None, form="split",
             n_slices=None, stream=None, error_est: torch.Tensor | None = None) -> torch.Tensor:
        """Evaluate ``mode_coefs[R, n_points]`` into ``out`` (targets of the
        plan's boxes overwritten; allocated if None).  float64 CUDA data for
        real tables, complex128 for complex (Helmholtz) tables; the split
        forms then slice the real and imaginary parts alike.  ``error_est``
        (``form="split_i8"``): a float64 CUDA tensor ``[R, 2]`` that receives
        the estimated relative error per density (rel Linf, rel L2 over the
        plan's targets)."""
        if n_slices is None:
            n_slices = DEFAULT_SLICES[form]
        f = mode_coefs
        if f.dim() == 1:
            f = f.unsqueeze(0)
        want = torch.complex128 if self.is_complex else torch.float64
        if f.dtype != want or f.device.type != "cuda" or f.stride(-1) != 1:
            raise ValueError(f"mode_coefs must be a contiguous-rows {want} CUDA tensor")
        if f.shape[1] != self.n_points:
            raise ValueError("mode_coefs has the wrong number of points")
        if out is None:
            out = torch.zeros_like(f)
        if out.dtype != want or out.stride(-1) != 1 or out.shape != f.shape:
            raise ValueError("out must match mode_coefs in shape and dtype")
        # a one-row tensor may report any stride for its first dimension
        ldf_el = f.stride(0) if f.shape[0] > 1 else f.shape[1]
        ldy_el = out.stride(0) if out.shape[0] > 1 else out.shape[1]
        # complex data go through the C ABI as interleaved doubles
        fr = torch.view_as_real(f) if self.is_complex else f
        yr = torch.view_as_real(out) if self.is_complex else out
        scale = 2 if self.is_complex else 1
        opts = _capi.SplitOpts(n_slices=int(n_slices), n_products=0)
        if error_est is not None:
            if form != "split_i8":
                raise ValueError("error estimates come with form='split_i8'")
            if (error_est.dtype != torch.float64 or error_est.device.type != "cuda"
                    or tuple(error_est.shape) != (int(f.shape[0]), 2) or not error_est.is_contiguous()):
                raise ValueError("error_est must be a contiguous float64 CUDA tensor [R, 2]")
            opts.error_est = error_est.data_ptr()
        _capi.check(
            _capi.lib().vl1_plan_eval(
                self.handle, FORMS[form], fr.data_ptr(), int(ldf_el) * scale, int(f.shape[0]),
                yr.data_ptr(), int(ldy_el) * scale, opts, _capi.stream_handle(stream),
            ),
            "vl1_plan_eval",
        )
        return out

    VARIANT_BITS = {1: "one_row_tile_per_wave", 2: "morton_order", 4: "no_phase_alignment"}

    def tune(self, mode_coefs: torch.Tensor, form="split", n_slices=None, reps=3, stream=None):
        """Time every valid tile-kernel variant on ``mode_coefs`` and keep the
        fastest for later evaluations of this form / slice count / RHS count.
        Returns ``{"variant": v, "flags": [...], "times_ms": {v: ms}}``."""
        if n_slices is None:
            n_slices = DEFAULT_SLICES[form]
        f, fr, ldf, scale = self._prep(mode_coefs)
        out = torch.zeros(f.shape, dtype=f.dtype, device=f.device)  # compact rows
        yr = torch.view_as_real(out) if self.is_complex else out
        ldy = int(f.shape[1])
        opts = _capi.SplitOpts(n_slices=int(n_slices), n_products=0)
        chosen = _capi.i32(0)
        times = (_capi.f64 * 8)()
        _capi.check(
            _capi.lib().vl1_plan_tune(
                self.handle, FORMS[form], fr.data_ptr(), ldf * scale, int(f.shape[0]),
                yr.data_ptr(), ldy * scale, opts, int(reps), _capi.C.byref(chosen), times,
                _capi.stream_handle(stream)),
            "vl1_plan_tune",
        )
        v = int(chosen.value)
        return {"variant": v, "flags": [n for b, n in self.VARIANT_BITS.items() if v & b],
                "times_ms": {i: float(t) for i, t in enumerate(times) if t >= 0}}

    def set_variant(self, variant, form="split", n_slices=None, n_rhs=64):
        """Pin the tile-kernel variant (bits as in ``VARIANT_BITS``) used by
        later evaluations of this form / slice count / RHS count, without
        timing; ``variant=None`` returns to the default."""
        v = -1 if variant is None else int(variant)
        if n_slices is None:
            n_slices = DEFAULT_SLICES[form]
        _capi.check(
            _capi.lib().vl1_plan_set_variant(self.handle, FORMS[form], int(n_slices), int(n_rhs), v),
            "vl1_plan_set_variant")

    def _prep(self, mode_coefs):
        f = mode_coefs if mode_coefs.dim() == 2 else mode_coefs.unsqueeze(0)
        want = torch.complex128 if self.is_complex else torch.float64
        if f.dtype != want or f.device.type != "cuda" or f.stride(-1) != 1:
            raise ValueError(f"mode_coefs must be a contiguous-rows {want} CUDA tensor")
        if f.shape[1] != self.n_points:
            raise ValueError("mode_coefs has the wrong number of points")
        fr = torch.view_as_real(f) if self.is_complex else f
        ldf = f.stride(0) if f.shape[0] > 1 else f.shape[1]
        return f, fr, int(ldf), 2 if self.is_complex else 1

    def capture(self, mode_coefs: torch.Tensor, out: torch.Tensor, form="split", n_slices=None):
        """Record one evaluation (density staging + tile contraction) of
        ``mode_coefs`` into ``out`` as a HIP graph and return it as a
        ``PlanGraph``; each ``replay()`` re-evaluates whatever ``mode_coefs``
        then holds.  One eager evaluation runs first: it allocates the staging
        buffers and uploads the tile schedule, so the captured launches
        allocate nothing.

        Lifetime: the graph holds this plan (and the tensors it was captured
        with), so the plan outlives it.  Later evaluations of the plan with
        more densities or another slice count allocate new staging / slice
        buffers; the ones the graph reads are retired, not freed, until the
        plan is destroyed, so replays stay valid."""
        self.eval(mode_coefs, out=out, form=form, n_slices=n_slices)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.eval(mode_coefs, out=out, form=form, n_slices=n_slices)
        return PlanGraph(graph, self, (mode_coefs, out))

    def stats(self):
        arr = (_capi.i64 * 5)()
        _capi.check(_capi.lib().vl1_plan_stats(self.handle, arr), "vl1_plan_stats")
        return dict(zip(["pairs", "tiles", "steps", "blocks", "device_bytes"], list(arr)))

    def close(self):
        """Destroy the plan now (graphs captured from it become invalid)."""
        if getattr(self, "handle", None) is not None and self.handle.value:
            _capi.lib().vl1_plan_destroy(self.handle)
            self.handle = _capi.vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
